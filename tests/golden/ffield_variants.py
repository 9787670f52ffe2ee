"""Deterministic variants of the stand-in C/H/O force field (tests/golden/ffield.cho_standin) that reach the branches the
stand-in itself leaves unused, and malformed files the reader must reject. Shared by the fixture writer
(tests/golden/make_variant_golden.py, which runs the reference's own reader and FP64 engine on them) and by the tests, which
rebuild the same texts from the same recipes.

The stand-in is a shielding-only field (van der Waals method 1: gamma_w > 0.5 and no inner wall) with the C2 correction on
(general parameter 6 > 0.001) and triple-bond stabilisation for the C-O mass pair only (general parameter 38 is 0). The reader
decides the method from the first element (shielding: gamma_w > 0.5; inner wall: rcore2 > 0.01 and acore2 > 0.01;
reference src/io/ffield_parser.cpp):

* inner_wall:   every element gets an inner wall and gamma_w 0.3 (no shielding)      -> method 2
* shield_inner: every element gets an inner wall, shielding kept                    -> method 3
* lg:           method 3 plus the low-gradient dispersion line of each element       -> method 3 with lgvdw yes
* triple_all:   general parameter 38 = 2: triple-bond stabilisation for every pair
* no_c2:        general parameter 6 = 0: C2 correction off
* numeric_prefix: a value with trailing characters ("62.48x9"), read as its numeric prefix by both readers

Rejected (the reference's reader refuses them, and so must the engine's):
* truncated:     the file ends inside the bond block
* short_line:    an element line one column short
* lg_missing:    the lgvdw variant read without its dispersion lines (lgvdw yes on the plain stand-in)
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
STANDIN = os.path.join(HERE, "ffield.cho_standin")

# inner-wall parameters (rcore2, ecore2, acore2) and low-gradient parameters (lgcij, lgre) per element; plausible magnitudes
INNER = {"C": (1.8903, 0.1838, 10.0), "H": (1.3550, 0.0930, 8.0), "O": (2.3890, 0.1000, 12.0)}
LG = {"C": (1.0564, 1.9510), "H": (0.4047, 1.5904), "O": (1.2032, 1.7010)}


def standin_text():
    with open(STANDIN) as fh:
        return fh.read()


def _fmt(v):
    return "%9.4f" % v


def _element_blocks(lines):
    """(first line index of each element block, element name) of the stand-in's atom section"""
    ngp = int(lines[1].split()[0])
    k = 2 + ngp
    nel = int(lines[k].split()[0])
    start = k + 4   # three continuation header lines after the count line
    out = []
    for e in range(nel):
        i = start + 4 * e
        out.append((i, lines[i].split()[0]))
    return out, start + 4 * nel


def _set_fields(line, values_from):
    """replaces trailing numeric fields of an element line; values_from: {field index: value} over the whitespace fields"""
    f = line.split()
    for k, v in values_from.items():
        f[k] = _fmt(v)
    lead = "" if line[:1] != " " else " "
    return lead + " ".join(x.rjust(9) if j else x for j, x in enumerate(f))


def _off_diagonal(lines):
    """index of the off-diagonal count line"""
    for k, ln in enumerate(lines):
        if "off-diagonal" in ln:
            return k
    raise ValueError("no off-diagonal block")


def variant(name):
    lines = standin_text().split("\n")
    blocks, _ = _element_blocks(lines)
    if name in ("inner_wall", "shield_inner", "lg"):
        for i, el in blocks:
            rc, ec, ac = INNER[el]
            lines[i + 3] = _set_fields(lines[i + 3], {5: rc, 6: ec, 7: ac})   # line 4: ..., rcore2, ecore2, acore2
            if name == "inner_wall":
                lines[i + 1] = _set_fields(lines[i + 1], {1: 0.3})               # line 2: alfa, gamma_w, ...
        if name == "lg":   # a fifth line per element: lgcij lgre (inserted from the last block backwards), and the pair lgcij
            names = [el for _, el in blocks]             # as the last column of every off-diagonal line
            for i, el in reversed(blocks):
                lines.insert(i + 4, "   " + _fmt(LG[el][0]) + _fmt(LG[el][1]))
            k = _off_diagonal(lines)
            for m in range(int(lines[k].split()[0])):
                f = lines[k + 1 + m].split()
                a, b = names[int(f[0]) - 1], names[int(f[1]) - 1]
                lines[k + 1 + m] = lines[k + 1 + m] + _fmt((LG[a][0] * LG[b][0]) ** 0.5)
    elif name == "triple_all":
        lines[2 + 37] = "    2.0000 !Molecular energy (not used)"
    elif name == "no_c2":
        lines[2 + 5] = "    0.0000 !C2-correction"
    elif name == "truncated":
        _, after = _element_blocks(lines)
        lines = lines[:after + 3]
    elif name == "numeric_prefix":   # a value with trailing garbage: the numeric prefix is read (LAMMPS' atof rule)
        i = blocks[1][0]
        lines[i + 2] = lines[i + 2].replace("62.4879", "62.48x9")
    elif name == "short_line":       # an element line one column short
        i = blocks[2][0]
        lines[i] = lines[i].rsplit(None, 1)[0]
    elif name == "lg_missing":
        pass
    else:
        raise KeyError(name)
    return "\n".join(lines)


ACCEPTED = {"inner_wall": False, "shield_inner": False, "lg": True, "triple_all": False, "no_c2": False, "numeric_prefix": False}   # -> lgvdw
REJECTED = {"truncated": False, "short_line": False, "lg_missing": True}
