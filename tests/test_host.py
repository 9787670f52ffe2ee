"""Host-side logic of the engine, through the C ABI, without a GPU: the run-time force-field reader (checked against the oracle's
independent reader and the reference's quirks), the pair_style keyword surface, and the library's exported symbols."""
import os
import re

import numpy as np
import pytest

import reaxamd
from oracle.oracle import Oracle

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def both_dumps(text, portable=False):
    return reaxamd.ffield_table_dump(text, portable=portable), Oracle(text).table_dump(portable)


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(REPO, "include", "reaxamd.h")).read()
    names = sorted(set(re.findall(r"\b(rxa_\w+)\s*\(", header)))
    assert len(names) >= 30
    L = reaxamd.lib()
    for nm in names:
        assert hasattr(L, nm), nm
    assert L.rxa_abi_version() == 100


def test_create_without_device_fails_cleanly():
    import ctypes
    h = ctypes.c_void_p()
    rc = reaxamd.lib().rxa_create(ctypes.byref(h), 0)
    if rc == 0:   # a GPU is present: the context works and is destroyed again
        reaxamd.lib().rxa_destroy(h)
    else:
        assert rc == -3 and not h.value


def test_reader_matches_oracle(ffield_text):
    a, b = both_dumps(ffield_text)
    assert a == b
    a, b = both_dumps(ffield_text, portable=True)
    assert a == b


def _blocks(text):
    """line indices of the block headers of an ffield text: (general, atoms, bonds, offdiag, angles, torsions, hbonds)"""
    lines = text.splitlines()
    ng = int(lines[1].split()[0])
    i_at = 2 + ng
    nat = int(lines[i_at].split()[0])
    i_b = i_at + 4 + 4 * nat
    nb = int(lines[i_b].split()[0])
    i_o = i_b + 2 + 2 * nb
    no = int(lines[i_o].split()[0])
    i_a = i_o + 1 + no
    na = int(lines[i_a].split()[0])
    i_t = i_a + 1 + na
    nt = int(lines[i_t].split()[0])
    i_h = i_t + 1 + nt
    return lines, dict(gp=1, at=i_at, b=i_b, o=i_o, a=i_a, t=i_t, h=i_h)


def _set_count(lines, idx, k):
    parts = lines[idx].split(None, 1)
    lines[idx] = " %d    %s" % (k, parts[1] if len(parts) > 1 else "")


def _dump_dict(d):
    out = {}
    for line in d.splitlines():
        p = line.split()
        out.setdefault(p[0], []).append(p[1:])
    return out


def test_angle_sets_for_jl_are_counted_twice(ffield_text):
    """Three sets for an H-O-H (j == l) triple overrun the reference's 5-slot array (each j == l set takes two slots): rejected by
    both readers; two sets are accepted and kept in file order."""
    lines, b = _blocks(ffield_text)
    hoh = "  2  3  2  77.0645  10.4737   1.2895   0.0000   0.9924   0.0000   1.1043"
    for extra, ok in ((1, True), (2, False)):
        L = list(lines)
        na = int(L[b["a"]].split()[0])
        for _ in range(extra):
            L.insert(b["a"] + 1 + na, hoh)
        _set_count(L, b["a"], na + extra)
        text = "\n".join(L) + "\n"
        if ok:
            d, o = both_dumps(text)
            assert d == o
            h3 = [r for r in _dump_dict(d)["H3"] if r[:3] == ["1", "2", "1"]]
            assert int(h3[0][3]) == 2
        else:
            with pytest.raises(reaxamd.ReaxError):
                reaxamd.ffield_table_dump(text)
            with pytest.raises(ValueError):
                Oracle(text)


def test_compact_torsions_fill_and_explicit_wins(ffield_text):
    """0-X-Y-0 entries fill every outer pair not set explicitly (both orientations); later compact entries overwrite earlier ones;
    explicit entries win whatever their position in the file (reference Q-10)."""
    lines, b = _blocks(ffield_text)
    nt = int(lines[b["t"]].split()[0])
    L = list(lines)
    new = ["  1  2  2  3   0.1111   0.2222   0.3333  -0.4444  -0.5555   0.0000   0.0000",
           "  0  2  2  0   9.0000   8.0000   7.0000  -6.0000  -5.0000   0.0000   0.0000",
           "  0  2  2  0   1.5000   2.5000   3.5000  -4.5000  -5.5000   0.0000   0.0000"]
    for k, l in enumerate(new):
        L.insert(b["t"] + 1 + nt + k, l)
    _set_count(L, b["t"], nt + len(new))
    text = "\n".join(L) + "\n"
    d, o = both_dumps(text)
    assert d == o
    h4 = {tuple(r[:4]): [float(v) for v in r[4:]] for r in _dump_dict(d)["H4"]}
    assert h4[("0", "1", "1", "2")] == [0.1111, 0.2222, 0.3333, -0.4444, -0.5555]   # explicit C-H-H-O, placed before the compacts
    assert h4[("2", "1", "1", "0")] == [0.1111, 0.2222, 0.3333, -0.4444, -0.5555]   # and its mirror
    assert h4[("0", "1", "1", "0")] == [1.5, 2.5, 3.5, -4.5, -5.5]                 # last compact entry wins
    assert h4[("2", "1", "1", "2")] == [1.5, 2.5, 3.5, -4.5, -5.5]


def test_missing_hbond_block_is_accepted_truncation_is_not(ffield_text):
    lines, b = _blocks(ffield_text)
    text = "\n".join(lines[:b["h"]]) + "\n"
    d, o = both_dumps(text)
    assert d == o and "HB " not in d
    text2 = "\n".join(lines[:b["t"] + 3]) + "\n"
    with pytest.raises(reaxamd.ReaxError):
        reaxamd.ffield_table_dump(text2)
    with pytest.raises(ValueError):
        Oracle(text2)


def test_derived_values(ffield_text):
    """eta stored as 2 x the file value, nlp_opt = (val_e - val) / 2, light-element valency override, mixing rules."""
    d = _dump_dict(reaxamd.ffield_table_dump(ffield_text))
    s = {r[1]: [float(v) for v in r[2:]] for r in d["S"]}
    # S fields: r_s valency mass r_vdw epsilon gamma r_pi valency_e nlp_opt alpha gamma_w valency_boc p_ovun5 p_hbond chi eta ...
    h = s["H"]
    assert h[8] == 0.5 * (h[7] - h[1])
    assert h[15] == 2 * 9.6093
    o = s["O"]
    t = {tuple(r[:2]): [float(v) for v in r[2:]] for r in d["T"]}
    ho = t[("1", "2")]
    assert ho[6] == 0.9215           # r_s overridden by the off-diagonal block
    assert ho[20] == 2 * 1.2885      # r_vdW = 2 x off-diagonal value
    assert ho[19] == 10.9190         # alpha
    hh = t[("1", "1")]
    assert hh[18] == np.sqrt(0.0930 * 0.0930)


def test_pair_style_keywords():
    assert reaxamd.check_pair_style(["NULL"])["precision"] == "fp64"            # backend cpu64 is the default
    assert reaxamd.check_pair_style(["NULL", "backend", "gpu"])["precision"] == "fp32"
    assert reaxamd.check_pair_style(["NULL", "backend", "metal"])["backend"] == "gpu"
    assert reaxamd.check_pair_style(["NULL", "backend", "gpu", "bonded", "cpu64"])["precision"] == "mixed"
    assert reaxamd.check_pair_style(["NULL", "precision", "mixed", "checkqeq", "no"])["checkqeq"] == "no"
    s = reaxamd.check_pair_style(["NULL", "safezone", "1.6", "mincap", "100", "list/blocking", "yes"])
    assert len(s["notices"]) == 3
    for bad in (["NULL", "backend", "cuda"], ["NULL", "checkqeq"], ["NULL", "foo", "1"], ["NULL", "mincap", "-1"], []):
        with pytest.raises(reaxamd.ReaxError):
            reaxamd.check_pair_style(bad)


def test_control_file():
    s = reaxamd.check_pair_style(["control.reax", "backend", "gpu"], "nbrhood_cutoff 4.5\nhbond_cutoff 6.0\nthb_cutoff 0.002\nqeq_freq 1\n")
    assert float(s["bond_cut"]) == 4.5 and float(s["hbond_cut"]) == 6.0 and float(s["thb_cut"]) == 0.002
    assert any("qeq_freq" in n for n in s["notices"])
    with pytest.raises(reaxamd.ReaxError):
        reaxamd.check_pair_style(["control.reax"], "bogus_key 1\n")


def test_pvector_maps_terms_to_lammps_slots():
    """rxa_pvector against the 14 slots LAMMPS reported beside the 13 terms, on all 58 reference goldens (tests/golden/pvector.json,
    data of the reference's tests/fixtures/reference/*.json): eb, ea = e_ov + e_un, elp, emol = 0, ev, epen, ecoa, ehb, et, eco, ew,
    ep = e_ele, efi = 0, eqeq = e_pol (Q-31: eng_coul is e_ele + e_pol). A wrong or transposed slot fails on the cases whose terms
    differ; the sum of the slots equals the total potential energy."""
    import ctypes
    import json
    with open(os.path.join(REPO, "tests", "golden", "pvector.json")) as fh:
        g = json.load(fh)
    L = reaxamd.lib()
    nonzero = np.zeros(14, dtype=bool)
    for c in g["cases"]:
        e = np.ascontiguousarray(c["terms"], dtype=np.float64)
        pv = np.full(14, np.nan)
        L.rxa_pvector(e.ctypes.data_as(ctypes.c_void_p), pv.ctypes.data_as(ctypes.c_void_p))
        want = np.array(c["slots"])
        assert np.allclose(pv, want, rtol=1e-12, atol=1e-12), (c["id"], pv.tolist(), want.tolist())
        assert abs(pv.sum() - e.sum()) <= 1e-9 * max(1.0, abs(e.sum()))
        nonzero |= want != 0
    # the goldens exercise every slot the 13 terms feed (emol and efi stay zero)
    assert nonzero.sum() == 12 and not nonzero[3] and not nonzero[12]


def test_ghost_native_input_checks(golden):
    """rxa_check_ghosts (the checks of rxa_set_atoms_ghosted, without a device) on LAMMPS-style owned + ghost arrays of a periodic
    fixture: accepted as generated; refused, with the reason, for more than one rank, a ghost whose owner is not on this rank
    (what a sub-domain's ghosts are under MPI), a ghost off the lattice by 1e-4 A, a type mismatch, a shift along a non-periodic
    direction and a duplicate owned id."""
    from helpers import case_inputs, lammps_ghosts
    case = next(c for c in golden["cases"] if c["id"] == "cho_water_box_8")
    x, types, tags, lo, hi, tilt, per, q, _ = case_inputs(case)
    n = len(types)
    T, G, X = lammps_ghosts(x, types, tags, lo, hi, tilt, per, shell=10.0)
    assert len(T) > 4 * n
    rc, msg = reaxamd.check_ghosts(lo, hi, tilt, per, n, T, G, X)
    assert rc == 0, msg
    rc, msg = reaxamd.check_ghosts(lo, hi, tilt, per, n, T, G, X, nranks=2)
    assert rc == -1 and "multi-rank" in msg
    G2 = G.copy(); G2[n + 3] = 10 ** 6
    rc, msg = reaxamd.check_ghosts(lo, hi, tilt, per, n, T, G2, X)
    assert rc == -1 and "no owned atom" in msg
    X2 = X.copy(); X2[n + 5, 1] += 1e-4
    rc, msg = reaxamd.check_ghosts(lo, hi, tilt, per, n, T, G, X2)
    assert rc == -1 and "lattice image" in msg
    T2 = T.copy(); T2[n + 7] = 1 + (T2[n + 7] % 3)
    rc, msg = reaxamd.check_ghosts(lo, hi, tilt, per, n, T2, G, X)
    assert rc == -1 and "type" in msg
    rc, msg = reaxamd.check_ghosts(lo, hi, tilt, [1, 1, 0], n, T, G, X)
    assert rc == -1 and "non-periodic" in msg
    G3 = G.copy(); G3[1] = G3[0]
    rc, msg = reaxamd.check_ghosts(lo, hi, tilt, per, n, T, G3, X)
    assert rc == -1 and "duplicate" in msg


def test_compiled_c_caller_builds_against_the_header(golden, tmp_path):
    """tests/abi_caller.c includes include/reaxamd.h from C99 (compiled with -Wall -Wextra -Werror by __graft_entry__.build and
    again here); without a device it reads its inputs and stops at rxa_create with RXA_EDEVICE (exit status 3)."""
    import subprocess
    import sys
    sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
    from make_golden import case_text
    exe = str(tmp_path / "abi_caller")
    subprocess.run(["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(REPO, "include"),
                    os.path.join(REPO, "tests", "abi_caller.c"), "-L", os.path.join(REPO, "metal-reax_amd"), "-lreaxamd",
                    "-Wl,-rpath," + os.path.join(REPO, "metal-reax_amd"), "-o", exe], check=True)
    case = next(c for c in golden["cases"] if c["id"] == "cho_water_box_8")
    path = tmp_path / "case.txt"
    path.write_text(case_text(case["origin"], case["vectors"], case["periodic"], case["positions"], case["types"], case["charges"]))
    res = subprocess.run([exe, os.path.join(REPO, "tests", "golden", "ffield.cho_standin"), "C,H,O", str(path)], capture_output=True,
                         text=True, timeout=60)
    import ctypes
    h = ctypes.c_void_p()
    if reaxamd.lib().rxa_create(ctypes.byref(h), 0) == 0:   # a GPU is present: the caller runs through
        reaxamd.lib().rxa_destroy(h)
        assert res.returncode == 0 and res.stdout.count("grad ") == len(case["types"]), res.stderr
    else:
        assert res.returncode == 3 and "no device" in res.stderr
