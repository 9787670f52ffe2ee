"""Writes tests/golden/ffield.C2013: the carbon ReaxFF "Reactive MD-force field: 2013 C" (S. G. Srinivasan, A. C. T. van Duin,
P. Ganesh, J. Phys. Chem. A 119, 571 (2015)), from the paper's Supporting Information PDF that the reference repository carries
(jp510274e_si_001.pdf). The reference turns that PDF into an ffield with `pdftotext -layout` and checks the title line
(examples/graphitization/make_ffield.sh:9-11); pdftotext is not available here, so this script reads the PDF's text directly:

* only two objects are decoded: the page's content stream and the font's ToUnicode CMap (zlib); the embedded font program
  is never read;
* the text operators of the content stream (BT/ET, Tm, Td, TD, T*, Tf, Tj, TJ) are interpreted with the font's fixed glyph
  advance (the PDF uses one monospaced font, /W widths all 602/1000 em), each glyph is mapped to Unicode through the CMap
  and placed on a character grid, and each text line becomes one line of the file (trailing blanks removed), which is
  what `pdftotext -layout` gives for a monospaced page;
* the result must contain the reference's check string and parse as a LAMMPS ffield (one element, C).

Run in a container that has the reference at /root/reference:  python tests/golden/make_carbon_ffield.py
"""
import os
import re
import sys
import zlib

REF_PDF = "/root/reference/jp510274e_si_001.pdf"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ffield.C2013")
CHECK = "Reactive MD-force field: 2013 C"


def objects(pdf):
    return {int(m.group(1)): m.group(2) for m in re.finditer(rb"(?:^|\n)(\d+) 0 obj(.*?)endobj", pdf, re.S)}


def stream_of(body):
    s = body.index(b"stream") + len(b"stream")
    s += 2 if body[s:s + 2] == b"\r\n" else (1 if body[s:s + 1] == b"\n" else 0)
    raw = body[s:body.rindex(b"endstream")]
    if b"/FlateDecode" not in body[:body.index(b"stream")]:
        raise ValueError("expected a FlateDecode stream")
    return zlib.decompress(raw)


def ref(body, key):
    m = re.search(rb"/" + key + rb"\s+(\d+) 0 R", body)
    return int(m.group(1)) if m else None


def cmap_table(text):
    """bfchar and bfrange entries of a ToUnicode CMap: glyph code -> Unicode string"""
    out = {}
    for blk in re.findall(r"beginbfchar(.*?)endbfchar", text, re.S):
        for a, b in re.findall(r"<([0-9a-fA-F]+)>\s*<([0-9a-fA-F]+)>", blk):
            out[int(a, 16)] = bytes.fromhex(b).decode("utf-16-be")
    for blk in re.findall(r"beginbfrange(.*?)endbfrange", text, re.S):
        for a, b, c in re.findall(r"<([0-9a-fA-F]+)>\s*<([0-9a-fA-F]+)>\s*<([0-9a-fA-F]+)>", blk):
            u = int(c, 16)
            for k in range(int(a, 16), int(b, 16) + 1):
                out[k] = chr(u + k - int(a, 16))
    return out


def tokens(content):
    """PDF content tokens: numbers, names, hex strings, literal strings, arrays' brackets, operators"""
    pat = re.compile(rb"\s*(<[0-9a-fA-F\s]*>|\((?:\\.|[^\\)])*\)|\[|\]|/[^\s/\[\]<>()]+|[-+]?(?:\d+\.?\d*|\.\d+)|[A-Za-z'\"*]+)", re.S)
    pos = 0
    while True:
        m = pat.match(content, pos)
        if not m:
            rest = content[pos:].strip()
            if rest:
                raise ValueError("unparsed content near %r" % rest[:40])
            return
        pos = m.end()
        yield m.group(1)


def extract(pdf_path):
    pdf = open(pdf_path, "rb").read()
    obj = objects(pdf)
    pages = [k for k, b in obj.items() if re.search(rb"/Type\s*/Page\b", b)]
    lines_out = []
    for pg in sorted(pages):
        page = obj[pg]
        content = stream_of(obj[ref(page, b"Contents")])
        res = obj[ref(page, b"Resources")] if ref(page, b"Resources") else page
        fonts = {}   # resource name -> (cmap, advance per glyph in text-space units at size 1)
        for name, num in re.findall(rb"/([\w.+-]+)\s+(\d+) 0 R", res[res.index(b"/Font"):]):
            fb = obj.get(int(num), b"")
            if b"/Font" not in fb:
                continue
            cmap = cmap_table(stream_of(obj[ref(fb, b"ToUnicode")]).decode("latin1"))
            desc = obj[int(re.search(rb"/DescendantFonts\s*\[\s*(\d+) 0 R", fb).group(1))]
            widths = {int(w) for w in re.findall(rb"\b(\d+)\b", desc[desc.index(b"/W"):].split(b"]]")[0])} - {0}
            adv = max(widths) / 1000.0   # monospaced: one width
            fonts[name.decode()] = (cmap, adv)
        grid = {}   # (line y rounded, column) -> char
        tm = [1, 0, 0, 1, 0, 0]
        lx = ly = 0.0   # text line matrix origin (text space)
        tx = 0.0        # advance along the line since the last line origin
        font, size = None, 1.0
        stack = []
        for tok in tokens(content):
            if tok in (b"q", b"Q", b"BT", b"ET", b"gs", b"rg", b"RG", b"g", b"G", b"cm", b"re", b"f", b"n", b"W"):
                if tok == b"BT":
                    tm, lx, ly, tx = [1, 0, 0, 1, 0, 0], 0.0, 0.0, 0.0
                stack = []
                continue
            if tok == b"Tm":
                tm = [float(v) for v in stack[-6:]]
                lx = ly = tx = 0.0
            elif tok in (b"Td", b"TD"):
                lx += float(stack[-2])   # Td moves the line origin (not the current point)
                ly += float(stack[-1])
                tx = 0.0
            elif tok == b"Tf":
                font, size = stack[-2][1:].decode(), float(stack[-1])
            elif tok in (b"Tj", b"TJ", b"'"):
                cmap, adv = fonts[font]
                items = stack if tok == b"TJ" else stack[-1:]
                for it in items:
                    if isinstance(it, (bytes, bytearray)) and it[:1] == b"<":
                        hx = re.sub(rb"\s", b"", it[1:-1])
                        for k in range(0, len(hx), 4):
                            code = int(hx[k:k + 4], 16)
                            col = round((lx + tx) / adv)
                            key = (round((tm[5] + ly * tm[3]) * 100), col)
                            grid[key] = cmap.get(code, "?")
                            tx += adv * size
                    elif isinstance(it, (bytes, bytearray)) and it[:1] not in (b"[", b"]", b"("):
                        tx -= float(it) / 1000.0 * size   # TJ kerning
                    elif isinstance(it, (bytes, bytearray)) and it[:1] == b"(":
                        raise ValueError("literal strings are not expected in this document")
            elif tok == b"T*":
                raise ValueError("T* is not expected in this document")
            stack = [] if re.fullmatch(rb"[A-Za-z'\"*]+", tok) else stack + [tok]
        rows = sorted({k[0] for k in grid}, reverse=True)
        for y in rows:
            cols = sorted(c for (yy, c) in grid if yy == y)
            line = [" "] * (cols[-1] + 1)
            for c in cols:
                line[c] = grid[(y, c)]
            lines_out.append("".join(line).rstrip())
    return "\n".join(lines_out) + "\n"


def main():
    text = extract(REF_PDF)
    if CHECK not in text:
        sys.exit("unexpected content: this is not the 2013 C force field")
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))), "metal-reax_amd"))
    import reaxamd
    reaxamd.ffield_table_dump(text)   # parses as a LAMMPS ffield, or raises
    with open(OUT, "w") as fh:
        fh.write(text)
    print("wrote %s (%d lines)" % (OUT, text.count("\n")))


if __name__ == "__main__":
    main()
