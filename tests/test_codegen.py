"""Code generation of the engine's kernels: no hot kernel may place locals in scratch (private) memory.

A scratch-resident local turns every access into a memory round trip; it happened once silently when a helper selected between a
reference into global memory and a reference to a local (k_qeq_build went from 117 to 185 us per 24k-atom step). The device
assembly of every source is generated here (hipcc, gfx950, the build's flags) and each kernel's private segment size read
from its descriptor; a frame that no scratch or buffer instruction of the kernel touches is reserved, not used, and passes.
Allowed: a reserved frame (FP32-term per-atom terms), spills of the FP64-mode H-bond donor, the FP64 valence and torsion
kernels held to 2 waves per SIMD (two 8-byte values each, reloaded once per outer partner, outside the innermost loop), and
the per-atom variants (only a force call that asks for per-atom tallies launches them).
"""
import concurrent.futures
import os
import re
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "metal-reax_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
# FP64 valence and torsion at 2 waves per SIMD (the torsion's census-recording variant Lb1 too)
ALLOWED = {"k_atom_termsIfdE", "k_atom_termsIffE", "k_hbond_donorIddLb0EE"} | \
    {"k_valence_dILi%dELb0EE" % g for g in (4, 8, 16, 32, 64)} | \
    {"k_torsion_dILi%dELb%dELb0EE" % (g, c) for g in (4, 8, 16, 32, 64) for c in (0, 1)}
MAX_ALLOWED_BYTES = 72
# the per-atom variants (last template flag set: the reference-split virial shares), launched only by a force call that asks
# for per-atom tallies, may use scratch
PERATOM = re.compile(r"^(k_valence(_d)?I.*Lb1EE|k_torsion_[fd]I.*ELb1EE|k_hbond_donorI.*Lb1EE)$")


def _asm(src, out):
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "--offload-device-only", "-S", src, "-o", out]
    subprocess.run(cmd, check=True, cwd=CSRC, capture_output=True, timeout=600)
    return open(out).read()


def _touches_scratch(asm, name):
    """Whether the kernel's code (from its label to its end marker) has a scratch or buffer load or store."""
    i = asm.index("\n" + name + ":") + 1
    j = asm.index(".Lfunc_end", i)
    return re.search(r"^\s*(scratch|buffer)_(load|store)", asm[i:j], re.M) is not None


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_no_scratch_in_engine_kernels(tmp_path):
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        asms = list(ex.map(lambda f: _asm(f, str(tmp_path / (f + ".s"))), srcs))
    bad = []
    seen = 0
    for asm in asms:
        for m in re.finditer(r"\.amdhsa_kernel (\S+)(.*?)\.end_amdhsa_kernel", asm, re.S):
            name, body = m.group(1), m.group(2)
            if not name.startswith("_ZN3rxa"):
                continue   # library kernels (hipCUB radix sort)
            seen += 1
            size = int(re.search(r"amdhsa_private_segment_fixed_size (\d+)", body).group(1))
            if size and not _touches_scratch(asm, name):
                size = 0   # a frame reserved by the compiler that no instruction of the kernel reads or writes
            short = re.sub(r"(ENS_|EvNS_).*", "", name.replace("_ZN3rxa12_GLOBAL__N_1", ""))
            short = re.sub(r"^\d+", "", short)
            if size and not PERATOM.match(short) and (short not in ALLOWED or size > MAX_ALLOWED_BYTES):
                bad.append((short, size))
    assert seen > 30
    assert not bad, bad
