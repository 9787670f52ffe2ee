"""Pins, against the reference's own reader and FP64 engine (golden vectors in tests/golden/variants.json, written by
tests/golden/make_variant_golden.py), what the C/H/O fixtures leave open:

* the engine's run-time ffield reader and the oracle's: the canonical table dump (plain and portable) of the stand-in field,
  the carbon field of the reference's SI PDF and every variant equals the reference reader's (SHA-256), and the files the
  reference rejects are rejected;
* the oracle's decision census (bonds, three- and four-body gates, H-bond triples, non-bonded pairs, SBO branch, lone-pair
  truncation) equals the reference engine's on the 41 C/H/O fixtures;
* the oracle's 13 energy terms and gradient equal the reference engine's with the van der Waals inner wall (method 2),
  shielding plus inner wall (3), low-gradient dispersion (lgvdw), triple-bond stabilisation of every pair, and without the C2
  correction, and on random-carbon cells with the carbon field.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from oracle.oracle import CENSUS, Oracle, TERMS

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import ffield_variants as fv   # noqa: E402


@pytest.fixture(scope="module")
def variants():
    with open(os.path.join(HERE, "golden", "variants.json")) as fh:
        return json.load(fh)


def carbon_text():
    with open(os.path.join(HERE, "golden", "ffield.C2013")) as fh:
        return fh.read()


def texts():
    out = {"standin": (fv.standin_text(), False), "C2013": (carbon_text(), False)}
    for name, lg in list(fv.ACCEPTED.items()) + list(fv.REJECTED.items()):
        out[name] = (fv.variant(name), lg)
    return out


def test_reader_dumps_match_reference(variants):
    import reaxamd
    for name, (text, lg) in texts().items():
        ref = variants["dumps"][name]
        if ref.get("rejected"):
            with pytest.raises(reaxamd.ReaxError):
                reaxamd.ffield_table_dump(text, lgvdw=lg)
            with pytest.raises(ValueError):
                Oracle(text, lgvdw=lg)
            continue
        for port, key in ((False, "sha256"), (True, "sha256_portable")):
            d = reaxamd.ffield_table_dump(text, lgvdw=lg, portable=port)
            assert hashlib.sha256(d.encode()).hexdigest() == ref[key], (name, key)
            assert hashlib.sha256(Oracle(text, lgvdw=lg).table_dump(port).encode()).hexdigest() == ref[key], (name, key, "oracle")


def test_oracle_census_matches_reference(golden, ffield_text, variants):
    from helpers import case_inputs
    O = Oracle(ffield_text)
    for c in golden["cases"]:
        if c["id"] not in variants["census"]:
            continue
        x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(c)
        r = O.compute(x, types - 1, tags, box12, per, q=q)
        for k, cat in enumerate(CENSUS):
            assert r["census"][cat].tolist() == variants["census"][c["id"]][k], (c["id"], cat)
    assert len(variants["census"]) == 41


def _close(r, e_ref, g_ref, label):
    for k in TERMS:
        assert abs(r["energy"][k] - e_ref[k]) <= 1e-12 * max(1.0, abs(e_ref[k])), (label, k, r["energy"][k], e_ref[k])
    assert np.abs(r["grad"] - np.array(g_ref)).max() <= 1e-10, label


def test_oracle_variants_match_reference(golden, variants):
    from helpers import case_inputs
    by_id = {c["id"]: c for c in golden["cases"]}
    for name, lg in fv.ACCEPTED.items():
        O = Oracle(fv.variant(name), lgvdw=lg)
        for cid, ref in variants["variants"][name].items():
            x, types, tags, lo, hi, tilt, per, q, box12 = case_inputs(by_id[cid])
            _close(O.compute(x, types - 1, tags, box12, per, q=q), ref["energy"], ref["grad"], (name, cid))
    assert len(variants["variants"]) == len(fv.ACCEPTED)


def test_oracle_carbon_matches_reference(variants):
    O = Oracle(carbon_text())
    for c in variants["carbon"]:
        L = c["L"]
        x = np.array(c["positions"])
        n = c["n"]
        r = O.compute(x, np.zeros(n, dtype=np.int32), np.arange(1, n + 1), [0, 0, 0, L, 0, 0, 0, L, 0, 0, 0, L], (1, 1, 1), q=np.zeros(n))
        _close(r, c["energy"], c["grad"], c["id"])
        for k, cat in enumerate(CENSUS):
            assert r["census"][cat].tolist() == c["census"][k], (c["id"], cat)
        assert sum(c["census"][0]) > n, c["id"]   # a bonded network, not a gas
