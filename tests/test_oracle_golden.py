"""The oracle reproduces the golden vectors written by the compiled reference
routines (tests/golden/make_golden.py) bitwise."""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

from musoc_solver_amd import synth
from tests.conftest import fields_equal

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return meta, {k: z[k] for k in z.files if k != "meta"}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes(order="F")).hexdigest()


def coefs_of(meta, d):
    if meta.get("operator") == "ideal":
        return synth.ideal_operator(meta["i2"], meta["j2"])
    return tuple(np.asfortranarray(d[k]) for k in ("al", "ab", "ac", "ar", "at"))


def test_golden_files_present():
    names = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
    assert names == ["bir_harness_192x96_h24", "bir_harness_192x96_h6", "bir_synth_48x24",
                     "rep1_synth_48x24", "rep_synth_48x24"]


@pytest.mark.parametrize("name", ["bir_harness_192x96_h24", "bir_harness_192x96_h6",
                                  "bir_synth_48x24"])
def test_bir_golden(oracle, name):
    meta, d = load(name)
    coefs = coefs_of(meta, d)
    ie = d["ie"]
    rinv, rinv1, zp = oracle.bir_prep(coefs, ie, meta["nbir"])
    assert zp == meta["zero_pivots"]
    if "rinv" in d:
        assert fields_equal(rinv, d["rinv"]) and fields_equal(rinv1, d["rinv1"])
    else:
        assert sha(rinv) == meta["rinv_sha256"]
        assert sha(rinv1[:, :len(ie) - 1]) == meta["rinv1_sha256"]
    x = np.asfortranarray(d["x0"]).copy(order="F")
    xx = np.zeros_like(d["xx"], order="F")
    oracle.bir_solve(coefs, d["s"], rinv, rinv1, x, xx, ie, meta["nbir"], meta["nsweep"],
                     meta["periodic"])
    assert fields_equal(x, d["x"])
    assert fields_equal(xx, d["xx"])


def test_harness_h24_overflows_in_the_reference():
    """The literal 24-row blocks of the 192x96 layout overflow the FP64
    march: the reference returns no finite interior value."""
    meta, d = load("bir_harness_192x96_h24")
    assert not np.isfinite(d["x"][1:-1, 1:-1]).any()


@pytest.mark.parametrize("name", ["rep1_synth_48x24", "rep_synth_48x24"])
def test_full_width_golden(oracle, name):
    meta, d = load(name)
    coefs = coefs_of(meta, d)
    ie = d["ie"]
    if meta["periodic"]:
        rinv, rinv1, zp = oracle.pre1(coefs, ie)
    else:
        rinv, rinv1, zp = oracle.pre(coefs, ie)
    assert fields_equal(rinv, d["rinv"]) and fields_equal(rinv1, d["rinv1"])
    x = np.asfortranarray(d["x0"]).copy(order="F")
    (oracle.rep1 if meta["periodic"] else oracle.rep)(coefs, d["rinv"], d["rinv1"], d["s"], x, ie)
    assert fields_equal(x, d["x"])
