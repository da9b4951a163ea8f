"""Golden fixtures (tests/golden/make_golden.py): the oracle reproduces them exactly
on the CPU; the HIP path reproduces them exactly on the GPU."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import CASES

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_golden(name):
    import oracle.pipeline as opipe
    g = _load(name)
    res = opipe.smart_interpolation(g["img"], g["labels"], **CASES[name]["kw"])
    np.testing.assert_array_equal(res["_hits_u32"], g["hits_u32"])
    assert res["_accepted"] == int(g["accepted"])
    for k in ("regular", "smooth", "uncertainty"):
        if k in g:
            np.testing.assert_array_equal(res[k], g[k], err_msg=k)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_hip_reproduces_golden(name):
    from biomedisa_amd import smart_interpolation
    g = _load(name)
    kw = dict(CASES[name]["kw"])
    res = smart_interpolation(g["img"], g["labels"], return_hits=kw.get("rng") != "reference", **kw)
    for k in ("regular", "smooth", "uncertainty"):
        if k in g:
            np.testing.assert_array_equal(res[k], g[k], err_msg=k)
    if "hits" in res:
        np.testing.assert_array_equal(res["hits"].astype(np.uint32), g["hits_u32"][:, 1:-1, 1:-1, 1:-1])
