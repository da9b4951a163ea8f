"""Writes tests/golden/*.npz: small input/output vectors of the smart-interpolation
path produced by the CPU oracle (oracle/), so that later changes to the oracle or to
the HIP path are checked against fixed data.  The reference repository holds no
golden vectors for this path; these are the project's own.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from biomedisa_amd.synth import ellipsoids_volume, structures_volume  # noqa: E402
import oracle.pipeline as opipe  # noqa: E402

CASES = {
    "xy_u8": dict(vol=lambda: ellipsoids_volume(shape=(20, 22, 24), seed=21, semi=(4, 7), every=4),
                  kw=dict(nbrw=3, sorw=120, smooth=2, uncertainty=True)),
    "allaxis_u8": dict(vol=lambda: structures_volume(shape=(22, 20, 24), seed=22, n_struct=3,
                                                     radius=(3, 6), every=5, allaxis=True),
                       kw=dict(nbrw=2, sorw=90, allaxis=True, smooth=1)),
    "xy_u8_reference_stream": dict(vol=lambda: ellipsoids_volume(shape=(18, 20, 22), seed=23,
                                                                 semi=(3, 6), every=5),
                                   kw=dict(nbrw=2, sorw=80, rng="reference")),
}


def make(name):
    c = CASES[name]
    v = c["vol"]()
    res = opipe.smart_interpolation(v.img, v.labels, **c["kw"])
    out = {"img": v.img, "labels": v.labels, "hits_u32": res["_hits_u32"],
           "accepted": np.int64(res["_accepted"]), "regular": res["regular"]}
    for k in ("smooth", "uncertainty"):
        if k in res:
            out[k] = res[k]
    return out


if __name__ == "__main__":
    for name in CASES:
        out = make(name)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, {k: getattr(v, "shape", v) for k, v in out.items()})
