"""Writes the golden fixtures of tests/golden/ from the literal reference.

Runs oracle/_ref/astro_ref_driver -- the reference's own Fortran modules
compiled from source with flang and driven like its *readModelInput + *solve
commands -- on small synthetic models written in the reference's input
format, and stores the arrays it dumps.  Only data is stored (inputs are
regenerated from the model parameters recorded in each file).

    python tests/golden/make_golden.py      (needs /root/reference and flang)
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from misc_amd.meshes import cantilever_block, thermal_block  # noqa: E402
from oracle import oracle as O  # noqa: E402

ELASTIC = [(2, 2, 1), (4, 4, 2), (6, 5, 3), (5, 3, 4)]
THERMAL = [(6, 5, 3)]
PREP_ONLY = [(24, 24, 16)]  # 10,625 nodes: 27*N*N wraps negative in 32-bit arithmetic
FULL_PREP = (64, 64, 32)    # the full-size block: 27*N*N wraps to a positive value

KEEP_ELASTIC = ["scalars", "currentRank", "nDofIndex", "elMatRange", "elMatLTRange", "elMatCols", "elasticMat",
                "elMPCMatRange0", "elMPCMatCols", "elMPCMat0", "elMPCMat", "elMPCRHS", "elMatLT", "elMatLTfact",
                "elasticLoad"]
KEEP_THERMAL = ["scalars", "currentRank", "thermMatRange", "thermMatLTRange", "thermMatCols", "thermMat",
                "thermMPCMat", "thermMPCRHS", "thermMatLT", "thermMatLTfact", "thermalLoad"]
KEEP_PREP = ["scalars", "currentRank", "elMatLTRange", "thermMatLTRange"]


def main():
    if not O.ref_available():
        sys.exit("build the reference driver first: make -C oracle ref")
    for d in ELASTIC:
        r = O.run_reference(cantilever_block(*d), "elastic")
        np.savez_compressed(os.path.join(HERE, "ref_elastic_%dx%dx%d.npz" % d), dims=np.array(d),
                            **{k: r[k] for k in KEEP_ELASTIC if k in r})
    for d in THERMAL:
        r = O.run_reference(thermal_block(*d), "thermal")
        np.savez_compressed(os.path.join(HERE, "ref_thermal_%dx%dx%d.npz" % d), dims=np.array(d),
                            **{k: r[k] for k in KEEP_THERMAL if k in r})
    for d in PREP_ONLY:
        r = O.run_reference(cantilever_block(*d), "prep")
        np.savez_compressed(os.path.join(HERE, "ref_prep_%dx%dx%d.npz" % d), dims=np.array(d),
                            **{k: r[k] for k in KEEP_PREP if k in r})
    r = O.run_reference(cantilever_block(*FULL_PREP), "prep")
    np.savez_compressed(os.path.join(HERE, "ref_prep_%dx%dx%d.npz" % FULL_PREP), dims=np.array(FULL_PREP),
                        scalars=r["scalars"], currentRank=r["currentRank"],
                        ltlen=np.diff(r["elMatLTRange"]).astype(np.uint16))
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
