"""Writes the golden fixtures of tests/golden/ from the FP64 oracle.

    python tests/golden/make_golden.py

The reference holds no known-answer data for this path (its tests pin only the
weight-table cache round trip), so the fixtures are the oracle's outputs on
small fixed workloads plus the published Philox4x32-10 known-answer vectors
(Salmon et al., SC'11; Random123 kat_vectors).  They freeze the oracle: a
change to it that moves any of these numbers fails test_golden, and the GPU
path is compared with them directly.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402
from tests import cases  # noqa: E402

PHILOX_KAT = [
    {"ctr": [0, 0, 0, 0], "key": [0, 0], "out": [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]},
    {"ctr": [0xFFFFFFFF] * 4, "key": [0xFFFFFFFF] * 2, "out": [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]},
    {"ctr": [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], "key": [0xA4093822, 0x299F31D0],
     "out": [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]},
]

PATHS = 256


def path_fixture(case, name):
    table, pos32 = cases.setup(case, pyoracle)
    r = pyoracle.simulate(pos32.astype(np.float64), case.N, case.L, case.K, PATHS, table, case.absorption,
                          seed=case.seed, want_curv=True, want_exit=True)
    np.savez_compressed(os.path.join(HERE, name), log10w=r["log10w"], valid=r["valid"], curv=r["curv"],
                        exit=r["exit"], positions=pos32, table_head=table.table[:64], table_tail=table.table[-64:],
                        table_sum=np.float64(table.table.astype(np.float64).sum()),
                        grid=np.array([table.ds, table.minK, table.maxK, table.step], np.float32),
                        case=np.array([case.N, case.K, case.bins, case.numStepsInt, case.seed], np.int64),
                        params=np.array([case.L, case.sigma, case.mu, case.scatter, case.absorption], np.float64),
                        log10_sum_w=np.float64(r["log10_sum_w_exact"]))


def adjoint_fixture(case, name, paths=2048):
    table, pos32 = cases.setup(case, pyoracle)
    r = pyoracle.adjoint(pos32.astype(np.float64), case.N, case.L, case.K, paths, table, case.absorption,
                         seed=case.seed, lerp_mode=0)
    np.savez_compressed(os.path.join(HERE, name), d_table=r["d_table"], d_seg=r["d_seg"],
                        log10_sum_w=np.float64(r["log10_sum_w"]), paths=np.int64(paths),
                        case=np.array([case.N, case.K, case.bins, case.numStepsInt, case.seed], np.int64))


def main():
    pyoracle.build()
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump(PHILOX_KAT, f, indent=1)
    path_fixture(cases.C1, "paths_c1.npz")
    path_fixture(cases.C2, "paths_c2.npz")
    path_fixture(cases.MID, "paths_n33.npz")
    adjoint_fixture(cases.SMALL, "adjoint_n4.npz")
    adjoint_fixture(cases.MID, "adjoint_n33.npz")


if __name__ == "__main__":
    main()
