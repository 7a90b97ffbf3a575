"""Writes tests/golden/vumps_xxz_d64.json: the oracle VUMPS (FP64 CPU, ARPACK eigensolves) on the
XXZ chain in reference convention Delta = -1 (the antiferromagnetic Heisenberg chain), J = 1,
hz = 0, two-site cell, QN = [0], bond 0 with 9 sectors of sizes [1, 3, 7, 12, 18, 12, 7, 3, 1]
(D = 64), seed 1001, iterated to a gradient norm F <= 1e-10 (the reference's F, norm_scaled)."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from oracle_vumps import OracleVUMPS  # noqa: E402

KEYS0 = [[2 * (i - 4)] for i in range(9)]
SIZES0 = [1, 3, 7, 12, 18, 12, 7, 3, 1]


def main():
    t0 = time.time()
    v = OracleVUMPS(0, 2, [0], np.array(KEYS0, np.int32), np.array(SIZES0, np.int32), [1.0, -1.0, 0.0], seed=1001)
    hist = []
    for it in range(400):
        e, F = v.sweep()
        hist.append([float(e), float(max(F))])
        print(it, "%.15f" % e, "%.3e" % max(F), flush=True)
        if max(F) <= 1e-10:
            break
    out = dict(model="XXZ", params=[1.0, -1.0, 0.0], N=2, QN=[0], keys0=KEYS0, sizes0=SIZES0, seed=1001,
               energy=hist[-1][0], F=hist[-1][1], sweeps=len(hist), bethe_ansatz=0.25 - np.log(2),
               seconds=time.time() - t0)
    with open(os.path.join(HERE, "vumps_xxz_d64.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
