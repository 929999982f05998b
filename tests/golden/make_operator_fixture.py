"""Write tests/golden/operator_checksums.json from the reference's shipped
operator tables (dim3/data/bin_mats/N*/{rst,M,Dr,...}.bin, read as data).

Two checksums per table (sum of squares and a cosine-weighted sum) pin the
tables without copying them.  Run in the build container:
    python tests/golden/make_operator_fixture.py
"""
import json
import os

import numpy as np

REF = "/root/reference/dim3/data/bin_mats"


def checksum(a):
    a = np.asarray(a, dtype=np.float64).ravel()
    w = np.cos(np.arange(a.size) * 0.7311)
    return [float(np.sum(a * a)), float(np.sum(a * w))]


def main():
    out = {}
    for N in range(1, 11):
        Np = (N + 1) * (N + 2) * (N + 3) // 6
        Nfp = (N + 1) * (N + 2) // 2
        hN = N + 2
        hNp = (hN + 1) * (hN + 2) * (hN + 3) // 6
        shapes = dict(rst=(Np, 3), M=(Np, Np), Dr=(Np, Np), Ds=(Np, Np), Dt=(Np, Np), WDr=(Np, Np), WDs=(Np, Np),
                      WDt=(Np, Np), Emat=(Np, 4 * Nfp), LIFT=(Np, 4 * Nfp), interpM2=(hNp, Np))
        out[str(N)] = {k: checksum(np.fromfile(os.path.join(REF, f"N{N}", f"{k}.bin")).reshape(s))
                       for k, s in shapes.items()}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "operator_checksums.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
