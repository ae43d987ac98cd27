"""Write tests/golden/ls_lstsq.npz: dense least-squares cases solved by
numpy's LAPACK (numpy.linalg.lstsq, numpy.linalg.qr) in this container.

Each case k stores A_k (m x n, well conditioned unless named otherwise), b_k,
x_k = lstsq(A_k, b_k) and R_k = qr(A_k, mode='r').  The oracle's Householder
QR uses LAPACK's sign convention (beta = -sign(alpha) ||x||), so R must match
numpy's R entry for entry, not only up to sign.

Run: python tests/golden/make_ls_fixtures.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.Generator(np.random.PCG64(7))
    shapes = [(1, 1), (2, 1), (5, 1), (5, 3), (13, 5), (25, 7), (63, 25), (64, 25), (65, 25),
              (130, 26), (200, 33), (365, 33), (40, 40)]
    out = {}
    for k, (m, n) in enumerate(shapes):
        A = rng.standard_normal((m, n))
        A[np.arange(min(m, n)), np.arange(min(m, n))] += 3.0  # keep it well conditioned
        b = np.zeros(m)
        b[rng.integers(0, m)] = 1.0
        x = np.linalg.lstsq(A, b, rcond=None)[0]
        R = np.linalg.qr(A, mode="r")
        out[f"A{k}"] = A
        out[f"b{k}"] = b
        out[f"x{k}"] = x
        out[f"R{k}"] = R
    # sparse-like SPAI submatrix: 2D Laplacian columns, interior column pattern
    out["ncase"] = np.array(len(shapes))
    np.savez_compressed(os.path.join(HERE, "ls_lstsq.npz"), **out)


if __name__ == "__main__":
    main()
