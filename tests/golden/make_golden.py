"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

The map fixtures cover the shapes of the reference's unit tests
(test/unit_tests_ipm.jl:4, :80, :155) plus odd shapes of the extension; the
even-size values are pinned bit-exactly by tests/test_oracle_maps.py against
three independent restatements of the reference.  The operator fixtures hold
M^T M, M_perp_beta and M_perpt_z on masked problems (FP64).

Run from the repo root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import maps, operator  # noqa: E402

EVEN = [(100,), (200,), (500,), (16, 16), (12, 18), (16, 24), (32, 32),
        (8, 8, 8), (2, 4, 6), (6, 10, 12), (14, 8, 4), (8, 6, 4), (16, 16, 16)]
ODD = [(9,), (255,), (15, 7), (9, 10), (10, 9), (5, 7, 9), (6, 5, 4), (7, 6, 8)]
OPS = [((1024,), 0.08), ((255,), 0.1), ((16, 24), 0.1), ((15, 17), 0.12),
       ((8, 6, 4), 0.1), ((9, 10, 7), 0.1)]


def name(dims):
    return "x".join(str(d) for d in dims)


def main():
    for dims in EVEN + ODD:
        rng = np.random.default_rng(sum(dims) * 7919 + len(dims))
        n = int(np.prod(dims))
        beta = rng.standard_normal(n)
        x = rng.standard_normal(n)
        vh = np.fft.rfftn(x.reshape(dims[::-1])).reshape(-1) / np.sqrt(n)
        out = dict(dims=np.asarray(dims), beta_in=beta, x_in=x, v_in=vh,
                   v_half=maps.beta_to_dft(beta, dims, rdft=True),
                   beta_out=maps.dft_to_beta(vh, dims))
        if n <= 1024:
            out["v_full"] = maps.beta_to_dft(beta, dims, rdft=False)
        np.savez_compressed(os.path.join(HERE, f"maps_{name(dims)}.npz"), **out)
    for dims, frac in OPS:
        n = int(np.prod(dims))
        rng = np.random.default_rng(n * 31 + len(dims))
        missing = np.sort(rng.choice(n, size=int(frac * n), replace=False)).astype(np.int64)
        x = rng.standard_normal(n)
        z = rng.standard_normal(n)
        z[missing] = 0.0
        np.savez_compressed(
            os.path.join(HERE, f"op_{name(dims)}.npz"),
            dims=np.asarray(dims), missing=missing, x=x, z=z,
            mtm=operator.mtm(x, dims, missing),
            m_perp_beta=operator.m_perp_beta(x, dims, missing),
            m_perpt_z=operator.m_perpt_z(z, dims))


if __name__ == "__main__":
    main()
