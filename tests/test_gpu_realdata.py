"""GPU parity of the real-data front end (benchmarks/mastodonte.jl, crystal.jl).

The reference's own real-data inputs (HDF5 files, remote artifacts) are not
available; the tests use synthetic inputs of the same kind: random masks and
weights, and a punched-lattice 3D signal built as crystal.jl builds its
problem (spheres of radius 0.2001 around every integer lattice point of a
0.02-spaced grid).

* csipm_mask_index_missing equals the oracle's index_missing bit for bit:
  sizes around the block size, empty and full masks, fractional weights.
* csipm_reconstruct equals the oracle's sqrt(N) real(ifft(beta_to_DFT(beta)))
  within 1e-12 relative on masked handles (the mask is ignored), and maps
  beta_true = DFT_to_beta(fft(x)/sqrt(N)) back to x.
* mastodonte(): the mask -> index_missing -> solve -> reconstruct pipeline on
  a punched crystal; objective within 1e-6 of the FP64 oracle IPM,
  recovered signal within 1e-12 of the oracle reconstruction of the same beta.
"""
import numpy as np
import pytest
import torch

from oracle import ipm
from oracle import operator as op
from oracle import realdata as ord_

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 7, 255, 4095, 4096, 4097, 65537, 300001])
def test_mask_index_missing(cuda, n):
    import csipm.realdata as rd
    print("mask", n, flush=True)
    rng = np.random.default_rng(n)
    for w in (rng.random(n) * (rng.random(n) > 0.3), np.zeros(n), np.ones(n),
              (rng.random(n) > 0.999).astype(float)):
        got = rd.index_missing_from_mask(torch.from_numpy(w).to(cuda)).cpu().numpy()
        assert np.array_equal(got, ord_.index_missing(w)), n


@pytest.mark.parametrize("dims", [(100,), (127,), (8192,), (64, 48), (31, 17), (12, 10, 8),
                                  (15, 9, 7)])
def test_reconstruct(cuda, dims):
    from csipm.fft_operator import FFTOperator
    import csipm.realdata as rd
    print("reconstruct", dims, flush=True)
    n = int(np.prod(dims))
    rng = np.random.default_rng(11)
    miss = np.sort(rng.choice(n, size=n // 7, replace=False))
    o = FFTOperator(n, len(dims), dims, miss)
    beta = rng.standard_normal(n)
    got = rd.reconstruct(o, torch.from_numpy(beta).to(cuda)).cpu().numpy()
    ref = ord_.reconstruct(beta, dims)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    x = rng.standard_normal(n)
    back = rd.reconstruct(o, torch.from_numpy(ord_.beta_of_signal(x, dims)).to(cuda))
    assert np.abs(back.cpu().numpy() - x).max() <= 1e-12 * np.abs(x).max()
    # the handle's own M_perp_beta still applies the mask afterwards
    from csipm.fft_operator import M_perp_beta
    y = M_perp_beta(o, torch.from_numpy(beta).to(cuda)).cpu().numpy()
    assert np.abs(y - op.m_perp_beta(beta, dims, miss)).max() <= 1e-12 * np.abs(ref).max()


def _crystal(shape=(26, 31, 31), spacing=0.1):
    """A punched-lattice problem in the manner of crystal.jl:30-80: a smooth
    periodic signal on a grid from -0.2 (spacing 0.1 here, 0.02 there, so
    that a small box holds several lattice points), zeroed in spheres of
    radius 0.2001 around the integer lattice points; mask = 0 there."""
    import csipm.realdata as rd
    x = -0.2 + spacing * np.arange(shape[0])
    y = -0.2 + spacing * np.arange(shape[1])
    z = -0.2 + spacing * np.arange(shape[2])
    X, Y, Zg = np.meshgrid(x, y, z, indexing="ij")
    sig = (np.cos(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.5 * np.cos(2 * np.pi * Zg)
           + 0.25 * np.sin(4 * np.pi * (X + Y)))
    centers = [(i, j, k) for i in range(3) for j in range(3) for k in range(3)]
    idx = rd.punch_lattice(x, y, z, centers, 0.2001)
    m = np.ones(int(np.prod(shape)))
    m[idx] = 0.0
    z = rd.to_column_major(sig).copy()
    z[idx] = 0.0
    return rd.from_column_major(z, shape), rd.from_column_major(m, shape), idx


@pytest.mark.parametrize("mixed", [False, True])
def test_mastodonte_pipeline(cuda, mixed):
    import csipm.realdata as rd
    print("mastodonte", "mixed" if mixed else "f64", flush=True)
    z0, mask, idx = _crystal()
    dims = mask.shape
    lam = 1e-2
    res = rd.mastodonte(z0, mask, lambda_=lam, rdft=False, max_iter=200,
                        tol=1e-7 if mixed else 1e-8, mixed=mixed)
    assert res.status == "solved"
    assert np.array_equal(res.index_missing, np.sort(idx))
    zc = rd.to_column_major(z0)
    b = op.m_perpt_z(zc, dims)
    mt = lambda v: op.mtm(v, dims, idx)    # noqa: E731
    ref = ipm.solve(b, lam, mt, ipm.IPMOptions(tol=1e-10, cg_rtol=1e-13))
    beta = res.beta.cpu().numpy()
    f = ipm.objective(beta, np.abs(beta), mt(beta), b, lam)
    f_ref = ipm.objective(ref.beta, np.abs(ref.beta), mt(ref.beta), b, lam)
    assert abs(f - f_ref) <= 1e-6 * abs(f_ref)
    rec = res.recovered.cpu().numpy()
    want = ord_.reconstruct(beta, dims)
    assert np.abs(rec - want).max() <= 1e-12 * np.abs(want).max()
    # the recovery fills the punched holes with the smooth signal's values
    assert np.abs(rec[idx]).max() > 0.0
