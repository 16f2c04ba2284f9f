"""CPU checks of the real-data front end and its oracle.

* index_missing: the literal loops of benchmarks/mastodonte.jl:38-50 equal
  the vectorised form on random, empty and full masks; a planted loop order
  (j fastest) is told apart.
* punch_3D_cart (crystal.jl:13-27): the host implementation equals the
  literal filter over every grid point, on the crystal.jl coordinate grid.
* reconstruct = sqrt(N) real(ifft(beta_to_DFT(beta))): equals the maskless
  oracle operator (R2C/C2R route) and inverts beta_true = DFT_to_beta(fft(x)
  / sqrt(N)) (test/ipm_example_3D.jl:82-83) for real x, even and odd sizes.
* array files round-trip (.npy / .npz, no pickles) and column-major helpers
  agree with the loop order.
"""
import numpy as np
import pytest

import csipm.realdata as rd
from oracle import operator as op
from oracle import realdata as ord_


@pytest.mark.parametrize("shape,p", [((5, 4, 3), 0.3), ((7, 6, 5), 0.0), ((4, 4, 4), 1.0),
                                     ((9, 2, 3), 0.5)])
def test_index_missing_loops(shape, p):
    m = (np.random.default_rng(1).random(shape) >= p).astype(float)
    ref = ord_.index_missing_loops(m)
    assert np.array_equal(ref, ord_.index_missing(rd.to_column_major(m)))
    if 0 < len(ref) < m.size:
        # a planted loop order with j fastest lists the same set in another order
        bad = [i + j * shape[0] + k * shape[0] * shape[1]
               for k in range(shape[2]) for i in range(shape[0]) for j in range(shape[1])
               if m[i, j, k] == 0]
        assert sorted(bad) == list(ref)
        assert np.all(np.diff(ref) > 0)
        if shape == (5, 4, 3):
            assert bad != list(ref)


def test_column_major_helpers():
    a = np.arange(24.0).reshape(2, 3, 4)
    v = rd.to_column_major(a)
    assert v[1] == a[1, 0, 0] and v[2] == a[0, 1, 0] and v[6] == a[0, 0, 1]
    assert np.array_equal(rd.from_column_major(v, (2, 3, 4)), a)


def test_punch_matches_filter():
    x = np.arange(-0.2, 4.01, 0.02)[:40]
    y = np.arange(-0.2, 6.01, 0.02)[:35]
    z = np.arange(-0.2, 6.01, 0.02)[:30]
    for c in ([0.0, 0.0, 0.0], [0.3, 0.2, 0.1], [0.5, 0.4, 0.3]):
        for r in (0.2001, (0.1, 0.2, 0.15)):
            ref = ord_.punch_loops(c, r, x, y, z)
            if len(ref) == 0:
                with pytest.raises(ValueError):
                    rd.punch_3D_cart(c, r, x, y, z)
                continue
            got = rd.punch_3D_cart(c, r, x, y, z)
            assert np.array_equal(np.sort(got), np.sort(ref))
            assert np.array_equal(got, ref)       # same (column-major) order


@pytest.mark.parametrize("dims", [(16,), (15,), (8, 6), (9, 7), (6, 5, 4), (7, 8, 5)])
def test_reconstruct_oracle(dims):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(int(np.prod(dims)))
    b = ord_.beta_of_signal(x, dims)
    xr = ord_.reconstruct(b, dims)
    assert np.abs(xr - x).max() <= 1e-13 * np.abs(x).max()
    beta = rng.standard_normal(int(np.prod(dims)))
    assert np.abs(ord_.reconstruct(beta, dims) - op.m_perp_beta(beta, dims)).max() <= 1e-13
    # a dropped sqrt(N) is caught
    assert np.abs(xr / np.sqrt(x.size) - x).max() > 1e-3


def test_array_files(tmp_path):
    a = np.random.default_rng(0).standard_normal((3, 4, 5))
    rd.save_array(tmp_path / "a.npy", a)
    assert np.array_equal(rd.load_array(tmp_path / "a.npy"), a)
    np.savez(tmp_path / "b.npz", data=a)
    assert np.array_equal(rd.load_array(tmp_path / "b.npz", "data"), a)
