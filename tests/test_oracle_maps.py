"""Pins the CPU map oracle (oracle/maps.py).

Even sizes: the oracle's block table must agree BIT-EXACTLY with three
independent restatements of the reference (kernel, loop and concatenation
forms, oracle/maps_ref_forms.py).  Odd sizes (a documented extension): exact
Hermitian output, round trip, norm preservation, and reduction to the even
layout is covered by the even tests.  A planted mistake shows the checks bite.
"""
import os

import numpy as np
import pytest

from oracle import maps
from oracle import maps_ref_forms as rf

EVEN = [(100,), (200,), (8,), (16, 16), (12, 18), (16, 24), (8, 8, 8), (2, 4, 6),
        (6, 10, 12), (14, 8, 4), (8, 6, 4)]
ODD = [(9,), (255,), (15, 7), (9, 10), (10, 9), (5, 7, 9), (6, 5, 4), (7, 6, 8), (3, 3, 3)]
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _spectra(dims, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    x = rng.standard_normal(n)
    vfull = np.fft.fftn(x.reshape(dims, order="F")) / np.sqrt(n)
    return x, vfull


def _same(a, b):
    return np.array_equal(np.asarray(a).reshape(-1), np.asarray(b).reshape(-1))


@pytest.mark.parametrize("dims", EVEN)
def test_dft_to_beta_three_forms(dims):
    x, vfull = _spectra(dims, 1)
    h1 = dims[0] // 2 + 1
    vhalf = vfull[:h1]
    ours = maps.dft_to_beta(vhalf.reshape(-1, order="F"), dims)
    assert _same(ours, rf.kernel_dft_to_beta(vhalf, dims))
    assert _same(ours, rf.loop_dft_to_beta(vhalf, dims))
    # the concatenation form reads second-half rows directly: identical on an
    # exactly Hermitian spectrum, equal to rounding on an FFT output
    herm = 0.5 * (vfull + np.conj(np.roll(np.flip(vfull), 1, axis=tuple(range(len(dims))))))
    assert _same(maps.dft_to_beta(herm[:h1].reshape(-1, order="F"), dims),
                 rf.wei_dft_to_beta(herm, dims))
    np.testing.assert_allclose(ours, rf.wei_dft_to_beta(vfull, dims), rtol=0, atol=1e-14)


@pytest.mark.parametrize("dims", EVEN)
def test_beta_to_dft_three_forms(dims):
    rng = np.random.default_rng(2)
    beta = rng.standard_normal(int(np.prod(dims)))
    h1 = dims[0] // 2 + 1
    half = maps.beta_to_dft(beta, dims, rdft=True).reshape((h1,) + dims[1:], order="F")
    full = maps.beta_to_dft(beta, dims, rdft=False).reshape(dims, order="F")
    assert _same(half, rf.kernel_beta_to_dft(beta, dims, rdft=True))
    assert _same(half, rf.loop_beta_to_dft(beta, dims, rdft=True))
    assert _same(full, rf.kernel_beta_to_dft(beta, dims, rdft=False))
    assert _same(full, rf.loop_beta_to_dft(beta, dims, rdft=False))
    assert _same(full, rf.wei_beta_to_dft(beta, dims))


@pytest.mark.parametrize("dims", EVEN + ODD)
def test_invariants(dims):
    n = int(np.prod(dims))
    x, vfull = _spectra(dims, 3)
    h1 = dims[0] // 2 + 1
    beta = maps.dft_to_beta(vfull[:h1].reshape(-1, order="F"), dims)
    # unitary: ||beta|| = ||x||
    assert abs(np.linalg.norm(beta) - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)
    # round trip to the full spectrum of the real signal
    back = maps.beta_to_dft(beta, dims, rdft=False).reshape(dims, order="F")
    assert np.abs(back - vfull).max() <= 1e-13 * np.abs(vfull).max()
    # beta -> DFT output is exactly Hermitian and DFT -> beta inverts it exactly
    rng = np.random.default_rng(4)
    b = rng.standard_normal(n)
    v = maps.beta_to_dft(b, dims, rdft=False).reshape(dims, order="F")
    mirror = np.conj(np.roll(np.flip(v), 1, axis=tuple(range(len(dims)))))
    assert _same(v, mirror)
    rt = maps.dft_to_beta(v[:h1].reshape(-1, order="F"), dims)
    assert np.abs(rt - b).max() <= 4.5e-16 * max(1.0, np.abs(b).max())
    # real signal from beta: imaginary part of the inverse FFT vanishes
    y = np.fft.ifftn(v)
    assert np.abs(y.imag).max() <= 1e-14 * np.abs(y).max()


def test_half_matches_full_rows():
    for dims in EVEN + ODD:
        b = np.random.default_rng(5).standard_normal(int(np.prod(dims)))
        h1 = dims[0] // 2 + 1
        half = maps.beta_to_dft(b, dims, rdft=True).reshape((h1,) + dims[1:], order="F")
        full = maps.beta_to_dft(b, dims, rdft=False).reshape(dims, order="F")
        assert _same(half, full[:h1])


@pytest.mark.parametrize("fname", sorted(f for f in os.listdir(GOLD) if f.startswith("maps_")))
def test_golden_fixtures(fname):
    d = np.load(os.path.join(GOLD, fname))
    dims = tuple(int(v) for v in d["dims"])
    assert _same(maps.beta_to_dft(d["beta_in"], dims, rdft=True), d["v_half"])
    assert _same(maps.dft_to_beta(d["v_in"], dims), d["beta_out"])
    if "v_full" in d:
        assert _same(maps.beta_to_dft(d["beta_in"], dims, rdft=False), d["v_full"])


def test_planted_mistakes_are_caught():
    """Plant two plausible mistakes in a copy of the table (a dropped minus sign
    on a mirror block, a transposed block offset) and check that the
    comparison with the reference's kernel form detects each."""
    dims = (6, 10, 12)
    x, vfull = _spectra(dims, 6)
    h1 = dims[0] // 2 + 1
    vhalf = vfull[:h1]
    ref = rf.kernel_dft_to_beta(vhalf, dims)
    t = maps.table(dims)
    saved = (t.coord.copy(), t.comp.copy())
    try:
        # (1) treat a mirror-read element as if its Im were read directly:
        # swap the comp of one Re/Im pair of a mirror block
        far = np.nonzero(t.coord[:, 0] > t.M[0])[0]
        i = far[len(far) // 2]
        t.comp[i] = 1 - t.comp[i]
        bad = maps.dft_to_beta(vhalf.reshape(-1, order="F"), dims)
        assert not _same(bad, ref)
        t.coord[:], t.comp[:] = saved
        # (2) transpose two coordinates inside an interior block
        j = t.N - 3
        t.coord[j], t.coord[j - 1] = t.coord[j - 1].copy(), t.coord[j].copy()
        bad = maps.dft_to_beta(vhalf.reshape(-1, order="F"), dims)
        assert not _same(bad, ref)
    finally:
        t.coord[:], t.comp[:] = saved
    assert _same(maps.dft_to_beta(vhalf.reshape(-1, order="F"), dims), ref)
