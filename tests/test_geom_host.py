"""CPU check of the device packing arithmetic.

libcsipm_geomcheck.so runs, on the host, the very functions the GPU kernels
use for the beta <-> DFT maps (beta_to_dft_elem, dft_to_beta_elem,
scatter_targets from csrc/csipm_device.hpp and csrc/csipm_kernels.hpp).  They
must reproduce the oracle bit-exactly, every index must fall inside its
buffer, and the scatter used by the fused operator must hit each beta index
exactly once per component and agree with the gather form.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import maps

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "compressedsensingipm.jl_amd", "csipm",
                   "libcsipm_geomcheck.so")
SHAPES = [(8,), (9,), (100,), (255,), (1024,), (2, 2), (16, 16), (12, 18), (15, 7), (9, 10),
          (10, 9), (255, 255), (63, 64), (2, 4, 6), (8, 6, 4), (6, 10, 12), (14, 8, 4),
          (5, 7, 9), (6, 5, 4), (7, 6, 8), (3, 3, 3), (2, 2, 2), (31, 30, 29), (17, 16, 15),
          (40, 32, 24)]


@pytest.fixture(scope="module")
def glib():
    if not os.path.exists(LIB):
        pytest.fail("libcsipm_geomcheck.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    p = ctypes.c_void_p
    L.geomchk_info.argtypes = [ctypes.c_int, p, p]
    L.geomchk_beta_to_dft.argtypes = [ctypes.c_int, p, p, p, ctypes.c_int]
    L.geomchk_dft_to_beta.argtypes = [ctypes.c_int, p, p, p, ctypes.c_int]
    L.geomchk_scatter.argtypes = [ctypes.c_int, p, p, p, p]
    L.geomchk_via_u.argtypes = [ctypes.c_int, p, p, p, p, p]
    return L


def _sizes(dims):
    return np.asarray(list(dims) + [1] * (3 - len(dims)), dtype=np.int64)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("dims", SHAPES)
def test_device_maps_match_oracle(glib, dims):
    n = int(np.prod(dims))
    sz = _sizes(dims)
    rng = np.random.default_rng(n)
    beta = rng.standard_normal(n)
    for full in (0, 1):
        rows = dims[0] if full else dims[0] // 2 + 1
        m = rows * (n // dims[0])
        v = np.full(2 * m, np.nan)
        glib.geomchk_beta_to_dft(len(dims), _p(sz), _p(beta), _p(v), full)
        want = maps.beta_to_dft(beta, dims, rdft=not full)
        assert np.array_equal(v[0::2], want.real) and np.array_equal(v[1::2], want.imag)
        # gather form DFT -> beta from the spectrum of a real signal
        x = rng.standard_normal(n)
        spec = np.fft.fftn(x.reshape(dims, order="F")) / np.sqrt(n)
        spec = spec[:rows].reshape(-1, order="F")
        vi = np.empty(2 * m)
        vi[0::2] = spec.real
        vi[1::2] = spec.imag
        out = np.full(n, np.nan)
        glib.geomchk_dft_to_beta(len(dims), _p(sz), _p(vi), _p(out), full)
        assert np.array_equal(out, maps.dft_to_beta(spec, dims))


@pytest.mark.parametrize("dims", SHAPES)
def test_scatter_is_inverse_of_gather(glib, dims):
    n = int(np.prod(dims))
    sz = _sizes(dims)
    h1 = dims[0] // 2 + 1
    m = h1 * (n // dims[0])
    re = np.empty(m, np.int64)
    im = np.empty(m, np.int64)
    fl = np.empty(m, np.int8)
    glib.geomchk_scatter(len(dims), _p(sz), _p(re), _p(im), _p(fl))
    hit = (fl & 1) == 1
    # every target inside [0, N), each beta index written exactly once
    targets = np.concatenate([re[hit], im[hit & (im >= 0)]])
    assert targets.min() >= 0 and targets.max() < n
    assert np.array_equal(np.sort(targets), np.arange(n))
    # same as the oracle's inverse view of DFT_to_beta
    ore, oim, osign = maps.scatter_targets(dims)
    assert np.array_equal(np.where(hit, re, -1), ore)
    assert np.array_equal(np.where(hit, im, -1), oim)
    neg = (fl & 4) == 4
    assert np.array_equal(neg & (oim >= 0), (osign < 0) & (oim >= 0))


@pytest.mark.parametrize("dims", SHAPES)
def test_spectrum_order_is_exact(glib, dims):
    """The spectrum-ordered copy u of beta used inside CG and the IPM: the FFT
    passes' load (u -> v) reproduces beta_to_DFT bit-exactly, and their store
    followed by the u -> beta conversion reproduces DFT_to_beta bit-exactly."""
    n = int(np.prod(dims))
    sz = _sizes(dims)
    rng = np.random.default_rng(n + 1)
    beta = rng.standard_normal(n)
    h1 = dims[0] // 2 + 1
    m = h1 * (n // dims[0])
    x = rng.standard_normal(n)
    spec = (np.fft.fftn(x.reshape(dims, order="F")) / np.sqrt(n))[:h1].reshape(-1, order="F")
    vin = np.empty(2 * m)
    vin[0::2] = spec.real
    vin[1::2] = spec.imag
    v = np.full(2 * m, np.nan)
    back = np.full(n, np.nan)
    glib.geomchk_via_u(len(dims), _p(sz), _p(beta), _p(v), _p(back), _p(vin))
    want = maps.beta_to_dft(beta, dims, rdft=True)
    assert np.array_equal(v[0::2], want.real) and np.array_equal(v[1::2], want.imag)
    assert np.array_equal(back, maps.dft_to_beta(spec, dims))
