"""GPU parity of the fused FFT operator through the C ABI.

FP64: within 1e-12 relative (max-norm) of the FP64 CPU oracle; the
reference's identities (M^T M = I without mask, norm preservation,
test/unit_tests_ipm.jl:7-30) hold to the same tolerance.  FP32: 2e-6.
"""
import os

import numpy as np
import pytest
import torch

from oracle import kkt, maps
from oracle import operator as op

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL64 = 1e-12
TOL32 = 2e-6


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def _make(dims, missing=(), dtype=torch.float64):
    from csipm.fft_operator import FFTOperator
    return FFTOperator(int(np.prod(dims)), len(dims), dims, np.asarray(missing, np.int64),
                       rdft=True, dtype=dtype)


UNIT = [(100,), (200,), (500,), (16, 16), (12, 18), (16, 24), (32, 32), (8, 8, 8), (2, 4, 6),
        (6, 10, 12), (14, 8, 4), (8, 6, 4), (16, 16, 16)]


@pytest.mark.parametrize("dims", UNIT + [(255,), (1024,), (255, 255), (15, 17, 9)])
def test_identities_no_mask(cuda, dims):
    from csipm.fft_operator import M_perp_beta, M_perpt_M_perp_vec, M_perpt_z
    print("identities", dims, flush=True)
    rng = np.random.default_rng(1)
    z = rng.uniform(size=int(np.prod(dims)))
    o = _make(dims)
    zt = torch.from_numpy(z).to(cuda)
    assert _rel(M_perpt_M_perp_vec(o, zt).cpu().numpy(), z) <= TOL64
    r1 = M_perpt_z(o, zt).cpu().numpy()
    r2 = M_perp_beta(o, zt).cpu().numpy()
    assert abs(np.linalg.norm(r1) - np.linalg.norm(z)) <= TOL64 * np.linalg.norm(z)
    assert abs(np.linalg.norm(r2) - np.linalg.norm(z)) <= TOL64 * np.linalg.norm(z)
    assert _rel(r1, op.m_perpt_z(z, dims)) <= TOL64
    assert _rel(r2, op.m_perp_beta(z, dims)) <= TOL64


@pytest.mark.parametrize("fname", sorted(f for f in os.listdir(GOLD) if f.startswith("op_")))
def test_operator_golden(cuda, fname):
    from csipm.fft_operator import M_perp_beta, M_perpt_M_perp_vec, M_perpt_z
    d = np.load(os.path.join(GOLD, fname))
    dims = tuple(int(v) for v in d["dims"])
    print("operator", dims, flush=True)
    o = _make(dims, d["missing"])
    x = torch.from_numpy(d["x"]).to(cuda)
    z = torch.from_numpy(d["z"]).to(cuda)
    assert _rel(M_perpt_M_perp_vec(o, x).cpu().numpy(), d["mtm"]) <= TOL64
    assert _rel(M_perp_beta(o, x).cpu().numpy(), d["m_perp_beta"]) <= TOL64
    assert _rel(M_perpt_z(o, z).cpu().numpy(), d["m_perpt_z"]) <= TOL64


# lengths whose prime factors are at most 31 factor directly (radices 19..31)
LARGE_RADIX = [(19, 23), (23, 19, 29), (31, 62), (57, 87)]
# a prime length whose L - 1 factors runs Rader's form (including dimension 1
# of a 2D/3D problem and the long-line kernel of 1D N = 4001)
RADER = [(127,), (4001,), (191, 127), (127, 60), (12, 10, 37), (37, 41, 47), (53, 59, 61)]
# other lengths with a prime factor above 31 run Bluestein's chirp-z form with
# a factorable convolution length M >= 2L - 1 (4007: M = 8064 on the 1024-thread kernel)
BLUESTEIN = [(83,), (4007,), (111, 20), (20, 111), (9, 8, 83), (107, 13, 10)]
# 1D lines above 4096 run as four-step transforms: direct when N factors into
# the supported radices, else with an outer Bluestein convolution
FOURSTEP = [(8192,), (50000,), (65536,), (4097,), (12289,), (65521,)]


@pytest.mark.parametrize("dims", [(1024,), (255, 255), (64, 48, 30), (51, 45, 35)] + LARGE_RADIX +
                         RADER + BLUESTEIN + FOURSTEP)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_operator_masked_vs_oracle(cuda, dims, dtype):
    from csipm.fft_operator import M_perp_beta, M_perpt_M_perp_vec, M_perpt_z
    print("operator", dims, dtype, flush=True)
    n = int(np.prod(dims))
    rng = np.random.default_rng(n)
    miss = np.sort(rng.choice(n, size=n // 9, replace=False))
    x = rng.standard_normal(n)
    tol = TOL64 if dtype == torch.float64 else TOL32
    o = _make(dims, miss, dtype)
    xt = torch.from_numpy(x).to(cuda, dtype)
    assert _rel(M_perpt_M_perp_vec(o, xt).double().cpu().numpy(), op.mtm(x, dims, miss)) <= tol
    assert _rel(M_perp_beta(o, xt).double().cpu().numpy(), op.m_perp_beta(x, dims, miss)) <= tol
    assert _rel(M_perpt_z(o, xt).double().cpu().numpy(), op.m_perpt_z(x, dims)) <= tol


# a tuned Bluestein convolution length (csipm_tune_bluestein) changes the
# plan, not the operator
@pytest.mark.parametrize("dims,L,M", [((83,), 83, 176), ((111, 20), 111, 240),
                                      ((9, 8, 83), 83, 171), ((37, 41, 47), 37, 80)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_operator_tuned_bluestein(cuda, dims, L, M, dtype):
    from csipm import _lib, tune
    from csipm.fft_operator import M_perp_beta, M_perpt_M_perp_vec, M_perpt_z
    print("tuned Bluestein", dims, L, M, dtype, flush=True)
    lib = _lib.lib()
    n = int(np.prod(dims))
    rng = np.random.default_rng(n + M)
    miss = np.sort(rng.choice(n, size=n // 9, replace=False))
    x = rng.standard_normal(n)
    tol = TOL64 if dtype == torch.float64 else TOL32
    assert lib.csipm_tune_bluestein(L, M) == 0
    try:
        o = _make(dims, miss, dtype)
        axis = list(dims).index(L)
        if (L, dims) == (37, (37, 41, 47)):
            assert tune.fft_shape(o, axis)[2] == 2       # a Rader length keeps Rader's form
        else:
            assert tune.fft_shape(o, axis) == (M, M, 1)
        xt = torch.from_numpy(x).to(cuda, dtype)
        assert _rel(M_perpt_M_perp_vec(o, xt).double().cpu().numpy(), op.mtm(x, dims, miss)) <= tol
        assert _rel(M_perp_beta(o, xt).double().cpu().numpy(), op.m_perp_beta(x, dims, miss)) <= tol
        assert _rel(M_perpt_z(o, xt).double().cpu().numpy(), op.m_perpt_z(x, dims)) <= tol
    finally:
        lib.csipm_tune_bluestein(L, 0)
        tune.load(lib)


@pytest.mark.parametrize("dims", [(1024,), (255, 255), (33, 30, 26)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_kkt_and_precond(cuda, dims, dtype):
    from csipm import _lib
    from csipm.fft_operator import _ptr
    print("kkt", dims, dtype, flush=True)
    n = int(np.prod(dims))
    rng = np.random.default_rng(3)
    miss = np.sort(rng.choice(n, size=n // 10, replace=False))
    o = _make(dims, miss, dtype)
    s1 = rng.uniform(0.05, 5.0, n)
    s2 = rng.uniform(0.05, 5.0, n)
    T = lambda a: torch.from_numpy(np.asarray(a)).to(cuda, dtype).contiguous()   # noqa: E731
    L1, L2, P11, P12, P22 = (torch.empty(n, device=cuda, dtype=dtype) for _ in range(5))
    s1t, s2t = T(s1), T(s2)       # keep the inputs alive while the kernels run
    _lib.call("csipm_precond_build", o.handle, _ptr(s1t), _ptr(s2t), _ptr(L1), _ptr(L2),
              _ptr(P11), _ptr(P12), _ptr(P22))
    ref = kkt.precond_build(s1, s2)
    tol = TOL64 if dtype == torch.float64 else TOL32
    for got, want in zip((L1, L2, P11, P12, P22), ref):
        assert _rel(got.double().cpu().numpy(), want) <= 10 * tol
    x = rng.standard_normal(2 * n)
    y0 = rng.standard_normal(2 * n)
    y = T(y0)
    xt = T(x)
    _lib.call("csipm_kkt_apply", o.handle, _ptr(xt), _ptr(y), _ptr(L1), _ptr(L2), 0.75, -0.5)
    want = kkt.kkt_mul(y0, x, ref[0], ref[1], lambda v: op.mtm(v, dims, miss), 0.75, -0.5)
    assert _rel(y.double().cpu().numpy(), want) <= 10 * tol
    y = T(y0)
    _lib.call("csipm_precond_apply", o.handle, _ptr(xt), _ptr(y), _ptr(P11), _ptr(P12),
              _ptr(P22), 2.0, 0.25)
    want = kkt.precond_mul(y0, x, ref[2], ref[3], ref[4], 2.0, 0.25)
    assert _rel(y.double().cpu().numpy(), want) <= 10 * tol


@pytest.mark.parametrize("dims", [(1024,), (255, 255), (33, 30, 26), (127, 60), (20000,),
                                  (12289,)])
def test_cg_solve(cuda, dims):
    import ctypes
    from csipm import _lib
    from csipm.fft_operator import _ptr
    print("cg", dims, flush=True)
    n = int(np.prod(dims))
    rng = np.random.default_rng(4)
    miss = np.sort(rng.choice(n, size=n // 10, replace=False))
    o = _make(dims, miss)
    s1 = rng.uniform(0.05, 5.0, n)
    s2 = rng.uniform(0.05, 5.0, n)
    L1, L2, P11, P12, P22 = kkt.precond_build(s1, s2)
    b = rng.standard_normal(2 * n)
    mt = lambda v: op.mtm(v, dims, miss)      # noqa: E731
    xr, itr, resr, str_ = kkt.pcg(lambda v: kkt.kkt_mul(np.zeros(2 * n), v, L1, L2, mt),
                                  b, lambda r: kkt.precond_mul(np.zeros(2 * n), r, P11, P12, P22),
                                  atol=1e-12, rtol=0.0)
    T = lambda a: torch.from_numpy(np.asarray(a)).to(cuda).contiguous()   # noqa: E731
    x = torch.zeros(2 * n, device=cuda, dtype=torch.float64)
    it = ctypes.c_int64()
    res = ctypes.c_double()
    st = ctypes.c_int()
    keep = [T(a) for a in (b, L1, L2, P11, P12, P22)]   # alive during the solve
    bt, L1t, L2t, P11t, P12t, P22t = keep
    _lib.call("csipm_cg_solve", o.handle, _ptr(bt), _ptr(x), _ptr(L1t), _ptr(L2t),
              _ptr(P11t), _ptr(P12t), _ptr(P22t), 1e-12, 0.0, 0, ctypes.byref(it),
              ctypes.byref(res), ctypes.byref(st))
    assert st.value == 0 and str_ == "solved"
    assert abs(it.value - itr) <= 1
    assert res.value <= 1e-12
    assert _rel(x.cpu().numpy(), xr) <= 1e-10


@pytest.mark.parametrize("dims", [(1000,), (16, 12), (15, 17), (9, 10, 7), (20000,)])
def test_complex_path_operator(cuda, dims):
    """An rdft=False handle (the reference's default) against the oracle's
    complex path; its maps default to the full spectrum."""
    from csipm.fft_operator import (FFTOperator, M_perp_beta, M_perpt_M_perp_vec, M_perpt_z,
                                    beta_to_DFT, DFT_to_beta)
    print("complex path", dims, flush=True)
    n = int(np.prod(dims))
    rng = np.random.default_rng(n + 2)
    miss = np.sort(rng.choice(n, size=n // 8, replace=False))
    o = FFTOperator(n, len(dims), dims, miss, rdft=False)
    x = rng.standard_normal(n)
    xt = torch.from_numpy(x).to(cuda)
    assert _rel(M_perpt_M_perp_vec(o, xt).cpu().numpy(), op.mtm(x, dims, miss, rdft=False)) <= TOL64
    assert _rel(M_perp_beta(o, xt).cpu().numpy(), op.m_perp_beta(x, dims, miss, rdft=False)) <= TOL64
    assert _rel(M_perpt_z(o, xt).cpu().numpy(), op.m_perpt_z(x, dims, rdft=False)) <= TOL64
    v = beta_to_DFT(o, xt).cpu().numpy()
    assert v.size == n
    assert np.array_equal(v, maps.beta_to_dft(x, dims, rdft=False))
    assert np.array_equal(DFT_to_beta(o, torch.from_numpy(v).to(cuda)).cpu().numpy(),
                          maps.dft_to_beta(v, dims))
