"""Shared checks of the parity tests (not a test module).

support_flips / assert_same_support: the support rule of the solves -- the
same support above rel * max|ref| (rel = 1e-6).  A flip is an index where
exactly one of the two solutions exceeds the threshold; the assertion lists
every flip with both values in units of the threshold instead of exempting
any band around it (except entries on the threshold itself, within 1e-3 of
it in both solutions, which are printed).  Given the reference solved to a
tighter tolerance, a flip on which beta agrees with that tighter solve is
the reference's own unconverged entry and is returned instead of failing.

certificate: the algorithm-free LASSO optimality certificate of oracle.ipm
(prox-gradient fixed-point residual and duality gap of the residual-scaled
dual point, P(beta) - P* <= gap), computed with the CPU oracle operator.
"""
import numpy as np

from oracle import ipm
from oracle import operator as op


def support_flips(beta, ref, rel=1e-6):
    beta = np.asarray(beta)
    ref = np.asarray(ref)
    thr = rel * float(np.abs(ref).max())
    idx = np.nonzero((np.abs(beta) > thr) != (np.abs(ref) > thr))[0]
    return idx, thr


BOUNDARY = 1e-3   # half-width of the threshold band of a boundary flip, in units of thr


def assert_same_support(beta, ref, name, rel=1e-6, tighter=None):
    """tighter: the reference solved to a smaller tolerance.  A flip against
    ref where beta falls on the same side of the threshold as `tighter` is
    the reference's own unconverged entry (an interior-point iterate at tol
    1e-8 can still hold ~1e-5 max|beta| off the support), not a difference
    of beta's: such flips are returned, the others fail.

    A flip with both values within BOUNDARY * thr of the threshold (an entry
    sitting on it, which two solutions equal to 1e-9 can straddle) is a
    boundary flip: printed with its values, not failed."""
    idx, thr = support_flips(beta, ref, rel)
    if len(idx):
        a = np.abs(np.asarray(beta)[idx])
        r = np.abs(np.asarray(ref)[idx])
        on = (np.minimum(a, r) >= (1 - BOUNDARY) * thr) & (np.maximum(a, r) <= (1 + BOUNDARY) * thr)
        if on.any():
            det = [(int(i), float(beta[i]) / thr, float(ref[i]) / thr) for i in idx[on][:12]]
            print(f"  {name}: {int(on.sum())} boundary flip(s) within {BOUNDARY:g} thr of the "
                  f"threshold (index, beta/thr, ref/thr): {det}", flush=True)
        idx = idx[~on]
    excused = []
    if tighter is not None and len(idx):
        t = np.abs(np.asarray(tighter)[idx]) > thr
        b = np.abs(np.asarray(beta)[idx]) > thr
        excused = [int(i) for i in idx[t == b]]
        idx = idx[t != b]
    if len(idx):
        det = [(int(i), round(float(beta[i]) / thr, 4), round(float(ref[i]) / thr, 4))
               for i in idx[:12]]
        raise AssertionError(f"{name}: {len(idx)} support flips above {rel:g}*max|ref| "
                             f"(index, beta/thr, ref/thr): {det}")
    if tighter is not None:
        return int(np.sum(np.abs(np.asarray(ref)) > thr)), excused
    return int(np.sum(np.abs(np.asarray(ref)) > thr))


def certificate(p, beta, workers=8):
    """(fixed-point residual, duality gap, primal value) of beta for problem p."""
    b = op.m_perpt_z(p.z0, p.dims, workers=workers)
    mt = lambda v: op.mtm(v, p.dims, p.missing, workers=workers)            # noqa: E731
    mp = lambda v: op.m_perp_beta(v, p.dims, p.missing, workers=workers)    # noqa: E731
    return ipm.lasso_certificate(np.asarray(beta), b, p.lam, mt, mp, p.z0)


def oracle_solve(p, tol, workers=8, cg_rtol=None):
    """Oracle FP64 solve and its objective f(beta, |beta|)."""
    b = op.m_perpt_z(p.z0, p.dims, workers=workers)
    mt = lambda v: op.mtm(v, p.dims, p.missing, workers=workers)            # noqa: E731
    opts = ipm.IPMOptions(tol=tol) if cg_rtol is None else ipm.IPMOptions(tol=tol, cg_rtol=cg_rtol)
    r = ipm.solve(b, p.lam, mt, opts)
    return r, objective(p, r.beta, workers)


def objective(p, beta, workers=8):
    beta = np.asarray(beta)
    b = op.m_perpt_z(p.z0, p.dims, workers=workers)
    return ipm.objective(beta, np.abs(beta), op.mtm(beta, p.dims, p.missing, workers=workers),
                         b, p.lam)
