"""The Gondzio formulation of the recovery problem (src/gondzio_model.jl,
src/gondzio_kkt.jl) over libcsipm.

x = [p; q; r] (3N), beta = p - q with p, q >= 0, r free; one equality block
c(x) = z0 - M (p - q) - r = 0; objective 1/2 r'r + lambda sum(p + q).  The
classes mirror the reference's GondzioNLPModel, CondensedGondzioKKTSystem,
GondzioPreconditioner and GondzioKKTSystem (initialize / build_kkt /
solve_kkt / mul / jtprod); `gondzio_solve` runs the library's interior-point
driver on the same problem (the reference runs MadNLP with
kkt_system=GondzioKKTSystem).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .fft_operator import FFTOperator, FFTParameters, M_perpt_z, _ptr


class GondzioNLPModel:
    """nvar = 2N + N (p, q >= 0; r free), ncon = N (gondzio_model.jl:12-61)."""

    def __init__(self, parameters: FFTParameters, krylov_solver="cg", rdft=False,
                 preconditioner=True, dtype=torch.float64, device="cuda", mixed=False):
        if krylov_solver != "cg":
            raise NotImplementedError("only krylov_solver='cg' is implemented")
        self.parameters = parameters
        self.nbeta = int(np.prod(parameters.DFTsize))
        self.nvar = 3 * self.nbeta
        self.ncon = self.nbeta
        self.preconditioner = preconditioner
        self.op_fft = FFTOperator(self.nbeta, parameters.DFTdim, parameters.DFTsize,
                                  parameters.index_missing, rdft, dtype=dtype, device=device,
                                  mixed=mixed)
        self.M_perpt_z0 = M_perpt_z(self.op_fft, parameters.z0)
        self.z0 = self.op_fft._vec(parameters.z0)
        self.counters = {"neval_obj": 0, "neval_grad": 0, "neval_cons": 0}

    @property
    def lam(self):
        return float(self.parameters.lambda_)

    def _v(self, x, n):
        return self.op_fft._vec(x, n=n)

    def obj(self, x):
        self.counters["neval_obj"] += 1
        x = self._v(x, self.nvar)
        f = ctypes.c_double()
        _lib.call("csipm_gondzio_obj", self.op_fft.handle, _ptr(x), self.lam, ctypes.byref(f))
        return f.value

    def grad(self, x, g=None):
        self.counters["neval_grad"] += 1
        x = self._v(x, self.nvar)
        g = self.op_fft.empty(self.nvar) if g is None else g
        _lib.call("csipm_gondzio_grad", self.op_fft.handle, _ptr(x), self.lam, _ptr(g))
        return g

    def cons(self, x, c=None):
        self.counters["neval_cons"] += 1
        x = self._v(x, self.nvar)
        c = self.op_fft.empty(self.ncon) if c is None else c
        _lib.call("csipm_gondzio_cons", self.op_fft.handle, _ptr(x), _ptr(self.z0), _ptr(c))
        return c


class CondensedGondzioKKTSystem:
    """K = [[M^T M + P^-1 Wp, -M^T M], [-M^T M, M^T M + Q^-1 Wq]] on 2N
    (gondzio_kkt.jl:1-53); InvP_Wp / InvQ_Wq hold the primal diagonals."""

    def __init__(self, nlp: GondzioNLPModel):
        self.nlp = nlp
        self.InvP_Wp = nlp.op_fft.empty().zero_()
        self.InvQ_Wq = nlp.op_fft.empty().zero_()

    def mul(self, x, alpha=1.0, beta=0.0, y=None):
        op = self.nlp.op_fft
        x = op._vec(x, n=2 * op.N)
        if y is None:
            y = op.empty(2 * op.N)
            beta = 0.0
        _lib.call("csipm_gondzio_kkt_apply", op.handle, _ptr(x), _ptr(y), _ptr(self.InvP_Wp),
                  _ptr(self.InvQ_Wq), float(alpha), float(beta))
        return y


class GondzioPreconditioner:
    """[[I + Wp, -I], [-I, I + Wq]]^-1 as (P11, P12, P22) (gondzio_kkt.jl:55-92)."""

    def __init__(self, nlp: GondzioNLPModel):
        op = nlp.op_fft
        self.op = op
        self.P11, self.P12, self.P22 = op.empty(), op.empty(), op.empty()

    def mul(self, x, alpha=1.0, beta=0.0, y=None):
        op = self.op
        x = op._vec(x, n=2 * op.N)
        if y is None:
            y = op.empty(2 * op.N)
            beta = 0.0
        _lib.call("csipm_precond_apply", op.handle, _ptr(x), _ptr(y), _ptr(self.P11),
                  _ptr(self.P12), _ptr(self.P22), float(alpha), float(beta))
        return y


class GondzioKKTSystem:
    """The MadNLP KKT-system hooks of gondzio_kkt.jl:94-389."""

    def __init__(self, nlp: GondzioNLPModel):
        self.nlp = nlp
        op = nlp.op_fft
        n = op.N
        self.K = CondensedGondzioKKTSystem(nlp)
        self.P = GondzioPreconditioner(nlp)
        self.pr_diag = op.empty(3 * n)
        self.krylov_iterations = []
        self.initialize()

    def initialize(self):
        """reg = pr_diag = 1, du_diag = 0 (gondzio_kkt.jl:175-184)."""
        self.pr_diag.fill_(1.0)

    def build_kkt(self):
        """Preconditioner from pr_diag's p and q blocks; the operator keeps
        them as its diagonals (gondzio_kkt.jl:272-296)."""
        op = self.nlp.op_fft
        n = op.N
        reg_p, reg_q = self.pr_diag[:n], self.pr_diag[n:2 * n]
        _lib.call("csipm_gondzio_precond_build", op.handle, _ptr(reg_p), _ptr(reg_q),
                  _ptr(self.P.P11), _ptr(self.P.P12), _ptr(self.P.P22))
        self.K.InvP_Wp.copy_(reg_p)
        self.K.InvQ_Wq.copy_(reg_q)

    def mul(self, x, alpha=1.0, beta=0.0, y=None):
        """The part of MadNLP.mul! the override adds on [p; q; r; y]
        (gondzio_kkt.jl:203-241)."""
        op = self.nlp.op_fft
        x = op._vec(x, n=4 * op.N)
        if y is None:
            y = op.empty(4 * op.N)
            beta = 0.0
        _lib.call("csipm_gondzio_kkt_mul_full", op.handle, _ptr(x), _ptr(y), float(alpha),
                  float(beta))
        return y

    def jtprod(self, v, out=None):
        op = self.nlp.op_fft
        v = op._vec(v)
        out = op.empty(3 * op.N) if out is None else out
        _lib.call("csipm_gondzio_jtprod", op.handle, _ptr(v), _ptr(out))
        return out

    def solve_kkt(self, w, atol=1e-12, rtol=0.0, itmax=0):
        """Condense [w_p; w_q; w_r; w_y], CG on K (atol 1e-12, rtol 0 as
        gondzio_kkt.jl:339), unpack in place (:310-362)."""
        op = self.nlp.op_fft
        n = op.N
        w = op._vec(w, n=4 * n) if not (isinstance(w, torch.Tensor) and w.dtype == op.dtype
                                         and w.is_contiguous()) else w
        rhs = op.empty(2 * n)
        _lib.call("csipm_gondzio_condense_rhs", op.handle, _ptr(w), _ptr(rhs))
        dx = op.empty(2 * n)
        its = ctypes.c_int64()
        res = ctypes.c_double()
        st = ctypes.c_int()
        _lib.call("csipm_gondzio_cg_solve", op.handle, _ptr(rhs), _ptr(dx), _ptr(self.K.InvP_Wp),
                  _ptr(self.K.InvQ_Wq), _ptr(self.P.P11), _ptr(self.P.P12), _ptr(self.P.P22),
                  float(atol), float(rtol), int(itmax), ctypes.byref(its), ctypes.byref(res),
                  ctypes.byref(st))
        self.krylov_iterations.append(int(its.value))
        _lib.call("csipm_gondzio_unpack_solution", op.handle, _ptr(w), _ptr(dx))
        return w


@dataclass
class GondzioStats:
    solution: torch.Tensor      # [p; q; r]
    multipliers: torch.Tensor   # y (N)
    objective: float
    iterations: int
    cg_iterations: int
    status: str
    dual_residual: float
    mu: float
    seconds: float

    @property
    def beta(self):
        n = self.multipliers.numel()
        return self.solution[:n] - self.solution[n:2 * n]


def gondzio_solve(nlp: GondzioNLPModel, tol=1e-8, max_iter=200, cg_rtol=1e-10, cg_itmax=0,
                  cg_atol=0.0):
    """The library's interior-point driver on the Gondzio problem
    (csipm_gondzio_solve); cg_atol / cg_rtol as csipm.model.ipm_solve."""
    op = nlp.op_fft
    if op.dtype != torch.float64 and not op.mixed:
        raise NotImplementedError("the driver runs on an FP64 or a mixed-precision operator")
    z0 = op._vec(nlp.parameters.z0, dtype=torch.float64)
    x = torch.empty(3 * op.N, dtype=torch.float64, device=op.device)
    y = torch.empty(op.N, dtype=torch.float64, device=op.device)
    info = (ctypes.c_double * 8)()
    _lib.call("csipm_gondzio_solve", op.handle, _ptr(z0), nlp.lam, float(tol), int(max_iter),
              float(cg_atol), float(cg_rtol), int(cg_itmax), _ptr(x), _ptr(y), info)
    status = {0: "solved", 1: "max_iter", 2: "failure"}[int(info[3])]
    return GondzioStats(x, y, info[0], int(info[1]), int(info[2]), status, info[4], info[5],
                        info[6])
