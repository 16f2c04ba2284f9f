"""FFTNLPModel, the condensed KKT operators and FFTKKTSystem over libcsipm.

Mirrors src/fft_model.jl and src/fft_kkt.jl: the same objects and methods
(Julia's mutating `f!` is `f` here), acting on device tensors through the C
ABI.  MadNLP itself is not part of this package; `ipm_solve` runs the
library's native interior-point driver (csipm_solve) on the same KKT system.

    FFTNLPModel(parameters; krylov_solver, rdft, preconditioner)  fft_model.jl:12-64
    obj / grad / cons                                              fft_model.jl:66-101
    CondensedFFTKKTSystem.mul                                      fft_kkt.jl:27-44
    FFTPreconditioner.mul                                          fft_kkt.jl:73-85
    FFTKKTSystem: initialize, build_kkt, solve_kkt, mul, jtprod    fft_kkt.jl:91-362
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .fft_operator import FFTOperator, FFTParameters, M_perpt_z, _ptr


class FFTNLPModel:
    """x = [beta; c], nvar = ncon = 2N, constraints -beta - c <= 0,
    beta - c <= 0 (fft_model.jl:20-34)."""

    def __init__(self, parameters: FFTParameters, krylov_solver="cg", rdft=False,
                 preconditioner=True, dtype=torch.float64, device="cuda", mixed=False):
        if krylov_solver != "cg":
            raise NotImplementedError("only krylov_solver='cg' is implemented")
        self.parameters = parameters
        self.nbeta = int(np.prod(parameters.DFTsize))
        self.nvar = 2 * self.nbeta
        self.ncon = 2 * self.nbeta
        self.krylov_solver = krylov_solver
        self.preconditioner = preconditioner
        self.op_fft = FFTOperator(self.nbeta, parameters.DFTdim, parameters.DFTsize,
                                  parameters.index_missing, rdft, dtype=dtype, device=device,
                                  mixed=mixed)
        dtype = self.op_fft.dtype
        self.M_perpt_z0 = M_perpt_z(self.op_fft, parameters.z0)
        self.counters = {"neval_obj": 0, "neval_grad": 0, "neval_cons": 0}
        dev = self.op_fft.device
        self.x0 = torch.zeros(self.nvar, dtype=dtype, device=dev)
        self.lvar = torch.full((self.nvar,), -np.inf, dtype=dtype, device=dev)
        self.uvar = torch.full((self.nvar,), np.inf, dtype=dtype, device=dev)
        self.lcon = torch.full((self.ncon,), -np.inf, dtype=dtype, device=dev)
        self.ucon = torch.zeros(self.ncon, dtype=dtype, device=dev)

    @property
    def lam(self):
        return float(self.parameters.lambda_)

    def obj(self, x):
        """fft_model.jl:66-75."""
        self.counters["neval_obj"] += 1
        x = self.op_fft._vec(x, self.nvar)
        f = ctypes.c_double()
        _lib.call("csipm_nlp_obj", self.op_fft.handle, _ptr(x), _ptr(self.M_perpt_z0),
                  self.lam, ctypes.byref(f))
        return f.value

    def grad(self, x, g=None):
        """fft_model.jl:77-89."""
        self.counters["neval_grad"] += 1
        x = self.op_fft._vec(x, self.nvar)
        g = self.op_fft.empty(self.nvar) if g is None else g
        _lib.call("csipm_nlp_grad", self.op_fft.handle, _ptr(x), _ptr(self.M_perpt_z0),
                  self.lam, _ptr(g))
        return g

    def cons(self, x, c=None):
        """fft_model.jl:91-101."""
        self.counters["neval_cons"] += 1
        x = self.op_fft._vec(x, self.nvar)
        c = self.op_fft.empty(self.ncon) if c is None else c
        _lib.call("csipm_nlp_cons", self.op_fft.handle, _ptr(x), _ptr(c))
        return c


class CondensedFFTKKTSystem:
    """K = [[M^T M + L1, L2], [L2, L1]], L1 = S1 + S2, L2 = S1 - S2 (fft_kkt.jl:9-44)."""

    def __init__(self, nlp: FFTNLPModel):
        self.nlp = nlp
        n = nlp.nbeta
        self.L1 = nlp.op_fft.empty(n)
        self.L2 = nlp.op_fft.empty(n)
        self.shape = (2 * n, 2 * n)

    def mul(self, y, x, alpha=1.0, beta=0.0):
        op = self.nlp.op_fft
        x = op._vec(x, 2 * op.N)
        _lib.call("csipm_kkt_apply", op.handle, _ptr(x), _ptr(y), _ptr(self.L1), _ptr(self.L2),
                  float(alpha), float(beta))
        return y


class FFTPreconditioner:
    """[[I + L1, L2], [L2, L1]]^-1 = [[P11, P12], [P12, P22]] (fft_kkt.jl:54-85)."""

    def __init__(self, nlp: FFTNLPModel):
        n = nlp.nbeta
        self.op = nlp.op_fft
        self.P11 = self.op.empty(n)
        self.P12 = self.op.empty(n)
        self.P22 = self.op.empty(n)
        self.shape = (2 * n, 2 * n)

    def mul(self, y, x, alpha=1.0, beta=0.0):
        x = self.op._vec(x, 2 * self.op.N)
        _lib.call("csipm_precond_apply", self.op.handle, _ptr(x), _ptr(y), _ptr(self.P11),
                  _ptr(self.P12), _ptr(self.P22), float(alpha), float(beta))
        return y


@dataclass
class KrylovStats:
    iterations: int
    residual: float
    status: str


class FFTKKTSystem:
    """Reduced KKT system of FFTNLPModel (fft_kkt.jl:91-362).  The primal
    diagonal pr_diag has n = 4N entries ([beta; c; s1; s2]); its slack part
    holds Sigma_1 = pr_diag[2N:3N] and Sigma_2 = pr_diag[3N:4N] (:289-290)."""

    def __init__(self, nlp: FFTNLPModel, cg_atol=1e-12, cg_rtol=0.0, cg_itmax=0):
        self.nlp = nlp
        n = nlp.nbeta
        op = nlp.op_fft
        self.K = CondensedFFTKKTSystem(nlp)
        self.P = FFTPreconditioner(nlp)
        self.pr_diag = op.empty(4 * n)
        self.du_diag = op.empty(2 * n)
        self.reg = op.empty(4 * n)
        self.z1 = op.empty(n)
        self.z2 = op.empty(2 * n)
        self.solution = op.empty(2 * n)
        self.cg_atol, self.cg_rtol, self.cg_itmax = cg_atol, cg_rtol, cg_itmax
        self.krylov_iterations = []
        self.krylov_timer = []

    def initialize(self):
        """fft_kkt.jl:182-191."""
        self.reg.fill_(1.0)
        self.pr_diag.fill_(1.0)
        self.du_diag.fill_(0.0)

    def sigmas(self):
        n = self.nlp.nbeta
        return self.pr_diag[2 * n:3 * n], self.pr_diag[3 * n:4 * n]

    def build_kkt(self):
        """fft_kkt.jl:285-307."""
        s1, s2 = self.sigmas()
        op = self.nlp.op_fft
        _lib.call("csipm_precond_build", op.handle, _ptr(s1), _ptr(s2), _ptr(self.K.L1),
                  _ptr(self.K.L2), _ptr(self.P.P11), _ptr(self.P.P12), _ptr(self.P.P22))

    def solve_kkt(self, w):
        """fft_kkt.jl:309-354 on the 6N KKT vector w (in place): condense,
        preconditioned CG (Krylov.jl cg semantics), unpack."""
        import time
        op = self.nlp.op_fft
        n = self.nlp.nbeta
        if w.numel() != 6 * n:
            raise ValueError("w must have 6N entries")
        s1, s2 = self.sigmas()
        _lib.call("csipm_condense_rhs", op.handle, _ptr(w), _ptr(s1), _ptr(s2), _ptr(self.z2))
        if self.nlp.preconditioner:
            P11, P12, P22 = self.P.P11, self.P.P12, self.P.P22
        else:
            P11 = torch.ones(n, dtype=op.dtype, device=op.device)
            P12 = torch.zeros(n, dtype=op.dtype, device=op.device)
            P22 = P11
        it = ctypes.c_int64()
        res = ctypes.c_double()
        st = ctypes.c_int()
        t0 = time.perf_counter()
        _lib.call("csipm_cg_solve", op.handle, _ptr(self.z2), _ptr(self.solution),
                  _ptr(self.K.L1), _ptr(self.K.L2), _ptr(P11), _ptr(P12), _ptr(P22),
                  float(self.cg_atol), float(self.cg_rtol), int(self.cg_itmax),
                  ctypes.byref(it), ctypes.byref(res), ctypes.byref(st))
        self.krylov_timer.append(time.perf_counter() - t0)
        self.krylov_iterations.append(int(it.value))
        _lib.call("csipm_unpack_solution", op.handle, _ptr(w), _ptr(self.solution), _ptr(s1),
                  _ptr(s2))
        self.last_stats = KrylovStats(int(it.value), float(res.value),
                                      ["solved", "itmax", "zero-curvature"][st.value])
        return True   # like the reference, CG status does not fail the step

    def mul(self, y, x, alpha=1.0, beta=0.0):
        """fft_kkt.jl:213-251 (without MadNLP._kktmul!'s diagonal terms)."""
        op = self.nlp.op_fft
        _lib.call("csipm_kkt_mul_full", op.handle, _ptr(x), _ptr(y), float(alpha), float(beta))
        return y

    def jtprod(self, y, x):
        """fft_kkt.jl:253-275: y (4N) = J^T x, x = [y1; y2] (2N)."""
        op = self.nlp.op_fft
        _lib.call("csipm_jtprod", op.handle, _ptr(x), _ptr(y))
        return y


@dataclass
class ExecutionStats:
    solution: torch.Tensor
    objective: float
    iterations: int
    cg_iterations: int
    status: str
    dual_residual: float
    mu: float
    seconds: float


# PCG stopping rules of the Newton systems: "relative" stops at
# sqrt(r'Pr) <= 1e-10 sqrt(r0'P r0); "reference" is the rule the reference
# passes to Krylov.krylov_solve!, atol = 1e-12, rtol = 0 (src/fft_kkt.jl:339)
CG_RULES = {"relative": (0.0, 1e-10), "reference": (1e-12, 0.0)}


def ipm_solve(nlp: FFTNLPModel, tol=1e-8, max_iter=200, cg_rtol=None, cg_itmax=0,
              cg_atol=None, cg_rule="relative"):
    """Solve the model with the library's interior-point driver
    (csipm_solve; the reference runs MadNLP, src/CompressedSensingIPM.jl:31-36).
    Returns the primal beta (device, FP64) and the driver's statistics.

    The PCG of every Newton system stops at sqrt(r'Pr) <= cg_atol +
    cg_rtol sqrt(r0'P r0); cg_rule picks the pair ("relative": 0, 1e-10;
    "reference": 1e-12, 0 as src/fft_kkt.jl:339), explicit cg_atol / cg_rtol
    override it.  On a mixed-precision model the FP32 CG runs to a relative
    max(cg_rtol, 1e-5) and every Newton direction gets one FP64 refinement
    step."""
    a, r = CG_RULES[cg_rule]
    cg_atol = a if cg_atol is None else cg_atol
    cg_rtol = r if cg_rtol is None else cg_rtol
    op = nlp.op_fft
    if op.dtype != torch.float64 and not op.mixed:
        raise NotImplementedError("the driver runs on an FP64 or a mixed-precision operator")
    z0 = op._vec(nlp.parameters.z0, dtype=torch.float64)
    beta = torch.empty(op.N, dtype=torch.float64, device=op.device)
    info = (ctypes.c_double * 8)()
    _lib.call("csipm_solve", op.handle, _ptr(z0), nlp.lam, float(tol), int(max_iter),
              float(cg_atol), float(cg_rtol), int(cg_itmax), _ptr(beta), info)
    status = {0: "solved", 1: "max_iter", 2: "failure"}[int(info[3])]
    return ExecutionStats(beta, info[0], int(info[1]), int(info[2]), status, info[4], info[5],
                          info[6])
