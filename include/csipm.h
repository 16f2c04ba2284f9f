/*
 * csipm.h -- C ABI of libcsipm.so, the MI355X (gfx950) implementation of the
 * matrix-free interior-point hot path of CompressedSensingIPM.jl.
 *
 * Every vector argument is a DEVICE pointer owned by the caller (a Julia
 * ROCArray under GC.@preserve, a torch tensor, hipMalloc memory).  Calls are
 * asynchronous on the handle's stream unless stated otherwise.  A handle owns
 * only its internal state (FFT twiddles, the sampling mask, scratch spectra,
 * CG scalars) and is not re-entrant: like the reference's FFTOperator buffers
 * (src/fft_utils.jl:16-54), its scratch is overwritten by every apply.
 *
 * Layout: Julia column-major, dimension 1 contiguous.  beta, z and the real
 * signal have N = N1*N2*N3 entries.  A spectrum ("v") is complex interleaved
 * (re, im) with dimension 1 of length M1+1 (rdft) or N1 (full), M1 = N1/2.
 * "T" below is double for CSIPM_F64 handles and float for CSIPM_F32 handles;
 * the beta/DFT maps always take double.
 *
 * Status: 0 on success; otherwise a CSIPM_ERR_* code with a message available
 * from csipm_last_error() (thread-local).
 *
 * Entry points that synchronise or allocate (handle creation/destruction,
 * the solves, the batch calls, csipm_mask_index_missing, the synchronous
 * objective evaluations) return CSIPM_ERR_UNSUPPORTED when their stream is
 * being captured into a HIP graph, instead of invalidating the capture.
 */
#ifndef CSIPM_H
#define CSIPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CSIPM_OK = 0,
  CSIPM_ERR_ARG = 1,          /* bad argument (null pointer, bad enum)      */
  CSIPM_ERR_SHAPE = 2,        /* sizes / lengths inconsistent               */
  CSIPM_ERR_ALLOC = 3,        /* device allocation failed                   */
  CSIPM_ERR_HIP = 4,          /* HIP runtime error                          */
  CSIPM_ERR_UNSUPPORTED = 5,  /* size or option not implemented            */
  CSIPM_ERR_NOT_CONVERGED = 6 /* iterative method stopped without success */
};

/* Precision of a handle.  CSIPM_MIXED: operator and CG in FP32 (complex64)
 * inside an FP64 interior-point driver (FP64 residuals, complementarity,
 * step lengths, and one FP64 refinement step of every Newton direction);
 * its operator entry points run in FP32. */
enum { CSIPM_F64 = 0, CSIPM_F32 = 1, CSIPM_MIXED = 2 };

typedef struct csipm_op* csipm_op_t;

/* Version string of the library. */
const char* csipm_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* csipm_last_error(void);

/* Create an operator handle.  Replaces FFTOperator{VT}(nβ, DFTdim, DFTsize,
 * index_missing, rdft) and its plan_rfft (src/fft_utils.jl:16-54).
 *   dim        1, 2 or 3 (FFTParameters.DFTdim, src/fft_utils.jl:1-14)
 *   sizes      N1, N2, N3 (unused trailing entries ignored)
 *   rdft       1: the real-to-complex path; 0: the complex path (the
 *              reference's default).  Both compute the same operator (the
 *              complex FFT of a Hermitian spectrum / of real samples equals
 *              the C2R / R2C one) with the same kernels; rdft only selects
 *              the spectrum form the maps use by default (half or full)
 *   precision  CSIPM_F64, CSIPM_F32 or CSIPM_MIXED (see above)
 *   missing    HOST array of n_missing 0-based linear indices (index_missing
 *              of FFTParameters, converted from Julia's 1-based Int), may be
 *              NULL when n_missing == 0
 *   stream     hipStream_t to run on (NULL = default stream)            */
int csipm_op_create(csipm_op_t* h, int dim, const int64_t sizes[3], int rdft,
                    int precision, const int64_t* missing, int64_t n_missing,
                    void* stream);
int csipm_op_destroy(csipm_op_t h);
/* index_missing of a sampling mask / weights array: the 0-based linear
 * (column-major) indices of its zero entries, in increasing order, as the
 * loops of benchmarks/mastodonte.jl:38-50 build them (1-based there).
 *   weights    device, n doubles (the mask; nonzero = observed)
 *   index_out  device, room for n int64 (receives *count entries)
 *   count      host: number of missing samples
 * Synchronous on stream.  The result feeds csipm_op_create's `missing`
 * after a copy to the host.                                               */
int csipm_mask_index_missing(const double* weights, int64_t n, int64_t* index_out,
                             int64_t* count, void* stream);
/* csipm_mask_index_missing without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (serving loops that must not hipMalloc, harnesses that forbid it).  Same
 * arguments, result and synchronisation as csipm_mask_index_missing; its
 * count workspace is a fixed array of CSIPM_MASK_NOALLOC_WORDS int64 in each
 * device's global memory (512 KiB, held until the process ends), so n is
 * limited to CSIPM_MASK_NOALLOC_MAX samples (CSIPM_ERR_ARG beyond).  Every
 * call on a device shares that device's array: calls that overlap on one
 * device (from several threads or streams) race.                          */
#define CSIPM_MASK_NOALLOC_WORDS 65536
#define CSIPM_MASK_NOALLOC_MAX ((int64_t)(CSIPM_MASK_NOALLOC_WORDS - 1) * 4096)
int csipm_mask_index_missing_noalloc(const double* weights, int64_t n, int64_t* index_out,
                                     int64_t* count, void* stream);

/* Device bytes handle h holds itself: plans, sampling and slot masks, the
 * scratch spectrum, partial sums, Bluestein tables, lazily made entry-point
 * scratch and its batch-of-one engine (csipm_solve / csipm_cg_solve); the
 * FFT twiddle tables are shared between handles and cached, not counted.
 * FFTOperator's buffers (src/fft_utils.jl:16-54) are what a Julia caller
 * would count; the serving loop budgets device memory with it.            */
int csipm_op_bytes(csipm_op_t h, int64_t* bytes);
/* Footprint of a handle before it is made: the batch-pool slots an instance
 * of it occupies (csipm_batch_slots) and csipm_op_bytes of the new handle
 * with n_missing missing samples, from the host-side planning alone (no
 * device work).  Lets an admission loop reject a request that can never
 * fit its budget before building anything.                                */
int csipm_op_footprint(int dim, const int64_t sizes[3], int precision, int64_t n_missing,
                       int64_t* slots, int64_t* bytes);

/* Change the stream later calls on h are issued to. */
int csipm_op_set_stream(csipm_op_t h, void* stream);
/* Query N (length of beta) and Nc (complex entries of the half spectrum). */
int csipm_op_info(csipm_op_t h, int64_t* n, int64_t* nc, int* precision);
/* The line transform of axis `axis` (0-based) of h: FFT length (the line
 * length; Bluestein's convolution length; L - 1 for Rader's form; the
 * four-step length of long 1D lines), LDS rows per line, and the form
 * (0 direct, 1 Bluestein, 2 Rader, 3 four-step).  No counterpart in the
 * reference (rocFFT/FFTW plans are opaque there); the autotuner keys its
 * table on (FFT length, rows).                                            */
int csipm_op_fft_shape(csipm_op_t h, int axis, int* fft_len, int* rows, int* mode);

/* Autotuned launch shapes of the batch engine's FFT passes (no counterpart
 * in the reference): for pass kind 0 (line pass) or 1 (dimension-1 pass),
 * complex element size 8 (FP32) or 16 (FP64), FFT length and LDS rows, cap
 * the tile at 2^lg_cap lines (line pairs) per workgroup and run it on nt
 * threads (128 or 256; 0 keeps the built-in choice); lg_cap < 0 removes the
 * entry.  csipm.tune measures the candidates and writes the committed
 * table the package loads at start-up.  Takes effect at the next table
 * rebuild of an engine (a new batch, or the next change of its instances). */
int csipm_tune_set(int kind, int elem_bytes, int fft_len, int rows, int lg_cap, int nt);
/* Removes every launch-shape and Bluestein-length entry.                  */
int csipm_tune_clear(void);
/* Autotuned Bluestein convolution length (no counterpart in the reference):
 * 2D/3D axes and 1D lines of length line_len that take Bluestein's form
 * convolve at length conv_len (2 line_len - 1 <= conv_len <= 4 line_len,
 * factorable into the supported radices) instead of the smallest such
 * length; conv_len 0 removes the entry.  Fewer, cheaper radix stages can
 * outweigh a longer convolution; csipm.tune --bluestein measures the
 * candidates.  Takes effect for handles created afterwards.              */
int csipm_tune_bluestein(int line_len, int conv_len);

/* beta -> DFT map.  beta_to_DFT! (src/fft_utils.jl:195-204, per-element form
 * src/mapping_cpu_kernels.jl:353-738).  v: (M1+1)*N2*N3 complex when full==0,
 * N1*N2*N3 complex when full==1.  Bit-exact for even sizes.               */
int csipm_beta_to_dft(csipm_op_t h, const double* beta, double* v, int full);
/* DFT -> beta map.  DFT_to_beta! (src/fft_utils.jl:144-153, per-element
 * form src/mapping_cpu_kernels.jl:1-351); v laid out as above.          */
int csipm_dft_to_beta(csipm_op_t h, const double* v, double* beta, int full);

/* y = M beta: unitary inverse DFT of beta_to_DFT(beta), zeroed at missing
 * samples.  M_perp_beta (src/fft_utils.jl:95-117).  T* in/out, N each.   */
int csipm_m_perp_beta(csipm_op_t h, const void* beta, void* y);
/* Recovered signal x = sqrt(N) * real(ifft(beta_to_DFT(beta))): M_perp_beta
 * of this handle with no sample missing, the reconstruction step the
 * reference's real-data driver performs after a solve
 * (benchmarks/mastodonte.jl:109-116).  T* in/out, N each.               */
int csipm_reconstruct(csipm_op_t h, const void* beta, void* x);
/* beta = M^T z = DFT_to_beta(R2C(z)/sqrt(N)).  M_perpt_z (src/fft_utils.jl:73-93). */
int csipm_m_perpt_z(csipm_op_t h, const void* z, void* beta);
/* y = M^T M x.  M_perpt_M_perp_vec (src/fft_utils.jl:119-123).          */
int csipm_mtm(csipm_op_t h, const void* x, void* y);

/* y = alpha*K*x + beta*y with K = [[M^T M + L1, L2], [L2, L1]] on 2N-vectors.
 * CondensedFFTKKTSystem mul! (src/fft_kkt.jl:27-44).                    */
int csipm_kkt_apply(csipm_op_t h, const void* x, void* y, const void* L1,
                    const void* L2, double alpha, double beta);

/* Preconditioner data from the slack diagonals sigma1, sigma2 (N each):
 * L1 = s1+s2, L2 = s1-s2 and P = [[1+L1, L2], [L2, L1]]^-1, i.e. the
 * reference's Minv = 1/(1+L1), P22 = S = 1/(L1 - L2^2 Minv), P12 = -L2 S Minv,
 * P11 = Minv + Minv L2 S L2 Minv, evaluated in the equivalent closed form
 * P11 = L1/det, P12 = -L2/det, P22 = (1+L1)/det, det = L1 + 4 s1 s2, which
 * avoids the cancellation in L1 - L2^2 Minv.
 * build_kkt! (src/fft_kkt.jl:285-307).  All T*.                          */
int csipm_precond_build(csipm_op_t h, const void* sigma1, const void* sigma2,
                        void* L1, void* L2, void* P11, void* P12, void* P22);
/* y = alpha*P*x + beta*y, P = [[P11, P12], [P12, P22]] on 2N-vectors.
 * FFTPreconditioner mul! (src/fft_kkt.jl:73-85).                          */
int csipm_precond_apply(csipm_op_t h, const void* x, void* y, const void* P11,
                        const void* P12, const void* P22, double alpha,
                        double beta);

/* Preconditioned CG on K x = b (2N), x0 = 0, with Krylov.jl's cg stopping
 * rule sqrt(r'Pr) <= atol + rtol*sqrt(r0'P r0), itmax (<= 0: 2*2N) and the
 * zero-curvature exit.  Replaces Krylov.krylov_solve!(workspace, K, b; M=P,
 * atol, rtol) at src/fft_kkt.jl:339.  Synchronous: returns after the solve.
 *   iters   number of iterations (host)      resid  final sqrt(r'Pr) (host)
 *   status  0 solved, 1 itmax, 2 zero curvature (host)                    */
int csipm_cg_solve(csipm_op_t h, const void* b, void* x, const void* L1,
                   const void* L2, const void* P11, const void* P12,
                   const void* P22, double atol, double rtol, int64_t itmax,
                   int64_t* iters, double* resid, int* status);

/* FFTNLPModel (src/fft_model.jl).  x = [beta; c] (2N, T), Mtz0 = M^T z0 (N, T).
 * obj   f = 1/2 ||M beta||^2 - beta'Mtz0 + lambda*sum(c)   (:66-75)  synchronous
 * grad  g = [M^T M beta - Mtz0; lambda]                      (:77-89)
 * cons  c = [-beta - c; beta - c]                             (:91-101)     */
int csipm_nlp_obj(csipm_op_t h, const void* x, const void* Mtz0, double lambda, double* f);
int csipm_nlp_grad(csipm_op_t h, const void* x, const void* Mtz0, double lambda, void* g);
int csipm_nlp_cons(csipm_op_t h, const void* x, void* c);

/* Full KKT product on 6N vectors [beta; c; s1; s2; y1; y2] without MadNLP's
 * own diagonal terms (MadNLP._kktmul!): the part FFTKKTSystem's mul! adds
 * (src/fft_kkt.jl:239-246).  y = alpha*A*x + beta*y.                       */
int csipm_kkt_mul_full(csipm_op_t h, const void* x, void* y, double alpha, double beta);
/* J^T y for y = [y1; y2] (2N) -> 4N (src/fft_kkt.jl:253-275).              */
int csipm_jtprod(csipm_op_t h, const void* y, void* out);
/* solve_kkt! pieces (src/fft_kkt.jl:309-354): condense the 6N right-hand side
 * w into b (2N), and unpack the condensed solution x into w in place.      */
int csipm_condense_rhs(csipm_op_t h, const void* w, const void* sigma1, const void* sigma2,
                       void* b);
int csipm_unpack_solution(csipm_op_t h, void* w, const void* x, const void* sigma1,
                          const void* sigma2);

/* Per-pass timing of the fused K apply (alpha=1, beta=0), the counterpart of
 * FFTOperator's fft_timer / mapping_timer (src/fft_utils.jl:24-25): HIP
 * events around every pass, averaged over nrep applies.
 *   layout 0: csipm_kkt_apply on reference-ordered vectors (2N, N);
 *          1: the apply CG runs internally, on spectrum-ordered vectors
 *             (2*U, U with U = 2*Nc from csipm_op_info), with its fused p'q;
 *          2: the reduced apply y = (M^T M + D) x the mixed-precision CG
 *             runs (x, y, D = L1 of U entries; L2 unused), with its p'q.
 *   pass_ms / pass_bytes: arrays of >= 8; pass_bytes receives the algorithmic
 *   HBM bytes of each pass (may be NULL).  Synchronous.                   */
int csipm_kkt_apply_timed(csipm_op_t h, const void* x, void* y, const void* L1,
                          const void* L2, int64_t nrep, int layout, double* pass_ms,
                          double* pass_bytes, int* npass);
/* Name of pass `pass` of the apply of that layout ("" if out of range). */
const char* csipm_kkt_pass_name(csipm_op_t h, int layout, int pass);

/* Full solve of FFTNLPModel (src/fft_model.jl:66-101) by the library's
 * primal-dual interior-point driver on the condensed FFT-KKT system (handle
 * precision CSIPM_F64 or CSIPM_MIXED).  The reference delegates this loop to
 * MadNLP (ipm_solve!, src/CompressedSensingIPM.jl:31-36).
 *   z0        device, N doubles (observed samples, zero at missing)
 *   lambda    l1 weight;  tol  IPM tolerance;  max_iter  IPM iteration cap
 *   cg_atol, cg_rtol  PCG stopping rule of every Newton system (FP64):
 *             sqrt(r'Pr) <= cg_atol + cg_rtol*sqrt(r0'P r0).  The reference
 *             passes cg_atol = 1e-12, cg_rtol = 0 to Krylov.krylov_solve!
 *             (src/fft_kkt.jl:339).  CSIPM_MIXED handles run their FP32 CG
 *             to max(cg_rtol, 1e-5) relative and ignore cg_atol (accuracy
 *             comes from the FP64 refinement step).  At least one > 0.
 *   cg_itmax  CG iteration cap per Newton system (<= 0: 4N)
 *   beta_out  device, N doubles: the recovered beta
 *   info      host array of 8 doubles: [objective, ipm iterations, total cg
 *             iterations, status (0 solved, 1 max_iter, 2 failure),
 *             final dual residual, final mu, seconds, reserved]
 * Synchronous.                                                           */
int csipm_solve(csipm_op_t h, const double* z0, double lambda, double tol,
                int64_t max_iter, double cg_atol, double cg_rtol, int64_t cg_itmax,
                double* beta_out, double* info);

/* ADMM for the same problem (CompressedSensingADMM, src/admm.jl:8-61):
 * x-update by CG on (M^T M + rho I) (Krylov.jl cg defaults: atol = rtol =
 * sqrt(eps), 2N iterations), z = soft threshold of x + y/rho at lambda/rho,
 * y += rho (x - z); stops when max(||dx||, ||dy||, ||dz||) <= tol or after
 * maxt iterations.  Returns z in beta_out (device, N doubles); info[8] =
 * [0, iterations, total CG iterations, status (0 converged, 1 maxt,
 * 2 failure), last err, 0, seconds, 0].  Handle precision CSIPM_F64 (the
 * reference's arithmetic) or CSIPM_MIXED (FP32 CG).  Synchronous.        */
int csipm_admm(csipm_op_t h, const double* z0, double lambda, double rho,
               int64_t maxt, double tol, double* beta_out, double* info);

/* ------------------------------------------------------------------------
 * Gondzio formulation (src/gondzio_model.jl, src/gondzio_kkt.jl): variables
 * x = [p; q; r] (3N), beta = p - q, p, q >= 0, r free, one equality block
 * c(x) = z0 - M (p - q) - r = 0, objective 1/2 r'r + lambda sum(p + q).
 * M = csipm_m_perp_beta, M^T = csipm_m_perpt_z (no mask, as the reference).
 * All vectors T* on the device (T = the handle's operator precision).
 * ------------------------------------------------------------------------ */
/* f = 1/2 r'r + lambda sum(p + q)  (gondzio_model.jl:67-79), synchronous.  */
int csipm_gondzio_obj(csipm_op_t h, const void* x, double lambda, double* f);
/* g = [lambda; lambda; r]  (gondzio_model.jl:81-94).                      */
int csipm_gondzio_grad(csipm_op_t h, const void* x, double lambda, void* g);
/* c = z0 - M (p - q) - r  (gondzio_model.jl:96-110).                      */
int csipm_gondzio_cons(csipm_op_t h, const void* x, const void* z0, void* c);
/* J^T v = [-M^T v; M^T v; -v]  (gondzio_kkt.jl:243-262).                  */
int csipm_gondzio_jtprod(csipm_op_t h, const void* v, void* out);
/* y = alpha K x + beta y on 2N vectors, K = [[M^T M + Wp, -M^T M],
 * [-M^T M, M^T M + Wq]]: CondensedGondzioKKTSystem mul!
 * (gondzio_kkt.jl:28-53).                                                  */
int csipm_gondzio_kkt_apply(csipm_op_t h, const void* x, void* y, const void* Wp, const void* Wq,
                            double alpha, double beta);
/* The part of MadNLP.mul! that GondzioKKTSystem adds on 4N vectors
 * [p; q; r; y] (gondzio_kkt.jl:203-241, without MadNLP._kktmul!'s
 * diagonal terms).                                                        */
int csipm_gondzio_kkt_mul_full(csipm_op_t h, const void* x, void* y, double alpha, double beta);
/* build_kkt! (gondzio_kkt.jl:272-296): P = [[1 + Wp, -1], [-1, 1 + Wq]]^-1
 * from the primal diagonals Wp, Wq (pr_diag), evaluated in the closed form
 * P11 = (1 + Wq)/det, P12 = 1/det, P22 = (1 + Wp)/det, det = Wp + Wq + Wp Wq
 * (the reference's inv/S sequence forms the same matrix).  The
 * preconditioner product (GondzioPreconditioner mul!, :80-92) is
 * csipm_precond_apply.                                                    */
int csipm_gondzio_precond_build(csipm_op_t h, const void* Wp, const void* Wq, void* P11,
                                void* P12, void* P22);
/* solve_kkt! pieces (gondzio_kkt.jl:310-362): rhs (2N) = [w_p - M^T(w_r +
 * w_y); w_q + M^T(w_r + w_y)] from w = [w_p; w_q; w_r; w_y] (4N); unpack:
 * w_x = dx, dr = -w_y - M (dp - dq), dy = dr - w_r, in place.            */
int csipm_gondzio_condense_rhs(csipm_op_t h, const void* w, void* rhs);
int csipm_gondzio_unpack_solution(csipm_op_t h, void* w, const void* dx);
/* PCG on the condensed Gondzio K (Krylov.krylov_solve! at
 * gondzio_kkt.jl:339, the reference passes atol = 1e-12, rtol = 0), with
 * csipm_cg_solve's conventions and outputs.  Runs as csipm_cg_solve on the
 * congruent condensed FFT system (x = A y, A = 1/2 [[I, I], [-I, I]]):
 * the same iterates and stopping point in exact arithmetic.  Synchronous. */
int csipm_gondzio_cg_solve(csipm_op_t h, const void* b, void* x, const void* Wp, const void* Wq,
                           const void* P11, const void* P12, const void* P22, double atol,
                           double rtol, int64_t itmax, int64_t* iters, double* resid,
                           int* status);
/* Full solve of GondzioNLPModel (ipm_solve! with kkt_system =
 * GondzioKKTSystem, src/CompressedSensingIPM.jl:31-36) by the library's
 * interior-point driver: the Gondzio problem is the FFT model under
 * p = (c + beta)/2, q = (c - beta)/2, with the same central path, so the
 * driver solves that and returns
 *   x_out  device, 3N doubles [p; q; r], r = z0 - M (p - q)
 *   y_out  device, N doubles: the multiplier of c(x) = 0 (= r at a
 *          solution, MadNLP's sign convention grad f + J^T y = z)
 *   info   as csipm_solve, info[0] = 1/2 r'r + lambda sum(p + q).
 * Handle precision CSIPM_F64 or CSIPM_MIXED.  Synchronous.               */
int csipm_gondzio_solve(csipm_op_t h, const double* z0, double lambda, double tol,
                        int64_t max_iter, double cg_atol, double cg_rtol, int64_t cg_itmax,
                        double* x_out, double* y_out, double* info);

/* ------------------------------------------------------------------------
 * Batch engine: many solves of uneven sizes and dimensionalities in the same
 * kernel launches (one launch per pass / vector step for all instances),
 * with per-instance convergence masks, and continuous batching: instances
 * are admitted and retired between interior-point iterations while the
 * others keep running.  No counterpart in the reference, which solves one
 * problem per MadNLP call (src/CompressedSensingIPM.jl:31-36); every
 * instance follows csipm_solve's algorithm and stopping test.
 *
 * precision: CSIPM_F64 (FP64 condensed-K CG, as csipm_solve on an F64
 * handle) or CSIPM_MIXED (reduced system, FP32 operator and CG, FP64
 * control and refinement).  Admitted handles must have that precision.
 * slots: pool size in spectrum-order slots (csipm_batch_slots per instance);
 * device memory is about csipm_batch_bytes(precision, slots).
 * All calls run on `stream` and return when their work is done.
 * ------------------------------------------------------------------------ */
typedef struct csipm_batch* csipm_batch_t;

int csipm_batch_create(csipm_batch_t* out, int precision, int64_t slots,
                       int max_instances, void* stream);
int csipm_batch_destroy(csipm_batch_t b);
/* pool slots instance h occupies, and pool bytes of a batch of `slots` */
int csipm_batch_slots(csipm_op_t h, int64_t* slots);
int csipm_batch_bytes(int precision, int64_t slots, int64_t* bytes);
/* Admit a problem: z0 (device, N doubles) and beta_out (device, N doubles)
 * stay owned by the caller and must live until the instance is harvested.
 * CSIPM_ERR_ALLOC when the pool has no free range or all ids are taken.
 * A handle serves one live instance at a time, in any batch (an instance's
 * passes run in its handle's scratch spectrum): admitting a handle that is
 * already live fails with CSIPM_ERR_ARG, and so does csipm_op_destroy of
 * an admitted handle.                                                     */
int csipm_batch_admit(csipm_batch_t b, csipm_op_t h, const double* z0,
                      double lambda, double* beta_out, int64_t cg_itmax, int* id);
/* One interior-point iteration of every admitted, unfinished instance
 * (newly admitted ones are initialised first).  *running: instances still
 * iterating.  Finished ones are listed by csipm_batch_poll.              */
int csipm_batch_step(csipm_batch_t b, double tol, int64_t max_iter,
                     double cg_atol, double cg_rtol, int* running);
int csipm_batch_poll(csipm_batch_t b, int* ids, int cap, int* n);
/* Write a finished instance's beta into its beta_out, fill info[8] as
 * csipm_solve does (seconds = 0) and free its slots and id.  beta_out is
 * written asynchronously, in order on the batch's stream (the host does not
 * wait); info comes from the flags the last step read.                   */
int csipm_batch_harvest(csipm_batch_t b, int id, double* info);
/* The same for n problems / instances in one call (a serving loop's turn):
 * one stream hand-over for all of them instead of one per call.
 * admit_many: rcs[k] = what csipm_batch_admit would return for problem k
 * (CSIPM_ERR_ALLOC: it does not fit now; later, smaller problems are still
 * tried; CSIPM_ERR_ARG also when a handle appears twice), ids[k] its id or
 * -1.  Returns CSIPM_OK unless a HIP call fails (problems before it stay
 * admitted).  harvest_many: every id must name a finished instance (checked
 * before any is harvested); info: n * 8 doubles.                         */
int csipm_batch_admit_many(csipm_batch_t b, int n, const csipm_op_t* ops,
                           const double* const* z0, const double* lambda,
                           double* const* beta_out, int64_t cg_itmax, int* ids, int* rcs);
int csipm_batch_harvest_many(csipm_batch_t b, int n, const int* ids, double* info);
/* Convenience: admit n problems, iterate until all finish, harvest all.
 * info: n * 8 doubles.                                                   */
int csipm_batch_solve(csipm_batch_t b, int n, const csipm_op_t* ops,
                      const double* const* z0, const double* lambda, double tol,
                      int64_t max_iter, double cg_atol, double cg_rtol, int64_t cg_itmax,
                      double* const* beta_out, double* info);
/* Timing of the batch's CG operator apply (one launch per pass and
 * dimensionality over all admitted instances, vectors zeroed): HIP events
 * around every launch, averaged over nrep applies; pass_bytes receives the
 * algorithmic HBM bytes of each launch.  cap: length of the output arrays.
 * csipm_batch_launch_name names launch k ("<dim>d:<pass>").              */
int csipm_batch_apply_timed(csipm_batch_t b, int64_t nrep, double* pass_ms,
                            double* pass_bytes, int cap, int* nlaunch);
const char* csipm_batch_launch_name(csipm_batch_t b, int k);
/* Host-side accounting of the engine since creation, out[8] = seconds in
 * csipm_batch_step, of which in launch-table builds, graph captures,
 * initialisation of admitted instances and CG solves; then the numbers of
 * table builds, graph captures and steps.                                */
int csipm_batch_stats(csipm_batch_t b, double* out);

#ifdef __cplusplus
}
#endif

#endif /* CSIPM_H */
