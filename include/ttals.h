/*
 * ttals.h -- C ABI of libttals.so, the MI355X (gfx950) TT-ALS engine.
 *
 * The reference (PGelss/scikit_tt) is pure Python/numpy; its "FFI" for the
 * hot path is the Python call surface scikit_tt.solvers.sle.als
 * (scikit_tt/solvers/sle.py:10), scikit_tt.solvers.evp.als
 * (scikit_tt/solvers/evp.py:17-27) and the TT methods they use.  Each entry
 * point below replaces one of those functions (cited per declaration); the
 * Python package scikit_tt_amd binds them with ctypes.
 *
 * Conventions
 *   - All numeric arrays are float64, C-order, in DEVICE memory (HIP), except
 *     the small integer shape arrays, which are HOST memory.
 *   - TT cores have shape (r_i, m_i, n_i, r_{i+1}) like the reference
 *     (scikit_tt/tensor_train.py:14-21); a TT "vector" has n_i = 1.
 *   - `stream` is a hipStream_t (NULL = default stream).  Calls are
 *     asynchronous on that stream unless documented as synchronous.
 *   - Return value: 0 on success, a negative TTALS_E* code on an argument or
 *     runtime error (message in ttals_last_error()), or a positive
 *     TTALS_S* status of the numerical result.
 *   - Thread safety: one ttals_ctx per host thread; a context owns its
 *     device workspace and must not be used concurrently.
 */
#ifndef TTALS_H
#define TTALS_H

#ifdef __cplusplus
extern "C" {
#endif

/* argument / runtime errors (negative) */
#define TTALS_E_ARG -1      /* bad argument (ValueError in Python) */
#define TTALS_E_SHAPE -2    /* inconsistent TT shapes (ValueError) */
#define TTALS_E_HIP -3      /* HIP runtime error */
#define TTALS_E_OOM -4      /* device allocation failed (MemoryError) */
#define TTALS_E_UNSUPPORTED -5 /* option not implemented (NotImplementedError) */

/* numerical status of a solve (positive) */
#define TTALS_S_SINGULAR 1    /* zero pivot in a local LU (numpy.linalg.LinAlgError) */
#define TTALS_S_NOCONV 2      /* some local Krylov solve hit its iteration cap */
#define TTALS_S_UNSUPPORTED 3 /* evp: sigma inside the spectrum of a large local problem */
#define TTALS_S_NONFINITE 4   /* NaN/Inf in a local solution */
#define TTALS_S_NONSYM 5      /* evp: operator not symmetric (NotImplementedError; the reference's
                                 'eig' would use a general eigensolver, scikit_tt/solvers/evp.py:424-432) */
#define TTALS_S_DEADLINE 7    /* the system's time budget ran out (ttals_set_time_budget) */

typedef struct ttals_ctx ttals_ctx;

/* A tensor train.  Shape arrays are host memory; cores are device pointers.
 * For outputs, `ranks` receives the final ranks and `cores` must have the
 * capacity of the corresponding input cores. */
typedef struct ttals_tt {
  int order;
  int* row_dims;   /* [order] */
  int* col_dims;   /* [order] */
  int* ranks;      /* [order + 1] */
  double** cores;  /* [order] device pointers */
} ttals_tt;

typedef struct ttals_result {
  int status;              /* 0 or a TTALS_S_* code */
  int sweeps;              /* ALS iterations performed */
  int unconverged;         /* local Krylov solves that missed the tolerance */
  int max_iters;           /* most Krylov iterations in one local solve */
  long long krylov_iters;  /* total Krylov iterations (matvecs) */
  long long local_solves;  /* local problems solved */
  double lambda;           /* evp: eigenvalue returned */
  double worst_residual;   /* sle: worst final relative residual of a local Krylov solve */
  double flops;            /* algorithmic flops of the solve (from shapes and iteration counts) */
  double cycles[7];        /* shader clock cycles: total, matvec, Gram-Schmidt, tridiagonal eigen,
                              QR + warm start, stack updates, on-chip set-up (diagnostics; only in
                              a library built with -DTTALS_PROFILE=1, else 0) */
  double device_ms;        /* device time of this system's solve inside its launch (the constant
                              clock, first to last step of its workgroup or gang's member 0) */
  double bytes;            /* algorithmic bytes of the local solves: per matvec the stacks,
                              operator core and iterate read once (+ the input and output vectors
                              where they live in global memory), one element per vector flop on
                              global-memory vectors; dense local solves: the operator pieces once
                              + the micro matrix written and read once */
} ttals_result;

/* ---- context ------------------------------------------------------------ */
int ttals_ctx_create(int device, ttals_ctx** out);
void ttals_ctx_destroy(ttals_ctx* ctx);
const char* ttals_last_error(ttals_ctx* ctx);
int ttals_version(void);

/* Step trace (diagnostics): when cap > 0, every later whole-solver call
 * records 6 doubles per local solve of its FIRST system into the device
 * buffer buf: position, forward(1)/backward(0), local size, Krylov
 * iterations, eigenvalue (evp), final relative residual (sle Krylov).
 * No counterpart in the reference (its solvers have no tracing). */
int ttals_set_trace(ttals_ctx* ctx, double* buf, long long cap);

/* ---- whole solvers (synchronous: return when the results are on the host) */

/* Replaces scikit_tt.solvers.sle.als (scikit_tt/solvers/sle.py:10-95) for
 * `count` independent systems in ONE launch (one workgroup per system).
 * solver: 0 = 'solve', 1 = 'lu' (both: LU for local size <= 96, matrix-free
 * CG/GMRES above).  tol/maxit: local Krylov relative tolerance / cap.
 * x_out[s] receives the solution (cores in place, final ranks). */
int ttals_sle_als_batch(ttals_ctx* ctx, void* stream, int count, const ttals_tt* A, const ttals_tt* x0,
                        const ttals_tt* b, const ttals_tt* x_out, int repeats, int solver, double tol, int maxit,
                        ttals_result* res);

/* Allocation-free form of ttals_sle_als_batch: same arguments, results and
 * kernels, but every device buffer the engine needs beyond the caller's
 * inputs and outputs (interface stacks, Krylov basis, QR and LU scratch,
 * descriptors, result rows) lives in a static device array of 64 MiB per GPU
 * (`scratch`, held by each GPU the library runs on until the process ends)
 * and the host staging in a static 4 MiB host buffer.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc,
 * hipFree, hipHostMalloc or event creation.  A batch needing more than the
 * static sizes is rejected with TTALS_E_ARG (message: the bytes needed).
 * Calls are serialised by a process-wide lock (they share the arrays); the
 * arrays keep the last call's contents, and the engine reads nothing of them
 * it has not written in the same call.  Synchronous. */
int ttals_sle_als_batch_static(ttals_ctx* ctx, void* stream, int count, const ttals_tt* A, const ttals_tt* x0,
                               const ttals_tt* b, const ttals_tt* x_out, int repeats, int solver, double tol,
                               int maxit, ttals_result* res);

/* Replaces scikit_tt.solvers.sle.als (scikit_tt/solvers/sle.py:10-95) for
 * large systems: each system is solved by a GANG of `gang` workgroups (<= 0:
 * automatic, one per row of the widest bond up to 32) that split every
 * phase -- the local matvec over the row-side rank index, the interface
 * stacks over their first index, the Krylov vector operations over the
 * same rows -- and meet at gang barriers; all gangs fit the device at once
 * and are launched together (one plain launch; workgroups are dispatched in
 * index order).  precision 0: FP64 throughout; 1: every Krylov
 * iteration in FP32 on an FP32 copy of the local operator, refined with the
 * FP64 local residual f - M v until it meets `tol` (cores, stacks, QR and
 * the residual FP64).  Local problems are solved by CG (operator symmetric
 * slice by slice; GMRES when p'Mp <= 0) or GMRES(40), never by a dense LU.
 * Synchronous; TTALS_E_ARG when the gangs do not fit the device at once. */
int ttals_sle_als_gang(ttals_ctx* ctx, void* stream, int count, const ttals_tt* A, const ttals_tt* x0,
                       const ttals_tt* b, const ttals_tt* x_out, int repeats, int solver, double tol, int maxit,
                       int precision, int gang, ttals_result* res);

/* Replaces scikit_tt.solvers.evp.als (scikit_tt/solvers/evp.py:17-179) with
 * number_ev=1, operator_gevp=None, previous=[] for `count` systems in ONE
 * launch.  solver: 0 = 'eig' (eigenvalue nearest sigma), 1 = 'eigh'
 * (largest).  The returned TT is the best-so-far iterate (evp.py:151-155). */
int ttals_evp_als_batch(ttals_ctx* ctx, void* stream, int count, const ttals_tt* A, const ttals_tt* x0,
                        const ttals_tt* x_out, int repeats, double conv_eps, int solver, double sigma, double tol,
                        int maxit, ttals_result* res);

/* ---- asynchronous form of the two solvers (serving) ----------------------
 * ttals_*_submit enqueue the whole solve on `stream` and return at once;
 * ttals_query returns 1 when the context's submission has finished (0 if
 * still running); ttals_wait blocks, writes x_out ranks and `count` results.
 * One submission may be pending per context; the x_out descriptors' rank
 * arrays must stay valid until ttals_wait.  No counterpart in the reference
 * (its solvers are synchronous Python). */
int ttals_sle_submit(ttals_ctx* ctx, void* stream, int count, const ttals_tt* A, const ttals_tt* x0,
                     const ttals_tt* b, const ttals_tt* x_out, int repeats, int solver, double tol, int maxit);
int ttals_evp_submit(ttals_ctx* ctx, void* stream, int count, const ttals_tt* A, const ttals_tt* x0,
                     const ttals_tt* x_out, int repeats, double conv_eps, int solver, double sigma, double tol,
                     int maxit);
int ttals_query(ttals_ctx* ctx);
int ttals_wait(ttals_ctx* ctx, ttals_result* res);

/* Size of the on-chip (LDS) arena in KB, 0..61 (default 61): local
 * problems whose operator and working vectors fit run LDS-resident on the
 * FP64 matrix cores, the others from HBM/L2; 0 forces the global path.
 * Applies to later submissions and pieces on this context.  No counterpart
 * in the reference: a tuning knob. */
int ttals_set_arena_kb(ttals_ctx* ctx, int kb);

/* Local tolerance policy of later evp submissions on this context:
 * 0 (default) every local eigenproblem to `tol`; k = 1..15 adaptive
 * (inexact ALS): sweeps 1-2 to max(tol, 1e-k), then to max(tol, min(1e-k,
 * 0.1 sqrt(|dlambda|/|lambda|))) of the last sweep-to-sweep change.  No
 * counterpart in the reference (its local eigensolves are dense). */
int ttals_set_local_tol_policy(ttals_ctx* ctx, int policy);

/* Form of the on-chip local matvec for later submissions and pieces on this
 * context: 0 (default) the two-stage form -- the operator core merged with
 * the right stack once per local problem, then two matrix-core GEMMs per
 * matvec -- wherever it fits the arena, else the three-stage form; 1 always
 * the three-stage form.  Results agree to rounding.  No counterpart in the
 * reference: a tuning and test knob. */
int ttals_set_matvec_variant(ttals_ctx* ctx, int variant);

/* Schedule of later sle batches (ttals_sle_als_batch / ttals_sle_submit) on
 * this context.  0 (default): one workgroup per system (k_als, two systems
 * per CU).  1: one mixed launch (k_hybrid) -- the systems the launch-time
 * model says would set the batch's time are solved by gangs of 2..32
 * workgroups (as ttals_sle_als_gang, precision `gang_precision`), the rest
 * by one workgroup each, largest first; when no gang shortens the modelled
 * time the batch runs as under 0.  Gang-bearing launches on one device are
 * ordered one after another (an event dependency, no host wait).  evp
 * batches always use 0.  No counterpart in the reference: a scheduling knob
 * of the batched engine. */
int ttals_set_schedule(ttals_ctx* ctx, int schedule, int gang_precision);

/* Time budget of every system of later submissions on this context, in ms
 * of device time from the system's start (0, the default: none).  A system
 * over budget stops before its next local step with status
 * TTALS_S_DEADLINE; its cores hold the last completed step.  The check runs
 * between local steps, so a step in progress finishes first.  No
 * counterpart in the reference: a serving control. */
int ttals_set_time_budget(ttals_ctx* ctx, double ms);

/* Launch-time model of the mixed schedule, `n` = 9 doubles: ms per on-chip
 * and per off-chip matvec flop of one sweep, ms per local solve, then the
 * time of a gang of 1, 2, 4, 8, 16, 32 workgroups relative to one workgroup
 * (see ttals_sle_cost).  Set fits it to the device (scikit_tt_amd.tune);
 * get reads the current one.  No counterpart in the reference. */
int ttals_set_cost_model(ttals_ctx* ctx, const double* params, int n);
int ttals_get_cost_model(ttals_ctx* ctx, double* params, int n);

/* Static cost of `count` sle systems, 3 doubles each: the flops of the
 * three-stage local matvec summed over the positions of one sweep (forward
 * and back) whose local problem fits the on-chip arena, the same for those
 * that do not, and the local solves per sweep.  Host arithmetic only (ctx
 * may be NULL: the default arena and matvec form).  No counterpart in the
 * reference. */
int ttals_sle_cost(ttals_ctx* ctx, int count, const ttals_tt* A, const ttals_tt* x0, double* out);

/* The split the mixed schedule would make of `count` sle systems on `slots`
 * resident workgroups (the device's k_hybrid capacity; 256 on MI355X): gang
 * size per system (0: one workgroup) and the modelled launch time.  Host
 * arithmetic only (ctx may be NULL: the default model).  No counterpart in
 * the reference. */
int ttals_sle_plan(ttals_ctx* ctx, int count, const ttals_tt* A, const ttals_tt* x0, int repeats, int slots,
                   int* gangs, double* makespan_ms);

/* Gang size the last sle batch on this context gave each of its `count`
 * systems (0: one workgroup).  No counterpart in the reference: a
 * diagnostic of the mixed schedule. */
int ttals_last_plan(ttals_ctx* ctx, int* gangs, int count);

/* Launch time (ms) the model predicted for the last mixed-schedule sle
 * batch on this context (0 when the last batch was not one).  No
 * counterpart in the reference: a diagnostic of the mixed schedule. */
int ttals_last_plan_ms(ttals_ctx* ctx, double* ms);

/* Device time (ms) of the engine kernel of the context's last finished
 * submission, from HIP events recorded on its stream immediately before and
 * after the launch (host-side preparation excluded).  TTALS_E_ARG while a
 * submission is pending or before the first one.  No counterpart in the
 * reference: a measurement hook for benchmarks. */
int ttals_last_launch_ms(ttals_ctx* ctx, double* ms);

/* ---- hot-path pieces (one workgroup; for unit parity) --------------------
 * Shapes: x core (a,n,b); operator core (P,n,n,Q); op stacks L (a,P,a),
 * R (b,Q,b); rhs core (p,n,q); rhs stacks Lb (p,a), Rb (q,b). */

/* scikit_tt/solvers/sle.py:194-219: Lout (b,Q,b) from L (a,P,a) */
int ttals_stack_left_op(ttals_ctx* ctx, void* stream, int a, int P, int n, int Q, int b, const double* L,
                        const double* Acore, const double* x, double* Lout);
/* scikit_tt/solvers/sle.py:250-276: Rout (a,P,a) from R (b,Q,b) */
int ttals_stack_right_op(ttals_ctx* ctx, void* stream, int a, int P, int n, int Q, int b, const double* R,
                         const double* Acore, const double* x, double* Rout);
/* scikit_tt/solvers/sle.py:222-247: Lbout (q,b) from Lb (p,a) */
int ttals_stack_left_rhs(ttals_ctx* ctx, void* stream, int a, int p, int n, int q, int b, const double* Lb,
                         const double* Bcore, const double* x, double* Lbout);
/* scikit_tt/solvers/sle.py:279-305: Rbout (p,a) from Rb (q,b) */
int ttals_stack_right_rhs(ttals_ctx* ctx, void* stream, int a, int p, int n, int q, int b, const double* Rb,
                          const double* Bcore, const double* x, double* Rbout);
/* scikit_tt/solvers/sle.py:308-347: dense local matrix (N x N, N = a*n*b),
 * rows (a',m,b'), cols (a,n,b) */
int ttals_micro_matrix(ttals_ctx* ctx, void* stream, int a, int P, int n, int Q, int b, const double* L,
                       const double* Acore, const double* R, double* M);
/* matrix-free y = M v with M as in ttals_micro_matrix */
int ttals_micro_matvec(ttals_ctx* ctx, void* stream, int a, int P, int n, int Q, int b, const double* L,
                       const double* Acore, const double* R, const double* v, double* y);
/* scikit_tt/solvers/sle.py:393-430: local rhs (a,n,b) */
int ttals_micro_rhs(ttals_ctx* ctx, void* stream, int a, int p, int n, int q, int b, const double* Lb,
                    const double* Bcore, const double* Rb, double* f);
/* scikit_tt/solvers/sle.py:502-509: solve M v = f (v holds the initial
 * guess on entry).  method: 0 auto (LU if N<=96, else CG if sym, GMRES),
 * 1 LU (N<=96), 2 CG, 3 GMRES.  iters (host) receives the Krylov count. */
int ttals_local_solve(ttals_ctx* ctx, void* stream, int a, int P, int n, int Q, int b, const double* L,
                      const double* Acore, const double* R, const double* f, double* v, int method, int sym,
                      double tol, int maxit, int* iters);
/* scikit_tt/solvers/evp.py:411-443, dense: the local eigenproblem
 * micro_op x = l micro_op_gevp x of order N <= 2048 on the device, with a
 * symmetric positive definite micro_op_gevp (M may be NULL: the standard
 * problem).  general = 0: micro_op symmetric; mode 0 ('eig'): the k
 * eigenpairs nearest sigma, ascending |l - sigma|; mode 1 ('eigh'): the k
 * largest, descending.  general = 1: any real micro_op, mode 0 only
 * (Hessenberg form and the double-shift QR algorithm, eigenvectors by
 * inverse iteration); a complex eigenvalue among the k nearest sigma is
 * TTALS_E_UNSUPPORTED.  unit != 0: eigenvectors of unit 2-norm
 * (scipy.linalg.eig), else M-orthonormal (scipy.linalg.eigh).  vals (k) and
 * vecs (N x k, row-major) are device memory.  TTALS_E_UNSUPPORTED if M is
 * not positive definite. */
int ttals_dense_eig(ttals_ctx* ctx, void* stream, int N, const double* A, const double* M, int k, int mode,
                    double sigma, int unit, int general, double* vals, double* vecs);
/* scikit_tt/solvers/evp.py:411-443 for symmetric local matrices: extreme
 * eigenpair (mode 0 lowest, 1 highest, 2 nearest sigma).  v: start vector
 * in, unit eigenvector out; lambda (host) receives the eigenvalue. */
int ttals_local_eig(ttals_ctx* ctx, void* stream, int a, int P, int n, int Q, int b, const double* L,
                    const double* Acore, const double* R, double* v, int mode, double sigma, double tol, int maxit,
                    double* lambda, int* iters);
/* scikit_tt/solvers/sle.py:517-525 (economic QR): Householder QR of the
 * row-major (m x c) matrix Y; Q (m x k) and R (k x c), k = min(m, c). */
int ttals_qr(ttals_ctx* ctx, void* stream, int m, int c, const double* Y, double* Q, double* R);

/* ---- TT algebra (replaces TT methods; asynchronous) ---------------------- */

/* scikit_tt/tensor_train.py:447-481: core product
 * (r1,m,n,r2) x (s1,n,p,s2) -> (r1*s1, m, p, r2*s2) */
int ttals_core_product(void* stream, int r1, int m, int n, int r2, int s1, int p, int s2, const double* A,
                       const double* X, double* out);
/* scikit_tt/tensor_train.py:313-333: write `src` (r0,m,n,r1) scaled by
 * `scale` into the block [o0:o0+r0, :, :, o1:o1+r1] of dst (R0,m,n,R1) */
int ttals_core_embed(void* stream, int r0, int mn, int r1, const double* src, double scale, int R0, int R1, int o0,
                     int o1, double* dst);
/* scikit_tt/tensor_train.py:1334-1430 (p=2): Euclidean/Frobenius norm of a
 * TT (operator cores are vectorised).  Synchronous; result on the host. */
int ttals_tt_norm(ttals_ctx* ctx, void* stream, const ttals_tt* t, double* out);

/* Residual norms of `count` systems in ONE launch (one workgroup per
 * system): res_norm[s] = ||A_s x_s - b_s|| and rhs_norm[s] = ||b_s|| (host
 * arrays), the quantities of the reference's rel.res check
 * (A.dot(x) - b).norm() / b.norm() (tests/test_sle.py:52-53; TT.dot
 * tensor_train.py:422-524, TT.__sub__ 346-366, TT.norm 1334-1430) and of
 * residual_error (tensor_train.py:2035-2074).  A x - b is never stored: its
 * block cores are generated into a right-to-left Householder QR sweep.
 * Synchronous. */
int ttals_residual_norm_batch(ttals_ctx* ctx, void* stream, int count, const ttals_tt* A, const ttals_tt* x,
                              const ttals_tt* b, double* res_norm, double* rhs_norm);

/* scikit_tt/utils.py:111-158 (truncated_svd) and the two-site split of
 * sle.mals (scikit_tt/solvers/sle.py:603-645): SVD of the row-major (m x c)
 * device matrix Y, keeping s/s0 > threshold (rel != 0) or s > threshold
 * (rel == 0) when threshold != 0, then at most max_rank (<= 0: no cap).
 * U (m x k), s (k), Vt (k x c) row-major device buffers with room for
 * k = min(m, c); k (host) receives the kept rank.  Householder QR of the
 * tall orientation, then a one-sided Jacobi SVD of R.  Synchronous. */
int ttals_truncated_svd(ttals_ctx* ctx, void* stream, int m, int c, const double* Y, double threshold, int rel,
                        int max_rank, double* U, double* s, double* Vt, int* k);

/* Supercore of two neighbouring cores (the two-site operator / rhs / vector
 * of sle.mals, scikit_tt/solvers/sle.py:376-388, 424-428):
 * out[r1, (m1,m2), (n1,n2), r2] = sum_s A[r1,m1,n1,s] B[s,m2,n2,r2]. */
int ttals_core_merge(void* stream, int r1, int m1, int n1, int s, int m2, int n2, int r2, const double* A,
                     const double* B, double* out);

/* scikit_tt/tensor_train.py:1092-1205 (dir 0, ortho_left over cores
 * start..end) and 1207-1310 (dir 1, ortho_right over cores start..end
 * descending), in place: the cores keep their capacity and t->ranks receives
 * the new ranks.  threshold == 0 and no max_ranks: Householder QR sweep (the
 * reference's thin-SVD ranks, same tensor).  threshold > 0 or max_ranks
 * (host int[order+1], <= 0 = unbounded): QR then one-sided Jacobi SVD of R,
 * keeping s/s0 > threshold then at most max_ranks[bond] triplets
 * (utils.py:145-156 rule).  Synchronous. */
int ttals_tt_ortho(ttals_ctx* ctx, void* stream, const ttals_tt* t, int dir, int start, int end, double threshold,
                   const int* max_ranks);

#ifdef __cplusplus
}
#endif
#endif /* TTALS_H */
