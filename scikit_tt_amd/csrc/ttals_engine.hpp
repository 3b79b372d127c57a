// The fused ALS engine: one workgroup runs a complete sle.als / evp.als
// solve (all sweeps, all core positions) for one TT system.  A launch with
// `count` workgroups solves `count` independent systems of uneven order and
// rank; the plain drop-in path is the same kernel with count = 1.
#pragma once
#include "ttals_solve.hpp"

namespace ttals {

constexpr int NDENSE = 96;  // local systems up to this size use dense LU

enum Status : int {
  ST_OK = 0,
  ST_SINGULAR = 1,      // exactly zero pivot in the local LU (LinAlgError)
  ST_NOCONV = 2,        // a local Krylov solve hit its iteration cap (warning)
  ST_UNSUPPORTED = 3,   // sigma inside the spectrum of a large local problem
  ST_NONFINITE = 4,     // NaN/Inf in a local solution
  ST_NONSYM = 5,        // evp: the operator is not symmetric (only symmetric local eigensolvers here)
  ST_DEADLINE = 7,      // the system's time budget ran out (checked before each local step)
};

// Uniform over the workgroup: has a system started at t0 (thread 0's clock)
// used up `budget` ticks of the 100 MHz constant clock?  0: no budget.
__device__ __forceinline__ bool over_budget(long long budget, unsigned long long t0, int* flag) {
  if (budget <= 0) return false;
  __syncthreads();
  if (threadIdx.x == 0) *flag = (wall_clock64() - t0 > (unsigned long long)budget) ? 1 : 0;
  __syncthreads();
  return *flag != 0;
}

// Relative asymmetry ||A - A^T||_F / ||A||_F above which evp rejects an
// operator.  Computed from Gram transfer matrices, whose rounding puts the
// value of an exactly symmetric operator near sqrt(eps) ~ 1e-8.
constexpr double SYM_TOL = 1e-6;

// Leading doubles of a system's compact result row (then its d+1 final
// ranks): lambda, sweeps, status, Krylov iterations, unconverged solves,
// local solves, most iterations in one solve, worst residual, flops, the
// seven cycle counters of profiling builds, constant-clock ticks of the solve,
// algorithmic bytes of the local solves.
constexpr int RES_HEAD = 18;

struct SysDesc {
  int d, kind, solver, repeats;
  int ortho_x0, sym, gm_restart, maxit;
  int adapt;                  // k > 0: evp local tolerance follows the eigenvalue change, capped at 1e-k
  int mvar;                   // on-chip matvec: 0 two-stage where it fits, 1 always three-stage
  int pad1[2];
  double tol, sigma, conv_eps, pad0;
  const int* n;               // [d]
  int* rx;                    // [d+1] x ranks (in/out)
  const int* ra;              // [d+1] operator ranks
  const int* rb;              // [d+1] rhs ranks
  double* const* xc;          // [d] x cores (capacity of the initial ranks)
  const double* const* ac;    // [d] operator cores (P, m, n, Q)
  const double* const* bc;    // [d] rhs cores (p, m, 1, q)
  double* const* bestc;       // [d] evp snapshot cores
  int* best_rx;               // [d+1]
  double* Ls;                 // op stacks, stack i at Ls + Loff[i]
  double* Rs;
  double* Lbs;
  double* Rbs;
  const long* Loff;
  const long* Roff;
  const long* Lboff;
  const long* Rboff;
  double* f;                  // nmax
  double* g;                  // nmax (warm start / solution)
  double* t1;                 // tmax
  double* t2;                 // tmax
  double* V;                  // (KMAX+1) * nmax
  double* dense;              // NDENSE^2 (+ NDENSE pivots as ints)
  double* qw;                 // nmax (QR work)
  double* qq;                 // nmax (explicit Q)
  double* tau;                // rmax
  double* rf;                 // rmax^2 (clean R factor)
  double* out;                // [16] lambda, sweeps, status, krylov iters, unconverged, local solves, ...
  const double* const* x0c;   // [d] initial-guess cores, copied into xc at the start (skipped where equal)
  double* res;                // compact result row: out[0..16], then the d+1 final ranks
  double* trace;              // optional: 6 doubles per local solve (nullptr = off)
  long trace_cap;             // capacity in local solves
  double* symw;               // evp: 4 * (max operator rank)^2 doubles for the symmetry test
  long long budget;           // 100 MHz clock ticks the system may run (0: no limit), checked per local step
};

// --- small helpers ---------------------------------------------------------

__device__ __forceinline__ void set_one(double* p) {
  if (threadIdx.x == 0) p[0] = 1.0;
}

// Buffers of one QR: W (m x c), Q (m x k), R (k x c), tau (k).  With `lds`
// set (global-memory panels: the arena is free), a tall panel is factored by
// the blocked Householder QR: lds / lds_cap its LDS, tws its global T
// storage.
struct QRBuf {
  double* W;
  double* Q;
  double* R;
  double* tau;
  double* lds = nullptr;
  long lds_cap = 0;
  double* tws = nullptr;
};

// On-chip QR buffers carved from the arena when m x c fits, else the global
// scratch of the system (then the arena is free for the blocked QR's block,
// and its T storage is the dense-LU buffer, which no QR overlaps).
__device__ __forceinline__ QRBuf qr_buffers(const SysDesc& D, int m, int c, double* arena, long cap) {
  const long k = m < c ? m : c;
  const long need = (long)m * c + (long)m * k + k * c + k;
  if (need <= cap) return QRBuf{arena, arena + (long)m * c, arena + (long)m * c + (long)m * k,
                                 arena + (long)m * c + (long)m * k + k * c};
  QRBuf B{D.qw, D.qq, D.rf, D.tau};
  if (cap >= GEMM_LDS_DOUBLES && ((long)c + 16) * 16 <= (long)NDENSE * NDENSE) {
    B.lds = arena;
    B.lds_cap = cap;
    B.tws = D.dense;
  }
  return B;
}

// Room for `len` more doubles in the arena after the on-chip QR buffers of an
// (m x c) factorisation, or nullptr (QR buffers in global memory, or no room).
__device__ __forceinline__ double* qr_stage(const QRBuf& B, int m, int c, long len, double* arena, long cap) {
  if (B.W != arena) return nullptr;
  const long k = m < c ? m : c;
  const long used = (long)m * c + (long)m * k + k * c + k;
  return used + len <= cap ? arena + used : nullptr;
}

// W <- the (m x c) view of v: v itself (row-major) or, transposed, v (c x m).
__device__ __forceinline__ void qr_load_panel(double* W, const double* v, int m, int c, bool transposed) {
  if (!transposed) {
    if ((long)m * c < (1L << 31)) {
      wg_copy4(v, W, m * c);
    } else {
      for (long e = threadIdx.x; e < (long)m * c; e += NT) W[e] = v[e];
    }
  } else {
    wg_transpose(v, c, m, W);  // W[i][j] = v[j][i]
  }
  __syncthreads();
}

// v is (m x c) row-major (forward) or (c x m) row-major transposed
// (backward).  Householder QR of the (m x c) view: explicit Q (m x k) into
// B.Q, the clean upper-triangular R (k x c) into B.R.  Returns k.  Small
// panels: the register QR (wg_qr_small); tall global-memory panels: the
// blocked form (wg_hhqr_blocked); else the column sweep.
__device__ int qr_factor(const double* v, int m, int c, bool transposed, const QRBuf& B, SolveLDS& S) {
  double* W = B.W;
  qr_load_panel(W, v, m, c, transposed);
  const int k = m < c ? m : c;
  if (m <= 64 && c <= 64) {
    wg_qr_small(W, m, c, B.tau, B.Q, B.R);
    return k;
  }
  if (B.lds && wg_hhqr_blocked(W, m, c, B.Q, B.R, B.tws, B.tau, B.lds, B.lds_cap, S.red)) return k;
  wg_householder(W, m, c, B.tau, S.red);
  wg_form_q(W, m, c, B.tau, B.Q);
  for (long e = threadIdx.x; e < (long)k * c; e += NT) {
    const int i = (int)(e / c), j = (int)(e % c);
    B.R[e] = (j >= i) ? W[(long)i * c + j] : 0.0;
  }
  __syncthreads();
  return k;
}

// --- algorithmic flop counts (from shapes; used for the roofline numerator) --

__device__ __forceinline__ double fl_op(double a, double P, double n, double Q, double b) {
  // stack update or local matvec: 2abn(aP + PnQ + bQ)
  return 2.0 * a * b * n * (a * P + P * n * Q + b * Q);
}
// algorithmic bytes of one local matvec at s bytes per element: the left and
// right stacks, the operator core and the iterate read once, plus the input
// and output vectors (2N) when they live in global memory.  The solvers'
// vector work (Gram-Schmidt, updates) on global-memory vectors is charged
// one element moved per vector flop on top.
__device__ __forceinline__ double by_op(double a, double P, double n, double Q, double b, double s, bool global_vec) {
  return s * (a * P * a + a * n * b + P * n * n * Q + b * Q * b + (global_vec ? 2.0 * a * n * b : 0.0));
}
__device__ __forceinline__ double fl_rhs(double a, double p, double n, double q, double b) {
  return 2.0 * a * n * q * (p + b);
}
__device__ __forceinline__ double fl_qr(double m, double c) {
  // Householder factorisation + explicit Q
  const double k = m < c ? m : c;
  return 4.0 * m * c * k - (4.0 / 3.0) * k * k * k;
}

// --- stacks ----------------------------------------------------------------

// Interface-stack updates.  When the operands (previous stack, core,
// operator core, rhs core), the two intermediates and the output fit the
// arena, they are staged in LDS and contracted with ds_gemm; otherwise the
// global-memory path (wg_gemm) runs.

template <class ACC>
__device__ double build_left(int i, const SysDesc& D, ACC* lds, long cap) {
  // L_i from L_{i-1} and core i-1
  const int j = i - 1;
  const int a = D.rx[j], n = D.n[j], b = D.rx[i], P = D.ra[j], Q = D.ra[i];
  const bool rhs = (D.kind == 0);
  const int p = D.rb[j], q = D.rb[i];
  const double fl = fl_op(a, P, n, Q, b) + (rhs ? fl_rhs(a, p, n, q, b) : 0.0);
  const long t1n = rhs && (long)a * n * q > ds_t1_size(a, P, n, b) ? (long)a * n * q : ds_t1_size(a, P, n, b);
  const long need = (long)a * P * a + (long)a * n * b + (long)P * n * n * Q + t1n + ds_t2_size(a, n, Q, b) +
                    (long)b * Q * b + (rhs ? (long)p * a + (long)p * n * q + (long)q * b : 0);
  if (need <= cap) {
    double* Lc = reinterpret_cast<double*>(lds);
    double* xc = Lc + (long)a * P * a;
    double* Ac = xc + (long)a * n * b;
    double* t1 = Ac + (long)P * n * n * Q;
    double* t2 = t1 + t1n;
    double* Lo = t2 + ds_t2_size(a, n, Q, b);
    double* Lb = Lo + (long)b * Q * b;
    double* Bc = Lb + (long)p * a;
    double* Lbo = Bc + (long)p * n * q;
    wg_copy(D.Ls + D.Loff[j], Lc, a * P * a);
    wg_copy(D.xc[j], xc, a * n * b);
    wg_copy_opcore_pnmq(D.ac[j], Ac, P, n, Q);
    if (rhs) {
      wg_copy(D.Lbs + D.Lboff[j], Lb, p * a);
      wg_copy(D.bc[j], Bc, p * n * q);
    }
    __syncthreads();
    stack_left_mf(a, P, n, Q, b, Lc, Ac, xc, Lo, t1, t2);
    if (rhs) {
      ds_gemm<double, double, true, true, true>(a, n * q, p, ctr<double>(Lb, a), crm<double>(Bc, n * q),
                              mview<double>(t1, plain(n), split(q, 1, a * n)));
      ds_gemm<double, double, true, true, true>(q, b, a * n, crm<double>(t1, a * n), crm<double>(xc, b), mrm<double>(Lbo, b));
      wg_copy(Lbo, D.Lbs + D.Lboff[i], q * b);
    }
    wg_copy(Lo, D.Ls + D.Loff[i], b * Q * b);
    __syncthreads();
    return fl;
  }
  stack_left_op<double, ACC>(a, P, n, Q, b, D.Ls + D.Loff[j], D.ac[j], D.xc[j], D.Ls + D.Loff[i], D.t1, D.t2, lds);
  if (rhs)
    stack_left_rhs<double, ACC>(a, p, n, q, b, D.Lbs + D.Lboff[j], D.bc[j], D.xc[j], D.Lbs + D.Lboff[i], D.t1, lds);
  return fl;
}

template <class ACC>
__device__ double build_right(int i, const SysDesc& D, ACC* lds, long cap) {
  // R_i from R_{i+1} and core i+1: the left-form update of the mirrored
  // problem (ranks a <-> b, operator ranks P <-> Q)
  const int j = i + 1;
  const int a = D.rx[j], n = D.n[j], b = D.rx[j + 1], P = D.ra[j], Q = D.ra[j + 1];
  const bool rhs = (D.kind == 0);
  const int p = D.rb[j], q = D.rb[j + 1];
  const double fl = fl_op(a, P, n, Q, b) + (rhs ? fl_rhs(a, p, n, q, b) : 0.0);
  const long t1n = rhs && (long)a * n * q > ds_t1_size(b, Q, n, a) ? (long)a * n * q : ds_t1_size(b, Q, n, a);
  const long need = (long)b * Q * b + (long)a * n * b * (rhs ? 2 : 1) + (long)P * n * n * Q + t1n +
                    ds_t2_size(b, n, P, a) + (long)a * P * a + (rhs ? (long)q * b + (long)p * n * q + (long)p * a : 0);
  if (need <= cap) {
    double* Rc = reinterpret_cast<double*>(lds);
    double* xt = Rc + (long)b * Q * b;
    double* Ac = xt + (long)a * n * b;
    double* t1 = Ac + (long)P * n * n * Q;
    double* t2 = t1 + t1n;
    double* Ro = t2 + ds_t2_size(b, n, P, a);
    double* xc = Ro + (long)a * P * a;  // rhs only: the core untransposed
    double* Rb = xc + (rhs ? (long)a * n * b : 0);
    double* Bc = Rb + (long)q * b;
    double* Rbo = Bc + (long)p * n * q;
    wg_copy(D.Rs + D.Roff[j], Rc, b * Q * b);
    wg_copy_core_t(D.xc[j], xt, a, n, b);
    wg_copy_opcore_qnmp(D.ac[j], Ac, P, n, Q);
    if (rhs) {
      wg_copy(D.xc[j], xc, a * n * b);
      wg_copy(D.Rbs + D.Rboff[j], Rb, q * b);
      wg_copy(D.bc[j], Bc, p * n * q);
    }
    __syncthreads();
    stack_left_mf(b, Q, n, P, a, Rc, Ac, xt, Ro, t1, t2);
    if (rhs) {
      ds_gemm<double, double, true, true, true>(a * n, q, b, crm<double>(xc, b), ctr<double>(Rb, b),
                              mview<double>(t1, split(n, 1, q * a), plain(a)));
      ds_gemm<double, double, true, true, true>(p, a, n * q, crm<double>(Bc, n * q), crm<double>(t1, a), mrm<double>(Rbo, a));
      wg_copy(Rbo, D.Rbs + D.Rboff[i], p * a);
    }
    wg_copy(Ro, D.Rs + D.Roff[i], a * P * a);
    __syncthreads();
    return fl;
  }
  stack_right_op<double, ACC>(a, P, n, Q, b, D.Rs + D.Roff[j], D.ac[j], D.xc[j], D.Rs + D.Roff[i], D.t1, D.t2, lds);
  if (rhs)
    stack_right_rhs<double, ACC>(a, p, n, q, b, D.Rbs + D.Rboff[j], D.bc[j], D.xc[j], D.Rbs + D.Rboff[i], D.t1,
                                 lds);
  return fl;
}

struct StepStats {
  long kit = 0;        // Krylov iterations (matvecs)
  int nunconv = 0;     // local solves that missed the tolerance
  int maxit_one = 0;   // most iterations in one local solve
  double worst = 0.0;  // worst final relative residual of a local Krylov solve
  double last_res = 0.0;
  long nsolve = 0;
  double flops = 0.0;  // algorithmic flops of the whole solve
  double mvb = 0.0;    // algorithmic bytes of the local solves (matvecs and vector work)
  double tol = 0.0;    // local Krylov tolerance of the current sweep
};

// Doubles of on-chip arena a local problem needs: the operator part
// (L, R, A', two intermediates) plus `nvec` working vectors of length N.
__host__ __device__ __forceinline__ long arena_need_vec(int a, int P, int n, int Q, int b, int nvec) {
  const long N = (long)a * n * b;
  return (long)a * P * a + (long)b * Q * b + (long)P * n * n * Q + ds_t1_size(a, P, n, b) +
         ds_t2_size(a, n, Q, b) + nvec * N;
}
// Two-stage form: L, M2, T1', the K-split scratch and the vectors; while M2
// is built, R and the operator core are staged where T1' and the vectors go.
__host__ __device__ __forceinline__ long arena_need_m2(int a, int P, int n, int Q, int b, int nvec) {
  const long N = (long)a * n * b, nb = (long)n * b, k2 = (long)P * n * b;
  const long body = (long)a * odd_stride((int)k2) + m2_red_need(a, (int)nb) + nvec * N;
  const long stage = (long)b * Q * b + (long)P * n * n * Q;
  return (long)a * P * a + k2 * nb + (body > stage ? body : stage);
}
// Arena doubles an on-chip local problem needs (mvar 0: the smaller form).
__host__ __device__ __forceinline__ long arena_need_onchip(int a, int P, int n, int Q, int b, int nvec, int mvar) {
  const long n3 = arena_need_vec(a, P, n, Q, b, nvec);
  if (mvar != 0) return n3;
  const long n2 = arena_need_m2(a, P, n, Q, b, nvec);
  return n2 < n3 ? n2 : n3;
}
__device__ __forceinline__ long arena_need(const SysDesc& D, int a, int P, int n, int Q, int b) {
  return arena_need_onchip(a, P, n, Q, b, D.kind == 0 ? 5 : 3, D.mvar);
}

// Sets up the on-chip local operator of `op` in the arena `ar` (`cap`
// doubles, `nvec` working vectors) and switches `op` to it: the two-stage
// form (L, M2) when mvar == 0 and it fits, else the three-stage form (L, R,
// the operator core as A'(P, n, m, Q), room for the two intermediates).
// Returns the first free double of the arena (the working vectors).  The
// caller has checked arena_need_onchip <= cap and puts a barrier between
// this and the first matvec.
__device__ __forceinline__ double* onchip_operator(LocalOp<double>& op, double* ar, long cap, int nvec, int mvar) {
  const int a = op.a, P = op.P, n = op.n, Q = op.Q, b = op.b;
  if (mvar == 0 && arena_need_m2(a, P, n, Q, b, nvec) <= cap) {
    const int nb = n * b, k2 = P * n * b;
    double* Lc = ar;
    double* M2 = Lc + (long)a * P * a;
    double* t1 = M2 + (long)k2 * nb;
    double* red = t1 + (long)a * odd_stride(k2);
    wg_copy(op.L, Lc, a * P * a);
    build_m2(op.A, op.R, P, n, Q, b, M2, t1);
    LocalOp<double> o{a, P, n, Q, b, Lc, nullptr, nullptr, t1, red, true};
    o.M2 = M2;
    op = o;
    return red + m2_red_need(a, nb);
  }
  double* Lc = ar;
  double* Rc = Lc + (long)a * P * a;
  double* Ac = Rc + (long)b * Q * b;
  double* t1 = Ac + (long)P * n * n * Q;
  double* t2 = t1 + ds_t1_size(a, P, n, b);
  wg_copy(op.L, Lc, a * P * a);
  wg_copy(op.R, Rc, b * Q * b);
  wg_copy_opcore_pnmq(op.A, Ac, P, n, Q);
  op = LocalOp<double>{a, P, n, Q, b, Lc, Ac, Rc, t1, t2, true};
  return t2 + ds_t2_size(a, n, Q, b);
}

// Working vectors of an on-chip local solve: sle (CG/GMRES) 5 vectors, evp
// (Lanczos) 3 vectors, from `vec` on.
__device__ __forceinline__ void onchip_work(Work& W, double* vec, int N, bool sle) {
  if (sle) {
    W.f = vec;
    W.x = vec + N;
    W.r = vec + 2 * N;
    W.p = vec + 3 * N;
    W.q = vec + 4 * N;
    W.cur = W.r;
    W.w = W.p;
  } else {
    W.cur = vec;
    W.w = vec + N;
    W.r = vec + 2 * N;
  }
}

// Solves the local problem at position i (warm start in D.g, solution left in
// D.g), orthonormalises and stores core i, and writes the warm start for the
// next position into D.g.  Returns a Status.  `lds` is the dynamic LDS
// region: GEMM staging tiles, or (during a local solve that fits) the
// on-chip arena of `arena` doubles.
template <class ACC>
__device__ int local_step(int i, bool forward, const SysDesc& D, bool sym, ACC* lds, long arena, SolveLDS& S,
                          double* lam, StepStats& st) {
  const int a = D.rx[i], n = D.n[i], b = D.rx[i + 1];
  const int P = D.ra[i], Q = D.ra[i + 1];
  const int N = a * n * b;
  LocalOp<double> op{a, P, n, Q, b, D.Ls + D.Loff[i], D.ac[i], D.Rs + D.Roff[i], D.t1, D.t2, false};
  int its = 0;
  const double fmv = fl_op(a, P, n, Q, b);
  const bool dense = (D.kind == 0 && N <= NDENSE);
  if (threadIdx.x == 0) S.vfl = 0.0;  // vector flops of this local solve, counted by the solvers
  if (D.kind == 0) {
    st.flops += fl_rhs(a, D.rb[i], n, D.rb[i + 1], b);
    micro_rhs<double, ACC>(a, D.rb[i], n, D.rb[i + 1], b, D.Lbs + D.Lboff[i], D.bc[i], D.Rbs + D.Rboff[i], D.f,
                           D.t1, lds);
  }
  // working vectors: on-chip arena when the local problem fits, else global
  const long long tsetup = ph_now();
  Work W;
  W.V = D.V;
  const bool onchip = !dense && arena_need(D, a, P, n, Q, b) <= arena;
  if (onchip) {
    double* vec = onchip_operator(op, reinterpret_cast<double*>(lds), arena, D.kind == 0 ? 5 : 3, D.mvar);
    onchip_work(W, vec, N, D.kind == 0);
    if (D.kind == 0) {
      wg_copy(D.f, W.f, N);
      wg_copy(D.g, W.x, N);
    }
    __syncthreads();
  } else {
    W.f = D.f;
    W.x = D.g;
    // sle: CG vectors in the (then unused) basis area; evp: Lanczos' previous
    // vector in the rhs buffer, which evp does not use
    W.r = (D.kind == 0) ? D.V : D.f;
    W.p = D.V + N;
    W.q = D.V + 2 * (long)N;
    W.cur = D.qw;
    W.w = D.qq;
  }
  ph_add(S, PH_SETUP, tsetup);
  if (D.kind == 0) {
    if (dense) {
      micro_dense<double>(a, P, n, Q, b, op.L, op.A, op.R, D.dense, N);
      wg_copy(D.f, D.g, N);
      __syncthreads();
      int* piv = reinterpret_cast<int*>(D.dense + (long)NDENSE * NDENSE);
      if (!wg_lu_solve(D.dense, N, D.g, piv, S.red)) return ST_SINGULAR;
      st.flops += 3.0 * (double)N * N * P * Q + (2.0 / 3.0) * (double)N * N * N + 2.0 * (double)N * N;
      // the operator pieces read once, the micro matrix written and read once
      st.mvb += by_op(a, P, n, Q, b, 8.0, false) + 8.0 * (2.0 * (double)N * N - N);
    } else {
      int r = 1;
      KrylovOut ko;
      if (sym) r = onchip ? wg_cg<ACC, true>(op, W, st.tol, D.maxit, lds, S, &its, &ko)
                          : wg_cg<ACC, false>(op, W, st.tol, D.maxit, lds, S, &its, &ko);
      if (r == 1) r = wg_gmres<ACC>(op, W, st.tol, D.maxit, D.gm_restart, lds, S, &its, &ko);
      if (r != 0) ++st.nunconv;
      // final true relative residual (diagnostic): the solvers' own final
      // check where they made one, else one more matvec
      double rn = ko.tres;
      if (rn < 0.0) {
        apply_op<double, ACC>(op, W.x, W.w, lds);
        ++ko.nmv;
        for (int e = threadIdx.x; e < N; e += NT) W.w[e] = W.f[e] - W.w[e];
        __syncthreads();
        rn = wg_nrm2(W.w, N, S.red);
      }
      const double fn = wg_nrm2(W.f, N, S.red);
      st.last_res = fn > 0.0 ? rn / fn : rn;
      st.worst = fmax(st.worst, st.last_res);
      if (onchip) {
        wg_copy(W.x, D.g, N);
        __syncthreads();
      }
      // the matvecs made + the vector work the solver counted as it ran
      // (Gram-Schmidt passes as made)
      st.flops += ko.nmv * fmv + S.vfl + 6.0 * N;
      st.mvb += ko.nmv * by_op(a, P, n, Q, b, 8.0, !onchip) + (onchip ? 0.0 : 8.0 * S.vfl);
    }
  } else {
    const int mode = (D.solver == 1) ? 1 : 2;
    const int r = onchip ? wg_lanczos<ACC, true>(op, D.g, W, mode, D.sigma, st.tol, D.maxit, LKMAX, lds, S, lam, &its)
                         : wg_lanczos<ACC, false>(op, D.g, W, mode, D.sigma, st.tol, D.maxit, LKMAX, lds, S, lam, &its);
    // matvecs + the vector work the solver counted as it ran (local passes
    // every step, full Gram-Schmidt only where it was made, Ritz vectors)
    st.flops += its * fmv + S.vfl;
    st.mvb += its * by_op(a, P, n, Q, b, 8.0, !onchip) + (onchip ? 0.0 : 8.0 * S.vfl);
    if (r == 3) return ST_UNSUPPORTED;
    if (r == 4) return ST_NONFINITE;
    if (r != 0) ++st.nunconv;
  }
  st.kit += its;
  st.maxit_one = its > st.maxit_one ? its : st.maxit_one;
  if (D.trace && st.nsolve < D.trace_cap && threadIdx.x == 0) {
    double* t = D.trace + 6 * st.nsolve;
    t[0] = i;
    t[1] = forward ? 1.0 : 0.0;
    t[2] = N;
    t[3] = its;
    t[4] = *lam;
    t[5] = st.last_res;
  }
  ++st.nsolve;
  if (D.kind == 0) {  // (the eigensolver checks its Ritz vector itself)
    const double nv = wg_nrm2(D.g, N, S.red);
    if (!isfinite(nv)) return ST_NONFINITE;
  }
  const long long tqr = ph_now();
  if (forward) {
    // QR of (a*n, b): core i = Q, warm start for i+1 = R * core_{i+1}
    const QRBuf B = qr_buffers(D, a * n, b, reinterpret_cast<double*>(lds), arena);
    const int k = qr_factor(D.g, a * n, b, false, B, S);
    wg_copy4(B.Q, D.xc[i], a * n * k);
    const int n1 = D.n[i + 1], c1 = D.rx[i + 2], nc = n1 * c1;
    st.flops += fl_qr(a * n, b) + 2.0 * k * n1 * c1 * b;
    // warm start g = R core_{i+1}: on the matrix cores when R is on-chip and
    // the next core, the product and the K-split scratch fit beside it
    double* xs = qr_stage(B, a * n, b, (long)b * nc + (long)k * nc + 768, reinterpret_cast<double*>(lds), arena);
    if (xs) {
      double* gs = xs + (long)b * nc;
      wg_copy(D.xc[i + 1], xs, b * nc);
      __syncthreads();
      mf_gemm_lin(k, nc, b, B.R, b, 1, xs, nc, 1, gs, fmap(1, nc, 0), fmap(1, 1, 0), gs + (long)k * nc, 768);
      wg_copy(gs, D.g, k * nc);
    } else if (B.lds) {
      wg_gemm<double, ACC>(k, nc, b, crm<double>(B.R, b), crm<double>(D.xc[i + 1], nc), mrm<double>(D.g, nc), lds);
    } else {
      ds_gemm<double, ACC>(k, nc, b, crm<double>(B.R, b), crm<double>(D.xc[i + 1], nc), mrm<double>(D.g, nc));
    }
    if (threadIdx.x == 0) D.rx[i + 1] = k;
    __syncthreads();
  } else if (i > 0) {
    // QR of the transpose of (a, n*b): core i = Q^T, warm start for i-1 =
    // core_{i-1} * R^T
    const QRBuf B = qr_buffers(D, n * b, a, reinterpret_cast<double*>(lds), arena);
    const int k = qr_factor(D.g, n * b, a, true, B, S);
    st.flops += fl_qr(n * b, a) + 2.0 * D.rx[i - 1] * D.n[i - 1] * k * a;
    // core i (k, n*b) = Q^T, Q is (n*b x k)
    wg_transpose(B.Q, n * b, k, D.xc[i]);
    const int c0 = D.rx[i - 1], n0 = D.n[i - 1], cn = c0 * n0;
    // g[(c0 n0), k] = core_{i-1}[(c0 n0), a] * R[k, a]^T (matrix cores as above)
    double* xs = qr_stage(B, n * b, a, (long)cn * a + (long)cn * k + 768, reinterpret_cast<double*>(lds), arena);
    if (xs) {
      double* gs = xs + (long)cn * a;
      wg_copy(D.xc[i - 1], xs, cn * a);
      __syncthreads();
      mf_gemm_lin(cn, k, a, xs, a, 1, B.R, 1, a, gs, fmap(1, k, 0), fmap(1, 1, 0), gs + (long)cn * k, 768);
      wg_copy(gs, D.g, cn * k);
    } else if (B.lds) {
      wg_gemm<double, ACC>(cn, k, a, crm<double>(D.xc[i - 1], a), ctr<double>(B.R, a), mrm<double>(D.g, k), lds);
    } else {
      ds_gemm<double, ACC>(cn, k, a, crm<double>(D.xc[i - 1], a), ctr<double>(B.R, a), mrm<double>(D.g, k));
    }
    if (threadIdx.x == 0) D.rx[i] = k;
    __syncthreads();
  } else {
    wg_copy(D.g, D.xc[0], N);
    __syncthreads();
  }
  ph_add(S, PH_QR, tqr);
  return ST_OK;
}

// Right-orthonormalisation of the initial guess (sle only): for i = d-1..1,
// core i = Q^T of the QR of its transpose, R^T pushed into core i-1.
template <class ACC>
__device__ void ortho_right_x(const SysDesc& D, ACC* lds, long arena, SolveLDS& S) {
  for (int i = D.d - 1; i >= 1; --i) {
    const int a = D.rx[i], n = D.n[i], b = D.rx[i + 1];
    const QRBuf B = qr_buffers(D, n * b, a, reinterpret_cast<double*>(lds), arena);
    const int k = qr_factor(D.xc[i], n * b, a, true, B, S);
    wg_transpose(B.Q, n * b, k, D.xc[i]);
    const int c0 = D.rx[i - 1], n0 = D.n[i - 1];
    if (B.lds)
      wg_gemm<double, ACC>(c0 * n0, k, a, crm<double>(D.xc[i - 1], a), ctr<double>(B.R, a), mrm<double>(D.g, k), lds);
    else
      ds_gemm<double, ACC>(c0 * n0, k, a, crm<double>(D.xc[i - 1], a), ctr<double>(B.R, a), mrm<double>(D.g, k));
    wg_copy(D.g, D.xc[i - 1], c0 * n0 * k);
    if (threadIdx.x == 0) D.rx[i] = k;
    __syncthreads();
  }
}

// --- the solve ---------------------------------------------------------------

__device__ bool operator_symmetric(const SysDesc& D, SolveLDS& S) {
  double bad = 0.0;
  for (int i = 0; i < D.d; ++i) {
    const int P = D.ra[i], n = D.n[i], Q = D.ra[i + 1];
    const long tot = (long)P * n * n * Q;
    const double* A = D.ac[i];
    for (long e = threadIdx.x; e < tot; e += NT) {
      const int q = (int)(e % Q);
      const long r = e / Q;
      const int nn = (int)(r % n), m = (int)((r / n) % n), p = (int)(r / ((long)n * n));
      if (A[e] != A[(((long)p * n + nn) * n + m) * Q + q]) bad = 1.0;
    }
  }
  return wg_sum(bad, S.red) == 0.0;
}

// ||A - A^T||_F / ||A||_F of the whole TT operator, for operators whose
// slices are not symmetric one by one (the full operator can still be).
// ||A - A^T||^2 = 2 <A, A> - 2 <A, A^T>; both inner products are chains of
// transfer matrices E'[q,q'] = sum E[p,p'] A[p,m,n,q] B[p',m,n,q'] with
// B = A or B[p',m,n,q'] = A[p',n,m,q'], rescaled after every core (the ratio
// is all that is needed).  D.symw holds the two pairs of matrices.
__device__ double operator_asymmetry(const SysDesc& D, SolveLDS& S) {
  int rmax = 1;
  for (int i = 0; i <= D.d; ++i) rmax = D.ra[i] > rmax ? D.ra[i] : rmax;
  const long r2 = (long)rmax * rmax;
  double* Ea = D.symw;            // <A, A> chain
  double* Et = D.symw + r2;       // <A, A^T> chain
  double* Fa = D.symw + 2 * r2;
  double* Ft = D.symw + 3 * r2;
  if (threadIdx.x == 0) {
    Ea[0] = 1.0;
    Et[0] = 1.0;
  }
  __syncthreads();
  for (int i = 0; i < D.d; ++i) {
    const int P = D.ra[i], n = D.n[i], Q = D.ra[i + 1];
    const double* A = D.ac[i];
    const long nn = (long)n * n;
    double mx = 0.0;
    for (int e = threadIdx.x; e < Q * Q; e += NT) {
      const int q = e / Q, q2 = e - q * Q;
      double sa = 0.0, st = 0.0;
      for (int p = 0; p < P; ++p)
        for (int p2 = 0; p2 < P; ++p2) {
          const double ea = Ea[p * P + p2], et = Et[p * P + p2];
          double ga = 0.0, gt = 0.0;
          for (int m = 0; m < n; ++m)
            for (int c = 0; c < n; ++c) {
              const double x = A[((long)p * nn + (long)m * n + c) * Q + q];
              ga += x * A[((long)p2 * nn + (long)m * n + c) * Q + q2];
              gt += x * A[((long)p2 * nn + (long)c * n + m) * Q + q2];
            }
          sa += ea * ga;
          st += et * gt;
        }
      Fa[e] = sa;
      Ft[e] = st;
      mx = fmax(mx, fabs(sa));
    }
    // rescale both chains by the same factor (the largest |<A,A>| entry)
    mx = wave_max(mx);
    __syncthreads();  // every F entry written, S.red free
    if ((threadIdx.x & 63) == 0) S.red[threadIdx.x >> 6] = mx;
    __syncthreads();
    double gm = 0.0;
    for (int w = 0; w < NWAVE; ++w) gm = fmax(gm, S.red[w]);
    const double sc = gm > 0.0 ? 1.0 / gm : 1.0;
    for (int e = threadIdx.x; e < Q * Q; e += NT) {
      Ea[e] = Fa[e] * sc;
      Et[e] = Ft[e] * sc;
    }
    __syncthreads();
  }
  const double aa = Ea[0], at = Et[0];
  __syncthreads();
  if (!(aa > 0.0)) return 0.0;  // the zero operator is symmetric
  return sqrt(fmax(0.0, 2.0 * (1.0 - at / aa)));
}

template <class ACC>
__device__ void als_solve(const SysDesc& D, ACC* lds, long arena, SolveLDS& S) {
  const int d = D.d;
  if (threadIdx.x < 8) S.prof[threadIdx.x] = 0;
  __syncthreads();
  const long long tstart = ph_now();
  const unsigned long long t_begin = wall_clock64();  // thread 0's value decides the time budget
  StepStats st;
  int status = ST_OK;
  double lam = 0.0;
  double best = INFINITY;
  double hist[3] = {INFINITY, INFINITY, INFINITY};
  int nhist = 1;  // history starts with one +inf row
  int sweeps = 0;
  long nsolves = 0;

  // operator symmetric slice by slice (A[P,m,n,Q] == A[P,n,m,Q]): enables CG
  bool sym = D.sym > 0;
  if (D.sym < 0) sym = operator_symmetric(D, S);
  // evp's local eigensolver (Lanczos) needs a symmetric operator: one whose
  // slices are not symmetric one by one is checked as a whole and rejected
  // when it is not symmetric (the reference's 'eig' would call dgeev)
  if (D.kind == 1 && !sym) {
    if (operator_asymmetry(D, S) > SYM_TOL) status = ST_NONSYM;
  }

  if (D.kind == 0 && D.ortho_x0) ortho_right_x<ACC>(D, lds, arena, S);

  // boundary stacks
  set_one(D.Ls + D.Loff[0]);
  set_one(D.Rs + D.Roff[d - 1]);
  if (D.kind == 0) {
    set_one(D.Lbs + D.Lboff[0]);
    set_one(D.Rbs + D.Rboff[d - 1]);
  }
  __syncthreads();
  if (status == ST_OK)
    for (int i = d - 2; i >= 0; --i) st.flops += build_right<ACC>(i, D, lds, arena);

  bool converged = false;
  double lam1 = 0.0, lam2 = 0.0;  // eigenvalue after the last and the second-to-last sweep
  st.tol = D.tol;
  while (sweeps < D.repeats && !converged && status == ST_OK) {
    if (D.kind == 1 && D.adapt) {
      // Inexact ALS: the local eigenproblems need only be solved to the
      // accuracy the sweep itself has reached.  The eigenvalue error of a
      // Lanczos pair with residual r is O(r^2), so the tolerance follows the
      // square root of the last relative eigenvalue change, 1e-adapt at most.
      const double cap = exp10(-(double)D.adapt);
      double t = cap;
      if (sweeps >= 2) t = fmin(cap, 0.1 * sqrt(fabs(lam1 - lam2) / fmax(fabs(lam1), 1e-300)));
      st.tol = fmax(D.tol, t);
    }
    // warm start for position 0: the current core 0
    wg_copy(D.xc[0], D.g, D.rx[0] * D.n[0] * D.rx[1]);
    __syncthreads();
    for (int i = 0; i < d && status == ST_OK; ++i) {
      const long long ts = ph_now();
      if (i > 0) st.flops += build_left<ACC>(i, D, lds, arena);
      ph_add(S, PH_STACK, ts);
      if (i < d - 1) {
        if (over_budget(D.budget, t_begin, &S.flag[3])) {
          status = ST_DEADLINE;
          break;
        }
        status = local_step<ACC>(i, true, D, sym, lds, arena, S, &lam, st);
        ++nsolves;
      }
    }
    for (int i = d - 1; i >= 0 && status == ST_OK; --i) {
      const long long ts = ph_now();
      if (i < d - 1) st.flops += build_right<ACC>(i, D, lds, arena);
      ph_add(S, PH_STACK, ts);
      if (over_budget(D.budget, t_begin, &S.flag[3])) {
        status = ST_DEADLINE;
        break;
      }
      status = local_step<ACC>(i, false, D, sym, lds, arena, S, &lam, st);
      ++nsolves;
    }
    if (status != ST_OK) break;
    ++sweeps;
    lam2 = lam1;
    lam1 = lam;
    if (D.kind == 1) {
      // best-so-far pair (strict improvement in |lambda - sigma|), then the
      // convergence test over the last <= 3 sweeps
      if (fabs(lam - D.sigma) < fabs(best - D.sigma)) {
        best = lam;
        for (int i = 0; i < d; ++i) wg_copy(D.xc[i], D.bestc[i], D.rx[i] * D.n[i] * D.rx[i + 1]);
        for (int i = threadIdx.x; i <= d; i += NT) D.best_rx[i] = D.rx[i];
        __syncthreads();
      }
      double mx = 0.0;
      const int cnt = nhist < 3 ? nhist : 3;
      for (int h = 0; h < cnt; ++h) mx = fmax(mx, fabs(hist[(nhist - 1 - h) % 3] - lam));
      if (mx < D.conv_eps) converged = true;
      hist[nhist % 3] = lam;
      ++nhist;
    }
  }
  if (D.kind == 1 && status == ST_OK && sweeps > 0) {
    for (int i = 0; i < d; ++i) wg_copy(D.bestc[i], D.xc[i], D.best_rx[i] * D.n[i] * D.best_rx[i + 1]);
    for (int i = threadIdx.x; i <= d; i += NT) D.rx[i] = D.best_rx[i];
    lam = best;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    D.out[0] = lam;
    D.out[1] = (double)sweeps;
    D.out[2] = (double)status;
    D.out[3] = (double)st.kit;
    D.out[4] = (double)st.nunconv;
    D.out[5] = (double)nsolves;
    D.out[6] = (double)st.maxit_one;
    D.out[7] = st.worst;
    D.out[8] = st.flops;
    D.out[9] = (double)(ph_now() - tstart);
    for (int k = 0; k < 6; ++k) D.out[10 + k] = (double)S.prof[k];
    D.out[16] = (double)(wall_clock64() - t_begin);  // constant-clock ticks of the solve
    D.out[17] = st.mvb;
  }
}

// One workgroup per system.  Dynamic LDS (`arena` doubles, at least the
// GEMM staging tiles) holds either the staging tiles or the on-chip copy of
// a local problem.
__device__ __forceinline__ void als_block(const SysDesc& D, double* dyn_lds, long arena, SolveLDS& S) {
  // x <- x0: the engine works in place on the output cores
  for (int i = 0; i < D.d; ++i) {
    const double* src = D.x0c[i];
    double* dst = D.xc[i];
    if (src == dst) continue;
    const long len = (long)D.rx[i] * D.n[i] * D.rx[i + 1];
    for (long e = threadIdx.x; e < len; e += NT) dst[e] = src[e];
  }
  __syncthreads();
  als_solve<double>(D, dyn_lds, arena, S);
  __syncthreads();
  // one compact row per system, so the host fetches every result with one copy
  for (int k = threadIdx.x; k < RES_HEAD + 1 + D.d; k += NT) D.res[k] = k < RES_HEAD ? D.out[k] : (double)D.rx[k - RES_HEAD];
}

}  // namespace ttals
