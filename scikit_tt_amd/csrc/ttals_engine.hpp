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
};

struct SysDesc {
  int d, kind, solver, repeats;
  int ortho_x0, sym, gm_restart, maxit;
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
  double* out;                // [8] lambda, sweeps, status, krylov iters, unconverged, local solves
  double* trace;              // optional: 6 doubles per local solve (nullptr = off)
  long trace_cap;             // capacity in local solves
};

// --- small helpers ---------------------------------------------------------

__device__ __forceinline__ void set_one(double* p) {
  if (threadIdx.x == 0) p[0] = 1.0;
}

// v is (m x c) row-major (forward) or (c x m) row-major transposed
// (backward).  Computes the Householder QR of the (m x c) view, writes the
// explicit Q (m x k) into qq, the clean upper-triangular R (k x c) into rf.
__device__ int qr_factor(const double* v, int m, int c, bool transposed, const SysDesc& D, SolveLDS& S) {
  double* W = D.qw;
  const long mc = (long)m * c;
  if (!transposed) {
    for (long e = threadIdx.x; e < mc; e += NT) W[e] = v[e];
  } else {
    // W[i][j] = v[j][i], v is (c x m)
    for (long e = threadIdx.x; e < mc; e += NT) {
      const int i = (int)(e / c), j = (int)(e % c);
      W[e] = v[(long)j * m + i];
    }
  }
  __syncthreads();
  wg_householder(W, m, c, D.tau, S.red);
  wg_form_q(W, m, c, D.tau, D.qq);
  const int k = m < c ? m : c;
  for (long e = threadIdx.x; e < (long)k * c; e += NT) {
    const int i = (int)(e / c), j = (int)(e % c);
    D.rf[e] = (j >= i) ? W[(long)i * c + j] : 0.0;
  }
  __syncthreads();
  return k;
}

// --- algorithmic flop counts (from shapes; used for the roofline numerator) --

__device__ __forceinline__ double fl_op(double a, double P, double n, double Q, double b) {
  // stack update or local matvec: 2abn(aP + PnQ + bQ)
  return 2.0 * a * b * n * (a * P + P * n * Q + b * Q);
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

template <class ACC>
__device__ double build_left(int i, const SysDesc& D, ACC* lds) {
  // L_i from L_{i-1} and core i-1
  const int j = i - 1;
  const int a = D.rx[j], n = D.n[j], b = D.rx[i];
  const double fl = fl_op(a, D.ra[j], n, D.ra[i], b) + (D.kind == 0 ? fl_rhs(a, D.rb[j], n, D.rb[i], b) : 0.0);
  stack_left_op<double, ACC>(a, D.ra[j], n, D.ra[i], b, D.Ls + D.Loff[j], D.ac[j], D.xc[j], D.Ls + D.Loff[i], D.t1,
                             D.t2, lds);
  if (D.kind == 0)
    stack_left_rhs<double, ACC>(a, D.rb[j], n, D.rb[i], b, D.Lbs + D.Lboff[j], D.bc[j], D.xc[j],
                                D.Lbs + D.Lboff[i], D.t1, lds);
  return fl;
}

template <class ACC>
__device__ double build_right(int i, const SysDesc& D, ACC* lds) {
  // R_i from R_{i+1} and core i+1
  const int j = i + 1;
  const int a = D.rx[j], n = D.n[j], b = D.rx[j + 1];
  const double fl = fl_op(a, D.ra[j], n, D.ra[j + 1], b) + (D.kind == 0 ? fl_rhs(a, D.rb[j], n, D.rb[j + 1], b) : 0.0);
  stack_right_op<double, ACC>(a, D.ra[j], n, D.ra[j + 1], b, D.Rs + D.Roff[j], D.ac[j], D.xc[j],
                              D.Rs + D.Roff[i], D.t1, D.t2, lds);
  if (D.kind == 0)
    stack_right_rhs<double, ACC>(a, D.rb[j], n, D.rb[j + 1], b, D.Rbs + D.Rboff[j], D.bc[j], D.xc[j],
                                 D.Rbs + D.Rboff[i], D.t1, lds);
  return fl;
}

// --- one local step ----------------------------------------------------------

struct StepStats {
  long kit = 0;        // Krylov iterations (matvecs)
  int nunconv = 0;     // local solves that missed the tolerance
  int maxit_one = 0;   // most iterations in one local solve
  double worst = 0.0;  // worst final relative residual of a local Krylov solve
  double last_res = 0.0;
  long nsolve = 0;
  double flops = 0.0;  // algorithmic flops of the whole solve
};

// Doubles of on-chip arena a local problem needs.
__device__ __forceinline__ long arena_need(const SysDesc& D, int a, int P, int n, int Q, int b) {
  const long N = (long)a * n * b;
  const long mx = P > Q ? P : Q;
  long need = (long)a * P * a + (long)b * Q * b + (long)P * n * n * Q + 2 * N * mx;
  need += (D.kind == 0) ? 5 * N : 2 * N;
  return need;
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
  if (D.kind == 0) {
    st.flops += fl_rhs(a, D.rb[i], n, D.rb[i + 1], b);
    micro_rhs<double, ACC>(a, D.rb[i], n, D.rb[i + 1], b, D.Lbs + D.Lboff[i], D.bc[i], D.Rbs + D.Rboff[i], D.f,
                           D.t1, lds);
  }
  // working vectors: on-chip arena when the local problem fits, else global
  Work W;
  W.V = D.V;
  const bool onchip = !dense && arena_need(D, a, P, n, Q, b) <= arena;
  if (onchip) {
    double* ar = reinterpret_cast<double*>(lds);
    const long mx = P > Q ? P : Q;
    double* Lc = ar;
    double* Rc = Lc + (long)a * P * a;
    double* Ac = Rc + (long)b * Q * b;
    double* t1 = Ac + (long)P * n * n * Q;
    double* t2 = t1 + (long)N * mx;
    double* vec = t2 + (long)N * mx;
    wg_copy(op.L, Lc, a * P * a);
    wg_copy(op.R, Rc, b * Q * b);
    wg_copy(op.A, Ac, P * n * n * Q);
    op = LocalOp<double>{a, P, n, Q, b, Lc, Ac, Rc, t1, t2, true};
    if (D.kind == 0) {
      W.f = vec;
      W.x = vec + N;
      W.r = vec + 2 * N;
      W.p = vec + 3 * N;
      W.q = vec + 4 * N;
      W.cur = W.r;
      W.w = W.p;
      wg_copy(D.f, W.f, N);
      wg_copy(D.g, W.x, N);
    } else {
      W.cur = vec;
      W.w = vec + N;
    }
    __syncthreads();
  } else {
    W.f = D.f;
    W.x = D.g;
    W.r = D.V;
    W.p = D.V + N;
    W.q = D.V + 2 * (long)N;
    W.cur = D.qw;
    W.w = D.qq;
  }
  if (D.kind == 0) {
    if (dense) {
      micro_dense<double>(a, P, n, Q, b, op.L, op.A, op.R, D.dense, N);
      wg_copy(D.f, D.g, N);
      __syncthreads();
      int* piv = reinterpret_cast<int*>(D.dense + (long)NDENSE * NDENSE);
      if (!wg_lu_solve(D.dense, N, D.g, piv, S.red)) return ST_SINGULAR;
      st.flops += 3.0 * (double)N * N * P * Q + (2.0 / 3.0) * (double)N * N * N + 2.0 * (double)N * N;
    } else {
      int r = 1;
      if (sym) r = wg_cg<ACC>(op, W, D.tol, D.maxit, lds, S, &its);
      if (r == 1) r = wg_gmres<ACC>(op, W, D.tol, D.maxit, D.gm_restart, lds, S, &its);
      if (r != 0) ++st.nunconv;
      // final true relative residual (diagnostic)
      apply_op<double, ACC>(op, W.x, W.w, lds);
      for (int e = threadIdx.x; e < N; e += NT) W.w[e] = W.f[e] - W.w[e];
      __syncthreads();
      const double rn = wg_nrm2(W.w, N, S.red), fn = wg_nrm2(W.f, N, S.red);
      st.last_res = fn > 0.0 ? rn / fn : rn;
      st.worst = fmax(st.worst, st.last_res);
      if (onchip) {
        wg_copy(W.x, D.g, N);
        __syncthreads();
      }
      // matvecs + vector work (Gram-Schmidt counted as two passes)
      const double kb = sym ? 10.0 : 8.0 * 0.5 * KMAX;
      st.flops += (its + 2) * fmv + (double)its * kb * N;
    }
  } else {
    const int mode = (D.solver == 1) ? 1 : 2;
    const int r = wg_lanczos<ACC>(op, D.g, W, mode, D.sigma, D.tol, D.maxit, D.gm_restart, lds, S, lam, &its);
    // matvecs + two Gram-Schmidt passes against on average half the basis
    st.flops += its * fmv + (double)its * 8.0 * N * 0.5 * (its < D.gm_restart ? its : D.gm_restart);
    if (r == 3) return ST_UNSUPPORTED;
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
  {
    const double nv = wg_nrm2(D.g, N, S.red);
    if (!isfinite(nv)) return ST_NONFINITE;
  }
  if (forward) {
    // QR of (a*n, b): core i = Q, warm start for i+1 = R * core_{i+1}
    const int k = qr_factor(D.g, a * n, b, false, D, S);
    wg_copy(D.qq, D.xc[i], a * n * k);
    const int n1 = D.n[i + 1], c1 = D.rx[i + 2];
    st.flops += fl_qr(a * n, b) + 2.0 * k * n1 * c1 * b;
    wg_gemm<double, ACC>(k, n1 * c1, b, crm<double>(D.rf, b), crm<double>(D.xc[i + 1], n1 * c1),
                         mrm<double>(D.g, n1 * c1), lds);
    if (threadIdx.x == 0) D.rx[i + 1] = k;
    __syncthreads();
  } else if (i > 0) {
    // QR of the transpose of (a, n*b): core i = Q^T, warm start for i-1 =
    // core_{i-1} * R^T
    const int k = qr_factor(D.g, n * b, a, true, D, S);
    st.flops += fl_qr(n * b, a) + 2.0 * D.rx[i - 1] * D.n[i - 1] * k * a;
    // core i (k, n*b) = Q^T, Q is (n*b x k)
    for (long e = threadIdx.x; e < (long)k * n * b; e += NT) {
      const int r = (int)(e / (n * b)), c = (int)(e % (n * b));
      D.xc[i][e] = D.qq[(long)c * k + r];
    }
    const int c0 = D.rx[i - 1], n0 = D.n[i - 1];
    // g[(c0 n0), k] = core_{i-1}[(c0 n0), a] * R[k, a]^T
    wg_gemm<double, ACC>(c0 * n0, k, a, crm<double>(D.xc[i - 1], a), ctr<double>(D.rf, a), mrm<double>(D.g, k),
                         lds);
    if (threadIdx.x == 0) D.rx[i] = k;
    __syncthreads();
  } else {
    wg_copy(D.g, D.xc[0], N);
    __syncthreads();
  }
  return ST_OK;
}

// Right-orthonormalisation of the initial guess (sle only): for i = d-1..1,
// core i = Q^T of the QR of its transpose, R^T pushed into core i-1.
template <class ACC>
__device__ void ortho_right_x(const SysDesc& D, ACC* lds, SolveLDS& S) {
  for (int i = D.d - 1; i >= 1; --i) {
    const int a = D.rx[i], n = D.n[i], b = D.rx[i + 1];
    const int k = qr_factor(D.xc[i], n * b, a, true, D, S);
    for (long e = threadIdx.x; e < (long)k * n * b; e += NT) {
      const int r = (int)(e / (n * b)), c = (int)(e % (n * b));
      D.xc[i][e] = D.qq[(long)c * k + r];
    }
    const int c0 = D.rx[i - 1], n0 = D.n[i - 1];
    wg_gemm<double, ACC>(c0 * n0, k, a, crm<double>(D.xc[i - 1], a), ctr<double>(D.rf, a), mrm<double>(D.g, k),
                         lds);
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

template <class ACC>
__device__ void als_solve(const SysDesc& D, ACC* lds, long arena, SolveLDS& S) {
  const int d = D.d;
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

  if (D.kind == 0 && D.ortho_x0) ortho_right_x<ACC>(D, lds, S);

  // boundary stacks
  set_one(D.Ls + D.Loff[0]);
  set_one(D.Rs + D.Roff[d - 1]);
  if (D.kind == 0) {
    set_one(D.Lbs + D.Lboff[0]);
    set_one(D.Rbs + D.Rboff[d - 1]);
  }
  __syncthreads();
  for (int i = d - 2; i >= 0; --i) st.flops += build_right<ACC>(i, D, lds);

  bool converged = false;
  while (sweeps < D.repeats && !converged && status == ST_OK) {
    // warm start for position 0: the current core 0
    wg_copy(D.xc[0], D.g, D.rx[0] * D.n[0] * D.rx[1]);
    __syncthreads();
    for (int i = 0; i < d && status == ST_OK; ++i) {
      if (i > 0) st.flops += build_left<ACC>(i, D, lds);
      if (i < d - 1) {
        status = local_step<ACC>(i, true, D, sym, lds, arena, S, &lam, st);
        ++nsolves;
      }
    }
    for (int i = d - 1; i >= 0 && status == ST_OK; --i) {
      if (i < d - 1) st.flops += build_right<ACC>(i, D, lds);
      status = local_step<ACC>(i, false, D, sym, lds, arena, S, &lam, st);
      ++nsolves;
    }
    if (status != ST_OK) break;
    ++sweeps;
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
  }
}

}  // namespace ttals
