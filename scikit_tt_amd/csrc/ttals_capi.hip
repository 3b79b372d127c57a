// libttals.so: kernels + C ABI of the MI355X TT-ALS engine (see include/ttals.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ttals_engine.hpp"
#include "ttals_resid.hpp"
#include "ttals_gang.hpp"
#include "ttals_hybrid.hpp"
#include "../../include/ttals.h"

using namespace ttals;

#define TTALS_VERSION 100
#define TTALS_DENSE_EIG_MAX 2048  // order of the largest dense local eigenproblem

// ---------------------------------------------------------------------------
// kernels

#ifndef TTALS_ALS_WAVES_PER_EU
#define TTALS_ALS_WAVES_PER_EU 2  // two resident systems per CU
#endif

// One workgroup per system (als_block).
template <class ACC>
__global__ __launch_bounds__(NT, TTALS_ALS_WAVES_PER_EU) void k_als(const SysDesc* descs, long arena) {
  static_assert(sizeof(ACC) == 8, "the engine computes in FP64");
  extern __shared__ double dyn_lds[];
  __shared__ SolveLDS S;
  als_block(descs[blockIdx.x], dyn_lds, arena, S);
}

struct PieceArgs {
  int op, a, P, n, Q, b;
  const double* L;
  const double* A;
  const double* R;   // or Lb/Rb/x depending on op
  const double* x;
  const double* f;
  double* out;
  double* v;
  double* t1;
  double* t2;
  double* V;
  double* res;       // [4] scalars out
  int method, sym, maxit, mode;
  double tol, sigma;
  long arena;        // dynamic LDS of the launch in doubles (on-chip path when the problem fits)
  int mvar;          // on-chip matvec form (SysDesc::mvar)
};

enum PieceOp { P_LEFT_OP, P_RIGHT_OP, P_LEFT_RHS, P_RIGHT_RHS, P_MICRO_DENSE, P_MATVEC, P_MICRO_RHS, P_SOLVE,
               P_EIG, P_QR };

// One workgroup.  The matvec, local solve and local eigensolve take the
// engine's on-chip path (operator and working vectors in the dynamic LDS
// arena) whenever the problem fits `A.arena`, else the global path.
__global__ __launch_bounds__(NT) void k_piece(PieceArgs A) {
  __shared__ double lds[GEMM_LDS_DOUBLES];
  __shared__ SolveLDS S;
  extern __shared__ double dyn_lds[];
  const long N = (long)A.a * A.n * A.b;
  switch (A.op) {
    case P_LEFT_OP:
    case P_RIGHT_OP: {
      // on-chip matrix-core update (the engine's) when it fits the arena; a
      // right stack is the left-form update of the mirrored problem
      const bool left = A.op == P_LEFT_OP;
      const int al = left ? A.a : A.b, Pl = left ? A.P : A.Q, Ql = left ? A.Q : A.P, bl = left ? A.b : A.a;
      const long need = (long)al * Pl * al + N + (long)A.P * A.n * A.n * A.Q + ds_t1_size(al, Pl, A.n, bl) +
                        ds_t2_size(al, A.n, Ql, bl) + (long)bl * Ql * bl;
      if (need <= A.arena) {
        double* Lc = dyn_lds;
        double* xc = Lc + (long)al * Pl * al;
        double* Ac = xc + N;
        double* t1 = Ac + (long)A.P * A.n * A.n * A.Q;
        double* t2 = t1 + ds_t1_size(al, Pl, A.n, bl);
        double* out = t2 + ds_t2_size(al, A.n, Ql, bl);
        if (left) {
          wg_copy(A.L, Lc, al * Pl * al);
          wg_copy(A.x, xc, (int)N);
          wg_copy_opcore_pnmq(A.A, Ac, A.P, A.n, A.Q);
        } else {
          wg_copy(A.R, Lc, al * Pl * al);
          wg_copy_core_t(A.x, xc, A.a, A.n, A.b);
          wg_copy_opcore_qnmp(A.A, Ac, A.P, A.n, A.Q);
        }
        __syncthreads();
        stack_left_mf(al, Pl, A.n, Ql, bl, Lc, Ac, xc, out, t1, t2);
        wg_copy(out, A.out, bl * Ql * bl);
      } else if (left) {
        stack_left_op<double, double>(A.a, A.P, A.n, A.Q, A.b, A.L, A.A, A.x, A.out, A.t1, A.t2, lds);
      } else {
        stack_right_op<double, double>(A.a, A.P, A.n, A.Q, A.b, A.R, A.A, A.x, A.out, A.t1, A.t2, lds);
      }
      break;
    }
    case P_LEFT_RHS:
      stack_left_rhs<double, double>(A.a, A.P, A.n, A.Q, A.b, A.L, A.A, A.x, A.out, A.t1, lds);
      break;
    case P_RIGHT_RHS:
      stack_right_rhs<double, double>(A.a, A.P, A.n, A.Q, A.b, A.R, A.A, A.x, A.out, A.t1, lds);
      break;
    case P_MICRO_DENSE:
      micro_dense<double>(A.a, A.P, A.n, A.Q, A.b, A.L, A.A, A.R, A.out, A.a * A.n * A.b);
      break;
    case P_MATVEC:
      if (arena_need_onchip(A.a, A.P, A.n, A.Q, A.b, 2, A.mvar) <= A.arena) {
        LocalOp<double> op{A.a, A.P, A.n, A.Q, A.b, A.L, A.A, A.R, A.t1, A.t2, false};
        double* vec = onchip_operator(op, dyn_lds, A.arena, 2, A.mvar);
        wg_copy(A.x, vec, (int)N);
        __syncthreads();
        apply_op<double, double>(op, vec, vec + N, lds);
        wg_copy(vec + N, A.out, (int)N);
      } else {
        local_matvec<double, double>(A.a, A.P, A.n, A.Q, A.b, A.L, A.A, A.R, A.x, A.out, A.t1, A.t2, lds);
      }
      break;
    case P_MICRO_RHS:
      micro_rhs<double, double>(A.a, A.P, A.n, A.Q, A.b, A.L, A.A, A.R, A.out, A.t1, lds);
      break;
    case P_SOLVE: {
      LocalOp<double> op{A.a, A.P, A.n, A.Q, A.b, A.L, A.A, A.R, A.t1, A.t2, false};
      int its = 0, r = 0;
      int method = A.method;
      if (method == 0) method = (N <= NDENSE) ? 1 : (A.sym ? 2 : 3);
      if (method == 1) {
        micro_dense<double>(A.a, A.P, A.n, A.Q, A.b, A.L, A.A, A.R, A.V, N);
        wg_copy(A.f, A.v, N);
        __syncthreads();
        r = wg_lu_solve(A.V, N, A.v, reinterpret_cast<int*>(A.V + (long)N * N), S.red) ? 0 : 1;
      } else if (arena_need_onchip(A.a, A.P, A.n, A.Q, A.b, 5, A.mvar) <= A.arena) {
        Work W;
        W.V = A.V;
        onchip_work(W, onchip_operator(op, dyn_lds, A.arena, 5, A.mvar), (int)N, true);
        wg_copy(A.f, W.f, (int)N);
        wg_copy(A.v, W.x, (int)N);
        __syncthreads();
        if (method == 2) r = wg_cg<double, true>(op, W, A.tol, A.maxit, lds, S, &its);
        if (method == 3 || r == 1) r = wg_gmres<double>(op, W, A.tol, A.maxit, KMAX, lds, S, &its);
        wg_copy(W.x, A.v, (int)N);
      } else {
        Work W;
        W.V = A.V;
        W.f = const_cast<double*>(A.f);
        W.x = A.v;
        W.r = A.V;
        W.p = A.V + N;
        W.q = A.V + 2 * (long)N;
        W.cur = A.V + (long)(KMAX + 1) * N;
        W.w = A.V + (long)(KMAX + 2) * N;
        if (method == 2) r = wg_cg<double, false>(op, W, A.tol, A.maxit, lds, S, &its);
        if (method == 3 || r == 1) r = wg_gmres<double>(op, W, A.tol, A.maxit, KMAX, lds, S, &its);
      }
      if (threadIdx.x == 0) {
        A.res[0] = r;
        A.res[1] = its;
      }
      break;
    }
    case P_EIG: {
      LocalOp<double> op{A.a, A.P, A.n, A.Q, A.b, A.L, A.A, A.R, A.t1, A.t2, false};
      Work W;
      W.V = A.V;
      int its = 0, r;
      double lam = 0.0;
      if (arena_need_onchip(A.a, A.P, A.n, A.Q, A.b, 3, A.mvar) <= A.arena) {
        onchip_work(W, onchip_operator(op, dyn_lds, A.arena, 3, A.mvar), (int)N, false);
        __syncthreads();
        r = wg_lanczos<double, true>(op, A.v, W, A.mode, A.sigma, A.tol, A.maxit, LKMAX, lds, S, &lam, &its);
      } else {
        W.cur = A.V + (long)(LKMAX + 1) * N;
        W.w = A.V + (long)(LKMAX + 2) * N;
        W.r = A.V + (long)(LKMAX + 3) * N;
        r = wg_lanczos<double, false>(op, A.v, W, A.mode, A.sigma, A.tol, A.maxit, LKMAX, lds, S, &lam, &its);
      }
      if (threadIdx.x == 0) {
        A.res[0] = r;
        A.res[1] = its;
        A.res[2] = lam;
      }
      break;
    }
    case P_QR: {
      // m = a, c = b; input x (m x c); out = Q (m x k); v = R (k x c); t1 = work
      // (W, then tau, then (c + 16) * 16 for the blocked QR's T); the
      // engine's choice of method for a global-memory panel (qr_factor)
      const int m = A.a, c = A.b;
      double* W = A.t1;
      for (long e = threadIdx.x; e < (long)m * c; e += NT) W[e] = A.x[e];
      __syncthreads();
      if (m <= 64 && c <= 64) {  // the engine's small-panel path
        wg_qr_small(W, m, c, A.t2, A.out, A.v);
        break;
      }
      if (wg_hhqr_blocked(W, m, c, A.out, A.v, A.t2 + (m > c ? m : c), A.t2, dyn_lds, A.arena, S.red)) break;
      wg_householder(W, m, c, A.t2, S.red);
      wg_form_q(W, m, c, A.t2, A.out);
      const int k = m < c ? m : c;
      for (long e = threadIdx.x; e < (long)k * c; e += NT) {
        const int i = (int)(e / c), j = (int)(e % c);
        A.v[e] = (j >= i) ? W[(long)i * c + j] : 0.0;
      }
      break;
    }
  }
}

// TT core product (r1,m,n,r2) x (s1,n,p,s2) -> (r1*s1, m, p, r2*s2):
// out[(i,k), mm, pp, (j,l)] = sum_nn A[i,mm,nn,j] X[k,nn,pp,l]
__global__ void k_core_product(int r1, int m, int n, int r2, int s1, int p, int s2, const double* __restrict__ A,
                               const double* __restrict__ X, double* __restrict__ out) {
  const long total = (long)r1 * s1 * m * p * r2 * s2;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    long t = e;
    const int l = (int)(t % s2); t /= s2;
    const int j = (int)(t % r2); t /= r2;
    const int pp = (int)(t % p); t /= p;
    const int mm = (int)(t % m); t /= m;
    const int k = (int)(t % s1); t /= s1;
    const int i = (int)t;
    double s = 0.0;
    for (int nn = 0; nn < n; ++nn)
      s += A[(((long)i * m + mm) * n + nn) * r2 + j] * X[(((long)k * n + nn) * p + pp) * s2 + l];
    out[e] = s;
  }
}

// dst[o0 + i, mn, o1 + j] = scale * src[i, mn, j]
__global__ void k_core_embed(int r0, int mn, int r1, const double* __restrict__ src, double scale, int R0, int R1,
                             int o0, int o1, double* __restrict__ dst) {
  const long total = (long)r0 * mn * r1;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int j = (int)(e % r1);
    const long t = e / r1;
    const int q = (int)(t % mn), i = (int)(t / mn);
    dst[((long)(o0 + i) * mn + q) * R1 + (o1 + j)] = scale * src[e];
  }
}

// ||t||_2 by a right-to-left QR sweep: only the R factors are needed
// (the norm is invariant under the orthogonal factors).  Cores are working
// copies with dims (r, mn, r'); they are overwritten.
struct NormArgs {
  int d;
  const int* mn;
  const int* r;
  const double* const* src;  // the caller's cores (read only)
  double* const* c;   // working cores: c[i-1] is written when core i is absorbed
  double* W;          // >= max_i mn_i * r_{i+1} * r_i
  double* Rc;         // >= r_i^2
  double* T;          // >= max_i r_{i-1} * mn_{i-1} * r_i
  double* tau;        // >= max rank
  double* out;
};
__global__ __launch_bounds__(NT) void k_tt_norm(NormArgs A) {
  __shared__ double lds[GEMM_LDS_DOUBLES];
  __shared__ SolveLDS S;
  int right = A.r[A.d];
  for (int i = A.d - 1; i >= 1; --i) {
    const int a = A.r[i], q = A.mn[i], b = right;
    const int m = q * b, c = a;
    // W (m x c) = core_i^T, core_i is (a x q*b); the last core is still the caller's
    const double* ci = (i == A.d - 1) ? A.src[i] : A.c[i];
    for (long e = threadIdx.x; e < (long)m * c; e += NT) {
      const int ii = (int)(e / c), jj = (int)(e % c);
      A.W[e] = ci[(long)jj * m + ii];
    }
    __syncthreads();
    wg_householder(A.W, m, c, A.tau, S.red);
    const int k = m < c ? m : c;
    for (long e = threadIdx.x; e < (long)k * c; e += NT) {
      const int ii = (int)(e / c), jj = (int)(e % c);
      A.Rc[e] = (jj >= ii) ? A.W[(long)ii * c + jj] : 0.0;
    }
    __syncthreads();
    // core_{i-1} (rows x a) <- core_{i-1} * R^T  (rows x k)
    const int rows = A.r[i - 1] * A.mn[i - 1];
    wg_gemm<double, double>(rows, k, a, crm<double>(A.src[i - 1], a), ctr<double>(A.Rc, c), mrm<double>(A.T, k),
                            lds);
    wg_copy(A.T, A.c[i - 1], rows * k);
    __syncthreads();
    right = k;
  }
  const long n0 = (long)A.r[0] * A.mn[0] * right;
  const double nrm = wg_nrm2(A.d == 1 ? A.src[0] : A.c[0], (int)n0, S.red);
  if (threadIdx.x == 0) A.out[0] = nrm;
}

// Left/right orthonormalisation sweep of a TT (one workgroup).  Cores are
// updated in place (capacity of the input ranks); ranks (device int[d+1]) are
// updated.  dims: mn[i] = m_i * n_i.
//   no truncation: Householder QR of the tall unfolding, Q kept, R pushed on;
//   truncation (threshold > 0 or a finite max rank): the same QR, then a
//   one-sided Jacobi SVD of R, singular triplets with s/s0 > threshold (then
//   at most max_rank[bond]) kept, S V^T (or U S) pushed on
//   (scikit_tt/tensor_train.py:1162-1189, 1268-1299).
struct OrthoArgs {
  int d, dir, start, end;  // dir 0: left (i = start..end), 1: right (i = start..end descending)
  int trunc;
  double thr;
  const int* maxr;  // [d+1], <= 0 = unbounded
  const int* mn;
  int* r;
  double* const* c;
  double* W;   // >= max core size
  double* Qm;  // >= max core size
  double* Rc;  // >= rmax^2
  double* T;   // >= max core size
  double* tau; // >= rmax
  double* Bm;  // >= rmax^2
  double* Vm;  // >= rmax^2
  double* sig; // >= rmax
  int* perm;   // >= rmax
};

// T (mt x ct) tall unfolding in A.W; on return the kept left factor
// (mt x kk, row-major) is in A.Qm and the pushed factor in A.Bm:
// B[:, perm] (ct x kk, row-major, ld kk).  Returns kk.
__device__ int ortho_factor(const OrthoArgs& A, int mt, int ct, int cap, double* lds, SolveLDS& S) {
  wg_householder(A.W, mt, ct, A.tau, S.red);
  wg_form_q(A.W, mt, ct, A.tau, A.T);  // Q (mt x k) in T
  const int k = mt < ct ? mt : ct;
  if (!A.trunc) {
    for (long e = threadIdx.x; e < (long)k * ct; e += NT) {
      const int ii = (int)(e / ct), jj = (int)(e % ct);
      A.Rc[e] = (jj >= ii) ? A.W[(long)ii * ct + jj] : 0.0;
    }
    wg_copy(A.T, A.Qm, mt * k);
    // pushed factor R^T (ct x k)
    for (long e = threadIdx.x; e < (long)ct * k; e += NT) {
      const int ii = (int)(e / k), jj = (int)(e % k);
      A.Bm[e] = A.Rc[(long)jj * ct + ii];
    }
    __syncthreads();
    return k;
  }
  // B = R^T (ct x k); Jacobi: B = U_B S, V
  for (long e = threadIdx.x; e < (long)ct * k; e += NT) {
    const int ii = (int)(e / k), jj = (int)(e % k);
    A.Rc[e] = (ii >= jj) ? A.W[(long)jj * ct + ii] : 0.0;
  }
  __syncthreads();
  wg_jacobi_svd(A.Rc, ct, k, A.Vm, A.sig, A.perm, S.red);
  // number kept: s/s0 > thr, then the cap
  int kk = 0;
  const double s0 = A.sig[A.perm[0]];
  for (int j = 0; j < k; ++j)
    if (A.thr == 0.0 || A.sig[A.perm[j]] / s0 > A.thr) ++kk;
  if (cap > 0 && kk > cap) kk = cap;
  if (kk < 1) kk = 1;
  // left factor (Q V)[:, perm[:kk]] -> Qm (mt x kk)
  wg_gemm<double, double>(mt, k, k, crm<double>(A.T, k), crm<double>(A.Vm, k), mrm<double>(A.W, k), lds);
  for (long e = threadIdx.x; e < (long)mt * kk; e += NT) {
    const int ii = (int)(e / kk), jj = (int)(e % kk);
    A.Qm[e] = A.W[(long)ii * k + A.perm[jj]];
  }
  // pushed factor B[:, perm[:kk]] (ct x kk)
  for (long e = threadIdx.x; e < (long)ct * kk; e += NT) {
    const int ii = (int)(e / kk), jj = (int)(e % kk);
    A.Bm[e] = A.Rc[(long)ii * k + A.perm[jj]];
  }
  __syncthreads();
  return kk;
}

__global__ __launch_bounds__(NT) void k_tt_ortho(OrthoArgs A) {
  __shared__ double lds[GEMM_LDS_DOUBLES];
  __shared__ SolveLDS S;
  if (A.dir == 0) {
    for (int i = A.start; i <= A.end; ++i) {
      const int a = A.r[i], q = A.mn[i], b = A.r[i + 1];
      const int m = a * q, c = b;
      wg_copy(A.c[i], A.W, m * c);
      __syncthreads();
      const int k = ortho_factor(A, m, c, A.maxr[i + 1], lds, S);
      wg_copy(A.Qm, A.c[i], m * k);
      __syncthreads();
      // core_{i+1} (b x rest) <- Bm^T (k x b) * core_{i+1}
      const int rest = A.mn[i + 1] * A.r[i + 2];
      wg_gemm<double, double>(k, rest, b, ctr<double>(A.Bm, k), crm<double>(A.c[i + 1], rest),
                              mrm<double>(A.T, rest), lds);
      wg_copy(A.T, A.c[i + 1], k * rest);
      if (threadIdx.x == 0) A.r[i + 1] = k;
      __syncthreads();
    }
  } else {
    for (int i = A.start; i >= A.end; --i) {
      const int a = A.r[i], q = A.mn[i], b = A.r[i + 1];
      const int m = q * b, c = a;
      for (long e = threadIdx.x; e < (long)m * c; e += NT) {
        const int ii = (int)(e / c), jj = (int)(e % c);
        A.W[e] = A.c[i][(long)jj * m + ii];
      }
      __syncthreads();
      const int k = ortho_factor(A, m, c, A.maxr[i], lds, S);
      // core_i (k x m) = Qm^T
      for (long e = threadIdx.x; e < (long)k * m; e += NT) {
        const int rr = (int)(e / m), cc = (int)(e % m);
        A.c[i][e] = A.Qm[(long)cc * k + rr];
      }
      __syncthreads();
      const int rows = A.r[i - 1] * A.mn[i - 1];
      wg_gemm<double, double>(rows, k, a, crm<double>(A.c[i - 1], a), crm<double>(A.Bm, k), mrm<double>(A.T, k),
                              lds);
      wg_copy(A.T, A.c[i - 1], rows * k);
      if (threadIdx.x == 0) A.r[i] = k;
      __syncthreads();
    }
  }
}

// Truncated SVD of a row-major m x c matrix Y (one workgroup): Householder
// QR of the tall orientation (Y, or Y^T when m < c), one-sided Jacobi SVD of
// R^T, singular triplets kept by s/s0 > thr (rel) or s > thr (absolute) when
// thr != 0, then at most maxr (the rule of scikit_tt/utils.py:145-156).
// U (m x k), s (k), Vt (k x c) row-major; k written to *kout.
struct SvdArgs {
  int m, c, maxr, rel;
  double thr;
  const double* Y;
  double* U;
  double* s;
  double* Vt;
  double* W;    // mt x ct
  double* Qm;   // mt x ct
  double* B;    // ct x ct
  double* Vm;   // ct x ct
  double* L;    // mt x ct
  double* sig;  // ct
  double* tau;  // ct
  int* perm;    // ct
  int* kout;
};
__global__ __launch_bounds__(NT) void k_svd_trunc(SvdArgs A) {
  __shared__ double lds[GEMM_LDS_DOUBLES];
  __shared__ double red[4 * NWAVE];
  const bool tall = A.m >= A.c;
  const int mt = tall ? A.m : A.c, ct = tall ? A.c : A.m;
  for (long e = threadIdx.x; e < (long)mt * ct; e += NT) {
    const int i = (int)(e / ct), j = (int)(e % ct);
    A.W[e] = tall ? A.Y[e] : A.Y[(long)j * A.c + i];
  }
  __syncthreads();
  wg_householder(A.W, mt, ct, A.tau, red);
  wg_form_q(A.W, mt, ct, A.tau, A.Qm);
  // B = R^T (ct x ct)
  for (long e = threadIdx.x; e < (long)ct * ct; e += NT) {
    const int i = (int)(e / ct), j = (int)(e % ct);
    A.B[e] = (i >= j) ? A.W[(long)j * ct + i] : 0.0;
  }
  __syncthreads();
  wg_jacobi_svd(A.B, ct, ct, A.Vm, A.sig, A.perm, red);
  int kk = 0;
  const double s0 = A.sig[A.perm[0]];
  for (int j = 0; j < ct; ++j) {
    const double sj = A.sig[A.perm[j]];
    if (A.thr == 0.0 || (A.rel ? (s0 > 0.0 && sj / s0 > A.thr) : sj > A.thr)) ++kk;
  }
  if (A.maxr > 0 && kk > A.maxr) kk = A.maxr;
  // left factor of the tall orientation: (Q V)[:, perm[:kk]] (mt x kk)
  wg_gemm<double, double>(mt, ct, ct, crm<double>(A.Qm, ct), crm<double>(A.Vm, ct), mrm<double>(A.L, ct), lds);
  for (long e = threadIdx.x; e < (long)mt * kk; e += NT) {
    const int i = (int)(e / kk), j = (int)(e % kk);
    const double v = A.L[(long)i * ct + A.perm[j]];
    if (tall) A.U[(long)i * kk + j] = v;       // U = left factor
    else A.Vt[(long)j * mt + i] = v;            // Vt = (left factor)^T
  }
  // right factor: B[:, perm[j]] / s (ct x kk)
  for (long e = threadIdx.x; e < (long)ct * kk; e += NT) {
    const int i = (int)(e / kk), j = (int)(e % kk);
    const double sj = A.sig[A.perm[j]];
    const double v = sj > 0.0 ? A.B[(long)i * ct + A.perm[j]] / sj : 0.0;
    if (tall) A.Vt[(long)j * ct + i] = v;       // Vt = (right factor)^T
    else A.U[(long)i * kk + j] = v;             // U = right factor
  }
  for (int j = threadIdx.x; j < kk; j += NT) A.s[j] = A.sig[A.perm[j]];
  if (threadIdx.x == 0) *A.kout = kk;
}

// Householder reduction of C (N x N, row-major, global memory) to upper
// Hessenberg form, in place (Q is not kept: eigenvalues only).
__device__ void wg_hessenberg(double* C, int N, double* red) {
  for (int j = 0; j < N - 2; ++j) {
    double s2 = 0.0;
    for (int i = j + 2 + threadIdx.x; i < N; i += NT) s2 += C[(long)i * N + j] * C[(long)i * N + j];
    s2 = wg_sum(s2, red);
    if (s2 == 0.0) continue;
    const double alpha = C[(long)(j + 1) * N + j];
    const double beta = alpha >= 0.0 ? -sqrt(alpha * alpha + s2) : sqrt(alpha * alpha + s2);
    const double tau = (beta - alpha) / beta, sc = 1.0 / (alpha - beta);
    __syncthreads();
    for (int i = j + 2 + threadIdx.x; i < N; i += NT) C[(long)i * N + j] *= sc;  // v below v_{j+1} = 1
    __syncthreads();
    // C <- (I - tau v v^T) C on rows j+1.., columns j+1..
    for (int q = j + 1 + threadIdx.x; q < N; q += NT) {
      double w = C[(long)(j + 1) * N + q];
      for (int i = j + 2; i < N; ++i) w += C[(long)i * N + j] * C[(long)i * N + q];
      w *= tau;
      C[(long)(j + 1) * N + q] -= w;
      for (int i = j + 2; i < N; ++i) C[(long)i * N + q] -= C[(long)i * N + j] * w;
    }
    __syncthreads();
    // C <- C (I - tau v v^T) on all rows, columns j+1..
    for (int r = threadIdx.x; r < N; r += NT) {
      double w = C[(long)r * N + j + 1];
      for (int i = j + 2; i < N; ++i) w += C[(long)r * N + i] * C[(long)i * N + j];
      w *= tau;
      C[(long)r * N + j + 1] -= w;
      for (int i = j + 2; i < N; ++i) C[(long)r * N + i] -= w * C[(long)i * N + j];
    }
    __syncthreads();
    if (threadIdx.x == 0) C[(long)(j + 1) * N + j] = beta;
    for (int i = j + 2 + threadIdx.x; i < N; i += NT) C[(long)i * N + j] = 0.0;
    __syncthreads();
  }
}

// Eigenvalues (wr + i wi) of the upper Hessenberg H (N x N, global) by the
// implicit double-shift QR algorithm (Francis; Golub & Van Loan, Matrix
// Computations, Alg. 7.5.1-7.5.2) on the active block, deflating at
// subdiagonals below eps (|h_{l-1,l-1}| + |h_ll|); exceptional shifts at
// iterations 10 and 20 of an eigenvalue.  The 3-element reflectors are
// computed by thread 0 and applied by all threads.  Returns 0, or 1 when an
// eigenvalue needs more than 60 iterations.
__device__ int wg_hqr(double* H, int N, double* wr, double* wi, double* sd, int* si, double* red) {
  auto h = [&](int r, int c) -> double& { return H[(long)r * N + c]; };
  double nrm = 0.0;
  for (long e = threadIdx.x; e < (long)N * N; e += NT) nrm = fmax(nrm, fabs(H[e]));
  nrm = wave_max(nrm);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nrm;
  __syncthreads();
  nrm = 0.0;
  for (int q = 0; q < NWAVE; ++q) nrm = fmax(nrm, red[q]);
  if (nrm == 0.0) nrm = 1.0;
  const double eps = 0x1p-52;
  int hi = N - 1, its = 0;
  while (hi >= 0) {
    __syncthreads();
    if (threadIdx.x == 0) {
      int l = hi;
      while (l > 0) {
        double s = fabs(h(l - 1, l - 1)) + fabs(h(l, l));
        if (s == 0.0) s = nrm;
        if (fabs(h(l, l - 1)) <= eps * s) break;
        --l;
      }
      si[0] = l;
    }
    __syncthreads();
    const int l = si[0];
    if (l == hi) {
      if (threadIdx.x == 0) {
        wr[hi] = h(hi, hi);
        wi[hi] = 0.0;
      }
      --hi;
      its = 0;
      continue;
    }
    if (l == hi - 1) {
      if (threadIdx.x == 0) {
        const double a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1), d = h(hi, hi);
        const double p = 0.5 * (a - d), q = p * p + b * c;
        if (q >= 0.0) {
          const double z = p + copysign(sqrt(q), p);
          wr[hi - 1] = d + z;
          wr[hi] = z != 0.0 ? d - b * c / z : d + z;
          wi[hi - 1] = wi[hi] = 0.0;
        } else {
          wr[hi - 1] = wr[hi] = d + p;
          wi[hi - 1] = sqrt(-q);
          wi[hi] = -sqrt(-q);
        }
      }
      hi -= 2;
      its = 0;
      continue;
    }
    if (its >= 60) return 1;
    ++its;
    if (threadIdx.x == 0) {
      double sh, t;
      if (its == 10 || its == 20) {  // exceptional shift
        const double ex = fabs(h(hi, hi - 1)) + fabs(h(hi - 1, hi - 2));
        const double x0 = h(hi, hi) + 0.75 * ex;
        sh = 2.0 * x0;
        t = x0 * x0 - 0.4375 * ex * ex;
      } else {
        sh = h(hi - 1, hi - 1) + h(hi, hi);
        t = h(hi - 1, hi - 1) * h(hi, hi) - h(hi - 1, hi) * h(hi, hi - 1);
      }
      sd[0] = h(l, l) * h(l, l) + h(l, l + 1) * h(l + 1, l) - sh * h(l, l) + t;
      sd[1] = h(l + 1, l) * (h(l, l) + h(l + 1, l + 1) - sh);
      sd[2] = h(l + 1, l) * h(l + 2, l + 1);
    }
    __syncthreads();
    for (int k = l; k <= hi - 2; ++k) {
      if (threadIdx.x == 0) {
        const double x = sd[0], y = sd[1], z = sd[2];
        const double s2 = y * y + z * z;
        if (s2 == 0.0) {
          sd[3] = 0.0;
        } else {
          const double beta = x >= 0.0 ? -sqrt(x * x + s2) : sqrt(x * x + s2);
          sd[3] = (beta - x) / beta;
          sd[4] = y / (x - beta);
          sd[5] = z / (x - beta);
        }
      }
      __syncthreads();
      const double tau = sd[3], u1 = sd[4], u2 = sd[5];
      if (tau != 0.0) {
        for (int q = (k - 1 > l ? k - 1 : l) + threadIdx.x; q <= hi; q += NT) {
          const double w = tau * (h(k, q) + u1 * h(k + 1, q) + u2 * h(k + 2, q));
          h(k, q) -= w;
          h(k + 1, q) -= u1 * w;
          h(k + 2, q) -= u2 * w;
        }
        __syncthreads();
        const int rlast = k + 3 < hi ? k + 3 : hi;
        for (int r = l + threadIdx.x; r <= rlast; r += NT) {
          const double w = tau * (h(r, k) + u1 * h(r, k + 1) + u2 * h(r, k + 2));
          h(r, k) -= w;
          h(r, k + 1) -= u1 * w;
          h(r, k + 2) -= u2 * w;
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        sd[0] = h(k + 1, k);
        sd[1] = h(k + 2, k);
        sd[2] = k + 3 <= hi ? h(k + 3, k) : 0.0;
      }
      __syncthreads();
    }
    // the last 2-element reflector on rows / columns hi-1, hi
    if (threadIdx.x == 0) {
      const double x = sd[0], y = sd[1];
      if (y == 0.0) {
        sd[3] = 0.0;
      } else {
        const double beta = x >= 0.0 ? -sqrt(x * x + y * y) : sqrt(x * x + y * y);
        sd[3] = (beta - x) / beta;
        sd[4] = y / (x - beta);
      }
    }
    __syncthreads();
    const double tau = sd[3], u1 = sd[4];
    if (tau != 0.0) {
      for (int q = (hi - 2 > l ? hi - 2 : l) + threadIdx.x; q <= hi; q += NT) {
        const double w = tau * (h(hi - 1, q) + u1 * h(hi, q));
        h(hi - 1, q) -= w;
        h(hi, q) -= u1 * w;
      }
      __syncthreads();
      for (int r = l + threadIdx.x; r <= hi; r += NT) {
        const double w = tau * (h(r, hi - 1) + u1 * h(r, hi));
        h(r, hi - 1) -= w;
        h(r, hi) -= u1 * w;
      }
    }
  }
  __syncthreads();
  return 0;
}

// LU with partial pivoting of B (N x N, global) in place, then two solves of
// B x = x (inverse iteration) with x normalised in between and at the end.
// piv: N ints of scratch; vmax/imax: NT doubles/ints of LDS.
__device__ void wg_inverse_iteration(double* B, int N, double* x, int* piv, double* vmax, int* imax) {
  for (int j = 0; j < N; ++j) {
    double best = -1.0;
    int bi = j;
    for (int i = j + threadIdx.x; i < N; i += NT) {
      const double v = fabs(B[(long)i * N + j]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    vmax[threadIdx.x] = best;
    imax[threadIdx.x] = bi;
    __syncthreads();
    if (threadIdx.x == 0) {
      double bv = -1.0;
      int bk = j;
      for (int t = 0; t < NT; ++t)
        if (vmax[t] > bv || (vmax[t] == bv && imax[t] < bk)) {
          bv = vmax[t];
          bk = imax[t];
        }
      piv[j] = bk;
      if (B[(long)bk * N + j] == 0.0) B[(long)bk * N + j] = 0x1p-1000;  // exactly singular: tiny pivot
    }
    __syncthreads();
    const int p = piv[j];
    if (p != j)
      for (int q = threadIdx.x; q < N; q += NT) {
        const double t = B[(long)j * N + q];
        B[(long)j * N + q] = B[(long)p * N + q];
        B[(long)p * N + q] = t;
      }
    __syncthreads();
    const double inv = 1.0 / B[(long)j * N + j];
    for (int i = j + 1 + threadIdx.x; i < N; i += NT) B[(long)i * N + j] *= inv;
    __syncthreads();
    const int rest = N - j - 1;
    for (long e = threadIdx.x; e < (long)rest * rest; e += NT) {
      const int i = j + 1 + (int)(e / rest), q = j + 1 + (int)(e % rest);
      B[(long)i * N + q] -= B[(long)i * N + j] * B[(long)j * N + q];
    }
    __syncthreads();
  }
  for (int it = 0; it < 2; ++it) {
    // normalise, permute, L y = P x, U x = y
    double s2 = 0.0;
    for (int i = threadIdx.x; i < N; i += NT) s2 += x[i] * x[i];
    s2 = wg_sum(s2, vmax);
    const double sc = s2 > 0.0 ? 1.0 / sqrt(s2) : 1.0;
    for (int i = threadIdx.x; i < N; i += NT) x[i] *= sc;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int j = 0; j < N; ++j)
        if (piv[j] != j) {
          const double t = x[j];
          x[j] = x[piv[j]];
          x[piv[j]] = t;
        }
    __syncthreads();
    for (int j = 0; j < N; ++j) {
      const double xj = x[j];
      for (int i = j + 1 + threadIdx.x; i < N; i += NT) x[i] -= B[(long)i * N + j] * xj;
      __syncthreads();
    }
    for (int j = N - 1; j >= 0; --j) {
      if (threadIdx.x == 0) x[j] /= B[(long)j * N + j];
      __syncthreads();
      const double xj = x[j];
      for (int i = threadIdx.x; i < j; i += NT) x[i] -= B[(long)i * N + j] * xj;
      __syncthreads();
    }
  }
  double s2 = 0.0;
  for (int i = threadIdx.x; i < N; i += NT) s2 += x[i] * x[i];
  s2 = wg_sum(s2, vmax);
  const double sc = s2 > 0.0 ? 1.0 / sqrt(s2) : 1.0;
  for (int i = threadIdx.x; i < N; i += NT) x[i] *= sc;
  __syncthreads();
}

// Dense local eigenproblem of order N on one workgroup, for the reference's
// dense local eigensolves of evp.als (scikit_tt/solvers/evp.py:411-443):
// A x = l x, or A x = l M x with M = M^T positive definite (M = L L^T,
// Cholesky; C = L^{-1} A L^{-T}; x = L^{-T} y).
// * symmetric A: C shifted positive definite by its Gershgorin bound, its
//   one-sided Jacobi SVD is its eigendecomposition; mode 0 the k nearest
//   `sigma` (ascending |l - sigma|), mode 1 the k largest (descending);
// * general A (mode 0 only): Hessenberg reduction and the double-shift QR
//   algorithm for the eigenvalues, the k nearest sigma selected (a complex one
//   among them is reported), each eigenvector by inverse iteration on
//   C - (l + delta) I with an LU factorisation.
// unit: eigenvectors of unit 2-norm (scipy.linalg.eig), else M-orthonormal
// (scipy.linalg.eigh; symmetric path).
struct EigArgs {
  int N, k, mode, unit, general;
  double sigma;
  const double* A;
  const double* M;
  double* C;
  double* C0;
  double* Lm;
  double* V;
  double* sig;
  double* wr;
  double* wi;
  double* xv;
  int* perm;
  int* idx;
  double* vals;
  double* vecs;  // N x k, row-major
  int* status;   // 0 ok, 1 M not positive definite, 2 non-finite, 3 a selected eigenvalue is complex, 4 no QR convergence
};

__global__ __launch_bounds__(NT) void k_dense_eig(EigArgs E) {
  __shared__ double red[NT];
  __shared__ int ired[NT];
  __shared__ double s_tau;
  __shared__ double sd[8];
  const int N = E.N;
  const long NN = (long)N * N;
  for (long e = threadIdx.x; e < NN; e += NT) {
    const long i = e / N, j = e % N;
    E.C[e] = E.general ? E.A[e] : 0.5 * (E.A[i * N + j] + E.A[j * N + i]);
    if (E.M) E.Lm[e] = 0.5 * (E.M[i * N + j] + E.M[j * N + i]);
  }
  __syncthreads();
  if (E.M) {
    // M = L L^T, lower L in place (right-looking)
    for (int j = 0; j < N; ++j) {
      const double piv = E.Lm[(long)j * N + j];
      if (!(piv > 0.0) || !isfinite(piv)) {
        if (threadIdx.x == 0) *E.status = 1;
        return;
      }
      const double ljj = sqrt(piv), inv = 1.0 / ljj;
      __syncthreads();
      for (int i = j + 1 + threadIdx.x; i < N; i += NT) E.Lm[(long)i * N + j] *= inv;
      if (threadIdx.x == 0) E.Lm[(long)j * N + j] = ljj;
      __syncthreads();
      const int rest = N - j - 1;
      for (long e = threadIdx.x; e < (long)rest * rest; e += NT) {
        const int p = j + 1 + (int)(e / rest), q = j + 1 + (int)(e % rest);
        if (q <= p) E.Lm[(long)p * N + q] -= E.Lm[(long)p * N + j] * E.Lm[(long)q * N + j];
      }
      __syncthreads();
    }
    // C <- L^{-1} C (forward substitution, one column per thread)
    for (int col = threadIdx.x; col < N; col += NT)
      for (int i = 0; i < N; ++i) {
        double t = E.C[(long)i * N + col];
        for (int q = 0; q < i; ++q) t -= E.Lm[(long)i * N + q] * E.C[(long)q * N + col];
        E.C[(long)i * N + col] = t / E.Lm[(long)i * N + i];
      }
    __syncthreads();
    // C <- C L^{-T} (one row per thread)
    for (int row = threadIdx.x; row < N; row += NT)
      for (int j = 0; j < N; ++j) {
        double t = E.C[(long)row * N + j];
        for (int q = 0; q < j; ++q) t -= E.C[(long)row * N + q] * E.Lm[(long)j * N + q];
        E.C[(long)row * N + j] = t / E.Lm[(long)j * N + j];
      }
    __syncthreads();
    if (!E.general)
      for (long e = threadIdx.x; e < NN; e += NT) {
        const long i = e / N, j = e % N;
        if (j > i) {
          const double v = 0.5 * (E.C[i * N + j] + E.C[j * N + i]);
          E.C[i * N + j] = v;
          E.C[j * N + i] = v;
        }
      }
    __syncthreads();
  }
  if (E.general) {
    for (long e = threadIdx.x; e < NN; e += NT) E.C0[e] = E.C[e];
    __syncthreads();
    wg_hessenberg(E.C, N, red);
    if (wg_hqr(E.C, N, E.wr, E.wi, sd, ired, red)) {
      if (threadIdx.x == 0) *E.status = 4;
      return;
    }
    if (threadIdx.x == 0) {
      for (int j = 0; j < N; ++j) E.perm[j] = 0;
      for (int s = 0; s < E.k; ++s) {
        int best = -1;
        double bv = 0.0;
        for (int j = 0; j < N; ++j) {
          if (E.perm[j]) continue;
          const double key = hypot(E.wr[j] - E.sigma, E.wi[j]);
          if (best < 0 || key < bv) {
            best = j;
            bv = key;
          }
        }
        E.perm[best] = 1;
        E.idx[s] = best;
        E.vals[s] = E.wr[best];
        if (fabs(E.wi[best]) > 1e-12 * fmax(1.0, fabs(E.wr[best]))) *E.status = 3;
      }
    }
    __syncthreads();
    if (*E.status != 0) return;
    double cn = 0.0;
    for (long e = threadIdx.x; e < NN; e += NT) cn = fmax(cn, fabs(E.C0[e]));
    cn = wave_max(cn);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cn;
    __syncthreads();
    cn = 0.0;
    for (int q = 0; q < NWAVE; ++q) cn = fmax(cn, red[q]);
    for (int s = 0; s < E.k; ++s) {
      const double lam = E.wr[E.idx[s]];
      const double mu = lam + 0x1p-36 * fmax(cn, fabs(lam));
      // B = C0 - mu I in E.C (the Hessenberg form is no longer needed)
      for (long e = threadIdx.x; e < NN; e += NT) E.C[e] = E.C0[e] - ((e / N == e % N) ? mu : 0.0);
      for (int i = threadIdx.x; i < N; i += NT) E.xv[i] = 1.0;
      __syncthreads();
      wg_inverse_iteration(E.C, N, E.xv, E.perm, red, ired);
      for (int i = threadIdx.x; i < N; i += NT) E.vecs[(long)i * E.k + s] = E.xv[i];
      __syncthreads();
    }
  } else {
    // Gershgorin bound of |eigenvalues|, then the positive definite shift
    double rmax = 0.0;
    for (int i = threadIdx.x; i < N; i += NT) {
      double sm = 0.0;
      for (int j = 0; j < N; ++j) sm += fabs(E.C[(long)i * N + j]);
      rmax = fmax(rmax, sm);
    }
    rmax = wave_max(rmax);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rmax;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = 0.0;
      for (int q = 0; q < NWAVE; ++q) m = fmax(m, red[q]);
      s_tau = m > 0.0 ? 1.0001 * m : 1.0;
      if (!isfinite(m)) *E.status = 2;
    }
    __syncthreads();
    if (!isfinite(s_tau)) return;
    const double tau = s_tau;
    for (int i = threadIdx.x; i < N; i += NT) E.C[(long)i * N + i] += tau;
    __syncthreads();
    wg_jacobi_svd(E.C, N, N, E.V, E.sig, E.perm, red);
    if (threadIdx.x == 0) {
      for (int j = 0; j < N; ++j) E.perm[j] = 0;  // reused as a taken-mark
      for (int s = 0; s < E.k; ++s) {
        int best = -1;
        double bv = 0.0;
        for (int j = 0; j < N; ++j) {
          if (E.perm[j]) continue;
          const double lam = E.sig[j] - tau;
          const double key = E.mode == 0 ? fabs(lam - E.sigma) : -lam;
          if (best < 0 || key < bv) {
            best = j;
            bv = key;
          }
        }
        E.perm[best] = 1;
        E.idx[s] = best;
        E.vals[s] = E.sig[best] - tau;
      }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < E.k; s += NT) {
      const int j = E.idx[s];
      for (int i = 0; i < N; ++i) E.vecs[(long)i * E.k + s] = E.V[(long)i * N + j];
    }
    __syncthreads();
  }
  // x = L^{-T} y for the generalised problem, then the normalisation
  for (int s = threadIdx.x; s < E.k; s += NT) {
    if (E.M)
      for (int i = N - 1; i >= 0; --i) {
        double t = E.vecs[(long)i * E.k + s];
        for (int q = i + 1; q < N; ++q) t -= E.Lm[(long)q * N + i] * E.vecs[(long)q * E.k + s];
        E.vecs[(long)i * E.k + s] = t / E.Lm[(long)i * N + i];
      }
    if (E.unit) {
      double nn = 0.0;
      for (int i = 0; i < N; ++i) nn += E.vecs[(long)i * E.k + s] * E.vecs[(long)i * E.k + s];
      const double sc = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
      for (int i = 0; i < N; ++i) E.vecs[(long)i * E.k + s] *= sc;
    }
  }
}

// Supercore of two neighbouring cores:
// out[r1, (m1, m2), (n1, n2), r2] = sum_s A[r1, m1, n1, s] B[s, m2, n2, r2]
__global__ void k_core_merge(int r1, int m1, int n1, int sr, int m2, int n2, int r2, const double* __restrict__ A,
                             const double* __restrict__ B, double* __restrict__ out) {
  const long total = (long)r1 * m1 * m2 * n1 * n2 * r2;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    long t = e;
    const int j = (int)(t % r2); t /= r2;
    const int q2 = (int)(t % n2); t /= n2;
    const int q1 = (int)(t % n1); t /= n1;
    const int p2 = (int)(t % m2); t /= m2;
    const int p1 = (int)(t % m1); t /= m1;
    const int i = (int)t;
    double acc = 0.0;
    for (int k = 0; k < sr; ++k)
      acc += A[(((long)i * m1 + p1) * n1 + q1) * sr + k] * B[(((long)k * m2 + p2) * n2 + q2) * r2 + j];
    out[e] = acc;
  }
}

// ---------------------------------------------------------------------------
// Static scratch of the allocation-free entry point
// (ttals_sle_als_batch_static): every GPU the library runs on holds one copy
// of this array, at its own address, until the process ends.  Host code
// never passes `scratch` itself to a kernel or a copy (on the host the name
// is only a stand-in): ttals_ctx_create resolves each device's address once
// with hipGetSymbolAddress, and that address is what the entry point uses.
constexpr size_t TTALS_STATIC_DEV_BYTES = (size_t)64 << 20;   // device scratch per GPU
constexpr size_t TTALS_STATIC_HOST_BYTES = (size_t)4 << 20;   // host staging (metadata and results)
constexpr int TTALS_MAX_DEVICES = 64;
static __device__ __align__(256) char scratch[TTALS_STATIC_DEV_BYTES];

// ---------------------------------------------------------------------------
// host side

struct Pending {
  size_t r_res;        // staging offset of the system's result row (16 scalars, then d+1 ranks)
  int d;
  int* ranks;          // caller's output rank array
};

// Launch-time model of one sle system, used to split a batch between gangs
// and single workgroups.  A system's cost is the flops of its three-stage
// local matvec summed over the positions of one sweep (forward and back),
// split by whether the local problem fits the on-chip arena:
//   t1 = repeats * (on * F_on + off * F_off + solve * local_solves)   [ms]
// for one workgroup, and repeats * (off * (F_on + F_off) * gang[k] + solve *
// local_solves) for a gang of 2^k workgroups: a gang keeps its operands in
// HBM/L2 and splits the flops, not the per-solve latency.
// The defaults are MI355X measurements; the tuner (scikit_tt_amd.tune)
// refits them on the device.
struct CostModel {
  double on = 1.0e-7, off = 3.4e-7, solve = 0.1;
  double gang[6] = {1.0, 0.65, 0.40, 0.36, 0.32, 0.32};
};

struct ttals_ctx {
  int device = 0;
  void* dws = nullptr;
  size_t dws_bytes = 0;
  char* hst = nullptr;       // pinned staging
  size_t hst_bytes = 0;
  hipEvent_t done = nullptr;  // completion of the last submission
  hipEvent_t k0 = nullptr, k1 = nullptr;  // bracket the last k_als launch (timing events)
  bool timed = false;
  bool inflight = false;
  std::vector<Pending> pending;
  double* trace = nullptr;   // optional device trace buffer (system 0 of a batch)
  long trace_cap = 0;
  // on-chip arena of k_als in doubles: 61 KB, which with the ~18.4 KB of
  // static LDS keeps two systems resident per CU (160 KB of LDS)
  long arena = 61 * 1024 / 8;
  bool arena_set = false;    // the caller chose the arena (ttals_set_arena_kb)
  long arena_hybrid = 0;     // k_hybrid's arena in doubles (0: not yet sized)
  double wall_khz = 0.0;     // the device's constant clock (result rows' device times)
  int tol_policy = 0;        // k > 0: adaptive evp local tolerance capped at 1e-k (SysDesc::adapt)
  int matvec_variant = 0;    // SysDesc::mvar
  int schedule = 0;          // sle batches: 0 one workgroup per system, 1 gangs for the systems that set the launch time
  int gang_prec = 0;         // GangDesc::prec of the gangs of a mixed schedule
  long long budget = 0;      // SysDesc/GangDesc::budget of later submissions (100 MHz clock ticks; 0: none)
  CostModel cost;
  std::vector<int> last_plan;  // gang size per system of the last sle batch (0: one workgroup)
  double last_plan_ms = 0.0;   // modelled makespan of the last mixed-schedule launch (0: none)
  std::string err;
};

namespace {

// per-device address of `scratch`, resolved once per device; the host
// staging of the static entry point; the lock that serialises its calls
std::mutex g_static_mu;
void* g_scratch_dev[TTALS_MAX_DEVICES] = {};
alignas(256) char g_static_host[TTALS_STATIC_HOST_BYTES];

int fail(ttals_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(ttals_ctx* c, hipError_t e, const char* where) {
  return fail(c, TTALS_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, call)                                  \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
  } while (0)

int ensure_dws(ttals_ctx* c, size_t bytes) {
  if (bytes <= c->dws_bytes) return 0;
  if (c->inflight) return fail(c, TTALS_E_ARG, "workspace busy: a submission is pending on this context");
  if (c->dws) (void)hipFree(c->dws);
  c->dws = nullptr;
  c->dws_bytes = 0;
  size_t want = std::max(bytes, (size_t)1 << 20);
  hipError_t e = hipMalloc(&c->dws, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, TTALS_E_OOM, "device workspace allocation of " + std::to_string(want) + " bytes failed");
  }
  c->dws_bytes = want;
  return 0;
}

int ensure_hst(ttals_ctx* c, size_t bytes) {
  if (bytes <= c->hst_bytes) return 0;
  if (c->inflight) return fail(c, TTALS_E_ARG, "staging busy: a submission is pending on this context");
  if (c->hst) (void)hipHostFree(c->hst);
  c->hst = nullptr;
  c->hst_bytes = 0;
  size_t want = std::max(bytes, (size_t)1 << 16);
  hipError_t e = hipHostMalloc((void**)&c->hst, want, hipHostMallocDefault);
  if (e != hipSuccess) return hip_fail(c, e, "hipHostMalloc");
  c->hst_bytes = want;
  return 0;
}

inline size_t al8(size_t x) { return (x + 7) & ~(size_t)7; }
inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte layout planner: first pass counts, second pass hands out offsets.
struct Layout {
  size_t bytes = 0;
  size_t take(size_t n) {
    size_t o = bytes;
    bytes = al256(bytes + n);
    return o;
  }
};

struct SysPlan {
  int d = 0;
  std::vector<int> n, rx, ra, rb;
  long nmax = 0, tmax = 0, rmax = 0;
  size_t xsz_total = 0;
  size_t nL = 0, nR = 0, nLb = 0, nRb = 0;
  // metadata (host staging offsets)
  size_t m_n, m_rx, m_ra, m_rb, m_brx, m_xc, m_ac, m_bc, m_bestc, m_x0c, m_Loff, m_Roff, m_Lboff, m_Rboff, m_desc;
  // device workspace offsets (doubles region)
  size_t w_L, w_R, w_Lb, w_Rb, w_f, w_g, w_t1, w_t2, w_V, w_dense, w_qw, w_qq, w_tau, w_rf, w_best, w_out, w_sym;
  std::vector<long> Loff, Roff, Lboff, Rboff;
};

int check_tt(ttals_ctx* c, const ttals_tt& t, bool vector, const char* what) {
  if (t.order < 1 || !t.row_dims || !t.col_dims || !t.ranks || !t.cores)
    return fail(c, TTALS_E_ARG, std::string(what) + ": incomplete TT descriptor");
  for (int i = 0; i < t.order; ++i) {
    if (t.row_dims[i] < 1 || t.col_dims[i] < 1 || t.ranks[i] < 1 || !t.cores[i])
      return fail(c, TTALS_E_SHAPE, std::string(what) + ": bad dimension, rank or core pointer");
    if (vector && t.col_dims[i] != 1) return fail(c, TTALS_E_SHAPE, std::string(what) + ": expected a TT vector");
  }
  if (t.ranks[t.order] < 1) return fail(c, TTALS_E_SHAPE, std::string(what) + ": bad rank");
  return 0;
}

// kind 0: sle (b given), 1: evp
int plan_system(ttals_ctx* c, int kind, const ttals_tt& A, const ttals_tt& x0, const ttals_tt* b, SysPlan& P) {
  int rc;
  if ((rc = check_tt(c, A, false, "operator"))) return rc;
  if ((rc = check_tt(c, x0, true, "initial_guess"))) return rc;
  if (b && (rc = check_tt(c, *b, true, "right_hand_side"))) return rc;
  const int d = A.order;
  if (x0.order != d || (b && b->order != d)) return fail(c, TTALS_E_SHAPE, "orders of the TTs differ");
  if (x0.ranks[0] != 1 || x0.ranks[d] != 1 || A.ranks[0] != 1 || A.ranks[d] != 1 ||
      (b && (b->ranks[0] != 1 || b->ranks[d] != 1)))
    return fail(c, TTALS_E_SHAPE, "outer TT ranks must be 1");
  P.d = d;
  P.n.resize(d);
  P.rx.resize(d + 1);
  P.ra.resize(d + 1);
  P.rb.assign(d + 1, 1);
  for (int i = 0; i < d; ++i) {
    if (A.row_dims[i] != A.col_dims[i])
      return fail(c, TTALS_E_SHAPE, "operator must be square core by core (row_dims == col_dims)");
    if (x0.row_dims[i] != A.col_dims[i]) return fail(c, TTALS_E_SHAPE, "initial guess dims != operator col dims");
    if (b && b->row_dims[i] != A.row_dims[i]) return fail(c, TTALS_E_SHAPE, "rhs dims != operator row dims");
    P.n[i] = A.row_dims[i];
  }
  for (int i = 0; i <= d; ++i) {
    P.rx[i] = x0.ranks[i];
    P.ra[i] = A.ranks[i];
    if (b) P.rb[i] = b->ranks[i];
  }
  P.nmax = P.tmax = P.rmax = 1;
  P.Loff.resize(d);
  P.Roff.resize(d);
  P.Lboff.resize(d);
  P.Rboff.resize(d);
  // Stacks are sized by the initial ranks (ranks never grow).  The left
  // stack L_i lives on bond i: (r_i, R_i, r_i), Lb_i (p_i, r_i); the right
  // stack R_i lives on bond i+1: (r_{i+1}, R_{i+1}, r_{i+1}), Rb_i
  // (p_{i+1}, r_{i+1}).
  long lo = 0, ro = 0, lbo = 0, rbo = 0;
  P.xsz_total = 0;
  for (int i = 0; i < d; ++i) {
    const long a = P.rx[i], bb = P.rx[i + 1], n = P.n[i];
    const long N = a * n * bb;
    P.nmax = std::max(P.nmax, N);
    const long mx = std::max(std::max((long)P.ra[i], (long)P.ra[i + 1]), std::max((long)P.rb[i], (long)P.rb[i + 1]));
    P.tmax = std::max(P.tmax, N * mx);
    P.rmax = std::max(P.rmax, std::max(a, bb));
    P.Loff[i] = lo;
    lo += a * P.ra[i] * a;
    P.Roff[i] = ro;
    ro += bb * P.ra[i + 1] * bb;
    P.Lboff[i] = lbo;
    lbo += (long)P.rb[i] * a;
    P.Rboff[i] = rbo;
    rbo += (long)P.rb[i + 1] * bb;
    P.xsz_total += (size_t)N;
  }
  P.nL = (size_t)lo;  // element counts
  P.nR = (size_t)ro;
  P.nLb = (size_t)lbo;
  P.nRb = (size_t)rbo;
  (void)kind;
  return 0;
}

template <class T>
T* hptr(ttals_ctx* c, size_t off) { return reinterpret_cast<T*>(c->hst + off); }

// Sizes of one engine system, read straight from its descriptors (no
// per-system heap storage: the batch planner makes two passes instead).
struct SysSizes {
  int d = 0, ramax = 1;
  long nmax = 1, tmax = 1, rmax = 1;
  size_t nL = 0, nR = 0, nLb = 0, nRb = 0, xsz = 0;
};

// Host-metadata and device-workspace offsets of one system.
struct SysOff {
  size_t m_n, m_rx, m_ra, m_rb, m_brx, m_xc, m_ac, m_bc, m_bestc, m_x0c, m_Loff, m_Roff, m_Lboff, m_Rboff;
  size_t w_L, w_R, w_Lb, w_Rb, w_f, w_g, w_t1, w_t2, w_V, w_dense, w_qw, w_qq, w_tau, w_rf, w_best, w_out, w_sym;
};

// Checks system s (shapes as plan_system, plus the output descriptor's
// dims and rank capacity) and computes its sizes.
int size_system(ttals_ctx* c, int kind, const ttals_tt& A, const ttals_tt& x0, const ttals_tt* b,
                const ttals_tt& xo, int s, SysSizes& z) {
  int rc;
  if ((rc = check_tt(c, A, false, "operator")) || (rc = check_tt(c, x0, true, "initial_guess")) ||
      (b && (rc = check_tt(c, *b, true, "right_hand_side"))))
    return fail(c, rc, "system " + std::to_string(s) + ": " + c->err);
  const int d = A.order;
  if (x0.order != d || (b && b->order != d)) return fail(c, TTALS_E_SHAPE, "orders of the TTs differ");
  if (x0.ranks[0] != 1 || x0.ranks[d] != 1 || A.ranks[0] != 1 || A.ranks[d] != 1 ||
      (b && (b->ranks[0] != 1 || b->ranks[d] != 1)))
    return fail(c, TTALS_E_SHAPE, "outer TT ranks must be 1");
  if (check_tt(c, xo, true, "x_out") || xo.order != d)
    return fail(c, TTALS_E_ARG, "system " + std::to_string(s) + ": bad output descriptor");
  z = SysSizes();
  z.d = d;
  for (int i = 0; i < d; ++i) {
    if (A.row_dims[i] != A.col_dims[i])
      return fail(c, TTALS_E_SHAPE, "operator must be square core by core (row_dims == col_dims)");
    if (x0.row_dims[i] != A.col_dims[i]) return fail(c, TTALS_E_SHAPE, "initial guess dims != operator col dims");
    if (b && b->row_dims[i] != A.row_dims[i]) return fail(c, TTALS_E_SHAPE, "rhs dims != operator row dims");
    // the engine writes x0-sized cores into the output buffers
    if (xo.row_dims[i] != A.row_dims[i] || xo.ranks[i] < x0.ranks[i] || xo.ranks[i + 1] < x0.ranks[i + 1])
      return fail(c, TTALS_E_SHAPE, "system " + std::to_string(s) + ": output core " + std::to_string(i) +
                                        " has other dims or less rank capacity than the initial guess");
    const long a = x0.ranks[i], bb = x0.ranks[i + 1], n = A.row_dims[i];
    const long p0 = b ? b->ranks[i] : 1, p1 = b ? b->ranks[i + 1] : 1;
    const long N = a * n * bb;
    z.nmax = std::max(z.nmax, N);
    const long mx = std::max(std::max((long)A.ranks[i], (long)A.ranks[i + 1]), std::max(p0, p1));
    z.tmax = std::max(z.tmax, N * mx);
    z.rmax = std::max(z.rmax, std::max(a, bb));
    // stacks sized by the initial ranks (ranks never grow): L_i on bond i
    // (r_i, R_i, r_i), Lb_i (p_i, r_i); R_i on bond i+1, Rb_i (p_{i+1}, r_{i+1})
    z.nL += (size_t)(a * A.ranks[i] * a);
    z.nR += (size_t)(bb * A.ranks[i + 1] * bb);
    z.nLb += (size_t)(p0 * a);
    z.nRb += (size_t)(p1 * bb);
    z.xsz += (size_t)N;
  }
  for (int i = 0; i <= d; ++i) z.ramax = std::max(z.ramax, A.ranks[i]);
  return 0;
}

void take_offsets(int kind, const SysSizes& z, Layout& hl, Layout& dl, SysOff& o) {
  const int d = z.d;
  o.m_n = hl.take(sizeof(int) * d);
  o.m_rx = hl.take(sizeof(int) * (d + 1));
  o.m_ra = hl.take(sizeof(int) * (d + 1));
  o.m_rb = hl.take(sizeof(int) * (d + 1));
  o.m_brx = hl.take(sizeof(int) * (d + 1));
  o.m_xc = hl.take(sizeof(void*) * d);
  o.m_ac = hl.take(sizeof(void*) * d);
  o.m_bc = hl.take(sizeof(void*) * d);
  o.m_bestc = hl.take(sizeof(void*) * d);
  o.m_x0c = hl.take(sizeof(void*) * d);
  o.m_Loff = hl.take(sizeof(long) * d);
  o.m_Roff = hl.take(sizeof(long) * d);
  o.m_Lboff = hl.take(sizeof(long) * d);
  o.m_Rboff = hl.take(sizeof(long) * d);
  o.w_L = dl.take(8 * std::max<size_t>(z.nL, 1));
  o.w_R = dl.take(8 * std::max<size_t>(z.nR, 1));
  o.w_Lb = dl.take(8 * std::max<size_t>(z.nLb, 1));
  o.w_Rb = dl.take(8 * std::max<size_t>(z.nRb, 1));
  o.w_f = dl.take(8 * z.nmax);
  o.w_g = dl.take(8 * z.nmax);
  o.w_t1 = dl.take(8 * z.tmax);
  o.w_t2 = dl.take(8 * z.tmax);
  o.w_V = dl.take(8 * (size_t)(LKMAX + 1) * z.nmax);
  o.w_dense = dl.take(8 * (size_t)(NDENSE * NDENSE + NDENSE));
  o.w_qw = dl.take(8 * z.nmax);
  o.w_qq = dl.take(8 * z.nmax);
  o.w_tau = dl.take(8 * z.rmax);
  o.w_rf = dl.take(8 * z.rmax * z.rmax);
  o.w_best = dl.take(kind == 1 ? 8 * z.xsz : 8);
  o.w_out = dl.take(8 * RES_HEAD);
  o.w_sym = dl.take(kind == 1 ? 8 * 4 * (size_t)z.ramax * z.ramax : 8);
}

// Host metadata of system s at the offsets `o`: dims, ranks, core pointers
// and stack offsets (metadata lives at hbase on the host and dmeta on the
// device, the workspace at dbase).
void fill_meta(int kind, const ttals_tt& A, const ttals_tt& x0, const ttals_tt* b, const ttals_tt& xo,
               const SysOff& o, char* hbase, char* dbase) {
  auto hp = [&](size_t off) { return reinterpret_cast<void*>(hbase + off); };
  const int d = A.order;
  int* hn = (int*)hp(o.m_n);
  int* hrx = (int*)hp(o.m_rx);
  int* hra = (int*)hp(o.m_ra);
  int* hrb = (int*)hp(o.m_rb);
  int* hbrx = (int*)hp(o.m_brx);
  for (int i = 0; i < d; ++i) hn[i] = A.row_dims[i];
  for (int i = 0; i <= d; ++i) {
    hrx[i] = hbrx[i] = x0.ranks[i];
    hra[i] = A.ranks[i];
    hrb[i] = b ? b->ranks[i] : 1;
  }
  double** xcp = (double**)hp(o.m_xc);
  const double** acp = (const double**)hp(o.m_ac);
  const double** bcp = (const double**)hp(o.m_bc);
  double** bestp = (double**)hp(o.m_bestc);
  const double** x0p = (const double**)hp(o.m_x0c);
  long* Lo = (long*)hp(o.m_Loff);
  long* Ro = (long*)hp(o.m_Roff);
  long* Lbo = (long*)hp(o.m_Lboff);
  long* Rbo = (long*)hp(o.m_Rboff);
  double* best = reinterpret_cast<double*>(dbase + o.w_best);
  size_t boff = 0;
  long lo = 0, ro = 0, lbo = 0, rbo = 0;
  for (int i = 0; i < d; ++i) {
    const long a = x0.ranks[i], bb = x0.ranks[i + 1], n = A.row_dims[i];
    x0p[i] = x0.cores[i];
    xcp[i] = xo.cores[i];
    acp[i] = A.cores[i];
    bcp[i] = b ? b->cores[i] : nullptr;
    bestp[i] = best + (kind == 1 ? boff : 0);
    boff += (size_t)(a * n * bb);
    Lo[i] = lo;
    lo += a * A.ranks[i] * a;
    Ro[i] = ro;
    ro += bb * A.ranks[i + 1] * bb;
    Lbo[i] = lbo;
    lbo += (long)hrb[i] * a;
    Rbo[i] = rbo;
    rbo += (long)hrb[i + 1] * bb;
  }
}

struct BatchArgs {
  int kind, count, repeats, solver, maxit;
  double tol, sigma, conv_eps;
  const ttals_tt *A, *x0, *b, *xo;
};

// Metadata and the single-workgroup descriptor of system s.
void fill_sys(ttals_ctx* c, const BatchArgs& B, int s, const SysOff& o, char* hbase, char* dmeta, char* dbase,
              double* dres, SysDesc& D) {
  const int kind = B.kind;
  const ttals_tt& A = B.A[s];
  const ttals_tt* b = kind == 0 ? &B.b[s] : nullptr;
  fill_meta(kind, A, B.x0[s], b, B.xo[s], o, hbase, dbase);
  auto dptr = [&](size_t off) { return reinterpret_cast<void*>(dmeta + off); };
  auto wptr = [&](size_t off) { return reinterpret_cast<double*>(dbase + off); };
  std::memset(&D, 0, sizeof(D));
  D.d = A.order;
  D.kind = kind;
  D.solver = B.solver;
  D.repeats = B.repeats;
  D.ortho_x0 = (kind == 0) ? 1 : 0;
  D.sym = -1;
  D.gm_restart = KMAX;
  D.adapt = (kind == 1) ? c->tol_policy : 0;
  D.mvar = c->matvec_variant;
  D.maxit = B.maxit;
  D.tol = B.tol;
  D.sigma = B.sigma;
  D.conv_eps = B.conv_eps;
  D.n = (const int*)dptr(o.m_n);
  D.rx = (int*)dptr(o.m_rx);
  D.ra = (const int*)dptr(o.m_ra);
  D.rb = (const int*)dptr(o.m_rb);
  D.best_rx = (int*)dptr(o.m_brx);
  D.xc = (double* const*)dptr(o.m_xc);
  D.ac = (const double* const*)dptr(o.m_ac);
  D.bc = (const double* const*)dptr(o.m_bc);
  D.bestc = (double* const*)dptr(o.m_bestc);
  D.Ls = wptr(o.w_L);
  D.Rs = wptr(o.w_R);
  D.Lbs = wptr(o.w_Lb);
  D.Rbs = wptr(o.w_Rb);
  D.Loff = (const long*)dptr(o.m_Loff);
  D.Roff = (const long*)dptr(o.m_Roff);
  D.Lboff = (const long*)dptr(o.m_Lboff);
  D.Rboff = (const long*)dptr(o.m_Rboff);
  D.f = wptr(o.w_f);
  D.g = wptr(o.w_g);
  D.t1 = wptr(o.w_t1);
  D.t2 = wptr(o.w_t2);
  D.V = wptr(o.w_V);
  D.dense = wptr(o.w_dense);
  D.qw = wptr(o.w_qw);
  D.qq = wptr(o.w_qq);
  D.tau = wptr(o.w_tau);
  D.rf = wptr(o.w_rf);
  D.out = wptr(o.w_out);
  D.x0c = (const double* const*)dptr(o.m_x0c);
  D.res = dres;
  D.trace = (c->trace && s == 0) ? c->trace : nullptr;
  D.trace_cap = c->trace_cap;
  D.symw = wptr(o.w_sym);
  D.budget = c->budget;
}

// Where a batch's metadata, workspace and results live.
struct Placement {
  char* host;       // staging: metadata, then the compact results
  char* dev;        // workspace, then the metadata mirror
  size_t host_cap;  // 0: the context's grow-only buffers (allocated here)
  size_t dev_cap;
};

// Plans and enqueues one k_als launch for a batch (descriptors, one metadata
// copy, the launch, the result copy).  On return *r_res is the staging offset
// of the compact results (res_stride doubles per system).
int enqueue_batch(ttals_ctx* c, hipStream_t stream, const BatchArgs& B, Placement& pl, size_t* r_res,
                  size_t* res_stride_out, bool timed) {
  const int kind = B.kind, count = B.count;
  if (count < 1 || !B.A || !B.x0 || !B.xo) return fail(c, TTALS_E_ARG, "bad batch arguments");
  if (kind == 0 && !B.b) return fail(c, TTALS_E_ARG, "missing right-hand side");
  if (B.repeats < 0) return fail(c, TTALS_E_ARG, "repeats must be >= 0");
  if (!(B.tol > 0.0)) return fail(c, TTALS_E_ARG, "tol must be > 0");
  if (B.maxit < 1) return fail(c, TTALS_E_ARG, "maxit must be >= 1");
  HIPCHK(c, hipSetDevice(c->device));
  // pass 1: checks and totals
  Layout hl, dl;  // host metadata layout (mirrored on device), device workspace layout
  const size_t m_descs = hl.take(sizeof(SysDesc) * count);
  int dmax = 0;
  SysSizes z;
  SysOff o;
  for (int s = 0; s < count; ++s) {
    int rc = size_system(c, kind, B.A[s], B.x0[s], kind == 0 ? &B.b[s] : nullptr, B.xo[s], s, z);
    if (rc) return rc;
    take_offsets(kind, z, hl, dl, o);
    dmax = std::max(dmax, z.d);
  }
  const size_t meta_bytes = hl.bytes;
  // compact results: one row of 16 scalars + (dmax+1) ranks per system, on
  // the device after the workspace and in the staging buffer after the
  // metadata
  const size_t res_stride = RES_HEAD + (size_t)dmax + 1;
  const size_t w_res = dl.take(8 * res_stride * count);
  Layout rl;
  rl.bytes = meta_bytes;
  *r_res = rl.take(8 * res_stride * count);
  *res_stride_out = res_stride;
  if (pl.host_cap == 0) {
    int rc;
    if ((rc = ensure_hst(c, rl.bytes))) return rc;
    if ((rc = ensure_dws(c, dl.bytes + meta_bytes))) return rc;
    pl.host = c->hst;
    pl.dev = reinterpret_cast<char*>(c->dws);
  } else if (rl.bytes > pl.host_cap || dl.bytes + meta_bytes > pl.dev_cap) {
    return fail(c, TTALS_E_ARG, "batch needs " + std::to_string(dl.bytes + meta_bytes) + " bytes of device scratch and " +
                                    std::to_string(rl.bytes) + " of host staging; the fixed buffers hold " +
                                    std::to_string(pl.dev_cap) + " and " + std::to_string(pl.host_cap));
  }
  char* hbase = pl.host;
  char* dbase = pl.dev;
  char* dmeta = dbase + dl.bytes;  // metadata mirrored after the workspace
  auto dptr = [&](size_t off) { return reinterpret_cast<void*>(dmeta + off); };
  auto wptr = [&](size_t off) { return reinterpret_cast<double*>(dbase + off); };
  SysDesc* hdesc = reinterpret_cast<SysDesc*>(hbase + m_descs);
  // pass 2: the same offsets again, descriptors written
  Layout hl2, dl2;
  (void)hl2.take(sizeof(SysDesc) * count);
  for (int s = 0; s < count; ++s) {
    (void)size_system(c, kind, B.A[s], B.x0[s], kind == 0 ? &B.b[s] : nullptr, B.xo[s], s, z);
    take_offsets(kind, z, hl2, dl2, o);
    fill_sys(c, B, s, o, hbase, dmeta, dbase, reinterpret_cast<double*>(dbase + w_res) + res_stride * s, hdesc[s]);
  }
  HIPCHK(c, hipMemcpyAsync(dmeta, hbase, meta_bytes, hipMemcpyHostToDevice, stream));
  const long arena = std::max<long>(c->arena, GEMM_LDS_DOUBLES);
  if (timed) {
    if (!c->k0) HIPCHK(c, hipEventCreate(&c->k0));
    if (!c->k1) HIPCHK(c, hipEventCreate(&c->k1));
    HIPCHK(c, hipEventRecord(c->k0, stream));
  }
  hipLaunchKernelGGL(k_als<double>, dim3(count), dim3(NT), (size_t)arena * 8, stream,
                     (const SysDesc*)dptr(m_descs), arena);
  HIPCHK(c, hipGetLastError());
  if (timed) {
    HIPCHK(c, hipEventRecord(c->k1, stream));
    c->timed = true;
  }
  HIPCHK(c, hipMemcpyAsync(hbase + *r_res, wptr(w_res), 8 * res_stride * count, hipMemcpyDeviceToHost, stream));
  return 0;
}

// Unpacks one system's compact result row (khz: the constant clock's rate);
// returns its status.
int unpack_result(const double* o, int d, int* ranks, ttals_result& R, double khz) {
  for (int i = 0; i <= d; ++i) ranks[i] = (int)o[RES_HEAD + i];
  R.device_ms = khz > 0.0 ? o[16] / khz : 0.0;
  R.bytes = o[17];
  R.lambda = o[0];
  R.sweeps = (int)o[1];
  R.status = (int)o[2];
  R.krylov_iters = (long long)o[3];
  R.unconverged = (int)o[4];
  R.local_solves = (long long)o[5];
  R.max_iters = (int)o[6];
  R.worst_residual = o[7];
  R.flops = o[8];
  for (int k = 0; k < 7; ++k) R.cycles[k] = o[9 + k];
  if (R.status == ST_OK && R.unconverged > 0) R.status = ST_NOCONV;
  return R.status;
}

// Worst status of a batch: the first non-NOCONV failure, else NOCONV, else 0.
int worst_status(int worst, int st) {
  if (st != ST_OK && (worst == 0 || (st != ST_NOCONV && worst == ST_NOCONV))) return st;
  return worst;
}

int enqueue_hybrid(ttals_ctx* c, hipStream_t stream, const BatchArgs& B, size_t* r_res, size_t* res_stride_out);

// Builds descriptors for `count` systems and enqueues k_als (one workgroup
// per system) and the result copy on the context's buffers; ttals_wait
// (finish_batch) unpacks.
int submit_batch(ttals_ctx* c, hipStream_t stream, int kind, int count, const ttals_tt* A, const ttals_tt* x0,
                 const ttals_tt* b, const ttals_tt* xo, int repeats, int solver, double tol, int maxit, double sigma,
                 double conv_eps) {
  if (c->inflight) return fail(c, TTALS_E_ARG, "a submission is still pending on this context (call ttals_wait)");
  BatchArgs B{kind, count, repeats, solver, maxit, tol, sigma, conv_eps, A, x0, b, xo};
  Placement pl{nullptr, nullptr, 0, 0};
  size_t r_res = 0, stride = 0;
  int rc = (kind == 0 && c->schedule == 1) ? enqueue_hybrid(c, stream, B, &r_res, &stride)
                                           : enqueue_batch(c, stream, B, pl, &r_res, &stride, true);
  if (!rc && !(kind == 0 && c->schedule == 1)) {
    c->last_plan.assign(kind == 0 ? count : 0, 0);
    c->last_plan_ms = 0.0;
  }
  if (rc) return rc;
  if (!c->done) HIPCHK(c, hipEventCreateWithFlags(&c->done, hipEventDisableTiming));
  HIPCHK(c, hipEventRecord(c->done, stream));
  c->pending.clear();
  for (int s = 0; s < count; ++s) c->pending.push_back(Pending{r_res + 8 * stride * s, A[s].order, xo[s].ranks});
  c->inflight = true;
  return 0;
}

// Waits for the context's submission and unpacks ranks and results.
int finish_batch(ttals_ctx* c, ttals_result* res) {
  if (!c->inflight) return fail(c, TTALS_E_ARG, "no submission pending on this context");
  c->inflight = false;
  HIPCHK(c, hipEventSynchronize(c->done));
  int worst = 0;
  for (size_t s = 0; s < c->pending.size(); ++s) {
    const Pending& pd = c->pending[s];
    const int st = unpack_result(hptr<double>(c, pd.r_res), pd.d, pd.ranks, res[s], c->wall_khz);
    if (st == ST_ABORT) return fail(c, TTALS_E_HIP, "gang barrier timed out (system " + std::to_string(s) + ")");
    worst = worst_status(worst, st);
  }
  return worst;
}

// scratch for the piece entry points: 2*t + (KMAX+2)*N doubles
int piece_scratch(ttals_ctx* c, long t, long N, double** t1, double** t2, double** V, double** res) {
  if (c->inflight) return fail(c, TTALS_E_ARG, "workspace busy: a submission is pending on this context");
  const long nv = std::max<long>((LKMAX + 4) * std::max<long>(N, 1), (long)NDENSE * NDENSE + NDENSE);
  const size_t bytes = 8 * (size_t)(2 * t + nv + 16) + 4096;
  int rc = ensure_dws(c, bytes);
  if (rc) return rc;
  double* base = reinterpret_cast<double*>(c->dws);
  *t1 = base;
  *t2 = base + t;
  *V = base + 2 * t;
  *res = base + 2 * t + nv;
  return 0;
}

int launch_piece(ttals_ctx* c, hipStream_t st, PieceArgs& a) {
  a.arena = c->arena;
  a.mvar = c->matvec_variant;
  hipLaunchKernelGGL(k_piece, dim3(1), dim3(NT), (size_t)a.arena * 8, st, a);
  HIPCHK(c, hipGetLastError());
  return 0;
}

bool dims_ok(std::initializer_list<int> v) {
  for (int x : v)
    if (x < 1) return false;
  return true;
}

// On-chip arena of k_hybrid's workgroups, in doubles.  The kernel holds one
// workgroup per CU (the gang path's registers), so its single-workgroup
// systems can use the CU's LDS less the static part (1 KB of slack, at most
// 136 KB): local problems of up to ~3000 unknowns stay on chip, against ~1500
// in k_als's 61 KB.  An arena the caller set is kept as is.
int hybrid_arena(ttals_ctx* c, long* out) {
  if (c->arena_hybrid == 0) {
    long doubles = std::max<long>(c->arena, GEMM_LDS_DOUBLES);
    if (!c->arena_set) {
      int max_lds = 0;
      size_t stat = 0;
      HIPCHK(c, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
      HIPCHK(c, hybrid_static_lds(&stat));
      const long bytes = std::min<long>((long)max_lds - (long)stat - 1024, 136L * 1024);
      doubles = std::max(doubles, bytes / 8);
    }
    c->arena_hybrid = doubles;
  }
  *out = c->arena_hybrid;
  return 0;
}

// Workgroups of k_hybrid (the kernel that runs every gang) that can be
// resident at once on the whole device with `dyn_lds` bytes of dynamic LDS.
int gang_capacity(ttals_ctx* c, size_t dyn_lds, int* cap) {
  int nb = 0, cus = 0;
  HIPCHK(c, hybrid_occupancy(&nb, dyn_lds));
  HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  *cap = nb * cus;
  return 0;
}

// Automatic gang size: one workgroup per row a' of the widest bond, at most 32.
int gang_auto_size(const ttals_tt& x0) {
  int r = 1;
  for (int i = 0; i <= x0.order; ++i) r = std::max(r, x0.ranks[i]);
  return std::max(1, std::min(32, r));
}

// Device-workspace offsets of one gang system beyond the single-workgroup ones.
struct GOff {
  SysOff o;
  size_t w_r, w_p, w_q, w_Vg, w_xt, w_At, w_L32, w_A32, w_R32, w_r32, w_p32, w_q32, w_x32, w_V32, w_t1, w_t2,
      w_part, w_gram, w_qrs;
  long rmax2, tmax;
};

void gang_offsets(const SysSizes& z, const ttals_tt& A, const ttals_tt& x0, int G, bool mixed, Layout& h,
                  Layout& dd, GOff& go) {
  take_offsets(0, z, h, dd, go.o);
  long kmax = 1, lmax = 1, rmax2 = 1;
  for (int i = 0; i < z.d; ++i) {
    const long a = x0.ranks[i], bb = x0.ranks[i + 1], n = A.row_dims[i];
    kmax = std::max(kmax, (long)A.ranks[i] * n * n * A.ranks[i + 1]);
    lmax = std::max(lmax, a * A.ranks[i] * a);
    rmax2 = std::max(rmax2, bb * A.ranks[i + 1] * bb);
  }
  go.tmax = z.tmax;
  const size_t nm = (size_t)z.nmax;
  go.w_r = dd.take(8 * nm);
  go.w_p = dd.take(8 * nm);
  go.w_q = dd.take(8 * nm);
  go.w_Vg = dd.take(8 * (size_t)(KMAX + 1) * nm);
  go.w_xt = dd.take(8 * nm);
  go.w_At = dd.take(8 * (size_t)kmax);
  go.w_L32 = dd.take(mixed ? 4 * (size_t)lmax : 4);
  go.w_A32 = dd.take(mixed ? 4 * (size_t)kmax : 4);
  go.w_R32 = dd.take(mixed ? 4 * (size_t)rmax2 : 4);
  go.w_r32 = dd.take(mixed ? 4 * nm : 4);
  go.w_p32 = dd.take(mixed ? 4 * nm : 4);
  go.w_q32 = dd.take(mixed ? 4 * nm : 4);
  go.w_x32 = dd.take(mixed ? 4 * nm : 4);
  go.w_V32 = dd.take(mixed ? 4 * (size_t)(KMAX + 1) * nm : 4);
  go.w_t1 = dd.take(8 * (size_t)G * z.tmax);
  go.w_t2 = dd.take(8 * (size_t)G * z.tmax);
  go.w_part = dd.take(8 * (size_t)G * GANG_LD);
  go.rmax2 = z.rmax * z.rmax;
  go.w_gram = dd.take(8 * (size_t)(G + 1) * go.rmax2);  // per-member partials, then the total
  go.w_qrs = dd.take(8 * (size_t)G * 3 * go.rmax2);
}

struct GangParams {
  int solver, repeats, maxit, prec;
  double tol;
  long long budget;
};

// Metadata and the gang descriptor of one system (members first .. first+G-1).
void gang_fill(const GOff& go, const ttals_tt& A, const ttals_tt& x0, const ttals_tt& b, const ttals_tt& xo, int G,
               int first, const GangParams& P, char* hbase, char* dmeta, char* dbase, GangSync* sync, double* dres,
               GangDesc& D) {
  const SysOff& o = go.o;
  fill_meta(0, A, x0, &b, xo, o, hbase, dbase);
  auto dptr = [&](size_t off) { return reinterpret_cast<void*>(dmeta + off); };
  auto wptr = [&](size_t off) { return reinterpret_cast<double*>(dbase + off); };
  auto fptr = [&](size_t off) { return reinterpret_cast<float*>(dbase + off); };
  std::memset(&D, 0, sizeof(D));
  D.d = A.order;
  D.solver = P.solver;
  D.repeats = P.repeats;
  D.maxit = P.maxit;
  D.prec = P.prec;
  D.G = G;
  D.first = first;
  D.maxit_inner = P.maxit;
  D.gm_restart = KMAX;
  D.tol = P.tol;
  D.tol_inner = 1e-5;
  D.n = (const int*)dptr(o.m_n);
  D.rx = (int*)dptr(o.m_rx);
  D.ra = (const int*)dptr(o.m_ra);
  D.rb = (const int*)dptr(o.m_rb);
  D.xc = (double* const*)dptr(o.m_xc);
  D.ac = (const double* const*)dptr(o.m_ac);
  D.bc = (const double* const*)dptr(o.m_bc);
  D.x0c = (const double* const*)dptr(o.m_x0c);
  D.Ls = wptr(o.w_L);
  D.Rs = wptr(o.w_R);
  D.Lbs = wptr(o.w_Lb);
  D.Rbs = wptr(o.w_Rb);
  D.Loff = (const long*)dptr(o.m_Loff);
  D.Roff = (const long*)dptr(o.m_Roff);
  D.Lboff = (const long*)dptr(o.m_Lboff);
  D.Rboff = (const long*)dptr(o.m_Rboff);
  D.f = wptr(o.w_f);
  D.g = wptr(o.w_g);
  D.r = wptr(go.w_r);
  D.p = wptr(go.w_p);
  D.q = wptr(go.w_q);
  D.V = wptr(go.w_Vg);
  D.xt = wptr(go.w_xt);
  D.At = wptr(go.w_At);
  D.L32 = fptr(go.w_L32);
  D.A32 = fptr(go.w_A32);
  D.R32 = fptr(go.w_R32);
  D.r32 = fptr(go.w_r32);
  D.p32 = fptr(go.w_p32);
  D.q32 = fptr(go.w_q32);
  D.x32 = fptr(go.w_x32);
  D.V32 = fptr(go.w_V32);
  D.t1 = wptr(go.w_t1);
  D.t2 = wptr(go.w_t2);
  D.tmax = go.tmax;
  D.part = wptr(go.w_part);
  D.qw = wptr(o.w_qw);
  D.qq = wptr(o.w_qq);
  D.tau = wptr(o.w_tau);
  D.rf = wptr(o.w_rf);
  D.gram = wptr(go.w_gram);
  D.qrs = wptr(go.w_qrs);
  D.rmax2 = go.rmax2;
  D.sync = sync;
  D.budget = P.budget;
  D.out = wptr(o.w_out);
  D.res = dres;
}

// Gang-bearing launches on one device run one at a time: two of them side by
// side could each hold part of the CUs with gangs that wait for members the
// other keeps out.  Each such launch waits on the previous one's event.
std::mutex g_gang_mu;
hipEvent_t g_gang_ev[TTALS_MAX_DEVICES] = {};

int gang_serialize_begin(ttals_ctx* c, hipStream_t stream) {
  if (c->device < 0 || c->device >= TTALS_MAX_DEVICES) return fail(c, TTALS_E_ARG, "device index out of range");
  if (g_gang_ev[c->device]) HIPCHK(c, hipStreamWaitEvent(stream, g_gang_ev[c->device], 0));
  return 0;
}

int gang_serialize_end(ttals_ctx* c, hipStream_t stream) {
  hipEvent_t& ev = g_gang_ev[c->device];
  if (!ev) HIPCHK(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  HIPCHK(c, hipEventRecord(ev, stream));
  return 0;
}

// Static cost of one sle system (see CostModel): matvec flops over one sweep,
// on-chip and off-chip, and the number of local solves per sweep.
void sle_cost(const ttals_tt& A, const ttals_tt& x0, long arena, int mvar, double* on, double* off, int* nsolve) {
  const int d = A.order;
  double fo = 0.0, ff = 0.0;
  for (int i = 0; i < d; ++i) {
    const int a = x0.ranks[i], n = A.row_dims[i], b = x0.ranks[i + 1], P = A.ranks[i], Q = A.ranks[i + 1];
    const double mv = 2.0 * ((double)a * P * a * n * b + (double)a * b * P * n * n * Q + (double)a * n * Q * b * b);
    const double visits = (i == d - 1) ? 1.0 : 2.0;  // forward half-sweep 0..d-1, backward d-2..0
    const long N = (long)a * n * b;
    if (N <= NDENSE || arena_need_onchip(a, P, n, Q, b, 5, mvar) <= arena)
      fo += visits * mv;
    else
      ff += visits * mv;
  }
  *on = fo;
  *off = ff;
  *nsolve = 2 * d - 1;
}

// Makespan of a split on `slots` resident workgroups: gang s holds G[s]
// slots until tg[s], then single-workgroup systems take the first free slot
// in the order given (largest first).
double simulate_split(const std::vector<double>& t1, const std::vector<double>& tgang, const std::vector<int>& G,
                      const std::vector<int>& order, int slots) {
  std::vector<double> heap;
  heap.reserve(slots);
  double end = 0.0;
  int used = 0;
  for (size_t s = 0; s < G.size(); ++s)
    if (G[s] > 0) {
      for (int k = 0; k < G[s]; ++k) heap.push_back(tgang[s]);
      used += G[s];
      end = std::max(end, tgang[s]);
    }
  for (int k = used; k < slots; ++k) heap.push_back(0.0);
  auto cmp = [](double x, double y) { return x > y; };
  std::make_heap(heap.begin(), heap.end(), cmp);
  for (int s : order) {
    if (G[s] > 0) continue;
    std::pop_heap(heap.begin(), heap.end(), cmp);
    const double e = heap.back() + t1[s];
    heap.back() = e;
    std::push_heap(heap.begin(), heap.end(), cmp);
    end = std::max(end, e);
  }
  return end;
}

// Gang sizes (0: one workgroup) that minimise the modelled makespan: for a
// range of target times T, every system slower than T alone gets the
// smallest power-of-two gang (at most gmax[s]) that brings it under T, where
// a gang of 2^k takes tmv[s] * gang[k] + tov[s].
double plan_split(const CostModel& cm, const std::vector<double>& t1, const std::vector<double>& tmv,
                  const std::vector<double>& tov, const std::vector<int>& gmax, int slots, std::vector<int>& best) {
  const int count = (int)t1.size();
  std::vector<int> order(count);
  for (int s = 0; s < count; ++s) order[s] = s;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return t1[x] > t1[y]; });
  best.assign(count, 0);
  std::vector<double> tgang(count, 0.0);
  double best_t = simulate_split(t1, tgang, best, order, slots);
  const double tmax = t1[order[0]];
  std::vector<int> G(count);
  for (double T = tmax * 0.97; T > tmax * 0.02; T *= 0.95) {
    std::fill(G.begin(), G.end(), 0);
    int used = 0;
    bool any = false;
    for (int s : order) {
      if (t1[s] <= T) break;
      int pick = 0;
      double tp = t1[s];
      for (int k = 1; k <= 5 && (1 << k) <= gmax[s]; ++k) {
        const double tk = tmv[s] * cm.gang[k] + tov[s];
        if (tk < tp) {
          pick = 1 << k;
          tp = tk;
        }
        if (tk <= T) break;
      }
      if (pick == 0) continue;  // no gang is faster: stays a single workgroup
      G[s] = pick;
      tgang[s] = tp;
      used += pick;
      any = true;
    }
    if (!any || used > slots) continue;
    const double m = simulate_split(t1, tgang, G, order, slots);
    if (m < best_t) {
      best_t = m;
      best = G;
    }
  }
  return best_t;
}

// Modelled one-workgroup times (t1), off-chip-priced times (tg1) and gang
// limits of an sle batch, then the split; returns the modelled makespan.
double plan_batch(const CostModel& cm, long arena, int mvar, int count, const ttals_tt* A, const ttals_tt* x0,
                  int repeats, int slots, std::vector<int>& G, std::vector<double>& t1) {
  std::vector<double> tmv(count), tov(count);
  std::vector<int> gmax(count);
  t1.assign(count, 0.0);
  for (int s = 0; s < count; ++s) {
    double on, off;
    int ns;
    sle_cost(A[s], x0[s], arena, mvar, &on, &off, &ns);
    t1[s] = repeats * (cm.on * on + cm.off * off + cm.solve * ns);
    tmv[s] = repeats * cm.off * (on + off);
    tov[s] = repeats * cm.solve * ns;
    gmax[s] = gang_auto_size(x0[s]);
  }
  return plan_split(cm, t1, tmv, tov, gmax, slots, G);
}

// One k_hybrid launch for an sle batch: gangs for the systems the cost model
// says set the launch time, one workgroup each for the rest (largest first).
// Falls back to the single-workgroup launch when no gang pays off.
int enqueue_hybrid(ttals_ctx* c, hipStream_t stream, const BatchArgs& B, size_t* r_res, size_t* res_stride_out) {
  const int count = B.count;
  if (count < 1 || !B.A || !B.x0 || !B.xo || !B.b) return fail(c, TTALS_E_ARG, "bad batch arguments");
  if (B.repeats < 0) return fail(c, TTALS_E_ARG, "repeats must be >= 0");
  if (!(B.tol > 0.0)) return fail(c, TTALS_E_ARG, "tol must be > 0");
  if (B.maxit < 1) return fail(c, TTALS_E_ARG, "maxit must be >= 1");
  HIPCHK(c, hipSetDevice(c->device));
  long arena = 0;
  int per_cu = 0, cus = 0, rc;
  if ((rc = hybrid_arena(c, &arena))) return rc;
  HIPCHK(c, hybrid_occupancy(&per_cu, (size_t)arena * 8));
  HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  const int slots = per_cu * cus;
  if (slots < 1) return fail(c, TTALS_E_HIP, "k_hybrid cannot be resident on this device");
  // pass 0: checks, then the split by the launch-time model
  SysSizes z;
  for (int s = 0; s < count; ++s)
    if ((rc = size_system(c, 0, B.A[s], B.x0[s], &B.b[s], B.xo[s], s, z))) return rc;
  std::vector<int> G;
  std::vector<double> t1;
  c->last_plan_ms = plan_batch(c->cost, arena, c->matvec_variant, count, B.A, B.x0, B.repeats, slots, G, t1);
  int ng = 0, wg = 0;
  for (int s = 0; s < count; ++s)
    if (G[s] > 0) {
      ++ng;
      wg += G[s];
    }
  if (ng == 0) {
    Placement pl{nullptr, nullptr, 0, 0};
    rc = enqueue_batch(c, stream, B, pl, r_res, res_stride_out, true);
    if (!rc) c->last_plan = G;
    return rc;
  }
  // gangs first, then single workgroups, each largest first
  std::vector<int> gs, fs;
  for (int s = 0; s < count; ++s) (G[s] > 0 ? gs : fs).push_back(s);
  auto by_time = [&](int x, int y) { return t1[x] > t1[y]; };
  std::stable_sort(gs.begin(), gs.end(), by_time);
  std::stable_sort(fs.begin(), fs.end(), by_time);
  const int nf = (int)fs.size();
  // pass 1: layout
  Layout hl, dl;
  const size_t m_gdesc = hl.take(sizeof(GangDesc) * ng);
  const size_t m_blocks = hl.take(sizeof(int) * wg);
  const size_t m_sdesc = hl.take(sizeof(SysDesc) * std::max(nf, 1));
  int dmax = 0;
  for (int s = 0; s < count; ++s) dmax = std::max(dmax, B.A[s].order);
  const size_t res_stride = RES_HEAD + (size_t)dmax + 1;
  const size_t w_sync = dl.take(sizeof(GangSync) * ng);
  const size_t w_res = dl.take(8 * res_stride * count);
  GOff go;
  SysOff o;
  auto lay = [&](Layout& h, Layout& d, int s) {
    (void)size_system(c, 0, B.A[s], B.x0[s], &B.b[s], B.xo[s], s, z);
    if (G[s] > 0)
      gang_offsets(z, B.A[s], B.x0[s], G[s], c->gang_prec == 1, h, d, go);
    else
      take_offsets(0, z, h, d, o);
  };
  for (int s : gs) lay(hl, dl, s);
  for (int s : fs) lay(hl, dl, s);
  const size_t meta_bytes = hl.bytes;
  Layout rl;
  rl.bytes = meta_bytes;
  *r_res = rl.take(8 * res_stride * count);
  *res_stride_out = res_stride;
  if ((rc = ensure_hst(c, rl.bytes))) return rc;
  if ((rc = ensure_dws(c, dl.bytes + meta_bytes))) return rc;
  char* hbase = c->hst;
  char* dbase = reinterpret_cast<char*>(c->dws);
  char* dmeta = dbase + dl.bytes;
  GangDesc* hg = reinterpret_cast<GangDesc*>(hbase + m_gdesc);
  int* hblocks = reinterpret_cast<int*>(hbase + m_blocks);
  SysDesc* hs = reinterpret_cast<SysDesc*>(hbase + m_sdesc);
  GangSync* syncs = reinterpret_cast<GangSync*>(dbase + w_sync);
  double* dres = reinterpret_cast<double*>(dbase + w_res);
  // pass 2: the same offsets again, descriptors written
  Layout hl2, dl2;
  (void)hl2.take(sizeof(GangDesc) * ng);
  (void)hl2.take(sizeof(int) * wg);
  (void)hl2.take(sizeof(SysDesc) * std::max(nf, 1));
  (void)dl2.take(sizeof(GangSync) * ng);
  (void)dl2.take(8 * res_stride * count);
  const GangParams P{B.solver, B.repeats, B.maxit, c->gang_prec, B.tol, c->budget};
  int first = 0;
  for (int k = 0; k < ng; ++k) {
    const int s = gs[k];
    lay(hl2, dl2, s);
    gang_fill(go, B.A[s], B.x0[s], B.b[s], B.xo[s], G[s], first, P, hbase, dmeta, dbase, syncs + k,
              dres + res_stride * s, hg[k]);
    for (int m = 0; m < G[s]; ++m) hblocks[first + m] = k;
    first += G[s];
  }
  for (int k = 0; k < nf; ++k) {
    const int s = fs[k];
    lay(hl2, dl2, s);
    fill_sys(c, B, s, o, hbase, dmeta, dbase, dres + res_stride * s, hs[k]);
  }
  HIPCHK(c, hipMemcpyAsync(dmeta, hbase, meta_bytes, hipMemcpyHostToDevice, stream));
  HIPCHK(c, hipMemsetAsync(syncs, 0, sizeof(GangSync) * ng, stream));
  if (!c->k0) HIPCHK(c, hipEventCreate(&c->k0));
  if (!c->k1) HIPCHK(c, hipEventCreate(&c->k1));
  {
    std::lock_guard<std::mutex> lk(g_gang_mu);
    if ((rc = gang_serialize_begin(c, stream))) return rc;
    HIPCHK(c, hipEventRecord(c->k0, stream));
    HIPCHK(c, hybrid_launch(wg + nf, (size_t)arena * 8, stream, (const GangDesc*)(dmeta + m_gdesc),
                            (const int*)(dmeta + m_blocks), wg, (const SysDesc*)(dmeta + m_sdesc), arena));
    HIPCHK(c, hipEventRecord(c->k1, stream));
    if ((rc = gang_serialize_end(c, stream))) return rc;
  }
  c->timed = true;
  HIPCHK(c, hipMemcpyAsync(hbase + *r_res, dres, 8 * res_stride * count, hipMemcpyDeviceToHost, stream));
  c->last_plan = G;
  return 0;
}

int gang_solve(ttals_ctx* c, hipStream_t stream, int count, const ttals_tt* A, const ttals_tt* x0, const ttals_tt* b,
               const ttals_tt* xo, int repeats, int solver, double tol, int maxit, int prec, int gang,
               ttals_result* res) {
  if (c->inflight) return fail(c, TTALS_E_ARG, "a submission is still pending on this context (call ttals_wait)");
  if (count < 1 || !A || !x0 || !b || !xo || !res) return fail(c, TTALS_E_ARG, "bad batch arguments");
  if (repeats < 0 || !(tol > 0.0) || maxit < 1) return fail(c, TTALS_E_ARG, "bad repeats/tol/maxit");
  if (prec != 0 && prec != 1) return fail(c, TTALS_E_ARG, "precision must be 0 (FP64) or 1 (FP32 + FP64 refinement)");
  if (gang > 256) return fail(c, TTALS_E_ARG, "gang size must be <= 256");
  HIPCHK(c, hipSetDevice(c->device));
  int cap = 0, rc;
  long arena = 0;  // GEMM tiles, then the split QR's factors
  if ((rc = hybrid_arena(c, &arena))) return rc;
  if ((rc = gang_capacity(c, (size_t)arena * 8, &cap))) return rc;
  // pass 1: checks, gang sizes, totals
  Layout hl, dl;
  const size_t m_descs = hl.take(sizeof(GangDesc) * count);
  int total = 0, dmax = 0;
  SysSizes z;
  for (int s = 0; s < count; ++s) {
    if ((rc = size_system(c, 0, A[s], x0[s], &b[s], xo[s], s, z))) return rc;
    dmax = std::max(dmax, z.d);
    total += gang > 0 ? gang : gang_auto_size(x0[s]);
  }
  if (total > cap)
    return fail(c, TTALS_E_ARG, "the gangs need " + std::to_string(total) + " resident workgroups; the device holds " +
                                    std::to_string(cap));
  const size_t m_blocks = hl.take(sizeof(int) * total);
  const size_t res_stride = RES_HEAD + (size_t)dmax + 1;
  const size_t w_sync = dl.take(sizeof(GangSync) * count);
  const size_t w_res = dl.take(8 * res_stride * count);
  GOff go;
  for (int s = 0; s < count; ++s) {
    (void)size_system(c, 0, A[s], x0[s], &b[s], xo[s], s, z);
    gang_offsets(z, A[s], x0[s], gang > 0 ? gang : gang_auto_size(x0[s]), prec == 1, hl, dl, go);
  }
  const size_t meta_bytes = hl.bytes;
  Layout rl;
  rl.bytes = meta_bytes;
  const size_t r_res = rl.take(8 * res_stride * count);
  if ((rc = ensure_hst(c, rl.bytes))) return rc;
  if ((rc = ensure_dws(c, dl.bytes + meta_bytes))) return rc;
  char* hbase = c->hst;
  char* dbase = reinterpret_cast<char*>(c->dws);
  char* dmeta = dbase + dl.bytes;
  GangDesc* hdesc = reinterpret_cast<GangDesc*>(hbase + m_descs);
  int* hblocks = reinterpret_cast<int*>(hbase + m_blocks);
  GangSync* syncs = reinterpret_cast<GangSync*>(dbase + w_sync);
  double* dres = reinterpret_cast<double*>(dbase + w_res);
  // pass 2: descriptors (same offsets again)
  Layout hl2, dl2;
  (void)hl2.take(sizeof(GangDesc) * count);
  (void)hl2.take(sizeof(int) * total);
  (void)dl2.take(sizeof(GangSync) * count);
  (void)dl2.take(8 * res_stride * count);
  const GangParams P{solver, repeats, maxit, prec, tol, c->budget};
  int first = 0;
  for (int s = 0; s < count; ++s) {
    (void)size_system(c, 0, A[s], x0[s], &b[s], xo[s], s, z);
    const int G = gang > 0 ? gang : gang_auto_size(x0[s]);
    gang_offsets(z, A[s], x0[s], G, prec == 1, hl2, dl2, go);
    gang_fill(go, A[s], x0[s], b[s], xo[s], G, first, P, hbase, dmeta, dbase, syncs + s, dres + res_stride * s,
              hdesc[s]);
    for (int k = 0; k < G; ++k) hblocks[first + k] = s;
    first += G;
  }
  HIPCHK(c, hipMemcpyAsync(dmeta, hbase, meta_bytes, hipMemcpyHostToDevice, stream));
  HIPCHK(c, hipMemsetAsync(syncs, 0, sizeof(GangSync) * count, stream));
  if (!c->k0) HIPCHK(c, hipEventCreate(&c->k0));
  if (!c->k1) HIPCHK(c, hipEventCreate(&c->k1));
  // every workgroup a gang member: k_hybrid with no single-workgroup part
  {
    std::lock_guard<std::mutex> lk(g_gang_mu);
    if ((rc = gang_serialize_begin(c, stream))) return rc;
    HIPCHK(c, hipEventRecord(c->k0, stream));
    HIPCHK(c, hybrid_launch(total, (size_t)arena * 8, stream, (const GangDesc*)(dmeta + m_descs),
                            (const int*)(dmeta + m_blocks), total, nullptr, arena));
    HIPCHK(c, hipEventRecord(c->k1, stream));
    if ((rc = gang_serialize_end(c, stream))) return rc;
  }
  c->timed = true;
  HIPCHK(c, hipMemcpyAsync(hbase + r_res, dres, 8 * res_stride * count, hipMemcpyDeviceToHost, stream));
  HIPCHK(c, hipStreamSynchronize(stream));
  int worst = 0;
  for (int s = 0; s < count; ++s) {
    const int st = unpack_result(reinterpret_cast<const double*>(hbase + r_res + 8 * res_stride * s), A[s].order,
                                 xo[s].ranks, res[s], c->wall_khz);
    if (st == ST_ABORT) return fail(c, TTALS_E_HIP, "gang barrier timed out (system " + std::to_string(s) + ")");
    worst = worst_status(worst, st);
  }
  return worst;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int ttals_version(void) { return TTALS_VERSION; }

int ttals_ctx_create(int device, ttals_ctx** out) {
  if (!out) return TTALS_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return TTALS_E_ARG;
  if (device >= TTALS_MAX_DEVICES) return TTALS_E_ARG;
  {
    // the static scratch's address on this device, once per device, with
    // the device current (the caller's current device is restored)
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (!g_scratch_dev[device]) {
      int prev = 0;
      if (hipGetDevice(&prev) != hipSuccess) return TTALS_E_HIP;
      if (hipSetDevice(device) != hipSuccess) return TTALS_E_HIP;
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      (void)hipSetDevice(prev);
      if (e != hipSuccess || !p) return TTALS_E_HIP;
      g_scratch_dev[device] = p;
    }
  }
  ttals_ctx* c = new ttals_ctx();
  c->device = device;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
    c->wall_khz = (double)khz;
  if (const char* e = std::getenv("TTALS_ARENA_KB")) {
    const long kb = std::atol(e);
    if (kb >= 0 && kb <= 61) c->arena = kb * 1024 / 8;
  }
  *out = c;
  return 0;
}

void ttals_ctx_destroy(ttals_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->done) {
    (void)hipEventSynchronize(c->done);
    (void)hipEventDestroy(c->done);
  }
  if (c->k0) (void)hipEventDestroy(c->k0);
  if (c->k1) (void)hipEventDestroy(c->k1);
  if (c->dws) (void)hipFree(c->dws);
  if (c->hst) (void)hipHostFree(c->hst);
  delete c;
}

const char* ttals_last_error(ttals_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ttals_sle_submit(ttals_ctx* c, void* stream, int count, const ttals_tt* A, const ttals_tt* x0, const ttals_tt* b,
                     const ttals_tt* x_out, int repeats, int solver, double tol, int maxit) {
  if (!c) return TTALS_E_ARG;
  if (solver != 0 && solver != 1) return fail(c, TTALS_E_UNSUPPORTED, "sle solver must be 'solve' or 'lu'");
  return submit_batch(c, (hipStream_t)stream, 0, count, A, x0, b, x_out, repeats, solver, tol, maxit, 0.0, 0.0);
}

int ttals_evp_submit(ttals_ctx* c, void* stream, int count, const ttals_tt* A, const ttals_tt* x0,
                     const ttals_tt* x_out, int repeats, double conv_eps, int solver, double sigma, double tol,
                     int maxit) {
  if (!c) return TTALS_E_ARG;
  if (solver != 0 && solver != 1) return fail(c, TTALS_E_UNSUPPORTED, "evp solver must be 'eig' or 'eigh'");
  return submit_batch(c, (hipStream_t)stream, 1, count, A, x0, nullptr, x_out, repeats, solver, tol, maxit, sigma,
                      conv_eps);
}

int ttals_query(ttals_ctx* c) {
  if (!c) return TTALS_E_ARG;
  if (!c->inflight) return 1;
  const hipError_t e = hipEventQuery(c->done);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return hip_fail(c, e, "hipEventQuery");
}

int ttals_wait(ttals_ctx* c, ttals_result* res) {
  if (!c || !res) return TTALS_E_ARG;
  return finish_batch(c, res);
}

int ttals_set_arena_kb(ttals_ctx* c, int kb) {
  if (!c || kb < 0 || kb > 61) return TTALS_E_ARG;
  if (c->inflight) return fail(c, TTALS_E_ARG, "a submission is pending on this context");
  c->arena = (long)kb * 1024 / 8;
  c->arena_set = true;
  c->arena_hybrid = 0;
  return 0;
}

int ttals_set_local_tol_policy(ttals_ctx* c, int policy) {
  if (!c || policy < 0 || policy > 15) return TTALS_E_ARG;
  if (c->inflight) return fail(c, TTALS_E_ARG, "a submission is pending on this context");
  c->tol_policy = policy;
  return 0;
}

int ttals_set_matvec_variant(ttals_ctx* c, int variant) {
  if (!c || variant < 0 || variant > 1) return TTALS_E_ARG;
  if (c->inflight) return fail(c, TTALS_E_ARG, "a submission is pending on this context");
  c->matvec_variant = variant;
  return 0;
}

int ttals_set_schedule(ttals_ctx* c, int schedule, int gang_precision) {
  if (!c) return TTALS_E_ARG;
  if (schedule != 0 && schedule != 1) return fail(c, TTALS_E_ARG, "schedule must be 0 or 1");
  if (gang_precision != 0 && gang_precision != 1) return fail(c, TTALS_E_ARG, "gang precision must be 0 or 1");
  c->schedule = schedule;
  c->gang_prec = gang_precision;
  return 0;
}

int ttals_set_time_budget(ttals_ctx* c, double ms) {
  if (!c) return TTALS_E_ARG;
  if (!(ms >= 0.0) || !std::isfinite(ms)) return fail(c, TTALS_E_ARG, "the time budget must be finite and >= 0");
  if (ms == 0.0) {
    c->budget = 0;
    return 0;
  }
  int khz = 0;
  HIPCHK(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
  if (khz <= 0) return fail(c, TTALS_E_HIP, "the device reports no constant clock rate");
  c->budget = std::max<long long>(1, (long long)(ms * khz));
  return 0;
}

int ttals_set_cost_model(ttals_ctx* c, const double* p, int n) {
  if (!c) return TTALS_E_ARG;
  if (!p || n != 9) return fail(c, TTALS_E_ARG, "the cost model has 9 parameters");
  for (int k = 0; k < 9; ++k)
    if (!(p[k] > 0.0) || !std::isfinite(p[k])) return fail(c, TTALS_E_ARG, "cost model parameters must be finite and > 0");
  c->cost.on = p[0];
  c->cost.off = p[1];
  c->cost.solve = p[2];
  for (int k = 0; k < 6; ++k) c->cost.gang[k] = p[3 + k];
  return 0;
}

int ttals_get_cost_model(ttals_ctx* c, double* p, int n) {
  if (!c) return TTALS_E_ARG;
  if (!p || n != 9) return fail(c, TTALS_E_ARG, "the cost model has 9 parameters");
  p[0] = c->cost.on;
  p[1] = c->cost.off;
  p[2] = c->cost.solve;
  for (int k = 0; k < 6; ++k) p[3 + k] = c->cost.gang[k];
  return 0;
}

int ttals_sle_cost(ttals_ctx* c, int count, const ttals_tt* A, const ttals_tt* x0, double* out) {
  if (count < 1 || !A || !x0 || !out) return c ? fail(c, TTALS_E_ARG, "bad arguments") : TTALS_E_ARG;
  const long arena = c ? std::max<long>(c->arena, GEMM_LDS_DOUBLES) : 61 * 1024 / 8;
  const int mvar = c ? c->matvec_variant : 0;
  for (int s = 0; s < count; ++s) {
    int rc;
    if ((rc = check_tt(c, A[s], false, "operator")) || (rc = check_tt(c, x0[s], true, "initial_guess"))) return rc;
    if (x0[s].order != A[s].order) return c ? fail(c, TTALS_E_SHAPE, "orders of the TTs differ") : TTALS_E_SHAPE;
    int ns = 0;
    sle_cost(A[s], x0[s], arena, mvar, &out[3 * s], &out[3 * s + 1], &ns);
    out[3 * s + 2] = ns;
  }
  return 0;
}

int ttals_sle_plan(ttals_ctx* c, int count, const ttals_tt* A, const ttals_tt* x0, int repeats, int slots,
                   int* gangs, double* makespan_ms) {
  if (count < 1 || !A || !x0 || !gangs || repeats < 0 || slots < 1)
    return c ? fail(c, TTALS_E_ARG, "bad arguments") : TTALS_E_ARG;
  for (int s = 0; s < count; ++s) {
    int rc;
    if ((rc = check_tt(c, A[s], false, "operator")) || (rc = check_tt(c, x0[s], true, "initial_guess"))) return rc;
    if (x0[s].order != A[s].order) return c ? fail(c, TTALS_E_SHAPE, "orders of the TTs differ") : TTALS_E_SHAPE;
  }
  const CostModel cm = c ? c->cost : CostModel();
  const long arena = c ? std::max<long>(c->arena, GEMM_LDS_DOUBLES) : 61 * 1024 / 8;
  std::vector<int> G;
  std::vector<double> t1;
  const double m = plan_batch(cm, arena, c ? c->matvec_variant : 0, count, A, x0, repeats, slots, G, t1);
  for (int s = 0; s < count; ++s) gangs[s] = G[s];
  if (makespan_ms) *makespan_ms = m;
  return 0;
}

int ttals_last_plan_ms(ttals_ctx* c, double* ms) {
  if (!c || !ms) return c ? fail(c, TTALS_E_ARG, "null argument") : TTALS_E_ARG;
  *ms = c->last_plan_ms;
  return 0;
}

int ttals_last_plan(ttals_ctx* c, int* gangs, int count) {
  if (!c) return TTALS_E_ARG;
  if (!gangs || count != (int)c->last_plan.size())
    return fail(c, TTALS_E_ARG, "count differs from the last sle batch (" + std::to_string(c->last_plan.size()) + ")");
  for (int s = 0; s < count; ++s) gangs[s] = c->last_plan[s];
  return 0;
}

int ttals_last_launch_ms(ttals_ctx* c, double* ms) {
  if (!c || !ms) return TTALS_E_ARG;
  if (c->inflight || !c->timed) return fail(c, TTALS_E_ARG, "no finished submission to time");
  float t = 0.0f;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipEventElapsedTime(&t, c->k0, c->k1));
  *ms = (double)t;
  return 0;
}

int ttals_sle_als_batch(ttals_ctx* c, void* stream, int count, const ttals_tt* A, const ttals_tt* x0,
                        const ttals_tt* b, const ttals_tt* x_out, int repeats, int solver, double tol, int maxit,
                        ttals_result* res) {
  if (!res) return c ? fail(c, TTALS_E_ARG, "null result array") : TTALS_E_ARG;
  int rc = ttals_sle_submit(c, stream, count, A, x0, b, x_out, repeats, solver, tol, maxit);
  return rc ? rc : finish_batch(c, res);
}

int ttals_sle_als_batch_static(ttals_ctx* c, void* stream, int count, const ttals_tt* A, const ttals_tt* x0,
                               const ttals_tt* b, const ttals_tt* x_out, int repeats, int solver, double tol, int maxit,
                               ttals_result* res) {
  if (!c) return TTALS_E_ARG;
  if (!res) return fail(c, TTALS_E_ARG, "null result array");
  if (solver != 0 && solver != 1) return fail(c, TTALS_E_UNSUPPORTED, "sle solver must be 'solve' or 'lu'");
  if (c->inflight) return fail(c, TTALS_E_ARG, "a submission is still pending on this context (call ttals_wait)");
  void* dev = g_scratch_dev[c->device];
  if (!dev) return fail(c, TTALS_E_HIP, "static scratch address not resolved for this device");
  // one call at a time: calls share the device scratch and the host staging
  std::lock_guard<std::mutex> lk(g_static_mu);
  BatchArgs B{0, count, repeats, solver, maxit, tol, 0.0, 0.0, A, x0, b, x_out};
  Placement pl{g_static_host, reinterpret_cast<char*>(dev), TTALS_STATIC_HOST_BYTES, TTALS_STATIC_DEV_BYTES};
  size_t r_res = 0, stride = 0;
  int rc = enqueue_batch(c, (hipStream_t)stream, B, pl, &r_res, &stride, false);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize((hipStream_t)stream));
  int worst = 0;
  for (int s = 0; s < count; ++s)
    worst = worst_status(worst, unpack_result(reinterpret_cast<const double*>(g_static_host + r_res + 8 * stride * s),
                                              A[s].order, x_out[s].ranks, res[s], c->wall_khz));
  return worst;
}

int ttals_sle_als_gang(ttals_ctx* c, void* stream, int count, const ttals_tt* A, const ttals_tt* x0,
                       const ttals_tt* b, const ttals_tt* x_out, int repeats, int solver, double tol, int maxit,
                       int precision, int gang, ttals_result* res) {
  if (!c) return TTALS_E_ARG;
  if (solver != 0 && solver != 1) return fail(c, TTALS_E_UNSUPPORTED, "sle solver must be 'solve' or 'lu'");
  return gang_solve(c, (hipStream_t)stream, count, A, x0, b, x_out, repeats, solver, tol, maxit, precision, gang,
                    res);
}

int ttals_evp_als_batch(ttals_ctx* c, void* stream, int count, const ttals_tt* A, const ttals_tt* x0,
                        const ttals_tt* x_out, int repeats, double conv_eps, int solver, double sigma, double tol,
                        int maxit, ttals_result* res) {
  if (!res) return c ? fail(c, TTALS_E_ARG, "null result array") : TTALS_E_ARG;
  int rc = ttals_evp_submit(c, stream, count, A, x0, x_out, repeats, conv_eps, solver, sigma, tol, maxit);
  return rc ? rc : finish_batch(c, res);
}

#define PIECE_PROLOGUE(tsize, Nsize)                                  \
  if (!c) return TTALS_E_ARG;                                         \
  double *t1, *t2, *V, *res;                                          \
  int rc = piece_scratch(c, (tsize), (Nsize), &t1, &t2, &V, &res);    \
  if (rc) return rc;                                                  \
  PieceArgs pa;                                                       \
  std::memset(&pa, 0, sizeof(pa));                                    \
  pa.t1 = t1;                                                         \
  pa.t2 = t2;                                                         \
  pa.V = V;                                                           \
  pa.res = res;

int ttals_stack_left_op(ttals_ctx* c, void* stream, int a, int P, int n, int Q, int b, const double* L,
                        const double* Acore, const double* x, double* Lout) {
  if (!dims_ok({a, P, n, Q, b})) return fail(c, TTALS_E_ARG, "bad dims");
  PIECE_PROLOGUE((long)a * n * b * std::max(P, Q), 1)
  pa.op = P_LEFT_OP;
  pa.a = a; pa.P = P; pa.n = n; pa.Q = Q; pa.b = b;
  pa.L = L; pa.A = Acore; pa.x = x; pa.out = Lout;
  return launch_piece(c, (hipStream_t)stream, pa);
}

int ttals_stack_right_op(ttals_ctx* c, void* stream, int a, int P, int n, int Q, int b, const double* R,
                         const double* Acore, const double* x, double* Rout) {
  if (!dims_ok({a, P, n, Q, b})) return fail(c, TTALS_E_ARG, "bad dims");
  PIECE_PROLOGUE((long)a * n * b * std::max(P, Q), 1)
  pa.op = P_RIGHT_OP;
  pa.a = a; pa.P = P; pa.n = n; pa.Q = Q; pa.b = b;
  pa.R = R; pa.A = Acore; pa.x = x; pa.out = Rout;
  return launch_piece(c, (hipStream_t)stream, pa);
}

int ttals_stack_left_rhs(ttals_ctx* c, void* stream, int a, int p, int n, int q, int b, const double* Lb,
                         const double* Bcore, const double* x, double* Lbout) {
  if (!dims_ok({a, p, n, q, b})) return fail(c, TTALS_E_ARG, "bad dims");
  PIECE_PROLOGUE((long)a * n * b * std::max(p, q), 1)
  pa.op = P_LEFT_RHS;
  pa.a = a; pa.P = p; pa.n = n; pa.Q = q; pa.b = b;
  pa.L = Lb; pa.A = Bcore; pa.x = x; pa.out = Lbout;
  return launch_piece(c, (hipStream_t)stream, pa);
}

int ttals_stack_right_rhs(ttals_ctx* c, void* stream, int a, int p, int n, int q, int b, const double* Rb,
                          const double* Bcore, const double* x, double* Rbout) {
  if (!dims_ok({a, p, n, q, b})) return fail(c, TTALS_E_ARG, "bad dims");
  PIECE_PROLOGUE((long)a * n * b * std::max(p, q), 1)
  pa.op = P_RIGHT_RHS;
  pa.a = a; pa.P = p; pa.n = n; pa.Q = q; pa.b = b;
  pa.R = Rb; pa.A = Bcore; pa.x = x; pa.out = Rbout;
  return launch_piece(c, (hipStream_t)stream, pa);
}

int ttals_micro_matrix(ttals_ctx* c, void* stream, int a, int P, int n, int Q, int b, const double* L,
                       const double* Acore, const double* R, double* M) {
  if (!dims_ok({a, P, n, Q, b})) return fail(c, TTALS_E_ARG, "bad dims");
  PIECE_PROLOGUE(1, 1)
  pa.op = P_MICRO_DENSE;
  pa.a = a; pa.P = P; pa.n = n; pa.Q = Q; pa.b = b;
  pa.L = L; pa.A = Acore; pa.R = R; pa.out = M;
  return launch_piece(c, (hipStream_t)stream, pa);
}

int ttals_micro_matvec(ttals_ctx* c, void* stream, int a, int P, int n, int Q, int b, const double* L,
                       const double* Acore, const double* R, const double* v, double* y) {
  if (!dims_ok({a, P, n, Q, b})) return fail(c, TTALS_E_ARG, "bad dims");
  PIECE_PROLOGUE((long)a * n * b * std::max(P, Q), 1)
  pa.op = P_MATVEC;
  pa.a = a; pa.P = P; pa.n = n; pa.Q = Q; pa.b = b;
  pa.L = L; pa.A = Acore; pa.R = R; pa.x = v; pa.out = y;
  return launch_piece(c, (hipStream_t)stream, pa);
}

int ttals_micro_rhs(ttals_ctx* c, void* stream, int a, int p, int n, int q, int b, const double* Lb,
                    const double* Bcore, const double* Rb, double* f) {
  if (!dims_ok({a, p, n, q, b})) return fail(c, TTALS_E_ARG, "bad dims");
  PIECE_PROLOGUE((long)a * n * std::max(p, q) * b, 1)
  pa.op = P_MICRO_RHS;
  pa.a = a; pa.P = p; pa.n = n; pa.Q = q; pa.b = b;
  pa.L = Lb; pa.A = Bcore; pa.R = Rb; pa.out = f;
  return launch_piece(c, (hipStream_t)stream, pa);
}

int ttals_local_solve(ttals_ctx* c, void* stream, int a, int P, int n, int Q, int b, const double* L,
                      const double* Acore, const double* R, const double* f, double* v, int method, int sym,
                      double tol, int maxit, int* iters) {
  if (!dims_ok({a, P, n, Q, b})) return fail(c, TTALS_E_ARG, "bad dims");
  const long N = (long)a * n * b;
  if (method < 0 || method > 3) return fail(c, TTALS_E_ARG, "bad method");
  if (method == 1 && N > NDENSE) return fail(c, TTALS_E_ARG, "dense LU limited to local size 96");
  if (!(tol > 0.0) || maxit < 1) return fail(c, TTALS_E_ARG, "bad tol/maxit");
  PIECE_PROLOGUE(N * std::max(P, Q), N)
  pa.op = P_SOLVE;
  pa.a = a; pa.P = P; pa.n = n; pa.Q = Q; pa.b = b;
  pa.L = L; pa.A = Acore; pa.R = R; pa.f = f; pa.v = v;
  pa.method = method; pa.sym = sym; pa.tol = tol; pa.maxit = maxit;
  if ((rc = launch_piece(c, (hipStream_t)stream, pa))) return rc;
  double h[2];
  HIPCHK(c, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream));
  HIPCHK(c, hipStreamSynchronize((hipStream_t)stream));
  if (iters) *iters = (int)h[1];
  const int r = (int)h[0];
  if (r == 0) return 0;
  if (method == 1 || (method == 0 && N <= NDENSE)) return ST_SINGULAR;
  return ST_NOCONV;
}

int ttals_local_eig(ttals_ctx* c, void* stream, int a, int P, int n, int Q, int b, const double* L,
                    const double* Acore, const double* R, double* v, int mode, double sigma, double tol, int maxit,
                    double* lambda, int* iters) {
  if (!dims_ok({a, P, n, Q, b})) return fail(c, TTALS_E_ARG, "bad dims");
  if (mode < 0 || mode > 2 || !(tol > 0.0) || maxit < 1) return fail(c, TTALS_E_ARG, "bad mode/tol/maxit");
  const long N = (long)a * n * b;
  PIECE_PROLOGUE(N * std::max(P, Q), N)
  pa.op = P_EIG;
  pa.a = a; pa.P = P; pa.n = n; pa.Q = Q; pa.b = b;
  pa.L = L; pa.A = Acore; pa.R = R; pa.v = v;
  pa.mode = mode; pa.sigma = sigma; pa.tol = tol; pa.maxit = maxit;
  if ((rc = launch_piece(c, (hipStream_t)stream, pa))) return rc;
  double h[3];
  HIPCHK(c, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream));
  HIPCHK(c, hipStreamSynchronize((hipStream_t)stream));
  if (iters) *iters = (int)h[1];
  if (lambda) *lambda = h[2];
  const int r = (int)h[0];
  return r == 0 ? 0 : (r == 3 ? ST_UNSUPPORTED : ST_NOCONV);
}

int ttals_qr(ttals_ctx* c, void* stream, int m, int cdim, const double* Y, double* Q, double* R) {
  if (!dims_ok({m, cdim})) return fail(c, TTALS_E_ARG, "bad dims");
  PIECE_PROLOGUE((long)m * cdim + std::max(m, cdim) + ((long)cdim + 16) * 16, 1)
  pa.op = P_QR;
  pa.a = m; pa.b = cdim;
  pa.x = Y; pa.out = Q; pa.v = R;
  pa.t2 = t1 + (long)m * cdim;  // tau after W
  return launch_piece(c, (hipStream_t)stream, pa);
}

int ttals_core_product(void* stream, int r1, int m, int n, int r2, int s1, int p, int s2, const double* A,
                       const double* X, double* out) {
  if (!dims_ok({r1, m, n, r2, s1, p, s2}) || !A || !X || !out) return TTALS_E_ARG;
  const long total = (long)r1 * s1 * m * p * r2 * s2;
  const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_core_product, dim3(blocks), dim3(256), 0, (hipStream_t)stream, r1, m, n, r2, s1, p, s2, A,
                     X, out);
  return hipGetLastError() == hipSuccess ? 0 : TTALS_E_HIP;
}

int ttals_core_embed(void* stream, int r0, int mn, int r1, const double* src, double scale, int R0, int R1, int o0,
                     int o1, double* dst) {
  if (!dims_ok({r0, mn, r1, R0, R1}) || o0 < 0 || o1 < 0 || o0 + r0 > R0 || o1 + r1 > R1 || !src || !dst)
    return TTALS_E_ARG;
  const long total = (long)r0 * mn * r1;
  const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_core_embed, dim3(blocks), dim3(256), 0, (hipStream_t)stream, r0, mn, r1, src, scale, R0, R1,
                     o0, o1, dst);
  return hipGetLastError() == hipSuccess ? 0 : TTALS_E_HIP;
}

int ttals_tt_norm(ttals_ctx* c, void* stream, const ttals_tt* t, double* out) {
  if (!c || !t || !out) return TTALS_E_ARG;
  if (c->inflight) return fail(c, TTALS_E_ARG, "workspace busy: a submission is pending on this context");
  int rc = check_tt(c, *t, false, "tt");
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  const int d = t->order;
  hipStream_t st = (hipStream_t)stream;
  // workspace: copies of the cores + W + Rc + T + tau + out + metadata
  Layout L;
  std::vector<size_t> coff(d);
  long wmax = 1, rmax = 1, tmax = 1;
  for (int i = 0; i < d; ++i) {
    const long mn = (long)t->row_dims[i] * t->col_dims[i];
    const long sz = (long)t->ranks[i] * mn * t->ranks[i + 1];
    coff[i] = L.take(8 * sz);
    wmax = std::max(wmax, sz);
    rmax = std::max(rmax, (long)std::max(t->ranks[i], t->ranks[i + 1]));
    tmax = std::max(tmax, sz);
  }
  const size_t oW = L.take(8 * wmax), oR = L.take(8 * rmax * rmax), oT = L.take(8 * tmax), otau = L.take(8 * rmax),
               oout = L.take(64);
  const size_t omn = L.take(sizeof(int) * d), orr = L.take(sizeof(int) * (d + 1)), optr = L.take(sizeof(void*) * d),
               osrc = L.take(sizeof(void*) * d);
  if ((rc = ensure_dws(c, L.bytes))) return rc;
  if ((rc = ensure_hst(c, L.bytes))) return rc;
  char* db = reinterpret_cast<char*>(c->dws);
  int* hmn = reinterpret_cast<int*>(c->hst + omn);
  int* hrr = reinterpret_cast<int*>(c->hst + orr);
  double** hp = reinterpret_cast<double**>(c->hst + optr);
  const double** hs = reinterpret_cast<const double**>(c->hst + osrc);
  for (int i = 0; i < d; ++i) {
    hmn[i] = t->row_dims[i] * t->col_dims[i];
    hp[i] = reinterpret_cast<double*>(db + coff[i]);
    hs[i] = t->cores[i];
  }
  for (int i = 0; i <= d; ++i) hrr[i] = t->ranks[i];
  HIPCHK(c, hipMemcpyAsync(db + omn, c->hst + omn, L.bytes - omn, hipMemcpyHostToDevice, st));
  NormArgs na;
  na.d = d;
  na.mn = reinterpret_cast<const int*>(db + omn);
  na.r = reinterpret_cast<const int*>(db + orr);
  na.c = reinterpret_cast<double* const*>(db + optr);
  na.src = reinterpret_cast<const double* const*>(db + osrc);
  na.W = reinterpret_cast<double*>(db + oW);
  na.Rc = reinterpret_cast<double*>(db + oR);
  na.T = reinterpret_cast<double*>(db + oT);
  na.tau = reinterpret_cast<double*>(db + otau);
  na.out = reinterpret_cast<double*>(db + oout);
  hipLaunchKernelGGL(k_tt_norm, dim3(1), dim3(NT), 0, st, na);
  HIPCHK(c, hipGetLastError());
  double h = 0.0;
  HIPCHK(c, hipMemcpyAsync(&h, na.out, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  *out = h;
  return 0;
}

int ttals_residual_norm_batch(ttals_ctx* c, void* stream, int count, const ttals_tt* A, const ttals_tt* x,
                              const ttals_tt* b, double* res_norm, double* rhs_norm) {
  if (!c) return TTALS_E_ARG;
  if (count < 1 || !A || !x || !b || !res_norm || !rhs_norm) return fail(c, TTALS_E_ARG, "bad batch arguments");
  if (c->inflight) return fail(c, TTALS_E_ARG, "workspace busy: a submission is pending on this context");
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  // plans: shapes checked like a solve (square operator, matching dims, outer ranks 1)
  struct RPlan {
    SysPlan sp;
    size_t m_n, m_ra, m_rx, m_rb, m_ac, m_xc, m_bc, w_W, w_U, w_R, w_tau, w_out;
  };
  std::vector<RPlan> plans(count);
  Layout hl, dl;
  const size_t m_descs = hl.take(sizeof(ResidDesc) * count);
  for (int s = 0; s < count; ++s) {
    RPlan& R = plans[s];
    int rc = plan_system(c, 0, A[s], x[s], &b[s], R.sp);
    if (rc) return fail(c, rc, "system " + std::to_string(s) + ": " + c->err);
    const SysPlan& P = R.sp;
    const int d = P.d;
    // residual ranks rho_i and the QR ranks k_i (k_d = 1, k_i = min(n_i k_{i+1}, rho_i))
    std::vector<long> rho(d + 1), k(d + 1);
    for (int i = 0; i <= d; ++i) rho[i] = (i == 0 || i == d) ? 1 : (long)P.ra[i] * P.rx[i] + P.rb[i];
    k[d] = 1;
    long wmax = 1, umax = 1, rmax = 1;
    for (int i = d - 1; i >= 0; --i) {
      const long mk = (long)P.n[i] * k[i + 1];
      wmax = std::max(wmax, mk * std::max(rho[i], (long)P.rb[i]));
      umax = std::max(umax, std::max((long)P.rx[i] * P.n[i] * P.ra[i + 1] * k[i + 1], mk));
      k[i] = std::min(mk, rho[i]);
      rmax = std::max(rmax, rho[i]);
    }
    rmax = std::max(rmax, rho[d]);
    R.m_n = hl.take(sizeof(int) * d);
    R.m_ra = hl.take(sizeof(int) * (d + 1));
    R.m_rx = hl.take(sizeof(int) * (d + 1));
    R.m_rb = hl.take(sizeof(int) * (d + 1));
    R.m_ac = hl.take(sizeof(void*) * d);
    R.m_xc = hl.take(sizeof(void*) * d);
    R.m_bc = hl.take(sizeof(void*) * d);
    R.w_W = dl.take(8 * (size_t)wmax);
    R.w_U = dl.take(8 * (size_t)umax);
    R.w_R = dl.take(8 * (size_t)rmax * rmax);
    R.w_tau = dl.take(8 * (size_t)rmax);
    R.w_out = dl.take(8 * 2);
  }
  const size_t meta_bytes = hl.bytes;
  Layout rl;
  rl.bytes = meta_bytes;
  const size_t r_out = rl.take(8 * 2 * (size_t)count);
  int rc;
  if ((rc = ensure_hst(c, rl.bytes))) return rc;
  if ((rc = ensure_dws(c, dl.bytes + meta_bytes + 8 * 2 * (size_t)count))) return rc;
  char* dbase = reinterpret_cast<char*>(c->dws);
  char* dmeta = dbase + dl.bytes;
  double* dres = reinterpret_cast<double*>(dmeta + meta_bytes);  // compact results
  ResidDesc* hd = hptr<ResidDesc>(c, m_descs);
  for (int s = 0; s < count; ++s) {
    const RPlan& R = plans[s];
    const SysPlan& P = R.sp;
    const int d = P.d;
    std::memcpy(hptr<int>(c, R.m_n), P.n.data(), sizeof(int) * d);
    std::memcpy(hptr<int>(c, R.m_ra), P.ra.data(), sizeof(int) * (d + 1));
    std::memcpy(hptr<int>(c, R.m_rx), P.rx.data(), sizeof(int) * (d + 1));
    std::memcpy(hptr<int>(c, R.m_rb), P.rb.data(), sizeof(int) * (d + 1));
    const double** ap = hptr<const double*>(c, R.m_ac);
    const double** xp = hptr<const double*>(c, R.m_xc);
    const double** bp = hptr<const double*>(c, R.m_bc);
    for (int i = 0; i < d; ++i) {
      ap[i] = A[s].cores[i];
      xp[i] = x[s].cores[i];
      bp[i] = b[s].cores[i];
    }
    ResidDesc& D = hd[s];
    std::memset(&D, 0, sizeof(D));
    D.d = d;
    D.n = reinterpret_cast<const int*>(dmeta + R.m_n);
    D.ra = reinterpret_cast<const int*>(dmeta + R.m_ra);
    D.rx = reinterpret_cast<const int*>(dmeta + R.m_rx);
    D.rb = reinterpret_cast<const int*>(dmeta + R.m_rb);
    D.ac = reinterpret_cast<const double* const*>(dmeta + R.m_ac);
    D.xc = reinterpret_cast<const double* const*>(dmeta + R.m_xc);
    D.bc = reinterpret_cast<const double* const*>(dmeta + R.m_bc);
    D.W = reinterpret_cast<double*>(dbase + R.w_W);
    D.U = reinterpret_cast<double*>(dbase + R.w_U);
    D.Rf = reinterpret_cast<double*>(dbase + R.w_R);
    D.tau = reinterpret_cast<double*>(dbase + R.w_tau);
    D.out = dres + 2 * s;
  }
  HIPCHK(c, hipMemcpyAsync(dmeta, c->hst, meta_bytes, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_residual_norm, dim3(count), dim3(NT), 0, st, (const ResidDesc*)dmeta);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->hst + r_out, dres, 8 * 2 * (size_t)count, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  const double* h = hptr<double>(c, r_out);
  for (int s = 0; s < count; ++s) {
    res_norm[s] = h[2 * s];
    rhs_norm[s] = h[2 * s + 1];
  }
  return 0;
}

int ttals_truncated_svd(ttals_ctx* c, void* stream, int m, int cdim, const double* Y, double threshold, int rel,
                        int max_rank, double* U, double* s, double* Vt, int* k) {
  if (!c) return TTALS_E_ARG;
  if (m < 1 || cdim < 1 || !Y || !U || !s || !Vt || !k || !(threshold >= 0.0))
    return fail(c, TTALS_E_ARG, "bad truncated_svd arguments");
  if (c->inflight) return fail(c, TTALS_E_ARG, "workspace busy: a submission is pending on this context");
  HIPCHK(c, hipSetDevice(c->device));
  const long mt = std::max(m, cdim), ct = std::min(m, cdim);
  Layout L;
  const size_t oW = L.take(8 * mt * ct), oQ = L.take(8 * mt * ct), oB = L.take(8 * ct * ct), oV = L.take(8 * ct * ct),
               oL = L.take(8 * mt * ct), osig = L.take(8 * ct), otau = L.take(8 * ct), operm = L.take(4 * ct),
               ok = L.take(8);
  int rc;
  if ((rc = ensure_dws(c, L.bytes))) return rc;
  char* db = reinterpret_cast<char*>(c->dws);
  SvdArgs a;
  a.m = m;
  a.c = cdim;
  a.maxr = max_rank;
  a.rel = rel;
  a.thr = threshold;
  a.Y = Y;
  a.U = U;
  a.s = s;
  a.Vt = Vt;
  a.W = reinterpret_cast<double*>(db + oW);
  a.Qm = reinterpret_cast<double*>(db + oQ);
  a.B = reinterpret_cast<double*>(db + oB);
  a.Vm = reinterpret_cast<double*>(db + oV);
  a.L = reinterpret_cast<double*>(db + oL);
  a.sig = reinterpret_cast<double*>(db + osig);
  a.tau = reinterpret_cast<double*>(db + otau);
  a.perm = reinterpret_cast<int*>(db + operm);
  a.kout = reinterpret_cast<int*>(db + ok);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_svd_trunc, dim3(1), dim3(NT), 0, st, a);
  HIPCHK(c, hipGetLastError());
  int h = 0;
  HIPCHK(c, hipMemcpyAsync(&h, a.kout, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  *k = h;
  return 0;
}

int ttals_dense_eig(ttals_ctx* c, void* stream, int N, const double* A, const double* M, int k, int mode,
                    double sigma, int unit, int general, double* vals, double* vecs) {
  if (!c) return TTALS_E_ARG;
  if (N < 1 || N > TTALS_DENSE_EIG_MAX || !A || k < 1 || k > N || (mode != 0 && mode != 1) || !vals || !vecs)
    return fail(c, TTALS_E_ARG, "bad dense_eig arguments (1 <= k <= N <= " + std::to_string(TTALS_DENSE_EIG_MAX) + ")");
  if (general && mode != 0)
    return fail(c, TTALS_E_UNSUPPORTED, "the largest eigenvalues ('eigh') of a non-symmetric matrix are not defined here");
  if (c->inflight) return fail(c, TTALS_E_ARG, "workspace busy: a submission is pending on this context");
  HIPCHK(c, hipSetDevice(c->device));
  const long NN = (long)N * N;
  Layout L;
  const size_t oC = L.take(8 * NN), oC0 = L.take(general ? 8 * NN : 8), oL = L.take(M ? 8 * NN : 8),
               oV = L.take(general ? 8 : 8 * NN), osig = L.take(8 * N), owr = L.take(8 * N), owi = L.take(8 * N),
               ox = L.take(8 * N), operm = L.take(4 * N), oidx = L.take(4 * k), ost = L.take(8);
  int rc;
  if ((rc = ensure_dws(c, L.bytes))) return rc;
  char* db = reinterpret_cast<char*>(c->dws);
  EigArgs e;
  e.N = N;
  e.k = k;
  e.mode = mode;
  e.unit = unit ? 1 : 0;
  e.general = general ? 1 : 0;
  e.sigma = sigma;
  e.A = A;
  e.M = M;
  e.C = reinterpret_cast<double*>(db + oC);
  e.C0 = reinterpret_cast<double*>(db + oC0);
  e.Lm = reinterpret_cast<double*>(db + oL);
  e.V = reinterpret_cast<double*>(db + oV);
  e.sig = reinterpret_cast<double*>(db + osig);
  e.wr = reinterpret_cast<double*>(db + owr);
  e.wi = reinterpret_cast<double*>(db + owi);
  e.xv = reinterpret_cast<double*>(db + ox);
  e.perm = reinterpret_cast<int*>(db + operm);
  e.idx = reinterpret_cast<int*>(db + oidx);
  e.vals = vals;
  e.vecs = vecs;
  e.status = reinterpret_cast<int*>(db + ost);
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(c, hipMemsetAsync(e.status, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_dense_eig, dim3(1), dim3(NT), 0, st, e);
  HIPCHK(c, hipGetLastError());
  int h = 0;
  HIPCHK(c, hipMemcpyAsync(&h, e.status, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  if (h == 1)
    return fail(c, TTALS_E_UNSUPPORTED, "the local matrix of operator_gevp is not symmetric positive definite "
                                        "(only symmetric-definite generalized eigenproblems are supported)");
  if (h == 2) return ST_NONFINITE;
  if (h == 3)
    return fail(c, TTALS_E_UNSUPPORTED, "a selected eigenvalue is complex (real=True keeps real eigenpairs only)");
  if (h == 4) return fail(c, TTALS_E_HIP, "the QR algorithm did not converge (60 iterations for one eigenvalue)");
  return 0;
}

int ttals_core_merge(void* stream, int r1, int m1, int n1, int sr, int m2, int n2, int r2, const double* A,
                     const double* B, double* out) {
  if (!dims_ok({r1, m1, n1, sr, m2, n2, r2}) || !A || !B || !out) return TTALS_E_ARG;
  const long total = (long)r1 * m1 * m2 * n1 * n2 * r2;
  const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_core_merge, dim3(blocks), dim3(256), 0, (hipStream_t)stream, r1, m1, n1, sr, m2, n2, r2, A,
                     B, out);
  return hipGetLastError() == hipSuccess ? 0 : TTALS_E_HIP;
}

int ttals_set_trace(ttals_ctx* c, double* buf, long long cap) {
  if (!c || cap < 0 || (cap > 0 && !buf)) return TTALS_E_ARG;
  c->trace = cap > 0 ? buf : nullptr;
  c->trace_cap = cap;
  return 0;
}

int ttals_tt_ortho(ttals_ctx* c, void* stream, const ttals_tt* t, int dir, int start, int end, double threshold,
                   const int* max_ranks) {
  if (!c || !t) return TTALS_E_ARG;
  if (c->inflight) return fail(c, TTALS_E_ARG, "workspace busy: a submission is pending on this context");
  int rc = check_tt(c, *t, false, "tt");
  if (rc) return rc;
  const int d = t->order;
  if (dir == 0 && (start < 0 || end > d - 2 || start > end + 1)) return fail(c, TTALS_E_ARG, "bad index range");
  if (dir == 1 && (start > d - 1 || end < 1 || start < end - 1)) return fail(c, TTALS_E_ARG, "bad index range");
  if (dir != 0 && dir != 1) return fail(c, TTALS_E_ARG, "bad direction");
  if (!(threshold >= 0.0)) return fail(c, TTALS_E_ARG, "threshold must be >= 0");
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  long cmax = 1, rmax = 1;
  bool trunc = threshold > 0.0;
  for (int i = 0; i < d; ++i) {
    cmax = std::max(cmax, (long)t->ranks[i] * t->row_dims[i] * t->col_dims[i] * t->ranks[i + 1]);
    rmax = std::max(rmax, (long)std::max(t->ranks[i], t->ranks[i + 1]));
  }
  if (max_ranks)
    for (int i = 0; i <= d; ++i) trunc = trunc || max_ranks[i] > 0;
  Layout L;
  const size_t oW = L.take(8 * cmax), oQ = L.take(8 * cmax), oR = L.take(8 * rmax * rmax), oT = L.take(8 * cmax),
               otau = L.take(8 * rmax), oB = L.take(8 * rmax * rmax), oV = L.take(8 * rmax * rmax),
               osig = L.take(8 * rmax), operm = L.take(sizeof(int) * rmax);
  const size_t omn = L.take(sizeof(int) * d), orr = L.take(sizeof(int) * (d + 1)), omx = L.take(sizeof(int) * (d + 1)),
               optr = L.take(sizeof(void*) * d);
  if ((rc = ensure_dws(c, L.bytes))) return rc;
  if ((rc = ensure_hst(c, L.bytes))) return rc;
  char* db = reinterpret_cast<char*>(c->dws);
  int* hmn = reinterpret_cast<int*>(c->hst + omn);
  int* hrr = reinterpret_cast<int*>(c->hst + orr);
  int* hmx = reinterpret_cast<int*>(c->hst + omx);
  double** hp = reinterpret_cast<double**>(c->hst + optr);
  for (int i = 0; i < d; ++i) {
    hmn[i] = t->row_dims[i] * t->col_dims[i];
    hp[i] = t->cores[i];
  }
  for (int i = 0; i <= d; ++i) {
    hrr[i] = t->ranks[i];
    hmx[i] = max_ranks ? max_ranks[i] : 0;
  }
  HIPCHK(c, hipMemcpyAsync(db + omn, c->hst + omn, L.bytes - omn, hipMemcpyHostToDevice, st));
  OrthoArgs oa;
  oa.d = d;
  oa.dir = dir;
  oa.start = start;
  oa.end = end;
  oa.trunc = trunc ? 1 : 0;
  oa.thr = threshold;
  oa.maxr = reinterpret_cast<const int*>(db + omx);
  oa.mn = reinterpret_cast<const int*>(db + omn);
  oa.r = reinterpret_cast<int*>(db + orr);
  oa.c = reinterpret_cast<double* const*>(db + optr);
  oa.W = reinterpret_cast<double*>(db + oW);
  oa.Qm = reinterpret_cast<double*>(db + oQ);
  oa.Rc = reinterpret_cast<double*>(db + oR);
  oa.T = reinterpret_cast<double*>(db + oT);
  oa.tau = reinterpret_cast<double*>(db + otau);
  oa.Bm = reinterpret_cast<double*>(db + oB);
  oa.Vm = reinterpret_cast<double*>(db + oV);
  oa.sig = reinterpret_cast<double*>(db + osig);
  oa.perm = reinterpret_cast<int*>(db + operm);
  hipLaunchKernelGGL(k_tt_ortho, dim3(1), dim3(NT), 0, st, oa);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->hst + orr, db + orr, sizeof(int) * (d + 1), hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  for (int i = 0; i <= d; ++i) t->ranks[i] = hrr[i];
  return 0;
}

}  // extern "C"
