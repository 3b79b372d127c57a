// Local solvers of the TT-ALS engine: matrix-free CG / GMRES(m) for the
// local linear system and restarted Lanczos for the extreme local
// eigenpair.  One workgroup per local problem.
//
// Two memory modes share the same solver code: when the local problem fits
// on-chip, the stacks, the operator core, the contraction intermediates and
// the Krylov working vectors live in LDS (op.ds = true, contractions by
// ds_gemm); otherwise everything is in global memory (L2-resident) and the
// contractions stage tiles through LDS (wg_gemm).  The Krylov basis is
// always in global memory.
#pragma once
#include "ttals_dev.hpp"

namespace ttals {

template <class T>
struct LocalOp {
  int a, P, n, Q, b;
  const T* L;
  const T* A;
  const T* R;
  T* t1;
  T* t2;
  bool ds;    // operands on-chip (operator core stored as (P, n, m, Q))
  const T* M2 = nullptr;  // on-chip two-stage form (L, M2; t1 = T1', t2 = K-split scratch)
  __device__ int size() const { return a * n * b; }
};

template <class T, class ACC>
__device__ __forceinline__ void apply_op(const LocalOp<T>& op, const T* v, T* y, ACC* lds) {
  if (op.ds && op.M2)
    local_matvec_m2(op.a, op.P, op.n, op.b, op.L, op.M2, v, y, op.t1, op.t2);
  else if (op.ds)
    local_matvec_mf<T>(op.a, op.P, op.n, op.Q, op.b, op.L, op.A, op.R, v, y, op.t1, op.t2);
  else
    local_matvec<T, ACC>(op.a, op.P, op.n, op.Q, op.b, op.L, op.A, op.R, v, y, op.t1, op.t2, lds);
}

constexpr int KMAX = 40;   // GMRES restart length (its Hessenberg matrix lives in LDS)
constexpr int LKMAX = 63;  // Lanczos restart length (basis in global memory, T in LDS)
static_assert(LKMAX < 64, "the tridiagonal eigensolver keeps one row of T per lane");
static_assert(KMAX <= LKMAX, "the basis and its scratch are sized for the longer restart");
struct SolveLDS {
  double red[4 * NWAVE];
  double red2[3 * NWAVE];           // Lanczos pass 1 only (pass 2 uses red)
  double h[LKMAX + 2];
  double H[(KMAX + 1) * KMAX];
  double cs[LKMAX + 1], sn[KMAX + 1], g[KMAX + 2];
  double alpha[LKMAX + 1], beta[LKMAX + 1], s[LKMAX + 1];
  double scal[4];
  double part[(LKMAX + 2) * NWAVE];  // per-wave partial dot products
  long long prof[8];                // cycle counters per phase (thread 0)
  double vfl;                       // vector flops of the current local solve (thread 0 only)
  int flag[4];
  int cnt[2 * NWAVE];               // per-wave ballot counts (double-buffered)
};

// Phase cycle counters (thread 0 accumulates; read back as diagnostics).
// Reading the shader clock (s_memtime) costs about 600 cycles, several times
// per Krylov step, so the counters are compiled in only with
// -DTTALS_PROFILE=1; otherwise they read 0.
#ifndef TTALS_PROFILE
#define TTALS_PROFILE 0
#endif
enum Phase { PH_MATVEC = 0, PH_ORTH, PH_TRI, PH_QR, PH_STACK, PH_SETUP, PH_OTHER };
__device__ __forceinline__ long long ph_now() {
#if TTALS_PROFILE
  return (long long)clock64();
#else
  return 0;
#endif
}
__device__ __forceinline__ void ph_add(SolveLDS& S, int ph, long long t0) {
#if TTALS_PROFILE
  if (threadIdx.x == 0) S.prof[ph] += ph_now() - t0;
#else
  (void)S;
  (void)ph;
  (void)t0;
#endif
}

// Working vectors of a local solve.  V (global): (KMAX+1) x N basis;
// cur, w (N each) and, for CG, r, p, q, x, f: on-chip when they fit.
struct Work {
  double* V;
  double* cur;
  double* w;
  double* r;
  double* p;
  double* q;
  double* x;
  double* f;
};

// h[i] = <V_i, w> for i < cnt (and h[cnt] = <w, w> if with_norm).  Each
// wave owns a contiguous slice of the elements and accumulates 8 basis
// vectors per pass over it (8 independent load streams per lane, so the L2
// latency of the basis overlaps); per-wave partial sums meet in LDS.  Fixed
// summation order: deterministic.
__device__ void wg_multidot(const double* V, long ldv, int cnt, const double* w, int N, double* h, double* part,
                            bool with_norm) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int chunk = (N + NWAVE - 1) / NWAVE;
  const int e0 = wv * chunk, e1 = (e0 + chunk < N) ? e0 + chunk : N;
  const int tot = cnt + (with_norm ? 1 : 0);
  for (int i0 = 0; i0 < tot; i0 += 8) {
    double acc[8];
#pragma unroll
    for (int g = 0; g < 8; ++g) acc[g] = 0.0;
    for (int e = e0 + lane; e < e1; e += 64) {
      const double we = w[e];
#pragma unroll
      for (int g = 0; g < 8; ++g) {
        const int i = i0 + g;
        if (i < cnt) acc[g] += V[(long)i * ldv + e] * we;
        else if (i == cnt && with_norm) acc[g] += we * we;
      }
    }
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      const double sg = wave_sum(acc[g]);
      if (lane == 0 && i0 + g < tot) part[(i0 + g) * NWAVE + wv] = sg;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < tot; i += NT) {
    double sg = 0.0;
#pragma unroll
    for (int q = 0; q < NWAVE; ++q) sg += part[i * NWAVE + q];
    h[i] = sg;
  }
  __syncthreads();
}

// w -= sum_i h[i] V_i   (no barrier)
__device__ __forceinline__ void wg_multiaxpy(const double* V, long ldv, int cnt, const double* h, double* w,
                                             int N) {
  for (int e = threadIdx.x; e < N; e += NT) {
    double sg = 0.0;
#pragma unroll 8
    for (int i = 0; i < cnt; ++i) sg += h[i] * V[(long)i * ldv + e];
    w[e] -= sg;
  }
}

// Iterated classical Gram-Schmidt with the DGKS criterion: a second (third,
// fourth) pass is made only while a pass shrinks w by more than a factor
// 0.7, so that even a near-breakdown residual made of rounding noise leaves
// orthogonal to the basis.  Accumulates the coefficients in hsum; returns
// ||w|| after orthogonalisation and the norm before it in *w0.
__device__ double wg_orth(const double* V, long ldv, int cnt, double* w, int N, double* h, double* hsum,
                          double* part, double* red, double* w0, double* vfl) {
  double prev = 0.0, nrm = 0.0;
  for (int pass = 0; pass < 4; ++pass) {
    // dots (and |w|^2 on the first pass), the update and the norm
    if (threadIdx.x == 0) *vfl += 2.0 * N * (2 * cnt + (pass == 0 ? 1 : 0) + 1);
    wg_multidot(V, ldv, cnt, w, N, h, part, pass == 0);
    if (pass == 0) {
      prev = sqrt(fmax(h[cnt], 0.0));
      *w0 = prev;
    }
    wg_multiaxpy(V, ldv, cnt, h, w, N);
    for (int i = threadIdx.x; i < cnt; i += NT) hsum[i] = (pass == 0 ? 0.0 : hsum[i]) + h[i];
    nrm = wg_nrm2u(w, N, red);  // barrier inside: w and hsum complete
    if (nrm > 0.7 * prev) break;
    prev = nrm;
  }
  return nrm;
}

// ---------------------------------------------------------------------------
// What a Krylov solve reports beyond its status: the norm of the true
// residual f - M x of the x it returns (every exit of CG, and of GMRES but
// one, has just computed it; -1 where it has not) and the matvecs it made.
struct KrylovOut {
  double tres = -1.0;
  int nmv = 0;
};

// Conjugate gradients on W.x (initial guess in W.x); vectors x, r, p, q, f
// on-chip when OC.  Returns 0 on convergence, 1 if p^T M p <= 0 was met (not
// SPD: caller switches to GMRES), 2 if maxit was reached.
template <class ACC, bool OC>
__device__ int wg_cg(const LocalOp<double>& op, const Work& W, double tol, int maxit, ACC* lds, SolveLDS& S,
                     int* iters, KrylovOut* out = nullptr) {
  using M = Mem<OC>;
  const int N = op.size();
  double *x = W.x, *r = W.r, *p = W.p, *q = W.q;
  const double* f = W.f;
  KrylovOut o;
  auto done = [&](int ret, double tres) {
    if (out) {
      out->tres = tres;
      out->nmv += o.nmv;
    }
    return ret;
  };
  double fn = 0.0, d1 = 0.0, d2 = 0.0;
  for (int e = threadIdx.x; e < N; e += NT) {
    const double fe = M::ld(f, e);
    fn += fe * fe;
  }
  wg_sum3(fn, d1, d2, S.red);
  fn = sqrt(fn);
  if (fn == 0.0) {
    for (int e = threadIdx.x; e < N; e += NT) M::st(x, e, 0.0);
    __syncthreads();
    return done(0, 0.0);
  }
  if (threadIdx.x == 0) S.vfl += 2.0 * N;
  double tr = -1.0;
  for (int restart = 0; restart < 3; ++restart) {
    apply_op<double, ACC>(op, x, q, lds);
    ++o.nmv;
    if (threadIdx.x == 0) S.vfl += 6.0 * N;  // this residual and the true-residual check below
    double rr = 0.0;
    for (int e = threadIdx.x; e < N; e += NT) {
      const double re = M::ld(f, e) - M::ld(q, e);
      M::st(r, e, re);
      M::st(p, e, re);
      rr += re * re;
    }
    wg_sum3(rr, d1, d2, S.red);
    if (sqrt(rr) <= tol * fn) return done(0, sqrt(rr));
    while (*iters < maxit) {
      long long t0 = ph_now();
      apply_op<double, ACC>(op, p, q, lds);
      ++o.nmv;
      ph_add(S, PH_MATVEC, t0);
      t0 = ph_now();
      double pq = 0.0;
#pragma unroll 8
      for (int e = threadIdx.x; e < N; e += NT) pq += M::ld(p, e) * M::ld(q, e);
      wg_sum3(pq, d1, d2, S.red);
      if (!(pq > 0.0)) return done(1, -1.0);
      const double al = rr / pq;
      double rn = 0.0;
#pragma unroll 4
      for (int e = threadIdx.x; e < N; e += NT) {
        M::st(x, e, M::ld(x, e) + al * M::ld(p, e));
        const double re = M::ld(r, e) - al * M::ld(q, e);
        M::st(r, e, re);
        rn += re * re;
      }
      wg_sum3(rn, d1, d2, S.red);
      ++*iters;
      if (threadIdx.x == 0) S.vfl += 10.0 * N;  // two dots, two updates, the direction
      if (sqrt(rn) <= tol * fn) break;
      const double be = rn / rr;
      rr = rn;
#pragma unroll 8
      for (int e = threadIdx.x; e < N; e += NT) M::st(p, e, M::ld(r, e) + be * M::ld(p, e));
      __syncthreads();
      ph_add(S, PH_ORTH, t0);
    }
    // true residual check (the recurrence drifts)
    apply_op<double, ACC>(op, x, q, lds);
    ++o.nmv;
    tr = 0.0;
    for (int e = threadIdx.x; e < N; e += NT) {
      const double re = M::ld(f, e) - M::ld(q, e);
      M::st(r, e, re);
      tr += re * re;
    }
    wg_sum3(tr, d1, d2, S.red);
    tr = sqrt(tr);
    if (tr <= 10.0 * tol * fn) return done(0, tr);
    if (*iters >= maxit) return done(2, tr);
  }
  return done(2, tr);
}

// ---------------------------------------------------------------------------
// Restarted GMRES(m) with DGKS-reorthogonalised Arnoldi; solution in W.x.
template <class ACC>
__device__ int wg_gmres(const LocalOp<double>& op, const Work& W, double tol, int maxit, int m, ACC* lds,
                        SolveLDS& S, int* iters, KrylovOut* out = nullptr) {
  const int N = op.size();
  double *V = W.V, *x = W.x, *cur = W.cur, *w = W.w;
  const double* f = W.f;
  KrylovOut o;
  auto done = [&](int ret, double tres) {
    if (out) {
      out->tres = tres;
      out->nmv += o.nmv;
    }
    return ret;
  };
  if (m > KMAX) m = KMAX;
  if (m > N) m = N;
  const double fn = wg_nrm2(f, N, S.red);
  if (fn == 0.0) {
    wg_fill(x, 0.0, N);
    __syncthreads();
    return done(0, 0.0);
  }
  for (;;) {
    apply_op<double, ACC>(op, x, w, lds);
    ++o.nmv;
    for (int e = threadIdx.x; e < N; e += NT) w[e] = f[e] - w[e];
    __syncthreads();
    const double beta = wg_nrm2(w, N, S.red);
    if (threadIdx.x == 0) S.vfl += 4.0 * N;
    if (beta <= tol * fn) return done(0, beta);
    if (*iters >= maxit) return done(2, beta);
    const double ib = 1.0 / beta;
    for (int e = threadIdx.x; e < N; e += NT) {
      const double t = w[e] * ib;
      cur[e] = t;
      V[e] = t;
    }
    if (threadIdx.x == 0) {
      S.g[0] = beta;
      for (int i = 1; i <= m; ++i) S.g[i] = 0.0;
    }
    __syncthreads();
    int jend = 0;
    double res = beta;
    for (int j = 0; j < m; ++j) {
      apply_op<double, ACC>(op, cur, w, lds);
      ++o.nmv;
      double w0;
      double hn = wg_orth(V, N, j + 1, w, N, S.h, S.H + j * (KMAX + 1), S.part, S.red, &w0, &S.vfl);
      if (!(hn > 1e-14 * w0)) hn = 0.0;  // (lucky) breakdown: Krylov space invariant
      if (threadIdx.x == 0) S.vfl += (double)N;
      if (hn > 0.0) {
        const double ih = 1.0 / hn;
        for (int e = threadIdx.x; e < N; e += NT) {
          const double t = w[e] * ih;
          cur[e] = t;
          V[(long)(j + 1) * N + e] = t;
        }
      }
      if (threadIdx.x == 0) {
        double* col = S.H + j * (KMAX + 1);  // column j, rows 0..j+1
        col[j + 1] = hn;
        for (int i = 0; i < j; ++i) {
          const double t = S.cs[i] * col[i] + S.sn[i] * col[i + 1];
          col[i + 1] = -S.sn[i] * col[i] + S.cs[i] * col[i + 1];
          col[i] = t;
        }
        const double h0 = col[j], h1 = col[j + 1];
        const double rr = sqrt(h0 * h0 + h1 * h1);
        const double c = rr == 0.0 ? 1.0 : h0 / rr, s = rr == 0.0 ? 0.0 : h1 / rr;
        S.cs[j] = c;
        S.sn[j] = s;
        col[j] = rr;
        col[j + 1] = 0.0;
        S.g[j + 1] = -s * S.g[j];
        S.g[j] = c * S.g[j];
        S.scal[0] = fabs(S.g[j + 1]);
      }
      __syncthreads();
      res = S.scal[0];
      ++*iters;
      jend = j + 1;
      if (res <= tol * fn || hn == 0.0 || *iters >= maxit) break;
    }
    // y = H^{-1} g  (upper triangular jend x jend), x += V y
    if (threadIdx.x == 0) {
      for (int i = jend - 1; i >= 0; --i) {
        double t = S.g[i];
        for (int k = i + 1; k < jend; ++k) t -= S.H[k * (KMAX + 1) + i] * S.h[k];
        const double d = S.H[i * (KMAX + 1) + i];
        S.h[i] = d != 0.0 ? t / d : 0.0;
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < N; e += NT) {
      double t = 0.0;
      for (int i = 0; i < jend; ++i) t += S.h[i] * V[(long)i * N + e];
      x[e] += t;
    }
    if (threadIdx.x == 0) S.vfl += 2.0 * N * jend + N;
    __syncthreads();
    double tres = -1.0;  // x changed: its true residual is unknown until checked
    // after one or two steps from an orthonormal basis the least-squares
    // residual is the true residual to rounding: no confirming matvec
    if (res <= tol * fn && jend <= 2) return done(0, res);
    if (res <= tol * fn) {
      // confirm with the true residual once
      apply_op<double, ACC>(op, x, w, lds);
      ++o.nmv;
      for (int e = threadIdx.x; e < N; e += NT) w[e] = f[e] - w[e];
      __syncthreads();
      tres = wg_nrm2(w, N, S.red);
      if (tres <= 10.0 * tol * fn) return done(0, tres);
    }
    if (*iters >= maxit) return done(2, tres);
  }
}

// ---------------------------------------------------------------------------
// Eigenvalues of a symmetric tridiagonal matrix (alpha, beta), m <= LKMAX.
// tri_eigval / tri_eigvec (interior eigenvalues): Sturm-count multisection by
// one wave (64 shifts per round), then inverse iteration for the eigenvector.
// tri_extreme (extreme eigenpairs, every convergence check): below.

// Sturm count #{eigenvalues < x} from the LDL^T pivots; the division by the
// previous pivot uses the hardware reciprocal refined by two Newton steps
// (full double precision, no IEEE division sequence).
__device__ __forceinline__ int sturm_count(const double* al, const double* be, int m, double x) {
  int c = 0;
  double q = al[0] - x;
  if (q < 0.0) ++c;
  for (int i = 1; i < m; ++i) {
    const double qq = (q == 0.0) ? 1e-300 : q;
    double r = __builtin_amdgcn_rcp(qq);
    r = r * (2.0 - qq * r);
    r = r * (2.0 - qq * r);
    q = (al[i] - x) - be[i - 1] * be[i - 1] * r;
    if (q < 0.0) ++c;
  }
  return c;
}

// The tridiagonal routines run only at convergence checks; they are kept out
// of line so that inlining them does not push wg_lanczos past the register
// budget of two workgroups per CU.
#define TT_COLD __attribute__((noinline))





// Sturm count (eigenvalues below x) of a tridiagonal T scaled so that
// |entries| and |x| are O(1): ta[k] = alpha_k, tb[k] = beta_{k-1}^2 (tb[0] =
// 0), k < m4, m4 a multiple of 4 (rows past m padded decoupled, above the
// spectrum).  Every lane reads the coefficients as LDS broadcasts four rows
// at a time; the three-term recurrence p_{k+1} = (alpha_k - x) p_k -
// beta_{k-1}^2 p_{k-1} is rescaled by powers of two between groups (per
// row the magnitude grows at most ~4x, so 2^600 leaves hundreds of binades).
// A zero p_k counts as positive: for an interior zero p_{k-1} and p_{k+1}
// have opposite signs, so the count does not depend on it; p_m = 0 means x
// is an eigenvalue, where either count keeps the bracket invariant.
__device__ __forceinline__ int sturm_count_lds4(const double* ta_, const double* tb_, int m4, double x) {
  const TT_LDS double* ta = (const TT_LDS double*)ta_;
  const TT_LDS double* tb = (const TT_LDS double*)tb_;
  double p0 = 0.0, p1 = 1.0;
  bool neg = false;
  int c = 0;
  for (int k = 0; k < m4; k += 4) {
    double av[4], bv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      av[q] = ta[k + q];
      bv[q] = tb[k + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double p2 = fma(av[q] - x, p1, -bv[q] * p0);
      const bool ng = p2 < 0.0;
      c += (ng != neg) ? 1 : 0;
      neg = ng;
      p0 = p1;
      p1 = p2;
    }
    const double ap = fabs(p1);
    if (ap > 0x1p+600) {
      p1 *= 0x1p-600;
      p0 *= 0x1p-600;
    } else if (ap < 0x1p-600 && fabs(p0) < 0x1p-600) {
      p1 *= 0x1p+600;
      p0 *= 0x1p+600;
    }
  }
  return c;
}

// p_m(x) / p_m'(x) for p_m = det(T - x I), T and x scaled as for
// sturm_count_lds4 (rows k < m only: no padding rows).  The derivative runs
// alongside, p'_{k+1} = (alpha_k - x) p'_k - beta_{k-1}^2 p'_{k-1} - p_k,
// both rescaled together (the ratio is unchanged).
__device__ __forceinline__ double newton_ratio_lds4(const double* ta_, const double* tb_, int m, double x) {
  const TT_LDS double* ta = (const TT_LDS double*)ta_;
  const TT_LDS double* tb = (const TT_LDS double*)tb_;
  double p0 = 0.0, p1 = 1.0, d0 = 0.0, d1 = 0.0;
  for (int k = 0; k < m; k += 4) {
    double av[4], bv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      av[q] = ta[k + q];
      bv[q] = tb[k + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (k + q < m) {
        const double ax = av[q] - x;
        const double p2 = fma(ax, p1, -bv[q] * p0);
        const double d2 = fma(ax, d1, -bv[q] * d0) - p1;
        p0 = p1;
        p1 = p2;
        d0 = d1;
        d1 = d2;
      }
    }
    const double mg = fabs(p1) + fabs(d1);
    if (mg > 0x1p+600) {
      p0 *= 0x1p-600;
      p1 *= 0x1p-600;
      d0 *= 0x1p-600;
      d1 *= 0x1p-600;
    } else if (mg < 0x1p-600 && fabs(p0) + fabs(d0) < 0x1p-600) {
      p0 *= 0x1p+600;
      p1 *= 0x1p+600;
      d0 *= 0x1p+600;
      d1 *= 0x1p+600;
    }
  }
  return p1 / d1;
}

// Extreme eigenpair of T (m <= 64; idx = 0 smallest or m-1 largest).
//  * eigenvalue: multisection over all NT threads (NT+1 subintervals a
//    round) from the Gershgorin interval (or a hint) to a relative width of
//    2^-26, invariant count(lo) <= idx < count(hi), then Newton to the
//    rounding level (multisection all the way if Newton does not settle).
//    Left in S.scal[1].
//  * unit eigenvector in s, by a twisted factorization: top-down LDL^T
//    pivots q+ (wave 0) and bottom-up UDU^T pivots q- (wave 1) run
//    concurrently; the twist index r minimises
//    |gamma_k| = |q+_k + q-_k - (alpha_k - th)|; then s_r = 1 and
//    s_k = -(beta_k / q+_k) s_{k+1} above r, s_k = -(beta_{k-1} / q-_k) s_{k-1}
//    below r.  Stable where one-sided recurrences are not: a converged Ritz
//    vector decays toward the bottom of T.
// The Sturm counts run on T scaled by a power of two (exact) to |T| ~ 1, its
// coefficients broadcast from tw; the pivot recurrences read al / be as
// broadcasts, one row ahead.
// tw: LDS scratch of 128 doubles.  Returns false (s untouched) if the
// eigenvector overflowed.
__device__ TT_COLD bool tri_extreme(const double* al, const double* be, int m_in, int idx_in, double* s, double* tw,
                                    SolveLDS& S, double hint_lo, double hint_hi) {
  const int m = __builtin_amdgcn_readfirstlane(m_in);
  const int idx = __builtin_amdgcn_readfirstlane(idx_in);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const bool row = lane < m;
  const double a_l = row ? al[lane] : 0.0;
  const double b_l = (lane < m - 1) ? be[lane] : 0.0;
  const double bm_l = (row && lane > 0) ? be[lane - 1] : 0.0;
  const double rad = fabs(b_l) + fabs(bm_l);
  double lo = wave_min(row ? a_l - rad : 1e300);
  double hi = wave_max(row ? a_l + rad : -1e300);
  const double pad = 1e-12 * fmax(hi - lo, 1e-300) + 1e-300;
  lo -= pad;
  hi += pad;
  // scaled coefficients for the counts: sc = 2^-e with |lo|, |hi| < 2^e
  const int e2 = ilogb(fmax(fabs(lo), fabs(hi))) + 1;
  const double sc = ldexp(1.0, -e2);
  const int m4 = (m + 3) & ~3;
  if (threadIdx.x < 64) {
    const double bs = (lane > 0 && lane < m) ? bm_l * sc : 0.0;
    tw[lane] = row ? a_l * sc : 4.0;  // padding rows: decoupled, above [-1, 1]
    tw[64 + lane] = bs * bs;
  }
  __syncthreads();
  // Warm start: a hint interval [hint_lo, hint_hi] (hint_lo < hint_hi) is
  // searched first with NT points that include both ends, so the first round
  // also tells whether the eigenvalue lies below or above it; the bracket then
  // continues from whichever side holds it.
  if (hint_lo < hint_hi && hint_lo > lo && hint_hi < hi) {
    const double h = (hint_hi - hint_lo) * (1.0 / (double)(NT - 1));
    const double x = (threadIdx.x == NT - 1) ? hint_hi : hint_lo + h * (double)threadIdx.x;
    const int c = sturm_count_lds4(tw, tw + 64, m4, x * sc);
    const unsigned long long ball = __ballot(c <= idx);
    int* cnt = S.cnt + NWAVE;
    if (lane == 0) cnt[wv] = __popcll(ball);
    __syncthreads();
    int nle = 0;
#pragma unroll
    for (int q = 0; q < NWAVE; ++q) nle += cnt[q];
    if (nle == 0) {
      hi = hint_lo;  // below the hint
    } else if (nle == NT) {
      lo = hint_hi;  // above the hint
    } else {
      lo = hint_lo + h * (double)(nle - 1);
      hi = (nle == NT - 1) ? hint_hi : hint_lo + h * (double)nle;
    }
  }
  // Multisection to a relative width of 2^-26, then Newton on det(T - x I)
  // from the bracket end outside the spectrum: every root is real, so the
  // iterates rise (fall) monotonically to the smallest (largest) eigenvalue,
  // quadratically once close.  Should Newton not settle within 8 steps (a
  // near-degenerate pair) or leave the bracket, multisection continues to the
  // rounding level.  The round counter runs on across both phases (it picks
  // the alternating count buffer).
  double th = 0.0;
  bool have_th = false;
  int round = 0;
  for (int phase = 0; phase < 2 && !have_th; ++phase) {
    const double relw = phase == 0 ? 0x1p-26 : 2.2e-16;
    for (; round < 128; ++round) {
      if (hi - lo <= relw * fmax(fabs(lo), fabs(hi)) + 1e-300) break;
      const double h = (hi - lo) * (1.0 / (double)(NT + 1));
      if (!(lo + h > lo) || !(lo + h * NT < hi)) break;  // spacing below the rounding of lo / hi
      const double x = lo + h * (double)(threadIdx.x + 1);
      const int c = sturm_count_lds4(tw, tw + 64, m4, x * sc);
      const unsigned long long ball = __ballot(c <= idx);
      int* cnt = S.cnt + (round & 1) * NWAVE;
      if (lane == 0) cnt[wv] = __popcll(ball);
      __syncthreads();
      int nle = 0;
#pragma unroll
      for (int q = 0; q < NWAVE; ++q) nle += cnt[q];
      // counts are monotone in x: the first nle points lie left of the eigenvalue
      const double nlo = nle > 0 ? lo + h * (double)nle : lo;
      const double nhi = nle < NT ? lo + h * (double)(nle + 1) : hi;
      lo = nlo;
      hi = nhi;
    }
    if (phase == 0 && (idx == 0 || idx == m - 1)) {
      // every thread runs the same iteration (no exchange needed)
      double x = (idx == 0 ? lo : hi) * sc;
      bool ok = false;
      for (int it = 0; it < 8; ++it) {
        const double dx = newton_ratio_lds4(tw, tw + 64, m, x);
        if (!isfinite(dx)) break;
        x -= dx;
        if (fabs(dx) <= 0x1p-50 * fabs(x) + 1e-300) {
          ok = true;
          break;
        }
      }
      const double t = x / sc;
      if (ok && t >= lo && t <= hi) {
        th = t;
        have_th = true;
      }
    }
  }
  if (!have_th) th = 0.5 * (lo + hi);
  double q_l = 0.0, r_l = 0.0;  // this lane's pivot and ratio
  __syncthreads();  // the counts are done with tw
  const TT_LDS double* la = (const TT_LDS double*)al;
  const TT_LDS double* lb = (const TT_LDS double*)be;
  if (wv == 0) {
    // q+_k = (alpha_k - th) - beta_{k-1} r_{k-1}, r_k = beta_k / q+_k
    double rprev = 0.0, bprev = 0.0;
    double ak = la[0], bk = (m > 1) ? lb[0] : 0.0;
    for (int k = 0; k < m; ++k) {
      const double an = (k + 1 < m) ? la[k + 1] : 0.0;
      const double bn = (k + 2 < m) ? lb[k + 1] : 0.0;
      double q = (ak - th) - bprev * rprev;
      if (q == 0.0) q = 1e-300;
      const double rk = bk * rcp_nr(q);
      if (lane == k) {
        q_l = q;
        r_l = rk;
      }
      rprev = rk;
      bprev = bk;
      ak = an;
      bk = bn;
    }
  } else if (wv == 1) {
    // q-_k = (alpha_k - th) - beta_k r_{k+1}, r_k = beta_{k-1} / q-_k
    double rnext = 0.0, bnext = 0.0;
    double ak = la[m - 1], bk = (m > 1) ? lb[m - 2] : 0.0;  // beta_{k-1}
    for (int k = m - 1; k >= 0; --k) {
      const double an = (k > 0) ? la[k - 1] : 0.0;
      const double bn = (k > 1) ? lb[k - 2] : 0.0;
      double q = (ak - th) - bnext * rnext;
      if (q == 0.0) q = 1e-300;
      const double rk = bk * rcp_nr(q);
      if (lane == k) {
        q_l = q;
        r_l = rk;
      }
      rnext = rk;
      bnext = bk;
      ak = an;
      bk = bn;
    }
    tw[lane] = q_l;
    tw[64 + lane] = r_l;
  }
  __syncthreads();
  if (wv == 0) {
    const double qm_l = tw[lane], rm_l = tw[64 + lane];
    const double g = row ? fabs(q_l + qm_l - (a_l - th)) : 1e308;
    const double gmin = wave_min(g);
    const unsigned long long hit = __ballot(g == gmin);
    const int r = hit ? (int)__builtin_ctzll(hit) : 0;
    double s_l = (lane == r) ? 1.0 : 0.0;
    double z = 1.0;
    for (int k = r - 1; k >= 0; --k) {
      z = -rdlane(r_l, k) * z;
      if (lane == k) s_l = z;
    }
    z = 1.0;
    for (int k = r + 1; k < m; ++k) {
      z = -rdlane(rm_l, k) * z;
      if (lane == k) s_l = z;
    }
    const double nn = wave_sum(s_l * s_l);
    const bool ok = isfinite(nn) && nn > 0.0;
    if (ok && row) s[lane] = s_l * (1.0 / sqrt(nn));
    if (lane == 0) {
      S.flag[1] = ok ? 1 : 0;
      S.scal[1] = th;
    }
  }
  __syncthreads();
  return S.flag[1] != 0;
}

// eigenvalue with index idx (0 = smallest) of T; executed by wave 0, result
// broadcast through S.scal[1].
__device__ TT_COLD void tri_eigval(const double* al, const double* be, int m, int idx,
                                                     SolveLDS& S) {
  if ((threadIdx.x >> 6) == 0) {
    const int lane = threadIdx.x & 63;
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < m; ++i) {
      const double r = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i < m - 1 ? fabs(be[i]) : 0.0);
      lo = fmin(lo, al[i] - r);
      hi = fmax(hi, al[i] + r);
    }
    const double span = fmax(hi - lo, 1e-300);
    lo -= 1e-12 * span + 1e-300;
    hi += 1e-12 * span + 1e-300;
    // invariant: count(lo) <= idx < count(hi)
    for (int round = 0; round < 40; ++round) {
      const double h = (hi - lo) / 65.0;
      const double xpt = lo + h * (lane + 1);
      const int c = sturm_count(al, be, m, xpt);
      const unsigned long long ball = __ballot(c <= idx);
      const int nle = __popcll(ball);  // counts are monotone in lane
      const double nlo = nle > 0 ? lo + h * nle : lo;
      const double nhi = nle < 64 ? lo + h * (nle + 1) : hi;
      lo = nlo;
      hi = nhi;
      if (hi - lo <= 2.2e-16 * fmax(fabs(lo), fabs(hi)) + 1e-300) break;
    }
    if (lane == 0) S.scal[1] = 0.5 * (lo + hi);
  }
  __syncthreads();
}

// eigenvector of T for eigenvalue th by inverse iteration (thread 0).
__device__ TT_COLD void tri_eigvec(const double* al, const double* be, int m, double th, double* s, double* tmpd) {
  if (threadIdx.x == 0) {
    double anorm = 0.0;
    for (int i = 0; i < m; ++i) anorm = fmax(anorm, fabs(al[i]) + (i < m - 1 ? fabs(be[i]) : 0.0));
    const double pert = 1e-15 * fmax(anorm, 1e-300);
    for (int i = 0; i < m; ++i) s[i] = 1.0 / sqrt((double)m) * (1.0 + 0.01 * (double)((i * 7) % 5));
    for (int it = 0; it < 3; ++it) {
      // solve (T - th I) z = s with an LDL^T sweep (pivots nudged away from 0)
      double dprev = 0.0, uprev = 0.0;
      for (int i = 0; i < m; ++i) {
        double di = al[i] - th;
        if (i > 0) {
          const double l = be[i - 1] * tmpd[i - 1];
          di -= l * be[i - 1];
          s[i] -= l * uprev;
        }
        if (fabs(di) < pert) di = (di < 0.0 ? -pert : pert);
        double rd = __builtin_amdgcn_rcp(di);
        rd = rd * (2.0 - di * rd);
        rd = rd * (2.0 - di * rd);
        tmpd[i] = rd;
        dprev = di;
        uprev = s[i];
      }
      for (int i = m - 1; i >= 0; --i) {
        double t = s[i];
        if (i < m - 1) t -= be[i] * s[i + 1];
        s[i] = t * tmpd[i];
      }
      double nn = 0.0;
      for (int i = 0; i < m; ++i) nn += s[i] * s[i];
      nn = 1.0 / sqrt(nn);
      for (int i = 0; i < m; ++i) s[i] *= nn;
    }
  }
  __syncthreads();
}

// Restarted Lanczos for one extreme eigenpair of the symmetric local
// operator.
//   mode 0: lowest, 1: highest, 2: nearest sigma (lowest/highest chosen
//   from the first Ritz spectrum; the choice is checked at the end).
// v (global): start vector on entry, unit eigenvector on exit.  The three
// working vectors (previous, current, next) rotate through W.r/W.cur/W.w
// (on-chip when OC).  Each step re-orthogonalises locally against the two
// previous vectors (two fused reductions); problems that fit the Krylov
// space (N <= KMAX, where the nearest-sigma rule needs the exact spectrum)
// and steps with severe cancellation (|w| < 1e-3 |A v|) use full DGKS
// Gram-Schmidt against the basis V (global).  Breakdowns (an invariant
// subspace, e.g. a symmetry sector) continue with a fresh orthogonal
// direction.  Returns 0 converged, 2 not converged, 3 sigma inside the
// spectrum of a problem larger than the Krylov space (unsupported), 4 a
// non-finite Ritz vector.  On 0 and 2, v holds the unit Ritz vector.
template <class ACC, bool OC>
__device__ int wg_lanczos(const LocalOp<double>& op, double* v, const Work& W, int mode, double sigma, double tol,
                          int maxit, int kmax, ACC* lds, SolveLDS& S, double* lambda, int* iters) {
  using M = Mem<OC>;
  const int N = op.size();
  double* V = W.V;
  double *prv = W.r, *cur = W.cur, *w = W.w;
  if (kmax > LKMAX) kmax = LKMAX;
  const bool full_orth = (N <= LKMAX);
  // start vector into cur (one pass over v, which may be in global memory);
  // every vector loop below gives element e to thread e mod NT
  double vn;
  {
    double ss = 0.0, d1 = 0.0, d2 = 0.0;
    for (int e = threadIdx.x; e < N; e += NT) {
      const double x = v[e];
      M::st(cur, e, x);
      ss += x * x;
    }
    wg_sum3(ss, d1, d2, S.red);
    vn = sqrt(ss);
    if (!(vn > 0.0) || !isfinite(vn)) {
      ss = 0.0;
      for (int e = threadIdx.x; e < N; e += NT) {
        const double x = 1.0 + 0.1 * (double)((e * 2654435761u) % 97) / 97.0;
        M::st(cur, e, x);
        ss += x * x;
      }
      wg_sum3(ss, d1, d2, S.red);
      vn = sqrt(ss);
    }
  }
  int target = (mode == 1) ? 1 : 0;
  bool decided = (mode != 2);
  double th_prev = 0.0, res_prev = 0.0;  // last check's Ritz value and residual estimate
  bool have_hint = false;
  int status = 2;
  double th = 0.0;
  for (int restart = 0; restart < 1000; ++restart) {
    const double iv = 1.0 / vn;
    for (int e = threadIdx.x; e < N; e += NT) {
      const double t = M::ld(cur, e) * iv;
      V[e] = t;
      M::st(cur, e, t);
      M::st(prv, e, 0.0);
    }
    __syncthreads();
    int m = 0;
    bool conv = false;
    bool full = false;
    double beta_prev = 0.0;
    int next_check = 2;
    bool have_vec = false;
    int m_prev = 0;
    double anorm = 0.0;
    const int kk = kmax < N ? kmax : N;
    for (int j = 0; j < kk; ++j) {
      long long t0 = ph_now();
      apply_op<double, ACC>(op, cur, w, lds);
      ph_add(S, PH_MATVEC, t0);
      t0 = ph_now();
      double alpha, w0, bn;
      bool dgks = full_orth;
      // the second local pass's correction w -= c1 cur + c0 prv is applied
      // lazily, fused into the normalisation at the end of the step
      double cc1 = 0.0, cc0 = 0.0;
      if (!full_orth) {
        // pass 1: alpha = <w, cur>, |w|^2
        double a1 = 0.0, ww = 0.0, dummy = 0.0;
        for (int e = threadIdx.x; e < N; e += NT) {
          const double we = M::ld(w, e);
          a1 += we * M::ld(cur, e);
          ww += we * we;
        }
        // the two passes use separate scratch, so each one's last readers have
        // passed the other's barrier: no leading barrier needed
        wg_sum3_nb(a1, ww, dummy, S.red2);
        w0 = sqrt(ww);
        // pass 2 fused with the update: c1 = <w', cur>, c0 = <w', prv>, |w'|^2
        double c1 = 0.0, c0 = 0.0, w2 = 0.0;
        for (int e = threadIdx.x; e < N; e += NT) {
          const double we = M::ld(w, e) - a1 * M::ld(cur, e) - beta_prev * M::ld(prv, e);
          M::st(w, e, we);
          c1 += we * M::ld(cur, e);
          c0 += we * M::ld(prv, e);
          w2 += we * we;
        }
        wg_sum3_nb(c1, c0, w2, S.red);
        if (threadIdx.x == 0) S.vfl += 14.0 * N;  // pass 1: two dots; pass 2: update and three dots
        cc1 = c1;
        cc0 = c0;
        alpha = a1 + c1;
        const double b2 = w2 - c1 * c1 - c0 * c0;
        bn = sqrt(fmax(b2, 0.0));
        if (!(bn > 1e-3 * w0)) {
          dgks = true;  // severe cancellation: restore orthogonality to the whole basis
          for (int e = threadIdx.x; e < N; e += NT)
            M::st(w, e, M::ld(w, e) - c1 * M::ld(cur, e) - c0 * M::ld(prv, e));
          cc1 = cc0 = 0.0;
          __syncthreads();
        }
        // every thread holds the same alpha (fixed-order sums) and writes it:
        // each one's later reads of row j then need no barrier.  (Each entry
        // is written once, with its final value: a provisional value could
        // land after another wave's final one.)
        if (!dgks) S.alpha[j] = alpha;
      }
      if (dgks) {
        // w = A cur - (alpha cur + beta prv) (local pass already applied when !full_orth)
        double h0;
        const double nb = wg_orth(V, N, j + 1, w, N, S.h, S.H, S.part, S.red, &h0, &S.vfl);
        if (full_orth) {
          w0 = h0;
          alpha = S.H[j];
        } else {
          alpha += S.H[j];
        }
        bn = nb;
        S.alpha[j] = alpha;
      }
      // no barrier: w is next read elementwise by its own threads (the
      // normalisation), row j of T by threads that wrote it themselves
      ph_add(S, PH_ORTH, t0);
      ++*iters;
      m = j + 1;
      // running max of |alpha_i| + |beta_{i-1}| (rows of T are final once written)
      anorm = fmax(anorm, fabs(S.alpha[j]) + (j > 0 ? fabs(S.beta[j - 1]) : 0.0));
      // breakdown: the residual is at the rounding level of A v_j
      const bool brk = !(bn > 1e-12 * fmax(w0, 1e-300));
      full = (m == N);
      if (brk && !full && m < kk) {
        // invariant subspace: continue with a fresh direction orthogonal to
        // the basis (T stays tridiagonal with a zero coupling)
        S.beta[j] = 0.0;
        for (int e = threadIdx.x; e < N; e += NT)
          M::st(w, e, (double)((((unsigned)e + 1u) * 2654435761u + ((unsigned)j + 1u) * 40503u) % 1009u) / 1009.0 -
                          0.5);
        __syncthreads();
        double r0;
        bn = wg_orth(V, N, j + 1, w, N, S.h, S.cs, S.part, S.red, &r0, &S.vfl);
        const double ib = bn > 0.0 ? 1.0 / bn : 0.0;
        double* t = prv;
        prv = cur;
        cur = w;
        w = t;
        for (int e = threadIdx.x; e < N; e += NT) {
          const double x = M::ld(cur, e) * ib;
          M::st(cur, e, x);
          V[(long)(j + 1) * N + e] = x;
        }
        beta_prev = 0.0;
        __syncthreads();
        continue;
      }
      S.beta[j] = brk ? 0.0 : bn;  // by every thread, as alpha
      // convergence checks at m = 2, then every 4 steps, spaced out while the
      // residual estimate is still far from the tolerance
      const bool check = full || brk || (m == kk) || m == next_check || *iters >= maxit;
      const long long tc = ph_now();
      if (check) {
        if (!decided) {
          // sigma outside the Gershgorin interval of T decides at once
          double glo = 1e300, ghi = -1e300;
          for (int i = 0; i < m; ++i) {
            const double rr = (i > 0 ? fabs(S.beta[i - 1]) : 0.0) + (i < m - 1 ? fabs(S.beta[i]) : 0.0);
            glo = fmin(glo, S.alpha[i] - rr);
            ghi = fmax(ghi, S.alpha[i] + rr);
          }
          if (sigma <= glo) {
            target = 0;
          } else if (sigma >= ghi) {
            target = 1;
          } else {
            tri_eigval(S.alpha, S.beta, m, 0, S);
            const double tlo = S.scal[1];
            tri_eigval(S.alpha, S.beta, m, m - 1, S);
            const double thi = S.scal[1];
            target = (sigma <= 0.5 * (tlo + thi)) ? 0 : 1;
          }
          decided = true;
        }
        int idx = target == 0 ? 0 : m - 1;
        if (full && mode == 2) {
          // the whole space is spanned: nearest sigma among all Ritz values;
          // the Sturm count of sigma splits the spectrum
          __syncthreads();
          if (threadIdx.x == 0) S.flag[0] = sturm_count(S.alpha, S.beta, m, sigma);
          __syncthreads();
          const int c = S.flag[0];
          double best = 0.0;
          int bidx = 0;
          if (c > 0) {
            tri_eigval(S.alpha, S.beta, m, c - 1, S);
            best = S.scal[1];
            bidx = c - 1;
          }
          if (c < m) {
            tri_eigval(S.alpha, S.beta, m, c, S);
            const double up = S.scal[1];
            if (c == 0 || fabs(up - sigma) < fabs(best - sigma)) bidx = c;
          }
          idx = bidx;
        }
        double res;
        if (full && mode == 2) {
          // interior eigenvalue: full precision and inverse iteration
          tri_eigval(S.alpha, S.beta, m, idx, S);
          th = S.scal[1];
          tri_eigvec(S.alpha, S.beta, m, th, S.s, S.cs);
          have_vec = true;
          res = 0.0;
        } else if (m == 2) {
          // 2x2 T (every local solve's first check): closed form.  The
          // eigenvalue of larger magnitude from the mean and the half-gap, the
          // other from the determinant divided by it (no cancellation); the
          // eigenvector from whichever row of T - th I gives the longer null
          // vector, its larger component made positive (the twisted
          // factorization's convention).
          const double a = S.alpha[0], c = S.alpha[1], b = S.beta[0];
          const double sm = a + c, rt = hypot(a - c, 2.0 * b);
          const double acmx = fabs(a) > fabs(c) ? a : c, acmn = fabs(a) > fabs(c) ? c : a;
          double rt1, rt2;
          if (sm != 0.0) {
            rt1 = 0.5 * (sm + copysign(rt, sm));
            rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
          } else {
            rt1 = 0.5 * rt;
            rt2 = -0.5 * rt;
          }
          th = (idx == 0) ? fmin(rt1, rt2) : fmax(rt1, rt2);
          double s0 = b, s1 = th - a;
          const double u0 = th - c, u1 = b;
          if (u0 * u0 + u1 * u1 > s0 * s0 + s1 * s1) {
            s0 = u0;
            s1 = u1;
          }
          double nn = sqrt(s0 * s0 + s1 * s1);
          if (!(nn > 0.0)) {  // a == c, b == 0: any vector
            s0 = 1.0;
            s1 = 0.0;
            nn = 1.0;
          }
          const double sg = (fabs(s0) >= fabs(s1) ? s0 : s1) < 0.0 ? -1.0 : 1.0;
          // written by every thread (identical values): a converged check
          // goes straight to the Ritz vector, which reads s without a barrier
          S.s[0] = sg * s0 / nn;
          S.s[1] = sg * s1 / nn;
          have_vec = true;
          res = (brk || full) ? 0.0 : bn * fabs(s1 / nn);
        } else {
          // extreme eigenvalue: multisection to full precision, eigenvector
          // by twisted factorization, residual estimate bn |s_m|
          // hint from the last check: by interlacing the extreme Ritz value only
          // moves outward, and a residual bound puts it within res of the old one
          double hlo = 0.0, hhi = -1.0;
          if (have_hint) {
            const double slack = 1.5 * res_prev + 1e-14 * fabs(th_prev) + 1e-300;
            if (target == 0) {
              hlo = th_prev - slack;
              hhi = th_prev + 4e-16 * fabs(th_prev) + 1e-300;
            } else {
              hlo = th_prev - 4e-16 * fabs(th_prev) - 1e-300;
              hhi = th_prev + slack;
            }
          }
          have_vec = tri_extreme(S.alpha, S.beta, m, idx, S.s, S.H + (KMAX + 1), S, hlo, hhi);
          th = S.scal[1];
          if (!have_vec) {
            tri_eigvec(S.alpha, S.beta, m, th, S.s, S.cs);
            have_vec = true;
          }
          res = (brk || full) ? 0.0 : bn * fabs(S.s[m - 1]);
        }
        ph_add(S, PH_TRI, tc);
        const double goal = tol * fmax(fabs(th), 1e-3 * anorm);
        if (res <= goal) { conv = true; break; }
        // next check: where the residual's geometric decay since the last
        // check reaches the goal (80 % of the way), 2..24 steps ahead (a
        // check costs about a Krylov step; caps of 10, 16, 24 and 32 measured,
        // 24 the fastest)
        int step = 4;
        if (have_hint && res < res_prev && m > m_prev && res > 0.0) {
          // steps to go = (m - m_prev) log(res / goal) / log(res_prev / res); an
          // estimate, so FP32 logarithms (overflow to inf only clamps to 10)
          const float lr = __log2f((float)(res_prev / res));
          const float need = __log2f((float)(res / goal)) * (float)(m - m_prev) / lr;
          step = (int)fminf(24.0f, fmaxf(2.0f, floorf(0.8f * need)));
        } else if (res > 1e6 * goal && m >= 10) {
          step = 8;
        }
        next_check = m + step;
        th_prev = th;
        res_prev = res;
        m_prev = m;
        have_hint = !full;
        if (*iters >= maxit) break;
      }
      const double ib = 1.0 / bn;
      double* t = prv;
      prv = cur;
      cur = w;
      w = t;
      for (int e = threadIdx.x; e < N; e += NT) {
        const double x = (M::ld(cur, e) - cc1 * M::ld(prv, e) - cc0 * M::ld(w, e)) * ib;
        M::st(cur, e, x);
        V[(long)(j + 1) * N + e] = x;
      }
      beta_prev = bn;
      if (threadIdx.x == 0) S.vfl += 5.0 * N;  // lazy correction and normalisation
      __syncthreads();
    }
    if (!have_vec) {
      const long long tv = ph_now();
      tri_eigvec(S.alpha, S.beta, m, th, S.s, S.cs);
      ph_add(S, PH_TRI, tv);
    }
    // Ritz vector into w (free now), basis rows read four at a time (the
    // basis may be in global memory), its norm from registers
    double ss = 0.0;
    for (int e = threadIdx.x; e < N; e += NT) {
      double t = 0.0;
      int i = 0;
      for (; i + 4 <= m; i += 4) {
        const double x0 = V[(long)i * N + e], x1 = V[(long)(i + 1) * N + e];
        const double x2 = V[(long)(i + 2) * N + e], x3 = V[(long)(i + 3) * N + e];
        t += S.s[i] * x0;
        t += S.s[i + 1] * x1;
        t += S.s[i + 2] * x2;
        t += S.s[i + 3] * x3;
      }
      for (; i < m; ++i) t += S.s[i] * V[(long)i * N + e];
      M::st(w, e, t);
      ss += t * t;
    }
    {
      double d1 = 0.0, d2 = 0.0;
      wg_sum3(ss, d1, d2, S.red);
    }
    vn = sqrt(ss);
    if (threadIdx.x == 0) S.vfl += 2.0 * N * m + 3.0 * N;  // Ritz vector, its norm and scaling
    if (!isfinite(vn) || !(vn > 0.0)) {
      status = 4;
      break;
    }
    if (conv || *iters >= maxit) {
      const double iv2 = 1.0 / vn;
      for (int e = threadIdx.x; e < N; e += NT) v[e] = M::ld(w, e) * iv2;
      __syncthreads();
      status = conv ? 0 : 2;
      if (conv && mode == 2 && !full) {
        // nearest-sigma with an extreme eigenvalue is only valid if sigma lies
        // outside the spectrum on that side
        if ((target == 0 && sigma > th) || (target == 1 && sigma < th)) status = 3;
      }
      break;
    }
    double* t = cur;  // restart from the Ritz vector
    cur = w;
    w = t;
  }
  *lambda = th;
  return status;
}

}  // namespace ttals
