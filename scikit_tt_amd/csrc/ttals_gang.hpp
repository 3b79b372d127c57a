// Multi-workgroup ("gang") ALS engine for large sle systems.
//
// One system is solved by a gang of G workgroups that run the whole
// sle.als (scikit_tt/solvers/sle.py:10-95) together: every phase is split
// over the gang and phases meet at a gang barrier.  The split follows the
// contraction structure, so no phase exchanges partial results:
//   - local matvec y = M v: over the row-side rank index a' of the output
//     (each workgroup computes all three stages for its rows, reading v
//     whole) -- the vector elements of a workgroup's rows are also the ones
//     it updates in the Krylov vector operations;
//   - left interface stack L_{i+1}[b, Q, b']: over b (its first index);
//     right stacks are the left update of the mirrored problem (core
//     transposed to (b, n, a), operator core to (Q, m, n, P));
//   - reductions (dot products, norms): per-workgroup partial sums in global
//     memory, read back by every workgroup in the same order, so every
//     workgroup takes the same decisions with the same numbers;
//   - small sequential pieces (rhs stacks, local rhs, the QR of the new core
//     and the warm start of the next position) run on workgroup 0.
// The gang workgroups of a launch come first in workgroup order and fit the
// device together (checked on the host); workgroups are dispatched in index
// order, so every gang eventually runs whole and the barrier's waits end.
// Each wait is also bounded in time, and a gang that times out raises an
// abort flag every member sees.
//
// Precision: prec 0 runs every Krylov iteration in FP64; prec 1 runs them in
// FP32 on FP32 copies of the local operator (L, A core, R) and refines the
// local solution with the FP64 residual f - M v (FP64 stacks and operator):
// solve M d = r in FP32 to a loose tolerance, v += d in FP64, until the FP64
// residual meets `tol`.  Cores, stacks, QR and the residual stay FP64.
#pragma once
#include "ttals_engine.hpp"

namespace ttals {

struct GangSync {
  unsigned count, gen, abort, pad;
};

struct GangDesc {
  int d, solver, repeats, maxit;
  int prec, G, first, maxit_inner;
  int gm_restart, pad0[3];
  double tol, tol_inner;
  const int* n;
  int* rx;
  const int* ra;
  const int* rb;
  double* const* xc;
  const double* const* ac;
  const double* const* bc;
  const double* const* x0c;
  double* Ls;
  double* Rs;
  double* Lbs;
  double* Rbs;
  const long* Loff;
  const long* Roff;
  const long* Lboff;
  const long* Rboff;
  double* f;     // local rhs (nmax)
  double* g;     // local solution / warm start (nmax)
  double* r;     // Krylov vectors (nmax each)
  double* p;
  double* q;
  double* V;     // GMRES basis (gm_restart + 1) x nmax
  double* xt;    // transposed core (right stacks), nmax
  double* At;    // transposed operator core, max P n n Q
  float* L32;    // FP32 local operator (mixed): max a P a
  float* A32;    // max P n n Q
  float* R32;    // max b Q b
  float* r32;    // FP32 Krylov vectors, nmax each
  float* p32;
  float* q32;
  float* x32;
  float* V32;    // FP32 GMRES basis (gm_restart + 1) x nmax
  double* t1;    // per workgroup: G x tmax
  double* t2;
  long tmax;
  double* part;  // G x GANG_LD partial sums
  double* qw;    // QR panel (nmax, shared); workgroup 0's Householder fallback: qq (nmax), tau, R (rmax^2)
  double* qq;
  double* tau;
  double* rf;
  double* gram;  // G x rmax^2 partial Gram matrices of the split QR
  double* qrs;   // per workgroup: 3 x rmax^2 (Cholesky factor, accumulated R, product)
  long rmax2;    // rmax^2
  GangSync* sync;
  double* out;   // [16] as SysDesc::out
  double* res;   // compact result row: out[0..15], then the d+1 final ranks
};

// --- gang barrier ---------------------------------------------------------------

// Longest wait at one barrier, in ticks of the 100 MHz constant clock
// (s_memrealtime): 20 s.  A member may wait for the others to be dispatched
// while other kernels still hold the CUs, so the bound is in time, not spins.
constexpr unsigned long long GANG_WAIT_TICKS = 20ull * 100000000ull;
// partial-sum slots per member: 4 scalars, then two regions of KMAX + 2 for
// the two Gram-Schmidt passes (a member may write the second pass's sums
// while another still reads the first's)
constexpr int GANG_LD = 4 + 2 * (KMAX + 2);

// Every workgroup of the gang calls this at the same points.  Returns false
// once the gang has aborted (a wait ran out): callers then unwind.  Thread 0
// makes this workgroup's writes visible with ONE release fence, arrives with a
// relaxed atomic, spins with relaxed loads (an acquire load in the spin would
// invalidate the caches on every iteration) and acquires once at the end.
__device__ bool gang_sync(GangSync* s, int G) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int good = 1;
    if (__hip_atomic_load(&s->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
      good = 0;
    } else {
      const unsigned gen = __hip_atomic_load(&s->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      const unsigned prev = __hip_atomic_fetch_add(&s->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (prev + 1 == (unsigned)G) {
        __hip_atomic_store(&s->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&s->gen, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        const unsigned long long t0 = wall_clock64();
        unsigned spins = 0;
        while (__hip_atomic_load(&s->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
          if (__hip_atomic_load(&s->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            good = 0;
            break;
          }
          if ((++spins & 255u) == 0 && wall_clock64() - t0 > GANG_WAIT_TICKS) {
            __hip_atomic_store(&s->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            good = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

// contiguous share [lo, hi) of `total` items for member g of G
__device__ __forceinline__ void gang_span(int total, int G, int g, int& lo, int& hi) {
  lo = (int)((long)total * g / G);
  hi = (int)((long)total * (g + 1) / G);
}

// Sum of the G partial sums of slot `k` (stride `ld` per member), same order everywhere.
__device__ __forceinline__ double gang_total(const double* part, int G, int ld, int k) {
  double s = 0.0;
  for (int h = 0; h < G; ++h) s += part[(long)h * ld + k];
  return s;
}

// out[k] = <V_k, w> over elements [e0, e1) for k < cnt, and out[cnt] = <w, w>
// (this member's partial sums; written by thread 0).  Eight basis vectors
// per pass, per-wave sums meeting in LDS (`wpart`, (cnt + 1) x NWAVE), fixed
// order.
template <class T>
__device__ void gang_multidot(const T* V, long ldv, int cnt, const T* w, long e0, long e1, double* out,
                              double* wpart) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tot = cnt + 1;
  for (int i0 = 0; i0 < tot; i0 += 8) {
    double acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = 0.0;
    for (long e = e0 + threadIdx.x; e < e1; e += NT) {
      const double we = (double)w[e];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int i = i0 + q;
        if (i < cnt) acc[q] += (double)V[(long)i * ldv + e] * we;
        else if (i == cnt) acc[q] += we * we;
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double sq = wave_sum(acc[q]);
      if (lane == 0 && i0 + q < tot) wpart[(i0 + q) * NWAVE + wv] = sq;
    }
  }
  __syncthreads();
  if (threadIdx.x < tot) {
    double sg = 0.0;
#pragma unroll
    for (int q = 0; q < NWAVE; ++q) sg += wpart[threadIdx.x * NWAVE + q];
    out[threadIdx.x] = sg;
  }
  __syncthreads();
}

// C[M x N] = A[M x K] B[K x N] for M <= SKINNY_M rows: A staged in LDS
// (M x K, K <= SKINNY_K), every thread owns whole output columns and keeps
// the M accumulators in registers; B rows are read with unit column stride
// (coalesced over the threads).  For the first matvec stage of a gang member
// (M = P times its rows), where 64 x 64 tiles would be mostly padding.
constexpr int SKINNY_M = 8, SKINNY_K = 256;
template <class T>
__device__ void wg_gemm_skinny(int M, int N, int K, CView<T> A, const T* B, int ldb, MView<T> C, T* lds) {
  for (int e = threadIdx.x; e < M * K; e += NT) {
    const int i = e / K, k = e - i * K;
    lds[e] = A.at(i, k);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < N; j += NT) {
    T acc[SKINNY_M];
#pragma unroll
    for (int i = 0; i < SKINNY_M; ++i) acc[i] = (T)0;
    const T* bp = B + j;
    for (int k = 0; k < K; ++k) {
      const T bv = bp[(long)k * ldb];
#pragma unroll
      for (int i = 0; i < SKINNY_M; ++i)
        if (i < M) acc[i] += lds[i * K + k] * bv;
    }
#pragma unroll
    for (int i = 0; i < SKINNY_M; ++i)
      if (i < M) C.at(i, j) = acc[i];
  }
  __syncthreads();
}

// --- split contractions -------------------------------------------------------------

// y[a', m, b'] for a' in [r0, r1) of the local matvec (all three stages
// for those rows; v read whole).  t1, t2: this member's scratch.
template <class T>
__device__ void gang_matvec(int a, int P, int n, int Q, int b, const T* L, const T* Ac, const T* R, const T* v,
                            T* y, int r0, int r1, T* t1, T* t2, T* lds) {
  const int s = r1 - r0;
  if (s <= 0) return;
  // T1[a'_l, b, P, n] = sum_a L[a, P, r0 + a'_l] v[a, n, b]
  if (P * s <= SKINNY_M && a <= SKINNY_K && (long)P * s * a * sizeof(T) <= GEMM_LDS_DOUBLES * 8)
    wg_gemm_skinny<T>(P * s, n * b, a, cview<T>(L + r0, split(s, a, 1), plain(P * a)), v, n * b,
                      mview<T>(t1, split(s, n, b * P * n), split(b, 1, P * n)), lds);
  else
    wg_gemm<T, T>(P * s, n * b, a, cview<T>(L + r0, split(s, a, 1), plain(P * a)), crm<T>(v, n * b),
                  mview<T>(t1, split(s, n, b * P * n), split(b, 1, P * n)), lds);
  // T2[a'_l, m, b, Q] = sum_{P,n} T1[(a'_l, b), (P, n)] A[P, m, n, Q]
  wg_gemm<T, T>(s * b, n * Q, P * n, crm<T>(t1, P * n), cview<T>(Ac, split(n, n * n * Q, Q), split(Q, n * Q, 1)),
                mview<T>(t2, split(b, n * b * Q, Q), split(Q, b * Q, 1)), lds);
  // y[(a', m), b'] = T2[(a'_l, m), (b, Q)] R[(b, Q), b']
  wg_gemm<T, T>(s * n, b, b * Q, crm<T>(t2, b * Q), crm<T>(R, b), mrm<T>(y + (long)r0 * n * b, b), lds);
}

// Rows [c0, c1) of the left stack update
//   Lout[b, Q, b'] = sum L[a, P, a'] x[a, n, b] A[P, m, n, Q] x[a', m, b']
// (scikit_tt/solvers/sle.py:214-219).  A right stack is the same update of
// the mirrored problem.
__device__ void gang_stack_left(int a, int P, int n, int Q, int b, const double* L, const double* Ac,
                                const double* x, double* Lout, int c0, int c1, double* t1, double* t2,
                                double* lds) {
  const int sb = c1 - c0;
  if (sb <= 0) return;
  // T1[a', b_l, P, n] = sum_a L[a, P, a'] x[a, n, c0 + b_l]
  wg_gemm<double, double>(P * a, n * sb, a, cview<double>(L, plain(1), plain(P * a)),
                          cview<double>(x + c0, plain(n * b), split(sb, b, 1)),
                          mview<double>(t1, split(a, n, sb * P * n), split(sb, 1, P * n)), lds);
  // T2[b_l, Q, a', m] = sum_{P,n} T1[(a', b_l), (P, n)] A[P, m, n, Q]
  wg_gemm<double, double>(a * sb, n * Q, P * n, crm<double>(t1, P * n),
                          cview<double>(Ac, split(n, n * n * Q, Q), split(Q, n * Q, 1)),
                          mview<double>(t2, split(sb, n, Q * a * n), split(Q, 1, a * n)), lds);
  // Lout[(c0 + b_l, Q), b'] = T2[(b_l, Q), (a', m)] x[(a', m), b']
  wg_gemm<double, double>(sb * Q, b, a * n, crm<double>(t2, a * n), crm<double>(x, b),
                          mrm<double>(Lout + (long)c0 * Q * b, b), lds);
}

// --- the gang solve ------------------------------------------------------------------

struct GangCtx {
  const GangDesc& D;
  int g, G;
  double* t1;
  double* t2;
  double* lds;
  SolveLDS& S;
  __device__ bool sync() {
    const long long t0 = ph_now();
    const bool r = gang_sync(D.sync, G);
    ph_add(S, PH_OTHER, t0);  // barrier waits (profiling builds)
    return r;
  }
};

// Local operator of position i: FP64 operands in global memory.
struct GangLocal {
  int a, P, n, Q, b, N;
  const double* L;
  const double* A;
  const double* R;
  int r0, r1;   // this member's rows a'
  long e0, e1;  // its vector elements
};

__device__ __forceinline__ GangLocal gang_local(const GangCtx& C, int i) {
  const GangDesc& D = C.D;
  GangLocal o;
  o.a = D.rx[i];
  o.n = D.n[i];
  o.b = D.rx[i + 1];
  o.P = D.ra[i];
  o.Q = D.ra[i + 1];
  o.N = o.a * o.n * o.b;
  o.L = D.Ls + D.Loff[i];
  o.A = D.ac[i];
  o.R = D.Rs + D.Roff[i];
  gang_span(o.a, C.G, C.g, o.r0, o.r1);
  o.e0 = (long)o.r0 * o.n * o.b;
  o.e1 = (long)o.r1 * o.n * o.b;
  return o;
}

// Conjugate gradients on M x = f (x holds the start), vectors of type T,
// matvec on operands (L, A, R) of type T, reductions in FP64.  Returns 0
// converged, 1 p^T M p <= 0 (not SPD), 2 maxit, -1 gang aborted.
template <class T>
__device__ int gang_cg(GangCtx& C, const GangLocal& o, const T* L, const T* A, const T* R, T* x, const T* f, T* r,
                       T* p, T* q, double tol, int maxit, int* iters, double* vfl) {
  const int G = C.G, g = C.g;
  const int ld = GANG_LD;
  double* part = C.D.part;
  T* t1 = reinterpret_cast<T*>(C.t1);
  T* t2 = reinterpret_cast<T*>(C.t2);
  T* lds = reinterpret_cast<T*>(C.lds);
  // r = f - M x, p = r; |f|, |r|
  gang_matvec<T>(o.a, o.P, o.n, o.Q, o.b, L, A, R, x, q, o.r0, o.r1, t1, t2, lds);
  double ff = 0.0, rr = 0.0;
  for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) {
    const double fe = (double)f[e], re = fe - (double)q[e];
    r[e] = (T)re;
    p[e] = (T)re;
    ff += fe * fe;
    rr += re * re;
  }
  double d1 = 0.0;
  wg_sum3(ff, rr, d1, C.S.red);
  if (threadIdx.x == 0) {
    part[(long)g * ld + 0] = ff;
    part[(long)g * ld + 1] = rr;
  }
  if (!C.sync()) return -1;
  ff = gang_total(part, G, ld, 0);
  rr = gang_total(part, G, ld, 1);
  const double fn = sqrt(ff);
  if (threadIdx.x == 0) *vfl += 4.0 * o.N;
  if (fn == 0.0 || sqrt(rr) <= tol * fn) return 0;
  while (*iters < maxit) {
    // q = M p (p complete: the barrier above / at the end of the last step)
    const long long tm = ph_now();
    gang_matvec<T>(o.a, o.P, o.n, o.Q, o.b, L, A, R, p, q, o.r0, o.r1, t1, t2, lds);
    ph_add(C.S, PH_MATVEC, tm);
    double pq = 0.0;
    for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) pq += (double)p[e] * (double)q[e];
    double u1 = 0.0, u2 = 0.0;
    wg_sum3(pq, u1, u2, C.S.red);
    if (threadIdx.x == 0) part[(long)g * ld + 2] = pq;
    if (!C.sync()) return -1;
    pq = gang_total(part, G, ld, 2);
    if (!(pq > 0.0)) return 1;
    const double al = rr / pq;
    double rn = 0.0;
    for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) {
      x[e] = (T)((double)x[e] + al * (double)p[e]);
      const double re = (double)r[e] - al * (double)q[e];
      r[e] = (T)re;
      rn += re * re;
    }
    wg_sum3(rn, u1, u2, C.S.red);
    if (threadIdx.x == 0) part[(long)g * ld + 3] = rn;
    if (!C.sync()) return -1;
    rn = gang_total(part, G, ld, 3);
    ++*iters;
    if (threadIdx.x == 0) *vfl += 10.0 * o.N;
    if (sqrt(rn) <= tol * fn) return 0;
    const double be = rn / rr;
    rr = rn;
    for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) p[e] = (T)((double)r[e] + be * (double)p[e]);
    if (!C.sync()) return -1;
  }
  return 2;
}

// Restarted GMRES(m) on M x = f with classical Gram-Schmidt applied twice
// (each pass one gang reduction).  Same conventions as gang_cg.  w is a
// work vector, V the (m + 1) x N basis.
template <class T>
__device__ int gang_gmres(GangCtx& C, const GangLocal& o, const T* L, const T* A, const T* R, T* x, const T* f,
                          T* w, T* V, double tol, int maxit, int m, int* iters, double* vfl) {
  const int G = C.G, g = C.g;
  const int ld = GANG_LD;
  double* part = C.D.part;
  SolveLDS& S = C.S;
  T* t1 = reinterpret_cast<T*>(C.t1);
  T* t2 = reinterpret_cast<T*>(C.t2);
  T* lds = reinterpret_cast<T*>(C.lds);
  const long N = o.N;
  if (m > KMAX) m = KMAX;
  if (m > o.N) m = o.N;
  double ff = 0.0;
  for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) ff += (double)f[e] * (double)f[e];
  double u1 = 0.0, u2 = 0.0;
  wg_sum3(ff, u1, u2, S.red);
  if (threadIdx.x == 0) part[(long)g * ld + 0] = ff;
  if (!C.sync()) return -1;
  const double fn = sqrt(gang_total(part, G, ld, 0));
  if (fn == 0.0) {
    for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) x[e] = (T)0;
    return C.sync() ? 0 : -1;
  }
  for (;;) {
    // w = f - M x; beta = |w|; v_0 = w / beta
    gang_matvec<T>(o.a, o.P, o.n, o.Q, o.b, L, A, R, x, w, o.r0, o.r1, t1, t2, lds);
    double bb = 0.0;
    for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) {
      const double we = (double)f[e] - (double)w[e];
      w[e] = (T)we;
      bb += we * we;
    }
    wg_sum3(bb, u1, u2, S.red);
    if (threadIdx.x == 0) part[(long)g * ld + 1] = bb;
    if (!C.sync()) return -1;
    const double beta = sqrt(gang_total(part, G, ld, 1));
    if (threadIdx.x == 0) *vfl += 4.0 * N;
    if (beta <= tol * fn) return 0;
    if (*iters >= maxit) return 2;
    const double ib = 1.0 / beta;
    for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) V[e] = (T)((double)w[e] * ib);
    if (threadIdx.x == 0) {
      S.g[0] = beta;
      for (int k = 1; k <= m; ++k) S.g[k] = 0.0;
    }
    if (!C.sync()) return -1;  // v_0 complete
    int jend = 0;
    double res = beta;
    bool brk = false;
    for (int j = 0; j < m; ++j) {
      const long long tm = ph_now();
      gang_matvec<T>(o.a, o.P, o.n, o.Q, o.b, L, A, R, V + (long)j * N, w, o.r0, o.r1, t1, t2, lds);
      ph_add(S, PH_MATVEC, tm);
      double* col = S.H + j * (KMAX + 1);  // column j (rows 0..j+1), identical in every member
      for (int k = threadIdx.x; k <= j + 1; k += NT) col[k] = 0.0;
      double w0 = 0.0, hn = 0.0;
      for (int pass = 0; pass < 2; ++pass) {
        // partial dots <V_k, w> over this member's elements, k <= j, and |w|^2
        const int base = 4 + pass * (KMAX + 2);
        gang_multidot<T>(V, N, j + 1, w, o.e0, o.e1, part + (long)g * ld + base, S.part);
        if (!C.sync()) return -1;
        for (int k = threadIdx.x; k <= j + 1; k += NT) S.h[k] = gang_total(part, G, ld, base + k);
        __syncthreads();
        if (pass == 0) w0 = sqrt(fmax(S.h[j + 1], 0.0));
        for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) {
          double we = (double)w[e];
          for (int k = 0; k <= j; ++k) we -= S.h[k] * (double)V[(long)k * N + e];
          w[e] = (T)we;
        }
        for (int k = threadIdx.x; k <= j; k += NT) col[k] += S.h[k];
        __syncthreads();
        if (threadIdx.x == 0) *vfl += 2.0 * N * (2 * (j + 1) + 1);
      }
      // |w| after the two passes
      double ww = 0.0;
      for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) ww += (double)w[e] * (double)w[e];
      ww = wg_sum(ww, S.red);
      if (threadIdx.x == 0) part[(long)g * ld + 2] = ww;
      if (!C.sync()) return -1;
      hn = sqrt(gang_total(part, G, ld, 2));
      if (!(hn > 1e-14 * w0)) {
        hn = 0.0;  // (lucky) breakdown: the Krylov space is invariant
        brk = true;
      }
      if (hn > 0.0) {
        const double ih = 1.0 / hn;
        T* vn = V + (long)(j + 1) * N;
        for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) vn[e] = (T)((double)w[e] * ih);
      }
      // Givens rotations on column j (every thread, identical values)
      __syncthreads();
      if (threadIdx.x == 0) {
        col[j + 1] = hn;
        for (int k = 0; k < j; ++k) {
          const double t = S.cs[k] * col[k] + S.sn[k] * col[k + 1];
          col[k + 1] = -S.sn[k] * col[k] + S.cs[k] * col[k + 1];
          col[k] = t;
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
      if (threadIdx.x == 0) *vfl += 3.0 * N;
      if (!C.sync()) return -1;  // v_{j+1} complete
      if (res <= tol * fn || brk || *iters >= maxit) break;
    }
    // y = H^{-1} g, x += V y (own elements)
    if (threadIdx.x == 0) {
      for (int k = jend - 1; k >= 0; --k) {
        double t = S.g[k];
        for (int l = k + 1; l < jend; ++l) t -= S.H[l * (KMAX + 1) + k] * S.h[l];
        const double dg = S.H[k * (KMAX + 1) + k];
        S.h[k] = dg != 0.0 ? t / dg : 0.0;
      }
    }
    __syncthreads();
    for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) {
      double t = (double)x[e];
      for (int k = 0; k < jend; ++k) t += S.h[k] * (double)V[(long)k * N + e];
      x[e] = (T)t;
    }
    if (threadIdx.x == 0) *vfl += 2.0 * N * jend;
    if (!C.sync()) return -1;  // x complete
    if (*iters >= maxit) {
      // final residual decides
      gang_matvec<T>(o.a, o.P, o.n, o.Q, o.b, L, A, R, x, w, o.r0, o.r1, t1, t2, lds);
      double tr = 0.0;
      for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) {
        const double we = (double)f[e] - (double)w[e];
        tr += we * we;
      }
      tr = wg_sum(tr, S.red);
      if (threadIdx.x == 0) part[(long)g * ld + 1] = tr;
      if (!C.sync()) return -1;
      return sqrt(gang_total(part, G, ld, 1)) <= tol * fn ? 0 : 2;
    }
  }
}


// --- the sweep ------------------------------------------------------------------------

constexpr int ST_ABORT = 6;  // a gang barrier wait ran out (reported as a runtime error)

// operator symmetric slice by slice (as operator_symmetric), member-local
__device__ bool gang_slices_symmetric(const GangDesc& D, SolveLDS& S) {
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

// L_i from L_{i-1} and core i-1 (member share: rows b of the op stack;
// member 0 also the rhs stack)
__device__ double gang_build_left(GangCtx& C, int i) {
  const GangDesc& D = C.D;
  const int j = i - 1;
  const int a = D.rx[j], n = D.n[j], b = D.rx[i], P = D.ra[j], Q = D.ra[i];
  const int p = D.rb[j], q = D.rb[i];
  int c0, c1;
  gang_span(b, C.G, C.g, c0, c1);
  gang_stack_left(a, P, n, Q, b, D.Ls + D.Loff[j], D.ac[j], D.xc[j], D.Ls + D.Loff[i], c0, c1, C.t1, C.t2, C.lds);
  if (C.g == 0)
    stack_left_rhs<double, double>(a, p, n, q, b, D.Lbs + D.Lboff[j], D.bc[j], D.xc[j], D.Lbs + D.Lboff[i], C.t1,
                                   C.lds);
  return fl_op(a, P, n, Q, b) + fl_rhs(a, p, n, q, b);
}

// R_i from R_{i+1} and core i+1: the transposed core and operator core are
// written first (split), then the left update of the mirrored problem
// (rows a of the output); member 0 also the rhs stack.  Two barriers.
__device__ double gang_build_right(GangCtx& C, int i, bool* ok) {
  const GangDesc& D = C.D;
  const int j = i + 1;
  const int a = D.rx[j], n = D.n[j], b = D.rx[j + 1], P = D.ra[j], Q = D.ra[j + 1];
  const int p = D.rb[j], q = D.rb[j + 1];
  const long N = (long)a * n * b, K = (long)P * n * n * Q;
  const double* x = D.xc[j];
  const double* A = D.ac[j];
  {
    const long lo = N * C.g / C.G, hi = N * (C.g + 1) / C.G;
    for (long e = lo + threadIdx.x; e < hi; e += NT) {
      // xt[bb, nn, aa] = x[aa, nn, bb]
      const long bb = e % b, r = e / b, nn = r % n, aa = r / n;
      D.xt[(bb * n + nn) * a + aa] = x[e];
    }
    const long klo = K * C.g / C.G, khi = K * (C.g + 1) / C.G;
    for (long e = klo + threadIdx.x; e < khi; e += NT) {
      // At[qq, m, nn, pp] = A[pp, m, nn, qq]
      const long qq = e % Q, r = e / Q, nn = r % n, r2 = r / n, m = r2 % n, pp = r2 / n;
      D.At[((qq * n + m) * n + nn) * P + pp] = A[e];
    }
  }
  if (!C.sync()) {
    *ok = false;
    return 0.0;
  }
  int c0, c1;
  gang_span(a, C.G, C.g, c0, c1);
  gang_stack_left(b, Q, n, P, a, D.Rs + D.Roff[j], D.At, D.xt, D.Rs + D.Roff[i], c0, c1, C.t1, C.t2, C.lds);
  if (C.g == 0)
    stack_right_rhs<double, double>(a, p, n, q, b, D.Rbs + D.Rboff[j], D.bc[j], x, D.Rbs + D.Rboff[i], C.t1, C.lds);
  *ok = C.sync();
  return fl_op(a, P, n, Q, b) + fl_rhs(a, p, n, q, b);
}

// Split QR of the new core's panel W (m x c, m >= c) over the gang by shifted
// Cholesky QR applied three times (Fukaya, Kannan, Nakatsukasa, Yamamoto,
// Yanagisawa: "Shifted Cholesky QR for computing the QR factorization of
// ill-conditioned matrices", SIAM J. Sci. Comput. 42 (2020)): each pass forms
// the Gram matrix from per-member partials (one barrier), every member
// factors it (identical inputs and order: identical R), and solves its own
// rows against R; the first pass shifts the Gram matrix by
// 11 (m c + c (c + 1)) u ||W||_F^2 so that it stays positive definite for
// condition numbers up to ~1/u.  The panel is copied from g into D.qw:
// forward W[row, col] = g[row c + col] (m = a n, c = b), backward
// W[row, col] = g[col m + row] (m = n b, c = a).  On success Q is in D.qw
// (own rows) and R (c x c, upper) in this member's qrs + rmax2; returns 0, or
// 1 when a pivot failed (every member alike), 2 non-finite, -1 aborted.
__device__ int gang_cholqr(GangCtx& C, const double* g, int m, int c, bool transposed, double* fl) {
  const GangDesc& D = C.D;
  SolveLDS& S = C.S;
  double* Wq = D.qw;
  int m0, m1;
  gang_span(m, C.G, C.g, m0, m1);
  const int rows = m1 - m0;
  double* Rp = D.qrs + (long)C.g * 3 * D.rmax2;  // this pass's factor
  double* Rt = Rp + D.rmax2;                       // accumulated R
  double* Tm = Rt + D.rmax2;                       // product temporary
  const long cc = (long)c * c;
  for (long e = (long)m0 * c + threadIdx.x; e < (long)m1 * c; e += NT) {
    const long row = e / c, col = e - row * c;
    Wq[e] = transposed ? g[col * m + row] : g[e];
  }
  for (long e = threadIdx.x; e < cc; e += NT) Rt[e] = (e / c == e % c) ? 1.0 : 0.0;
  __syncthreads();
  for (int pass = 0; pass < 3; ++pass) {
    // partial Gram of this member's rows
    double* slot = D.gram + (long)C.g * D.rmax2;
    wg_gemm<double, double>(c, c, rows, ctr<double>(Wq + (long)m0 * c, c), crm<double>(Wq + (long)m0 * c, c),
                            mrm<double>(slot, c), C.lds);
    if (!C.sync()) return -1;
    // total Gram into Rp (every member, same order)
    for (long e = threadIdx.x; e < cc; e += NT) {
      double sg = 0.0;
      for (int h = 0; h < C.G; ++h) sg += D.gram[(long)h * D.rmax2 + e];
      Rp[e] = sg;
    }
    __syncthreads();
    if (pass == 0) {
      double tr = 0.0;
      for (int j = 0; j < c; ++j) tr += Rp[(long)j * c + j];
      if (!isfinite(tr)) return 2;
      const double shift = 11.0 * ((double)m * c + (double)c * (c + 1)) * 0x1p-53 * tr;
      __syncthreads();
      for (int j = threadIdx.x; j < c; j += NT) Rp[(long)j * c + j] += shift;
      __syncthreads();
    }
    // Cholesky Gram = R^T R (upper R in place, right-looking)
    for (int j = 0; j < c; ++j) {
      const double piv = Rp[(long)j * c + j];
      if (!(piv > 0.0) || !isfinite(piv)) return 1;
      const double rjj = sqrt(piv), inv = 1.0 / rjj;
      __syncthreads();
      for (int q = j + 1 + threadIdx.x; q < c; q += NT) Rp[(long)j * c + q] *= inv;
      if (threadIdx.x == 0) Rp[(long)j * c + j] = rjj;
      __syncthreads();
      const int rest = c - j - 1;
      for (int e = threadIdx.x; e < rest * rest; e += NT) {
        const int p = j + 1 + e / rest, q = j + 1 + e % rest;
        if (q >= p) Rp[(long)p * c + q] -= Rp[(long)j * c + p] * Rp[(long)j * c + q];
      }
      __syncthreads();
    }
    // own rows: W <- W R^{-1} (forward substitution per row)
    for (int row = m0 + threadIdx.x; row < m1; row += NT) {
      double* w = Wq + (long)row * c;
      for (int j = 0; j < c; ++j) {
        double t = w[j];
        for (int q = 0; q < j; ++q) t -= w[q] * Rp[(long)q * c + j];
        w[j] = t / Rp[(long)j * c + j];
      }
    }
    // R_total <- R_pass R_total (upper triangular)
    for (long e = threadIdx.x; e < cc; e += NT) {
      const int p = (int)(e / c), q = (int)(e % c);
      double t = 0.0;
      for (int k2 = p; k2 <= q; ++k2) t += Rp[(long)p * c + k2] * Rt[(long)k2 * c + q];
      Tm[e] = (q >= p) ? t : 0.0;
    }
    __syncthreads();
    for (long e = threadIdx.x; e < cc; e += NT) Rt[e] = Tm[e];
    __syncthreads();
  }
  *fl += 3.0 * (4.0 * (double)m * c * c + (double)c * c * c);
  return 0;
}

// One local step at position i: local rhs, the split Krylov solve (FP64, or
// FP32 with FP64 refinement), then on member 0 the QR of the new core, the
// core store and the warm start of the next position.  Returns a Status.
__device__ int gang_step(GangCtx& C, int i, bool forward, bool sym, StepStats& st) {
  const GangDesc& D = C.D;
  SolveLDS& S = C.S;
  const GangLocal o = gang_local(C, i);
  const int a = o.a, n = o.n, b = o.b;
  const double fmv = fl_op(a, o.P, n, o.Q, b);
  if (C.g == 0) {
    micro_rhs<double, double>(a, D.rb[i], n, D.rb[i + 1], b, D.Lbs + D.Lboff[i], D.bc[i], D.Rbs + D.Rboff[i], D.f,
                              C.t1, C.lds);
  }
  st.flops += fl_rhs(a, D.rb[i], n, D.rb[i + 1], b);
  if (!C.sync()) return ST_ABORT;
  int its = 0, r = 0;
  double vfl = 0.0;
  if (D.prec == 0) {
    r = 1;
    if (sym) r = gang_cg<double>(C, o, o.L, o.A, o.R, D.g, D.f, D.r, D.p, D.q, D.tol, D.maxit, &its, &vfl);
    if (r == 1) r = gang_gmres<double>(C, o, o.L, o.A, o.R, D.g, D.f, D.q, D.V, D.tol, D.maxit, D.gm_restart, &its,
                                        &vfl);
    st.flops += (its + 2) * fmv + vfl;
  } else {
    // FP32 copies of the local operator
    {
      const long nl = (long)a * o.P * a, na = (long)o.P * n * n * o.Q, nr = (long)b * o.Q * b;
      const long tot = nl + na + nr;
      const long lo = tot * C.g / C.G, hi = tot * (C.g + 1) / C.G;
      for (long e = lo + threadIdx.x; e < hi; e += NT) {
        if (e < nl) D.L32[e] = (float)o.L[e];
        else if (e < nl + na) D.A32[e - nl] = (float)o.A[e - nl];
        else D.R32[e - nl - na] = (float)o.R[e - nl - na];
      }
    }
    if (!C.sync()) return ST_ABORT;
    r = 2;
    int inner_total = 0;
    for (int outer = 0; outer < 30; ++outer) {
      // FP64 residual of the current solution: q = f - M g
      gang_matvec<double>(a, o.P, n, o.Q, b, o.L, o.A, o.R, D.g, D.q, o.r0, o.r1, C.t1, C.t2, C.lds);
      double ff = 0.0, rr = 0.0, mx = 0.0;
      for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) {
        const double fe = D.f[e], re = fe - D.q[e];
        D.q[e] = re;
        ff += fe * fe;
        rr += re * re;
      }
      wg_sum3(ff, rr, mx, S.red);
      const int ld = GANG_LD;
      if (threadIdx.x == 0) {
        D.part[(long)C.g * ld + 0] = ff;
        D.part[(long)C.g * ld + 1] = rr;
      }
      if (!C.sync()) return ST_ABORT;
      const double fn = sqrt(gang_total(D.part, C.G, ld, 0)), rn = sqrt(gang_total(D.part, C.G, ld, 1));
      vfl += 5.0 * o.N;
      st.flops += fmv;
      if (fn == 0.0 || rn <= D.tol * fn) {
        r = 0;
        break;
      }
      if (inner_total >= D.maxit) break;
      // the correction in FP32: M d = r, d from 0, to the inner tolerance;
      // the residual scaled to unit norm (FP32 range)
      const double sc = 1.0 / rn;
      for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) {
        D.r32[e] = (float)(D.q[e] * sc);
        D.x32[e] = 0.0f;
      }
      if (!C.sync()) return ST_ABORT;
      int ii = 0, ri = 1;
      if (sym)
        ri = gang_cg<float>(C, o, D.L32, D.A32, D.R32, D.x32, D.r32, D.p32, D.q32, reinterpret_cast<float*>(D.V),
                            D.tol_inner, D.maxit_inner, &ii, &vfl);
      if (ri == 1)
        ri = gang_gmres<float>(C, o, D.L32, D.A32, D.R32, D.x32, D.r32, D.q32, D.V32, D.tol_inner, D.maxit_inner,
                               D.gm_restart, &ii, &vfl);
      if (ri < 0) return ST_ABORT;
      inner_total += ii;
      st.flops += (ii + 2) * fmv;
      for (long e = o.e0 + threadIdx.x; e < o.e1; e += NT) D.g[e] += rn * (double)D.x32[e];
      if (!C.sync()) return ST_ABORT;
    }
    its = inner_total;
    st.flops += vfl;
  }
  if (r < 0) return ST_ABORT;
  if (r != 0) ++st.nunconv;
  st.kit += its;
  st.maxit_one = its > st.maxit_one ? its : st.maxit_one;
  ++st.nsolve;
  // QR of the new core and the warm start of the next position: split
  // Cholesky QR over the gang for tall panels, else (or when a pivot fails)
  // Householder QR on member 0
  int status = ST_OK;
  const long long tq = ph_now();
  if (i > 0 || forward) {
    const int m = forward ? a * n : n * b, c = forward ? b : a;
    if (m >= c && (long)c * c <= D.rmax2) {
      double qfl = 0.0;
      const int q = gang_cholqr(C, D.g, m, c, !forward, &qfl);
      if (q < 0) return ST_ABORT;
      if (q == 2) status = ST_NONFINITE;
      if (q == 0) {
        const double* Rt = D.qrs + (long)C.g * 3 * D.rmax2 + D.rmax2;
        int m0, m1;
        gang_span(m, C.G, C.g, m0, m1);
        // Q (own rows) into the core
        for (long e = (long)m0 * c + threadIdx.x; e < (long)m1 * c; e += NT) {
          const long row = e / c, col = e - row * c;
          if (forward) D.xc[i][e] = D.qw[e];
          else D.xc[i][col * m + row] = D.qw[e];
        }
        if (!C.sync()) return ST_ABORT;  // every member done with g (its panel copy) before g is overwritten
        if (forward) {
          // g[(k), (n1 c1)] = R core_{i+1}: split over the columns
          const int n1 = D.n[i + 1], c1 = D.rx[i + 2], nc = n1 * c1;
          int j0, j1;
          gang_span(nc, C.G, C.g, j0, j1);
          if (j1 > j0)
            wg_gemm<double, double>(c, j1 - j0, b, crm<double>(Rt, c), crm<double>(D.xc[i + 1] + j0, nc),
                                    mrm<double>(D.g + j0, nc), C.lds);
          st.flops += qfl + 2.0 * c * nc * b;
        } else {
          // g[(c0 n0), k] = core_{i-1} R^T: split over the rows
          const int c0 = D.rx[i - 1], n0 = D.n[i - 1], cn = c0 * n0;
          int r0, r1;
          gang_span(cn, C.G, C.g, r0, r1);
          if (r1 > r0)
            wg_gemm<double, double>(r1 - r0, c, a, crm<double>(D.xc[i - 1] + (long)r0 * a, a), ctr<double>(Rt, c),
                                    mrm<double>(D.g + (long)r0 * c, c), C.lds);
          st.flops += qfl + 2.0 * cn * c * a;
        }
        ph_add(S, PH_QR, tq);
        if (!C.sync()) return ST_ABORT;
        return status;
      }
      // a pivot failed (or non-finite): every member got here alike; member 0 below
      if (!C.sync()) return ST_ABORT;
    }
  }
  if (C.g == 0) {
    const QRBuf B{D.qw, D.qq, D.rf, D.tau};
    const double nv = wg_nrm2(D.g, o.N, S.red);
    if (!isfinite(nv)) {
      status = ST_NONFINITE;
    } else if (forward) {
      const int k = qr_factor(D.g, a * n, b, false, B, S);
      wg_copy(B.Q, D.xc[i], a * n * k);
      const int n1 = D.n[i + 1], c1 = D.rx[i + 2], nc = n1 * c1;
      st.flops += fl_qr(a * n, b) + 2.0 * k * n1 * c1 * b;
      ds_gemm<double, double>(k, nc, b, crm<double>(B.R, b), crm<double>(D.xc[i + 1], nc), mrm<double>(D.g, nc));
      if (threadIdx.x == 0) D.rx[i + 1] = k;
    } else if (i > 0) {
      const int k = qr_factor(D.g, n * b, a, true, B, S);
      st.flops += fl_qr(n * b, a) + 2.0 * D.rx[i - 1] * D.n[i - 1] * k * a;
      for (long e = threadIdx.x; e < (long)k * n * b; e += NT) {
        const int rr = (int)(e / (n * b)), cc = (int)(e % (n * b));
        D.xc[i][e] = B.Q[(long)cc * k + rr];
      }
      const int c0 = D.rx[i - 1], n0 = D.n[i - 1];
      ds_gemm<double, double>(c0 * n0, k, a, crm<double>(D.xc[i - 1], a), ctr<double>(B.R, a), mrm<double>(D.g, k));
      if (threadIdx.x == 0) D.rx[i] = k;
    } else {
      wg_copy(D.g, D.xc[0], o.N);
    }
    __syncthreads();
    if (threadIdx.x == 0) D.part[0] = (double)status;
  }
  ph_add(S, PH_QR, tq);
  if (!C.sync()) return ST_ABORT;
  status = (int)D.part[0];
  // every member has read the status before member 0 writes part again
  if (!C.sync()) return ST_ABORT;
  return status;
}

__device__ void gang_als(GangCtx& C) {
  const GangDesc& D = C.D;
  SolveLDS& S = C.S;
  const int d = D.d;
  StepStats st;
  int status = ST_OK;
  int sweeps = 0;
  long nsolves = 0;
  if (threadIdx.x < 8) S.prof[threadIdx.x] = 0;
  __syncthreads();
  const long long tstart = ph_now();
  // x <- x0 (split copy of every core)
  for (int i = 0; i < d; ++i) {
    const double* src = D.x0c[i];
    double* dst = D.xc[i];
    if (src == dst) continue;
    const long len = (long)D.rx[i] * D.n[i] * D.rx[i + 1];
    const long lo = len * C.g / C.G, hi = len * (C.g + 1) / C.G;
    for (long e = lo + threadIdx.x; e < hi; e += NT) dst[e] = src[e];
  }
  bool ok = C.sync();
  bool sym = false;
  if (ok) {
    // member 0: slice symmetry, right-orthonormal start, boundary stacks
    if (C.g == 0) {
      sym = gang_slices_symmetric(D, S);
      SysDesc sd;
      sd.d = d;
      sd.n = D.n;
      sd.rx = D.rx;
      sd.xc = D.xc;
      sd.qw = D.qw;
      sd.qq = D.qq;
      sd.rf = D.rf;
      sd.tau = D.tau;
      sd.g = D.g;
      ortho_right_x<double>(sd, C.lds, 0, S);
      if (threadIdx.x == 0) {
        D.Ls[D.Loff[0]] = 1.0;
        D.Rs[D.Roff[d - 1]] = 1.0;
        D.Lbs[D.Lboff[0]] = 1.0;
        D.Rbs[D.Rboff[d - 1]] = 1.0;
        D.part[0] = sym ? 1.0 : 0.0;
      }
      __syncthreads();
    }
    ok = C.sync();
  }
  if (ok) {
    sym = D.part[0] > 0.5;
    ok = C.sync();
  }
  for (int i = d - 2; i >= 0 && ok; --i) st.flops += gang_build_right(C, i, &ok);
  while (ok && status == ST_OK && sweeps < D.repeats) {
    // warm start for position 0: the current core 0
    if (C.g == 0) wg_copy(D.xc[0], D.g, D.rx[0] * D.n[0] * D.rx[1]);
    ok = C.sync();
    for (int i = 0; i < d && ok && status == ST_OK; ++i) {
      if (i > 0) {
        const long long ts = ph_now();
        st.flops += gang_build_left(C, i);
        ph_add(S, PH_STACK, ts);
        ok = C.sync();
        if (!ok) break;
      }
      if (i < d - 1) {
        status = gang_step(C, i, true, sym, st);
        ++nsolves;
      }
    }
    for (int i = d - 1; i >= 0 && ok && status == ST_OK; --i) {
      const long long ts = ph_now();
      if (i < d - 1) st.flops += gang_build_right(C, i, &ok);
      ph_add(S, PH_STACK, ts);
      if (!ok) break;
      status = gang_step(C, i, false, sym, st);
      ++nsolves;
    }
    if (status == ST_ABORT) ok = false;
    if (ok && status == ST_OK) ++sweeps;
  }
  if (!ok) status = ST_ABORT;
  if (C.g == 0 && threadIdx.x == 0) {
    D.out[0] = 0.0;
    D.out[1] = (double)sweeps;
    D.out[2] = (double)status;
    D.out[3] = (double)st.kit;
    D.out[4] = (double)st.nunconv;
    D.out[5] = (double)nsolves;
    D.out[6] = (double)st.maxit_one;
    D.out[7] = 0.0;
    D.out[8] = st.flops;
    // profiling builds: member 0's cycles -- total, matvec, -, -, QR, stacks, -, barrier waits
    D.out[9] = (double)(ph_now() - tstart);
    for (int k = 0; k < 6; ++k) D.out[10 + k] = (double)S.prof[k];
    D.out[15] = (double)S.prof[PH_OTHER];
    for (int k = 0; k < 16; ++k) D.res[k] = D.out[k];
    for (int k = 0; k <= d; ++k) D.res[16 + k] = (double)D.rx[k];
  }
}

// Workgroup `blk` of a launch whose first workgroups form gangs of
// consecutive workgroups, one gang per system (block_sys: system of each).
// `lds` holds at least GEMM_LDS_DOUBLES.
__device__ __forceinline__ void gang_block(const GangDesc* descs, const int* block_sys, int blk, double* lds,
                                           SolveLDS& S) {
  const GangDesc& D = descs[block_sys[blk]];
  const int g = blk - D.first;
  GangCtx C{D, g, D.G, D.t1 + (long)g * D.tmax, D.t2 + (long)g * D.tmax, lds, S};
  gang_als(C);
}

}  // namespace ttals
