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
  double* gram;  // (G + 1) x rmax^2: partial Gram matrices of the split QR, then their total
  double* qrs;   // per workgroup: 3 x rmax^2 (Cholesky factor, accumulated R, product)
  long rmax2;    // rmax^2
  GangSync* sync;
  long long budget;  // as SysDesc::budget (member 0's clock decides, broadcast at a barrier)
  double* out;   // [16] as SysDesc::out
  double* res;   // compact result row: out[0..15], then the d+1 final ranks
};

// --- gang barrier ---------------------------------------------------------------

// Longest wait at one barrier, in ticks of the 100 MHz constant clock
// (s_memrealtime): 20 s.  A member may wait for the others to be dispatched
// while other kernels still hold the CUs, so the bound is in time, not spins.
constexpr unsigned long long GANG_WAIT_TICKS = 20ull * 100000000ull;
// partial-sum slots per member: 4 scalars, then two regions of 64 (the two
// Gram-Schmidt passes of GMRES, KMAX + 2 sums each; the split Householder
// QR's dots, alternating by column): a member may write one region's sums
// while another still reads the other's
constexpr int GANG_LD = 4 + 2 * 64;
static_assert(KMAX + 2 <= 64, "a Gram-Schmidt pass's sums fit one region");

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
// Sums of the G members' partials part[h ld + k0 + k], k < cnt, into out[k]
// (LDS).  Every load of a round is issued at once -- the partials sit in
// other members' L2 or in HBM, and G dependent loads would each pay that
// latency -- into buf (LDS, GEMM_LDS_DOUBLES doubles: the GEMM staging
// tiles, free between GEMMs); then each sum runs in member order, identical
// on every member.  Ends with a barrier of the workgroup.
__device__ void gang_totals(const double* part, int G, int ld, int k0, int cnt, double* buf, double* out) {
  const int chunk = GEMM_LDS_DOUBLES / G > 0 ? GEMM_LDS_DOUBLES / G : 1;
  for (int c0 = 0; c0 < cnt; c0 += chunk) {
    const int nk = cnt - c0 < chunk ? cnt - c0 : chunk;
#pragma unroll 4
    for (int idx = threadIdx.x; idx < nk * G; idx += NT) {
      const int k = idx / G, h = idx - k * G;
      buf[idx] = part[(long)h * ld + k0 + c0 + k];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nk; k += NT) {
      double sg = 0.0;
      for (int h = 0; h < G; ++h) sg += buf[k * G + h];
      out[c0 + k] = sg;
    }
    __syncthreads();
  }
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
  double* lds;   // GEMM staging tiles first, lds_cap doubles in all
  long lds_cap;
  SolveLDS& S;
  // set once a Cholesky QR of this system failed (a numerically
  // rank-deficient panel): its later panels go to Householder directly
  bool chol_off = false;
  __device__ bool sync() {
    const long long t0 = ph_now();
    const bool r = gang_sync(D.sync, G);
    ph_add(S, PH_OTHER, t0);  // barrier waits (profiling builds)
    return r;
  }
};

// QR buffers of member 0's (whole-panel) Householder QR: the blocked form
// when its block fits the member's LDS and its T storage the Gram slots
// (free outside the Cholesky QR), else the column sweep.
__device__ __forceinline__ QRBuf gang_qrbuf(const GangCtx& C, int c) {
  const GangDesc& D = C.D;
  QRBuf B{D.qw, D.qq, D.rf, D.tau};
  if (((long)c + 16) * 16 <= (long)(C.G + 1) * D.rmax2) {
    B.lds = C.lds;
    B.lds_cap = C.lds_cap;
    B.tws = D.gram;
  }
  return B;
}
__device__ __forceinline__ bool gang_blocked_qr_fits(const GangCtx& C, int m, int c) {
  const long lim = C.lds_cap < HHQR_LDS_CAP ? C.lds_cap : HHQR_LDS_CAP;
  return m >= c && HHBlockLDS<4>::need(m) <= lim && ((long)c + 16) * 16 <= (long)(C.G + 1) * C.D.rmax2;
}

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
  double* tot = C.S.scal + 2;  // two LDS slots for scalar totals
  gang_totals(part, G, ld, 0, 2, C.lds, tot);
  ff = tot[0];
  rr = tot[1];
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
    gang_totals(part, G, ld, 2, 1, C.lds, tot);
    pq = tot[0];
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
    gang_totals(part, G, ld, 3, 1, C.lds, tot);
    rn = tot[0];
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
// work vector, V the (m + 1) x N basis.  A run that converges in one or two
// steps returns without the confirming residual matvec (the least-squares
// residual of an orthonormal basis is the true one to rounding) and counts
// it in *unconfirmed.
template <class T>
__device__ int gang_gmres(GangCtx& C, const GangLocal& o, const T* L, const T* A, const T* R, T* x, const T* f,
                          T* w, T* V, double tol, int maxit, int m, int* iters, double* vfl, int* unconfirmed) {
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
  double* tot = S.scal + 2;  // two LDS slots for scalar totals (S.scal[0] is the Givens residual)
  gang_totals(part, G, ld, 0, 1, C.lds, tot);
  const double fn = sqrt(tot[0]);
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
    gang_totals(part, G, ld, 1, 1, C.lds, tot);
    const double beta = sqrt(tot[0]);
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
        gang_totals(part, G, ld, base, j + 2, C.lds, S.h);
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
      gang_totals(part, G, ld, 2, 1, C.lds, tot);
      hn = sqrt(tot[0]);
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
    if (res <= tol * fn && jend <= 2) {
      ++*unconfirmed;
      return 0;
    }
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
      gang_totals(part, G, ld, 1, 1, C.lds, tot);
      return sqrt(tot[0]) <= tol * fn ? 0 : 2;
    }
  }
}


// --- the sweep ------------------------------------------------------------------------

// Time budget, decided by member 0 and broadcast: 1 over budget, 0 not,
// -1 the gang aborted.  With a budget, one barrier: member 0 stores its
// decision in the sync record before arriving; every member reads it after
// the barrier (the next decision is stored only after further barriers of
// the step, which every member must pass first).
__device__ int gang_over_budget(GangCtx& C, unsigned long long t0) {
  const GangDesc& D = C.D;
  if (D.budget <= 0) return 0;
  if (C.g == 0 && threadIdx.x == 0) {
    const unsigned v = (wall_clock64() - t0 > (unsigned long long)D.budget) ? 1u : 0u;
    __hip_atomic_store(&D.sync->pad, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (!C.sync()) return -1;
  if (threadIdx.x == 0)
    C.S.flag[3] = (int)__hip_atomic_load(&D.sync->pad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return C.S.flag[3] != 0 ? 1 : 0;
}

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
// factors it (identical inputs and order: identical R), inverts the factor
// and multiplies its own rows by the inverse (one GEMM); the first pass
// shifts the Gram matrix by 11 (m c + c (c + 1)) u ||W||_F^2 so that it stays
// positive definite for condition numbers up to ~1/u.  The c x c work runs in
// LDS when 2 c^2 + c doubles fit after the GEMM tiles, else in the member's
// global slots.  The panel is copied from g: forward W[row, col] =
// g[row c + col] (m = a n, c = b), backward W[row, col] = g[col m + row]
// (m = n b, c = a).  On success *Q is the panel's Q (own rows valid, row
// stride c, in D.qw or D.qq) and *R the upper c x c factor; returns 0, or 1
// when a pivot failed (every member alike), 2 non-finite, -1 aborted.
__device__ int gang_cholqr(GangCtx& C, const double* g, int m, int c, bool transposed, double* fl,
                           const double** Qout, const double** Rout) {
  const GangDesc& D = C.D;
  int m0, m1;
  gang_span(m, C.G, C.g, m0, m1);
  const int rows = m1 - m0;
  const long cc = (long)c * c;
  double *Rp, *Rt, *tmp;
  if (GEMM_LDS_DOUBLES + 2 * cc + c <= C.lds_cap) {
    Rp = C.lds + GEMM_LDS_DOUBLES;
    Rt = Rp + cc;
    tmp = Rt + cc;
  } else {
    Rp = D.qrs + (long)C.g * 3 * D.rmax2;
    Rt = Rp + D.rmax2;
    tmp = Rt + D.rmax2;
  }
  double* Win = D.qw;
  double* Wout = D.qq;
  for (long e = (long)m0 * c + threadIdx.x; e < (long)m1 * c; e += NT) {
    const long row = e / c, col = e - row * c;
    Win[e] = transposed ? g[col * m + row] : g[e];
  }
  for (long e = threadIdx.x; e < cc; e += NT) Rt[e] = (e / c == e % c) ? 1.0 : 0.0;
  __syncthreads();
  for (int pass = 0; pass < 3; ++pass) {
    const long long tg = 0;
    // partial Gram of this member's rows
    double* slot = D.gram + (long)C.g * D.rmax2;
    if (rows > 0) {
      wg_gemm<double, double>(c, c, rows, ctr<double>(Win + (long)m0 * c, c), crm<double>(Win + (long)m0 * c, c),
                              mrm<double>(slot, c), C.lds);
    } else {
      for (long e = threadIdx.x; e < cc; e += NT) slot[e] = 0.0;
    }
    if (!C.sync()) return -1;
    // total Gram: each member sums its slice of the c^2 entries over the
    // members, one barrier, then every member reads the whole total into Rp
    {
      const long s0 = cc * C.g / C.G, s1 = cc * (C.g + 1) / C.G;
      double* total = D.gram + (long)C.G * D.rmax2;
      gang_totals(D.gram, C.G, (int)D.rmax2, (int)s0, (int)(s1 - s0), C.lds, total + s0);
      if (!C.sync()) return -1;
      for (long e = threadIdx.x; e < cc; e += NT) Rp[e] = total[e];
      __syncthreads();
    }
    (void)tg;
    if (pass == 0) {
      double tr = 0.0;
      for (int j = 0; j < c; ++j) tr += Rp[(long)j * c + j];
      if (!isfinite(tr)) return 2;
      const double shift = cholqr_shift(m, c, tr);
      __syncthreads();
      for (int j = threadIdx.x; j < c; j += NT) Rp[(long)j * c + j] += shift;
      __syncthreads();
    }
    // Cholesky, R accumulation and the inverse factor (identical in every member)
    if (!cholqr_small_step(Rp, Rt, tmp, c)) return 1;
    // own rows: W <- W R^{-1}
    if (rows > 0)
      wg_gemm<double, double>(rows, c, c, crm<double>(Win + (long)m0 * c, c), crm<double>(Rp, c),
                              mrm<double>(Wout + (long)m0 * c, c), C.lds);
    __syncthreads();
    double* t = Win;
    Win = Wout;
    Wout = t;
  }
  *fl += 3.0 * (4.0 * (double)m * c * c + 2.0 * (double)c * c * c);
  *Qout = Win;
  *Rout = Rt;
  return 0;
}

// out[q] = sum over rows i in [r_lo, r_hi) of v_i X[i][col0 + q], q < nq,
// with v_i = 1 at i = j and V[i][j] below it (a Householder vector stored in
// column j of V; rows of stride c).  Threads split (row group, column);
// buf is NT doubles of LDS.  Ends with a barrier.
__device__ void rows_dots(const double* V, const double* X, int c, int j, int r_lo, int r_hi, int col0, int nq,
                          double* buf, double* out) {
  const int groups = NT / nq;
  const int q = (int)threadIdx.x % nq, grp = (int)threadIdx.x / nq;
  double acc = 0.0;
  if (grp < groups)
    for (int i = r_lo + grp; i < r_hi; i += groups) {
      const double v = (i == j) ? 1.0 : V[(long)i * c + j];
      acc += v * X[(long)i * c + col0 + q];
    }
  buf[threadIdx.x] = grp < groups ? acc : 0.0;
  __syncthreads();
  for (int t = threadIdx.x; t < nq; t += NT) {
    double sg = 0.0;
    for (int h = 0; h < groups; ++h) sg += buf[h * nq + t];
    out[t] = sg;
  }
  __syncthreads();
}

// Split Householder QR of the panel (m x c, m >= c, c <= 64) over the gang,
// for panels whose Gram matrix is numerically singular (the Cholesky QR
// fails): the panel of the reference's own QR (numpy.linalg.qr, Householder)
// and the same completion of a rank-deficient basis, up to rounding.
// Column j: the norm below the diagonal and the pivot (one gang reduction),
// then the reflector's dots with the trailing columns (a second one); Q is
// formed by applying the reflectors to [I; 0] backwards (one reduction per
// column).  Rows as in gang_cholqr; *Q in D.qq (own rows), *R in D.rf.
// Returns 0, or -1 when the gang aborted.
__device__ int gang_hhqr(GangCtx& C, const double* g, int m, int c, bool transposed, double* fl, const double** Qout,
                         const double** Rout) {
  const GangDesc& D = C.D;
  SolveLDS& S = C.S;
  const int G = C.G, ld = GANG_LD;
  int m0, m1;
  gang_span(m, G, C.g, m0, m1);
  double* W = D.qw;
  double* Qm = D.qq;
  double* part = D.part;
  double* tot = S.scal + 2;
  double* wv = S.h;       // c <= 64 <= LKMAX + 2 values
  double* taus = S.alpha;  // c <= LKMAX + 1
  for (long e = (long)m0 * c + threadIdx.x; e < (long)m1 * c; e += NT) {
    const long row = e / c, col = e - row * c;
    W[e] = transposed ? g[col * m + row] : g[e];
  }
  __syncthreads();
  for (int j = 0; j < c; ++j) {
    double s2 = 0.0, al = 0.0, u = 0.0;
    for (int i = (m0 > j + 1 ? m0 : j + 1) + threadIdx.x; i < m1; i += NT) {
      const double t = W[(long)i * c + j];
      s2 += t * t;
    }
    if (threadIdx.x == 0 && j >= m0 && j < m1) al = W[(long)j * c + j];
    wg_sum3(s2, al, u, S.red);
    // slots alternate by column: when column j has no reflector there is no
    // second barrier before column j + 1 writes its sums
    const int ns = 2 * (j & 1);
    if (threadIdx.x == 0) {
      part[(long)C.g * ld + ns] = s2;
      part[(long)C.g * ld + ns + 1] = al;
    }
    if (!C.sync()) return -1;
    gang_totals(part, G, ld, ns, 2, C.lds, tot);
    s2 = tot[0];
    const double alpha = tot[1];
    double t = 0.0;
    if (s2 != 0.0) {  // else H_j = I (LAPACK convention for an already-reduced column)
      const double nrm = sqrt(alpha * alpha + s2);
      const double beta = alpha >= 0.0 ? -nrm : nrm;
      t = (beta - alpha) / beta;
      const double sc = 1.0 / (alpha - beta);
      for (int i = (m0 > j + 1 ? m0 : j + 1) + threadIdx.x; i < m1; i += NT) W[(long)i * c + j] *= sc;
      if (threadIdx.x == 0 && j >= m0 && j < m1) W[(long)j * c + j] = beta;
    }
    if (threadIdx.x == 0) taus[j] = t;
    __syncthreads();
    const int nq = c - j - 1;
    if (t != 0.0 && nq > 0) {
      const int lo = m0 > j ? m0 : j;
      rows_dots(W, W, c, j, lo, m1, j + 1, nq, C.lds, wv);
      const int rg = 4 + 64 * (j & 1);
      for (int q = threadIdx.x; q < nq; q += NT) part[(long)C.g * ld + rg + q] = wv[q];
      if (!C.sync()) return -1;
      gang_totals(part, G, ld, rg, nq, C.lds, wv);
      for (long e = threadIdx.x; e < (long)(m1 > lo ? m1 - lo : 0) * nq; e += NT) {
        const int i = lo + (int)(e / nq), q = (int)(e % nq);
        const double v = (i == j) ? 1.0 : W[(long)i * c + j];
        W[(long)i * c + j + 1 + q] -= t * v * wv[q];
      }
      __syncthreads();
    }
  }
  // R: the owners of rows 0 .. c-1 publish them
  for (long e = (long)m0 * c + threadIdx.x; e < (long)(m1 < c ? m1 : c) * c; e += NT) {
    const long row = e / c, col = e - row * c;
    D.rf[e] = col >= row ? W[e] : 0.0;
  }
  // Q = H_0 ... H_{c-1} [I; 0], own rows, backwards
  for (long e = (long)m0 * c + threadIdx.x; e < (long)m1 * c; e += NT) {
    const long row = e / c, col = e - row * c;
    Qm[e] = row == col ? 1.0 : 0.0;
  }
  __syncthreads();
  int qcount = 0;
  for (int j = c - 1; j >= 0; --j) {
    const double t = taus[j];
    if (t == 0.0) continue;
    const int nq = c - j;  // columns j .. c-1
    const int lo = m0 > j ? m0 : j;
    rows_dots(W, Qm, c, j, lo, m1, j, nq, C.lds, wv);
    // one barrier per column: consecutive columns use the two regions in turn
    const int rg = 4 + 64 * (qcount++ & 1);
    for (int q = threadIdx.x; q < nq; q += NT) part[(long)C.g * ld + rg + q] = wv[q];
    if (!C.sync()) return -1;
    gang_totals(part, G, ld, rg, nq, C.lds, wv);
    for (long e = threadIdx.x; e < (long)(m1 > lo ? m1 - lo : 0) * nq; e += NT) {
      const int i = lo + (int)(e / nq), q = (int)(e % nq);
      const double v = (i == j) ? 1.0 : W[(long)i * c + j];
      Qm[(long)i * c + j + q] -= t * v * wv[q];
    }
    __syncthreads();
  }
  if (!C.sync()) return -1;  // R published, Q complete
  *fl += 4.0 * (double)m * c * c - 4.0 / 3.0 * (double)c * c * c;
  *Qout = Qm;
  *Rout = D.rf;
  return 0;
}

// Split blocked Householder QR of the panel (m x c, m >= c) over the gang:
// the structure of wg_hhqr_blocked, with every member factoring each column
// block in its own LDS (identical inputs and order: identical reflectors, no
// exchange), the trailing update split over the members by columns (one gang
// barrier per block), and Q formed in one step from the compact WY form of
// all reflectors, each member computing its own rows.  Rows of the panel as
// in gang_cholqr; *Q in D.qq (every row), *R in D.rf.  Returns 0, or -1 when
// the gang aborted.  Callers check gang_split_qr_fits first.
__device__ __forceinline__ bool gang_split_qr_fits(const GangCtx& C, int m, int c) {
  const long lim = C.lds_cap < HHQR_LDS_CAP ? C.lds_cap : HHQR_LDS_CAP;
  return m >= c && c <= NT && m < (1 << 22) / 17 && HHBlockLDS<4>::need(m) <= lim &&
         GEMM_LDS_DOUBLES + 2L * c * c <= lim;
}

template <int NB>
__device__ int gang_hhqr_run(GangCtx& C, int m, int c) {
  const GangDesc& D = C.D;
  constexpr int ld = NB + 1;
  double* W = D.qw;
  double* Pn = C.lds;
  double* dred = Pn + (long)m * ld;
  double* Yb = dred + NWAVE * NB;
  double* T = Yb + NB * 64;
  double* tau = T + NB * NB;
  double* G = Yb + NB * NB + NB;  // the block's Gram rows (after the pivot rows and w_q)
  const int k = c;
  for (int jb = 0; jb < k; jb += NB) {
    const int kb = (k - jb) < NB ? (k - jb) : NB, rows = m - jb;
    hh_load_block<NB>(W, c, jb, rows, kb, Pn);
    __syncthreads();
    hh_block_factor<NB>(Pn, rows, kb, tau, G, dred, Yb);
    if (C.g == 0) {
      for (int i = threadIdx.x; i < rows; i += NT) {
        double* dst = W + (long)(jb + i) * c + jb;
#pragma unroll
        for (int q = 0; q < NB; ++q)
          if (q < kb) dst[q] = Pn[i * ld + q];
      }
      if (threadIdx.x < kb) D.tau[jb + threadIdx.x] = tau[threadIdx.x];
    }
    hh_block_t<NB>(tau, G, kb, T);
    const int c2 = c - jb - kb;
    if (c2 > 0) {
      int j0, j1;
      gang_span(c2, C.G, C.g, j0, j1);
      if (j1 > j0) hh_block_apply<NB>(W + (long)jb * c + jb + kb + j0, c, rows, j1 - j0, Pn, kb, T, true, Yb);
    }
    __syncthreads();
    if (!C.sync()) return -1;
  }
  // member 0: R, and W made the explicit V (unit diagonal, zeros above)
  if (C.g == 0) {
    for (long e = threadIdx.x; e < (long)k * c; e += NT) {
      const int i = (int)(e / c), j = (int)(e - (long)i * c);
      D.rf[e] = (j >= i) ? W[e] : 0.0;
      if (j >= i) W[e] = (j == i) ? 1.0 : 0.0;
    }
  }
  if (!C.sync()) return -1;
  // every member: G = V^T V, T (dlarft), Ms = -T V1^T; then its own rows of
  // Q = V Ms + [I; 0]
  double* Gs = C.lds + GEMM_LDS_DOUBLES;
  double* Ms = Gs + (long)k * k;
  wg_gemm_mf(k, k, m, ctr<double>(W, c), crm<double>(W, c), mrm<double>(Gs, k), C.lds);
  for (int i = 0; i < k; ++i) {
    const int j = threadIdx.x;
    double v = 0.0;
    if (j < i)
      for (int l = j; l < i; ++l) v += Gs[j * k + l] * Gs[l * k + i];
    __syncthreads();
    if (j < i) Gs[j * k + i] = -D.tau[i] * v;
    if (j == i) Gs[i * k + i] = D.tau[i];
    __syncthreads();
  }
  for (int e = threadIdx.x; e < k * k; e += NT) {
    const int i = e / k, j = e - i * k;
    if (i > j) Gs[e] = 0.0;
  }
  __syncthreads();
  wg_gemm_mf(k, k, k, crm<double>(Gs, k), ctr<double>(W, c), mrm<double>(Ms, k), C.lds);
  for (int e = threadIdx.x; e < k * k; e += NT) Ms[e] = -Ms[e];
  __syncthreads();
  int m0, m1;
  gang_span(m, C.G, C.g, m0, m1);
  if (m1 > m0)
    wg_gemm_mf(m1 - m0, k, k, crm<double>(W + (long)m0 * c, c), crm<double>(Ms, k), mrm<double>(D.qq + (long)m0 * k, k),
               C.lds);
  for (int j = m0 + (int)threadIdx.x; j < (m1 < k ? m1 : k); j += NT) D.qq[(long)j * k + j] += 1.0;
  __syncthreads();
  return 0;
}

__device__ int gang_hhqr_blocked(GangCtx& C, const double* g, int m, int c, bool transposed, double* fl,
                                 const double** Qout, const double** Rout) {
  const GangDesc& D = C.D;
  int m0, m1;
  gang_span(m, C.G, C.g, m0, m1);
  for (long e = (long)m0 * c + threadIdx.x; e < (long)m1 * c; e += NT) {
    const long row = e / c, col = e - row * c;
    D.qw[e] = transposed ? g[col * m + row] : g[e];
  }
  __syncthreads();
  if (!C.sync()) return -1;
  const long lim = C.lds_cap < HHQR_LDS_CAP ? C.lds_cap : HHQR_LDS_CAP;
  const int r = HHBlockLDS<8>::need(m) <= lim ? gang_hhqr_run<8>(C, m, c) : gang_hhqr_run<4>(C, m, c);
  if (r < 0) return -1;
  if (!C.sync()) return -1;  // Q rows and R complete
  *fl += 4.0 * (double)m * c * c - 4.0 / 3.0 * (double)c * c * c;
  *Qout = D.qq;
  *Rout = D.rf;
  return 0;
}

// Right-orthonormal start x0 (the sle engine solves from it): for i = d-1 .. 1
// the split QR of core i's transposed unfolding (n b x a), Q^T into core i,
// core i-1 <- core i-1 R^T (split over its rows, via g); a panel with fewer
// rows than columns, or a failed pivot, runs Householder QR on member 0
// (blocked where the panel fits; which may lower the rank, as the
// reference's QR does) or, for a rank-deficient panel too tall for that, the
// split column sweep.  After one failed pivot the system skips the Cholesky
// QR.  Returns false once the gang has aborted.
__device__ bool gang_ortho_right(GangCtx& C, double* fl) {
  const GangDesc& D = C.D;
  SolveLDS& S = C.S;
  for (int i = D.d - 1; i >= 1; --i) {
    const int a = D.rx[i], n = D.n[i], b = D.rx[i + 1];
    const int m = n * b, c = a;
    const int c0 = D.rx[i - 1], n0 = D.n[i - 1], cn = c0 * n0;
    if (m >= c && (long)c * c <= D.rmax2 && !C.chol_off) {
      const double *Q, *R;
      const int q = gang_cholqr(C, D.xc[i], m, c, true, fl, &Q, &R);
      if (q < 0) return false;
      if (q == 0) {
        int m0, m1;
        gang_span(m, C.G, C.g, m0, m1);
        for (long e = (long)m0 * c + threadIdx.x; e < (long)m1 * c; e += NT) {
          const long row = e / c, col = e - row * c;
          D.xc[i][col * m + row] = Q[e];
        }
        int r0, r1;
        gang_span(cn, C.G, C.g, r0, r1);
        if (r1 > r0)
          wg_gemm<double, double>(r1 - r0, c, a, crm<double>(D.xc[i - 1] + (long)r0 * a, a), ctr<double>(R, c),
                                  mrm<double>(D.g + (long)r0 * c, c), C.lds);
        __syncthreads();
        for (long e = (long)r0 * c + threadIdx.x; e < (long)r1 * c; e += NT) D.xc[i - 1][e] = D.g[e];
        *fl += 2.0 * cn * c * a;
        if (!C.sync()) return false;
        continue;
      }
      if (!C.sync()) return false;  // a pivot failed on every member alike
      if (q == 1) C.chol_off = true;
    }
    const bool split_hh = gang_split_qr_fits(C, m, c);
    if (C.chol_off && m >= c &&
        (split_hh || (c <= 64 && (long)c * c <= D.rmax2 && !gang_blocked_qr_fits(C, m, c)))) {
      // a numerically rank-deficient panel: the split blocked QR where it
      // fits, else the split column sweep
      const double *Q2, *R2;
      if ((split_hh ? gang_hhqr_blocked(C, D.xc[i], m, c, true, fl, &Q2, &R2)
                    : gang_hhqr(C, D.xc[i], m, c, true, fl, &Q2, &R2)) < 0)
        return false;
      int m0, m1;
      gang_span(m, C.G, C.g, m0, m1);
      for (long e = (long)m0 * c + threadIdx.x; e < (long)m1 * c; e += NT) {
        const long row = e / c, col = e - row * c;
        D.xc[i][col * m + row] = Q2[e];
      }
      int r0, r1;
      gang_span(cn, C.G, C.g, r0, r1);
      if (r1 > r0)
        wg_gemm<double, double>(r1 - r0, c, a, crm<double>(D.xc[i - 1] + (long)r0 * a, a), ctr<double>(R2, c),
                                mrm<double>(D.g + (long)r0 * c, c), C.lds);
      __syncthreads();
      for (long e = (long)r0 * c + threadIdx.x; e < (long)r1 * c; e += NT) D.xc[i - 1][e] = D.g[e];
      *fl += 2.0 * cn * c * a;
      if (!C.sync()) return false;
      continue;
    }
    if (C.g == 0) {
      const QRBuf B = gang_qrbuf(C, a);
      const int k = qr_factor(D.xc[i], m, a, true, B, S);
      wg_transpose(B.Q, m, k, D.xc[i]);
      if (B.lds)
        wg_gemm<double, double>(cn, k, a, crm<double>(D.xc[i - 1], a), ctr<double>(B.R, a), mrm<double>(D.g, k),
                                C.lds);
      else
        ds_gemm<double, double>(cn, k, a, crm<double>(D.xc[i - 1], a), ctr<double>(B.R, a), mrm<double>(D.g, k));
      wg_copy(D.g, D.xc[i - 1], cn * k);
      if (threadIdx.x == 0) D.rx[i] = k;
      *fl += fl_qr(m, a) + 2.0 * cn * k * a;
      __syncthreads();
    }
    if (!C.sync()) return false;
  }
  return true;
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
  int unconf = 0;  // GMRES runs that ended without the confirming matvec
  double vfl = 0.0;
  if (D.prec == 0) {
    r = 1;
    if (sym) r = gang_cg<double>(C, o, o.L, o.A, o.R, D.g, D.f, D.r, D.p, D.q, D.tol, D.maxit, &its, &vfl);
    if (r == 1) r = gang_gmres<double>(C, o, o.L, o.A, o.R, D.g, D.f, D.q, D.V, D.tol, D.maxit, D.gm_restart, &its,
                                        &vfl, &unconf);
    st.flops += (its + 2 - unconf) * fmv + vfl;
    st.mvb += (its + 2 - unconf) * by_op(a, o.P, n, o.Q, b, 8.0, true) + 8.0 * vfl;
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
      gang_totals(D.part, C.G, ld, 0, 2, C.lds, S.scal + 2);
      const double fn = sqrt(S.scal[2]), rn = sqrt(S.scal[3]);
      vfl += 5.0 * o.N;
      st.flops += fmv;
      st.mvb += by_op(a, o.P, n, o.Q, b, 8.0, true) + 40.0 * o.N;
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
      unconf = 0;
      const double vfl0 = vfl;
      if (sym)
        ri = gang_cg<float>(C, o, D.L32, D.A32, D.R32, D.x32, D.r32, D.p32, D.q32, reinterpret_cast<float*>(D.V),
                            D.tol_inner, D.maxit_inner, &ii, &vfl);
      if (ri == 1)
        ri = gang_gmres<float>(C, o, D.L32, D.A32, D.R32, D.x32, D.r32, D.q32, D.V32, D.tol_inner, D.maxit_inner,
                               D.gm_restart, &ii, &vfl, &unconf);
      if (ri < 0) return ST_ABORT;
      inner_total += ii;
      st.flops += (ii + 2 - unconf) * fmv;
      st.mvb += (ii + 2 - unconf) * by_op(a, o.P, n, o.Q, b, 4.0, true) + 4.0 * (vfl - vfl0);
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
  // Cholesky QR over the gang for tall panels, else (or when a pivot fails,
  // and for the rest of the system's panels after one has) Householder QR on
  // member 0, blocked where the panel fits, or the split column sweep for a
  // rank-deficient panel too tall for that
  int status = ST_OK;
  const long long tq = ph_now();
  if (i > 0 || forward) {
    const int m = forward ? a * n : n * b, c = forward ? b : a;
    const bool split_hh = gang_split_qr_fits(C, m, c);
    if (m >= c && (long)c * c <= D.rmax2 &&
        !(C.chol_off && !split_hh && (c > 64 || gang_blocked_qr_fits(C, m, c)))) {
      double qfl = 0.0;
      const double *Q = nullptr, *Rt = nullptr;
      const long long t_cq = ph_now();
      // Cholesky QR unless an earlier panel of this system was rank deficient
      const int q = C.chol_off ? 1 : gang_cholqr(C, D.g, m, c, !forward, &qfl, &Q, &Rt);
      ph_add(S, PH_ORTH, t_cq);  // profiling builds: the split QR itself
      const long long t_ws = ph_now();
      if (q < 0) return ST_ABORT;
      if (q == 2) status = ST_NONFINITE;
      if (q == 0) {
        int m0, m1;
        gang_span(m, C.G, C.g, m0, m1);
        // Q (own rows) into the core
        for (long e = (long)m0 * c + threadIdx.x; e < (long)m1 * c; e += NT) {
          const long row = e / c, col = e - row * c;
          if (forward) D.xc[i][e] = Q[e];
          else D.xc[i][col * m + row] = Q[e];
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
        ph_add(S, PH_TRI, t_ws);  // profiling builds: Q store and warm start
        ph_add(S, PH_QR, tq);
        if (!C.sync()) return ST_ABORT;
        return status;
      }
      // a pivot failed (or non-finite, or skipped): every member got here alike
      if (!C.sync()) return ST_ABORT;
      if (q == 1) C.chol_off = true;
      if (q == 1 && (split_hh || (c <= 64 && !gang_blocked_qr_fits(C, m, c)))) {
        const double* Q2 = nullptr;
        const double* R2 = nullptr;
        if ((split_hh ? gang_hhqr_blocked(C, D.g, m, c, !forward, &qfl, &Q2, &R2)
                      : gang_hhqr(C, D.g, m, c, !forward, &qfl, &Q2, &R2)) < 0)
          return ST_ABORT;
        int m0, m1;
        gang_span(m, C.G, C.g, m0, m1);
        for (long e = (long)m0 * c + threadIdx.x; e < (long)m1 * c; e += NT) {
          const long row = e / c, col = e - row * c;
          if (forward) D.xc[i][e] = Q2[e];
          else D.xc[i][col * m + row] = Q2[e];
        }
        if (!C.sync()) return ST_ABORT;
        if (forward) {
          const int n1 = D.n[i + 1], c1 = D.rx[i + 2], nc = n1 * c1;
          int j0, j1;
          gang_span(nc, C.G, C.g, j0, j1);
          if (j1 > j0)
            wg_gemm<double, double>(c, j1 - j0, b, crm<double>(R2, c), crm<double>(D.xc[i + 1] + j0, nc),
                                    mrm<double>(D.g + j0, nc), C.lds);
          st.flops += qfl + 2.0 * c * nc * b;
        } else {
          const int c0 = D.rx[i - 1], n0 = D.n[i - 1], cn = c0 * n0;
          int r0, r1;
          gang_span(cn, C.G, C.g, r0, r1);
          if (r1 > r0)
            wg_gemm<double, double>(r1 - r0, c, a, crm<double>(D.xc[i - 1] + (long)r0 * a, a), ctr<double>(R2, c),
                                    mrm<double>(D.g + (long)r0 * c, c), C.lds);
          st.flops += qfl + 2.0 * cn * c * a;
        }
        ph_add(S, PH_QR, tq);
        if (!C.sync()) return ST_ABORT;
        return status;
      }
    }
  }
  if (C.g == 0) {
    const QRBuf B = gang_qrbuf(C, forward ? b : a);
    const double nv = wg_nrm2(D.g, o.N, S.red);
    if (!isfinite(nv)) {
      status = ST_NONFINITE;
    } else if (forward) {
      const int k = qr_factor(D.g, a * n, b, false, B, S);
      wg_copy4(B.Q, D.xc[i], a * n * k);
      const int n1 = D.n[i + 1], c1 = D.rx[i + 2], nc = n1 * c1;
      st.flops += fl_qr(a * n, b) + 2.0 * k * n1 * c1 * b;
      if (B.lds)
        wg_gemm<double, double>(k, nc, b, crm<double>(B.R, b), crm<double>(D.xc[i + 1], nc), mrm<double>(D.g, nc),
                                C.lds);
      else
        ds_gemm<double, double>(k, nc, b, crm<double>(B.R, b), crm<double>(D.xc[i + 1], nc), mrm<double>(D.g, nc));
      if (threadIdx.x == 0) D.rx[i + 1] = k;
    } else if (i > 0) {
      const int k = qr_factor(D.g, n * b, a, true, B, S);
      st.flops += fl_qr(n * b, a) + 2.0 * D.rx[i - 1] * D.n[i - 1] * k * a;
      wg_transpose(B.Q, n * b, k, D.xc[i]);
      const int c0 = D.rx[i - 1], n0 = D.n[i - 1];
      if (B.lds)
        wg_gemm<double, double>(c0 * n0, k, a, crm<double>(D.xc[i - 1], a), ctr<double>(B.R, a), mrm<double>(D.g, k),
                                C.lds);
      else
        ds_gemm<double, double>(c0 * n0, k, a, crm<double>(D.xc[i - 1], a), ctr<double>(B.R, a),
                                mrm<double>(D.g, k));
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
  const unsigned long long t_begin = wall_clock64();  // member 0, thread 0 decides the time budget
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
    // right-orthonormal start (split QR over the gang)
    ok = gang_ortho_right(C, &st.flops);
  }
  if (ok) {
    // member 0: slice symmetry, boundary stacks
    if (C.g == 0) {
      sym = gang_slices_symmetric(D, S);
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
        const int ob = gang_over_budget(C, t_begin);
        if (ob < 0) ok = false;
        if (ob != 0) {
          if (ob > 0) status = ST_DEADLINE;
          break;
        }
        status = gang_step(C, i, true, sym, st);
        ++nsolves;
      }
    }
    for (int i = d - 1; i >= 0 && ok && status == ST_OK; --i) {
      const long long ts = ph_now();
      if (i < d - 1) st.flops += gang_build_right(C, i, &ok);
      ph_add(S, PH_STACK, ts);
      if (!ok) break;
      const int ob = gang_over_budget(C, t_begin);
      if (ob < 0) ok = false;
      if (ob != 0) {
        if (ob > 0) status = ST_DEADLINE;
        break;
      }
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
    D.out[16] = (double)(wall_clock64() - t_begin);  // constant-clock ticks of the gang's solve
    D.out[17] = st.mvb;
    for (int k = 0; k < RES_HEAD; ++k) D.res[k] = D.out[k];
    for (int k = 0; k <= d; ++k) D.res[RES_HEAD + k] = (double)D.rx[k];
  }
}

// Workgroup `blk` of a launch whose first workgroups form gangs of
// consecutive workgroups, one gang per system (block_sys: system of each).
// `lds` holds lds_cap >= GEMM_LDS_DOUBLES doubles.
__device__ __forceinline__ void gang_block(const GangDesc* descs, const int* block_sys, int blk, double* lds,
                                           long lds_cap, SolveLDS& S) {
  const GangDesc& D = descs[block_sys[blk]];
  const int g = blk - D.first;
  GangCtx C{D, g, D.G, D.t1 + (long)g * D.tmax, D.t2 + (long)g * D.tmax, lds, lds_cap, S};
  gang_als(C);
}

}  // namespace ttals
