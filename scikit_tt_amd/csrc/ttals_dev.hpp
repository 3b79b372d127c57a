// Device building blocks of the TT-ALS engine (gfx950, wave64).
//
// Every routine here is executed by ONE whole workgroup of NT threads and
// works on operands in global memory (HBM, L2-resident at these sizes),
// staging tiles through LDS.  Routines end with a workgroup barrier so the
// next routine sees their global writes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ttals {

constexpr int NT = 256;            // threads per workgroup: 4 x wave64
constexpr int NWAVE = NT / 64;

// ---------------------------------------------------------------------------
// index maps: a matrix row (or column) index i is split into
// (i / div, i % div) and mapped to  (i / div) * hi + (i % div) * lo.
// This expresses every reshape/transpose of the interface-stack
// contractions as a plain GEMM operand without copies.
struct IMap {
  int div, hi, lo;
};
__host__ __device__ inline IMap plain(int stride) { return IMap{1 << 30, 0, stride}; }
__host__ __device__ inline IMap split(int div, int hi, int lo) { return IMap{div, hi, lo}; }
__device__ __forceinline__ long imap(const IMap& m, int i) {
  int h = i / m.div;
  return (long)h * m.hi + (long)(i - h * m.div) * m.lo;
}

template <class T>
struct CView {  // read-only matrix view
  const T* p;
  IMap r, c;
  __device__ __forceinline__ T at(int i, int j) const { return p[imap(r, i) + imap(c, j)]; }
};
template <class T>
struct MView {  // writable matrix view
  T* p;
  IMap r, c;
  __device__ __forceinline__ T& at(int i, int j) const { return p[imap(r, i) + imap(c, j)]; }
};

template <class T>
__device__ __forceinline__ CView<T> cview(const T* p, IMap r, IMap c) { return CView<T>{p, r, c}; }
template <class T>
__device__ __forceinline__ MView<T> mview(T* p, IMap r, IMap c) { return MView<T>{p, r, c}; }
// row-major (rows x ld) matrix
template <class T>
__device__ __forceinline__ CView<T> crm(const T* p, int ld) { return CView<T>{p, plain(ld), plain(1)}; }
template <class T>
__device__ __forceinline__ MView<T> mrm(T* p, int ld) { return MView<T>{p, plain(ld), plain(1)}; }
// transpose of a row-major (rows x ld) matrix
template <class T>
__device__ __forceinline__ CView<T> ctr(const T* p, int ld) { return CView<T>{p, plain(1), plain(ld)}; }

// ---------------------------------------------------------------------------
// Memory-space accessors.  Mem<true> reads/writes through an LDS-qualified
// pointer (the address must lie in the workgroup's LDS) so the compiler
// emits ds_read/ds_write instead of flat accesses; Mem<false> is a plain
// (global or generic) access.
#define TT_LDS __attribute__((address_space(3)))
template <bool OnChip>
struct Mem {
  template <class T>
  static __device__ __forceinline__ T ld(const T* p, long i) { return p[i]; }
  template <class T>
  static __device__ __forceinline__ void st(T* p, long i, T v) { p[i] = v; }
};
template <>
struct Mem<true> {
  template <class T>
  static __device__ __forceinline__ T ld(const T* p, long i) { return ((const TT_LDS T*)p)[i]; }
  template <class T>
  static __device__ __forceinline__ void st(T* p, long i, T v) { ((TT_LDS T*)p)[i] = v; }
};

// ---------------------------------------------------------------------------
// reductions (fixed order: deterministic for a given shape)

// DPP lane moves of a double (two 32-bit halves); every control used here
// reads a lane of the same 16-lane row, so all lanes are valid sources and
// no "old" value is needed (none is materialised: two instructions a move).
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
// value of v in lane l (l uniform across the wave)
__device__ __forceinline__ double rdlane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
// Sum over the wave, returned in every lane.  Within each 16-lane row:
// quad_perm [1,0,3,2] and [2,3,0,1] (quad sums), row_half_mirror (8-lane
// sums), row_mirror (row sums) -- DPP moves, no LDS round trip; then the four
// row sums are read back with v_readlane.  Fixed order: deterministic.
// (About 200 cycles of latency against 530 for a shuffle tree.)
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x140>(v);
  return (rdlane(v, 0) + rdlane(v, 16)) + (rdlane(v, 32) + rdlane(v, 48));
}
// min / max over the wave in every lane (one of the inputs, exactly): DPP
// stages within each 16-lane row as in wave_sum, then the four row results
__device__ __forceinline__ double wave_min(double v) {
  v = fmin(v, dpp_mov<0xB1>(v));
  v = fmin(v, dpp_mov<0x4E>(v));
  v = fmin(v, dpp_mov<0x141>(v));
  v = fmin(v, dpp_mov<0x140>(v));
  return fmin(fmin(rdlane(v, 0), rdlane(v, 16)), fmin(rdlane(v, 32), rdlane(v, 48)));
}
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_mov<0xB1>(v));
  v = fmax(v, dpp_mov<0x4E>(v));
  v = fmax(v, dpp_mov<0x141>(v));
  v = fmax(v, dpp_mov<0x140>(v));
  return fmax(fmax(rdlane(v, 0), rdlane(v, 16)), fmax(rdlane(v, 32), rdlane(v, 48)));
}

// Sum over the workgroup; every thread gets the result.  `red` is LDS
// scratch of at least NWAVE doubles.
__device__ __forceinline__ double wg_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NWAVE; ++i) s += red[i];
  return s;
}

// Three workgroup sums with one pair of barriers.
__device__ __forceinline__ void wg_sum3(double& a, double& b, double& c, double* red) {
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[wv] = a;
    red[NWAVE + wv] = b;
    red[2 * NWAVE + wv] = c;
  }
  __syncthreads();
  a = b = c = 0.0;
#pragma unroll
  for (int q = 0; q < NWAVE; ++q) {
    a += red[q];
    b += red[NWAVE + q];
    c += red[2 * NWAVE + q];
  }
}

// wg_sum3 without the leading barrier: for a `red` that no thread can still
// be reading, i.e. one whose last readers have passed a barrier since.
__device__ __forceinline__ void wg_sum3_nb(double& a, double& b, double& c, double* red) {
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[wv] = a;
    red[NWAVE + wv] = b;
    red[2 * NWAVE + wv] = c;
  }
  __syncthreads();
  a = b = c = 0.0;
#pragma unroll
  for (int q = 0; q < NWAVE; ++q) {
    a += red[q];
    b += red[NWAVE + q];
    c += red[2 * NWAVE + q];
  }
}

template <class T>
__device__ double wg_nrm2(const T* x, int n, double* red) {
  // scaled to avoid overflow/underflow of the squares
  double mx = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) mx = fmax(mx, fabs((double)x[i]));
  mx = wave_max(mx);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = mx;
  __syncthreads();
  mx = 0.0;
#pragma unroll
  for (int i = 0; i < NWAVE; ++i) mx = fmax(mx, red[i]);
  if (mx == 0.0 || !isfinite(mx)) return mx == 0.0 ? 0.0 : mx;
  const double inv = 1.0 / mx;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) {
    double t = (double)x[i] * inv;
    s += t * t;
  }
  return mx * sqrt(wg_sum(s, red));
}

template <class T, class U>
__device__ __forceinline__ void wg_copy(const T* x, U* y, int n) {
  for (int i = threadIdx.x; i < n; i += NT) y[i] = (U)x[i];
}

// y <- x (n elements, global memory); four loads in flight per thread before
// their stores (x and y may alias as far as the compiler knows, which would
// otherwise serialise every element on its load latency).
template <class T, class U>
__device__ __forceinline__ void wg_copy4(const T* x, U* y, int n) {
  int i = threadIdx.x;
  for (; i + 3 * NT < n; i += 4 * NT) {
    const T a0 = x[i], a1 = x[i + NT], a2 = x[i + 2 * NT], a3 = x[i + 3 * NT];
    y[i] = (U)a0;
    y[i + NT] = (U)a1;
    y[i + 2 * NT] = (U)a2;
    y[i + 3 * NT] = (U)a3;
  }
  for (; i < n; i += NT) y[i] = (U)x[i];
}

// dst (cols x rows, row-major) <- the transpose of src (rows x cols): read
// in src's order (coalesced), four loads in flight per thread, stores
// scattered (they do not stall).
__device__ __forceinline__ void wg_transpose(const double* src, int rows, int cols, double* dst) {
  const long n = (long)rows * cols;
  if (n < (1L << 22)) {
    const float ic = 1.0f / (float)cols;
    int e = threadIdx.x;
    for (; e + 3 * NT < n; e += 4 * NT) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = src[e + u * NT];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int f = e + u * NT, i = (int)(((float)f + 0.5f) * ic), j = f - i * cols;
        dst[(long)j * rows + i] = v[u];
      }
    }
    for (; e < n; e += NT) {
      const int i = (int)(((float)e + 0.5f) * ic), j = e - i * cols;
      dst[(long)j * rows + i] = src[e];
    }
  } else {
    for (long e = threadIdx.x; e < n; e += NT) {
      const long i = e / cols, j = e - i * cols;
      dst[j * rows + i] = src[e];
    }
  }
}
template <class T>
__device__ __forceinline__ void wg_fill(T* y, T v, int n) {
  for (int i = threadIdx.x; i < n; i += NT) y[i] = v;
}

// ---------------------------------------------------------------------------
// Workgroup GEMM  C[M x N] = A[M x K] * B[K x N]  (T = storage/compute type,
// accumulation in ACC).  64x64 output tiles, K-steps of 16 staged in LDS;
// each thread owns a 4x4 register micro-tile.

constexpr int GBM = 64, GBN = 64, GBK = 16;
constexpr int GEMM_LDS_DOUBLES = GBK * (GBM + 1) + GBK * (GBN + 1);

// i -> (i / div) * hi + (i % div) * lo for 0 <= i < 2^22 with an FP32
// reciprocal quotient (exact there: the quotient sits at least 0.5/div from an
// integer) and ordinary 32-bit products (offsets may exceed 2^24; a system's
// arrays stay below 2^31 elements).
struct GMap {
  float inv;
  int div, hi, lo;
};
__device__ __forceinline__ GMap gmap(const IMap& m) {
  const int div = m.div < (1 << 22) ? m.div : (1 << 22);  // plain maps: quotient 0 below 2^22
  return GMap{1.0f / (float)div, div, m.hi, m.lo};
}
__device__ __forceinline__ int gmap_at(const GMap& g, int i) {
  const int q = (int)(((float)i + 0.5f) * g.inv);
  return q * g.hi + (i - q * g.div) * g.lo;
}

typedef double dx4_t __attribute__((ext_vector_type(4)));
typedef float fx4_t __attribute__((ext_vector_type(4)));

// One 16x16x4 matrix-core step in FP64 or FP32 (operands: lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]) and the row of result element e
// held by lane group lk = l / 16 (column l % 16): the FP64 instruction
// interleaves the rows (lk + 4e), the FP32 one gives each group four
// consecutive rows (4 lk + e).
__device__ __forceinline__ dx4_t mfma16x16x4(double a, double b, dx4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ fx4_t mfma16x16x4(float a, float b, fx4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <class T>
__device__ __forceinline__ int mfma16_row(int lk, int e) {
  return std::is_same<T, double>::value ? lk + 4 * e : 4 * lk + e;
}

// K-steps of operands in flight during a K-step's MFMAs: 1 where the kernel
// must stay within 256 registers for two workgroups per CU (k_als), 2 in a
// translation unit that defines TTALS_GEMM_PREFETCH for a one-workgroup-per-CU
// kernel (the same arithmetic either way).
#ifndef TTALS_GEMM_PREFETCH
#define TTALS_GEMM_PREFETCH 1
#endif
constexpr int GEMM_PREFETCH = TTALS_GEMM_PREFETCH;

// FP64 GEMM over global (L2/HBM) operands on the matrix cores.  64x64 output
// tiles, K-steps of 16 staged through LDS ([GBK][GBM+1] and [GBK][GBN+1]:
// the odd stride keeps the transposing stores of k-fast operands off a single
// bank); every wave owns a 32x32 quadrant of the tile (2x2
// v_mfma_f64_16x16x4f64 tiles).  The (tile, K-step) sequence is one
// pipeline: the operands of step s+1 -- across a tile boundary too -- are
// loaded into registers while the MFMAs of step s run, so the load latency
// of L2/HBM overlaps the arithmetic instead of adding to it.  Elements past
// M, N or K are staged as zeros (their products vanish; rows/columns past M,
// N are not written).
template <class T>
__device__ __forceinline__ void wg_gemm_mf(int M, int N, int K, CView<T> A, CView<T> B, MView<T> C, T* lds_) {
  typedef typename std::conditional<std::is_same<T, double>::value, dx4_t, fx4_t>::type acc_t;
  TT_LDS T* As = (TT_LDS T*)lds_;              // [GBK][GBM+1]
  TT_LDS T* Bs = (TT_LDS T*)lds_ + GBK * (GBM + 1);  // [GBK][GBN+1]
  static_assert(GBM == 64 && GBN == 64 && GBK == 16 && NT == 256, "tile shape assumed below");
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = (w & 1) << 5, wn = (w >> 1) << 5;
  const GMap Ar = gmap(A.r), Ac = gmap(A.c), Br = gmap(B.r), Bc = gmap(B.c);
  // contiguous direction of each operand decides which index is thread-fast (coalescing)
  const bool a_kfast = (A.c.lo == 1 && A.c.div >= K);
  const bool b_nfast = (B.c.lo == 1 && B.c.div >= N);
  // per-thread element coordinates in a (64 x 16) / (16 x 64) stage, j = 0..3
  const int a_i0 = a_kfast ? (tid >> 4) : (tid & 63), a_k0 = a_kfast ? (tid & 15) : (tid >> 6);
  const int a_di = a_kfast ? 16 : 0, a_dk = a_kfast ? 0 : 4;
  const int b_n0 = b_nfast ? (tid & 63) : (tid >> 4), b_k0 = b_nfast ? (tid >> 6) : (tid & 15);
  const int b_dn = b_nfast ? 0 : 16, b_dk = b_nfast ? 4 : 0;
  const int tn = (N + GBN - 1) / GBN, ks = (K + GBK - 1) / GBK;
  const int steps = ((M + GBM - 1) / GBM) * tn * ks;
  // loader position (tile origin, K offset) and compute position, advanced without division
  int lm0 = 0, ln0 = 0, lk0 = 0, ltj = 0;
  int cm0 = 0, cn0 = 0, ckk = 0, ctj = 0;
  // plain (single-stride) index maps -- most operands -- skip the split-map
  // arithmetic (a wave-uniform branch)
  const bool a_plain = A.r.div >= (1 << 22) && A.c.div >= (1 << 22);
  const bool b_plain = B.r.div >= (1 << 22) && B.c.div >= (1 << 22);
  auto load = [&](T (&ra)[4], T (&rb)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gm = lm0 + a_i0 + j * a_di, gk = lk0 + a_k0 + j * a_dk;
      const int ao = a_plain ? gm * A.r.lo + gk * A.c.lo : gmap_at(Ar, gm) + gmap_at(Ac, gk);
      ra[j] = (gm < M && gk < K) ? A.p[ao] : (T)0;
      const int gn = ln0 + b_n0 + j * b_dn, gk2 = lk0 + b_k0 + j * b_dk;
      const int bo = b_plain ? gk2 * B.r.lo + gn * B.c.lo : gmap_at(Br, gk2) + gmap_at(Bc, gn);
      rb[j] = (gn < N && gk2 < K) ? B.p[bo] : (T)0;
    }
    lk0 += GBK;
    if (lk0 >= K) {
      lk0 = 0;
      ln0 += GBN;
      if (++ltj == tn) {
        ltj = 0;
        ln0 = 0;
        lm0 += GBM;
      }
    }
  };
  auto stage = [&](const T (&ra)[4], const T (&rb)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      As[(a_k0 + j * a_dk) * (GBM + 1) + a_i0 + j * a_di] = ra[j];
      Bs[(b_k0 + j * b_dk) * (GBN + 1) + b_n0 + j * b_dn] = rb[j];
    }
  };
  acc_t acc[2][2];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int g = 0; g < 2; ++g) acc[h][g] = acc_t{0, 0, 0, 0};
  // the MFMAs of the staged K-step, and the tile's store after its last one
  auto compute = [&]() {
#pragma unroll
    for (int q = 0; q < GBK / 4; ++q) {
      const int kr = (4 * q + lk);
      const T a0 = As[kr * (GBM + 1) + wm + li], a1 = As[kr * (GBM + 1) + wm + 16 + li];
      const T b0 = Bs[kr * (GBN + 1) + wn + li], b1 = Bs[kr * (GBN + 1) + wn + 16 + li];
      acc[0][0] = mfma16x16x4(a0, b0, acc[0][0]);
      acc[0][1] = mfma16x16x4(a0, b1, acc[0][1]);
      acc[1][0] = mfma16x16x4(a1, b0, acc[1][0]);
      acc[1][1] = mfma16x16x4(a1, b1, acc[1][1]);
    }
    ckk += GBK;
    if (ckk >= K) {
      // tile done: lane l holds C[i0 + mfma16_row(l/16, e)][j0 + l%16] of each 16x16 block
      const GMap Cr = gmap(C.r), Cc = gmap(C.c);
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const int gn = cn0 + wn + 16 * g + li;
        const bool nv = gn < N;
        const int cc = nv ? gmap_at(Cc, gn) : 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int gm = cm0 + wm + 16 * h + mfma16_row<T>(lk, e);
            if (nv && gm < M) C.p[gmap_at(Cr, gm) + cc] = acc[h][g][e];
          }
          acc[h][g] = acc_t{0, 0, 0, 0};
        }
      }
      ckk = 0;
      cn0 += GBN;
      if (++ctj == tn) {
        ctj = 0;
        cn0 = 0;
        cm0 += GBM;
      }
    }
  };
  if constexpr (GEMM_PREFETCH == 2) {
    // two register sets: while step s computes, steps s+1 and s+2 are loading
    T ra0[4], rb0[4], ra1[4], rb1[4];
    if (steps > 0) {
      load(ra0, rb0);
      stage(ra0, rb0);
    }
    if (steps > 1) load(ra1, rb1);
    __syncthreads();
    for (int s = 0; s < steps; s += 2) {
      if (s + 2 < steps) load(ra0, rb0);
      compute();
      __syncthreads();
      if (s + 1 >= steps) break;
      stage(ra1, rb1);
      __syncthreads();
      if (s + 3 < steps) load(ra1, rb1);
      compute();
      __syncthreads();
      if (s + 2 < steps) {
        stage(ra0, rb0);
        __syncthreads();
      }
    }
  } else {
    T ra[4], rb[4];
    if (steps > 0) {
      load(ra, rb);
      stage(ra, rb);
    }
    __syncthreads();
    for (int s = 0; s < steps; ++s) {
      if (s + 1 < steps) load(ra, rb);
#pragma unroll
      for (int q = 0; q < GBK / 4; ++q) {
        const int kr = (4 * q + lk);
        const T a0 = As[kr * (GBM + 1) + wm + li], a1 = As[kr * (GBM + 1) + wm + 16 + li];
        const T b0 = Bs[kr * (GBN + 1) + wn + li], b1 = Bs[kr * (GBN + 1) + wn + 16 + li];
        acc[0][0] = mfma16x16x4(a0, b0, acc[0][0]);
        acc[0][1] = mfma16x16x4(a0, b1, acc[0][1]);
        acc[1][0] = mfma16x16x4(a1, b0, acc[1][0]);
        acc[1][1] = mfma16x16x4(a1, b1, acc[1][1]);
      }
      ckk += GBK;
      if (ckk >= K) {
        const GMap Cr = gmap(C.r), Cc = gmap(C.c);
#pragma unroll
        for (int g = 0; g < 2; ++g) {
          const int gn = cn0 + wn + 16 * g + li;
          const bool nv = gn < N;
          const int cc = nv ? gmap_at(Cc, gn) : 0;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int gm = cm0 + wm + 16 * h + mfma16_row<T>(lk, e);
              if (nv && gm < M) C.p[gmap_at(Cr, gm) + cc] = acc[h][g][e];
            }
            acc[h][g] = acc_t{0, 0, 0, 0};
          }
        }
        ckk = 0;
        cn0 += GBN;
        if (++ctj == tn) {
          ctj = 0;
          cn0 = 0;
          cm0 += GBM;
        }
      }
      __syncthreads();
      if (s + 1 < steps) {
        stage(ra, rb);
        __syncthreads();
      }
    }
  }
}

template <class T, class ACC>
__device__ void wg_gemm_valu(int M, int N, int K, CView<T> A, CView<T> B, MView<T> C, ACC* lds);

// GEMM over global operands: the matrix cores in FP64 or FP32 (accumulating
// in the operands' precision), the vector units (register tiles) otherwise.
template <class T, class ACC>
__device__ void wg_gemm(int M, int N, int K, CView<T> A, CView<T> B, MView<T> C, ACC* lds) {
  if constexpr ((std::is_same<T, double>::value || std::is_same<T, float>::value) && std::is_same<ACC, T>::value)
    wg_gemm_mf<T>(M, N, K, A, B, C, lds);
  else
    wg_gemm_valu<T, ACC>(M, N, K, A, B, C, lds);
}

template <class T, class ACC>
__device__ void wg_gemm_valu(int M, int N, int K, CView<T> A, CView<T> B, MView<T> C, ACC* lds) {
  ACC* As = lds;                      // [GBK][GBM+1]
  ACC* Bs = lds + GBK * (GBM + 1);    // [GBK][GBN+1]
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const GMap Ar = gmap(A.r), Ac = gmap(A.c), Br = gmap(B.r), Bc = gmap(B.c);
  // contiguous direction of each operand decides the load order (coalescing)
  const bool a_kfast = (A.c.lo == 1 && A.c.div >= K);
  const bool b_nfast = (B.c.lo == 1 && B.c.div >= N);
  for (int m0 = 0; m0 < M; m0 += GBM) {
    for (int n0 = 0; n0 < N; n0 += GBN) {
      ACC acc[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = (ACC)0;
      for (int k0 = 0; k0 < K; k0 += GBK) {
        for (int e = tid; e < GBM * GBK; e += NT) {
          int mi, ki;
          if (a_kfast) { ki = e % GBK; mi = e / GBK; } else { mi = e % GBM; ki = e / GBM; }
          const int gm = m0 + mi, gk = k0 + ki;
          As[ki * (GBM + 1) + mi] = (gm < M && gk < K) ? (ACC)A.p[gmap_at(Ar, gm) + gmap_at(Ac, gk)] : (ACC)0;
        }
        for (int e = tid; e < GBN * GBK; e += NT) {
          int ni, ki;
          if (b_nfast) { ni = e % GBN; ki = e / GBN; } else { ki = e % GBK; ni = e / GBK; }
          const int gn = n0 + ni, gk = k0 + ki;
          Bs[ki * (GBN + 1) + ni] = (gn < N && gk < K) ? (ACC)B.p[gmap_at(Br, gk) + gmap_at(Bc, gn)] : (ACC)0;
        }
        __syncthreads();
        const int kmax = (K - k0) < GBK ? (K - k0) : GBK;
        for (int kk = 0; kk < kmax; ++kk) {
          ACC a[4], b[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) a[r] = As[kk * (GBM + 1) + ty + 16 * r];
#pragma unroll
          for (int c = 0; c < 4; ++c) b[c] = Bs[kk * (GBN + 1) + tx + 16 * c];
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[r][c] += a[r] * b[c];
        }
        __syncthreads();
      }
      const GMap Cr = gmap(C.r), Cc = gmap(C.c);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + ty + 16 * r;
        if (gm >= M) continue;
        const int cr = gmap_at(Cr, gm);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int gn = n0 + tx + 16 * c;
          if (gn < N) C.p[cr + gmap_at(Cc, gn)] = (T)acc[r][c];
        }
      }
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Direct GEMM for operands that already sit on-chip (LDS): no staging, each
// thread owns 2x2 output micro-tiles, and the K index is walked with split
// counters (no integer division in the inner loop).

struct IWalk {
  long hi, lo, off;
  int div, l;
  __device__ __forceinline__ void start(const IMap& m) {
    div = m.div;
    hi = m.hi;
    lo = m.lo;
    l = 0;
    off = 0;
  }
  __device__ __forceinline__ void next() {
    if (++l == div) {
      l = 0;
      off += hi - (long)(div - 1) * lo;
    } else {
      off += lo;
    }
  }
};

template <class T, class ACC, bool OA = false, bool OB = false, bool OC = false>
__device__ void ds_gemm(int M, int N, int K, CView<T> A, CView<T> B, MView<T> C) {
  const int tN = (N + 1) >> 1;
  const int tiles = ((M + 1) >> 1) * tN;
  for (int t = threadIdx.x; t < tiles; t += NT) {
    const int ti = t / tN, tj = t - ti * tN;
    const int i0 = 2 * ti, j0 = 2 * tj;
    const bool i1 = i0 + 1 < M, j1 = j0 + 1 < N;
    const long ar0 = imap(A.r, i0), ar1 = i1 ? imap(A.r, i0 + 1) : ar0;
    const long bc0 = imap(B.c, j0), bc1 = j1 ? imap(B.c, j0 + 1) : bc0;
    IWalk ka, kb;
    ka.start(A.c);
    kb.start(B.r);
    ACC c00 = 0, c01 = 0, c10 = 0, c11 = 0;
    for (int k = 0; k < K; ++k) {
      const ACC a0 = (ACC)Mem<OA>::ld(A.p, ar0 + ka.off), a1 = (ACC)Mem<OA>::ld(A.p, ar1 + ka.off);
      const ACC b0 = (ACC)Mem<OB>::ld(B.p, kb.off + bc0), b1 = (ACC)Mem<OB>::ld(B.p, kb.off + bc1);
      c00 += a0 * b0;
      c01 += a0 * b1;
      c10 += a1 * b0;
      c11 += a1 * b1;
      ka.next();
      kb.next();
    }
    const long cr0 = imap(C.r, i0), cc0 = imap(C.c, j0);
    Mem<OC>::st(C.p, cr0 + cc0, (T)c00);
    if (j1) Mem<OC>::st(C.p, cr0 + imap(C.c, j0 + 1), (T)c01);
    if (i1) {
      const long cr1 = imap(C.r, i0 + 1);
      Mem<OC>::st(C.p, cr1 + cc0, (T)c10);
      if (j1) Mem<OC>::st(C.p, cr1 + imap(C.c, j0 + 1), (T)c11);
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// FP64 matrix-core GEMM with every operand in LDS.
//
// v_mfma_f64_16x16x4f64 operand layout (wave64): lane l supplies
// A[i0 + l%16][k + l/16] and B[k + l/16][j0 + l%16]; element e of its
// 4-wide accumulator is C[i0 + l/16 + 4*e][j0 + l%16] (checked on gfx950).  Each lane loads
// one A and one B value per 16 FMAs, against four loads per four FMAs for a
// 2x2 register tile: the local matvec is LDS-bound, so this is what counts.

// offsets of k, k+4, k+8, ... along an IMap, advanced without division
struct KStep4 {
  int off, r, div, dr, step, wrap;
  __device__ __forceinline__ void start(const IMap& m, int k) {
    div = m.div;
    const int q = k / div;
    r = k - q * div;
    off = q * m.hi + r * m.lo;
    const int dq = 4 / div;
    dr = 4 - dq * div;
    step = dq * m.hi + dr * m.lo;
    wrap = (int)((long)m.hi - (long)div * m.lo);
  }
  __device__ __forceinline__ void next() {
    r += dr;
    off += step;
    if (r >= div) {
      r -= div;
      off += wrap;
    }
  }
};

// unscaled 2-norm (for vectors of moderate magnitude in the hot loops)
template <class T>
__device__ __forceinline__ double wg_nrm2u(const T* x, int n, double* red) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) {
    const double t = (double)x[i];
    s += t * t;
  }
  return sqrt(wg_sum(s, red));
}

// ---------------------------------------------------------------------------
// Interface-stack contractions and the local matvec.
//
// Shapes: x core (a, n, b); operator core A (P, m, n, Q) with m == n;
// op stacks L (a, P, a') / R (b, Q, b'); rhs core B (p, m, q); rhs stacks
// Lb (p, a') / Rb (q, b').  All row-major.  t1/t2 are scratch of
// max(P*a*n*b, a*n*Q*b) elements.

// stage 1 shared by the matvec and the left stack:
//   T1[a', b, P, n] = sum_a L[a, P, a'] v[a, n, b]
template <class T, class ACC>
__device__ void stage_left1(int a, int P, int n, int b, const T* L, const T* v, T* t1, ACC* lds) {
  const int a2 = a;
  wg_gemm<T, ACC>(P * a2, n * b, a, cview<T>(L, plain(1), plain(P * a2)), crm<T>(v, n * b),
                  mview<T>(t1, split(a2, n, b * P * n), split(b, 1, P * n)), lds);
}

// y[a', m, b'] = sum L[a,P,a'] A[P,m,n,Q] R[b,Q,b'] v[a,n,b]
template <class T, class ACC>
__device__ void local_matvec(int a, int P, int n, int Q, int b, const T* L, const T* Ac, const T* R,
                             const T* v, T* y, T* t1, T* t2, ACC* lds) {
  stage_left1<T, ACC>(a, P, n, b, L, v, t1, lds);
  // T2[a', m, b, Q] = sum_{P,n} T1[(a',b),(P,n)] A[P,m,n,Q]
  wg_gemm<T, ACC>(a * b, n * Q, P * n, crm<T>(t1, P * n),
                  cview<T>(Ac, split(n, n * n * Q, Q), split(Q, n * Q, 1)),
                  mview<T>(t2, split(b, n * b * Q, Q), split(Q, b * Q, 1)), lds);
  // y[(a',m), b'] = T2[(a',m),(b,Q)] R[(b,Q), b']
  wg_gemm<T, ACC>(a * n, b, b * Q, crm<T>(t2, b * Q), crm<T>(R, b), mrm<T>(y, b), lds);
}

// ---------------------------------------------------------------------------
// On-chip local matvec (every operand in LDS) on the FP64 matrix cores.
//
// Layouts: L (a, P, a'), R (b, Q, b'), the operator core stored PERMUTED as
// A'(P, n, m, Q) (so that the second stage reads it with plain strides), v and
// y (a, n, b).  The two intermediates use odd row strides (P*n | 1, b*Q | 1
// doubles) so rows read by neighbouring lanes fall into different LDS banks:
//   T1[a', b | P, n] = sum_a L[a, P, a'] v[a, n, b]           (a*b*(P*n|1))
//   T2[a', m | b, Q] = sum_{P,n} T1[a', b | P, n] A'[P, n | m, Q]  (a*n*(b*Q|1))
//   y[a', m, b']     = sum_{b,Q} T2[a', m | b, Q] R[b, Q, b']
__host__ __device__ __forceinline__ int odd_stride(int s) { return s | 1; }
__host__ __device__ __forceinline__ long ds_t1_size(int a, int P, int n, int b) {
  return (long)a * b * odd_stride(P * n);
}
__host__ __device__ __forceinline__ long ds_t2_size(int a, int n, int Q, int b) {
  return (long)a * n * odd_stride(b * Q);
}

// i -> (i / div) * hi + (i % div) * lo for 0 <= i < 2^22 and values below
// 2^24, with full-rate arithmetic only (32-bit integer multiplies are
// quarter rate on CDNA): the quotient is floor((i + 0.5) * (1/div)) in FP32
// (the exact quotient lies at least 0.5/div from an integer, more than the
// two FP32 roundings can move it while i + 0.5 < 2^22) and the products are
// 24-bit multiplies.
struct FMap {
  float inv;
  int div, hi, lo;
};
__device__ __forceinline__ FMap fmap(int div, int hi, int lo) { return FMap{1.0f / (float)div, div, hi, lo}; }
__device__ __forceinline__ int fmap_at(const FMap& f, int i) {
  const int q = (int)(((float)i + 0.5f) * f.inv);
  return __mul24(q, f.hi) + __mul24(i - __mul24(q, f.div), f.lo);
}

// floor(i / div) for 0 <= i < 2^22 with the FP32 reciprocal of div (see FMap)
__device__ __forceinline__ int fquot(int i, float inv) { return (int)(((float)i + 0.5f) * inv); }

// Output maps of the first two stages (computed once per local problem).
struct MvPlan {
  FMap c1r, c1c, c2r, c2c;
};
__device__ __forceinline__ MvPlan mv_plan(int a, int P, int n, int Q, int b) {
  const int s1 = odd_stride(P * n), s2 = odd_stride(b * Q);
  return MvPlan{fmap(a, n, b * s1), fmap(b, 1, s1), fmap(b, n * s2, Q), fmap(Q, s2, 1)};
}

// The operator core (P, m, n, Q) copied into LDS as A'(P, n, m, Q).
__device__ __forceinline__ void wg_copy_opcore_pnmq(const double* src, double* dst, int P, int n, int Q) {
  const int tot = P * n * n * Q;
  const float iQ = 1.0f / (float)Q, in = 1.0f / (float)n;
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int r1 = fquot(e, iQ), q = e - r1 * Q;
    const int r2 = fquot(r1, in), nn = r1 - r2 * n;
    const int p = fquot(r2, in), mm = r2 - p * n;
    dst[((p * n + nn) * n + mm) * Q + q] = src[e];
  }
}

// C[M,N] = A[M,K] B[K,N] with every operand in LDS and plain operand strides:
//   A(i,k) = pa[i*ars + k*acs], B(k,j) = pb[k*brs + j*bcs],
//   C(i,j) = pc[fmap_at(cr, i) + fmap_at(cc, j)].
// 16x16 output tiles round-robin over the waves.  When there are at most
// half as many tiles as waves and `red` holds (g-1)*tiles*256 doubles
// (g = NWAVE / tiles), the K range of each tile is split over g waves and the
// partial sums meet in `red` in a fixed order.  Operands are loaded four
// K-steps ahead of their MFMAs; indices past M, N or K read a clamped
// in-range address and contribute zero.
__device__ void mf_gemm_lin(int M, int N, int K, const double* A_, int ars, int acs, const double* B_, int brs,
                            int bcs, double* C_, const FMap& cr, const FMap& cc, double* red_, int red_cap) {
  M = __builtin_amdgcn_readfirstlane(M);
  N = __builtin_amdgcn_readfirstlane(N);
  K = __builtin_amdgcn_readfirstlane(K);
  const TT_LDS double* pa = (const TT_LDS double*)A_;
  const TT_LDS double* pb = (const TT_LDS double*)B_;
  TT_LDS double* pc = (TT_LDS double*)C_;
  TT_LDS double* red = (TT_LDS double*)red_;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 15, lk = lane >> 4;
  // no integer divisions below (each expands to a long instruction sequence):
  // tile rows by an FP32-reciprocal quotient (exact for these small counts),
  // the K-split cases by shifts
  static_assert(NWAVE == 4, "the K-split cases assume four waves");
  const int tn = (N + 15) >> 4, tiles = ((M + 15) >> 4) * tn;
  const float itn = 1.0f / (float)tn;
  const int ks = (K + 3) >> 2;
  int gl = 0;  // log2 of the K groups: tiles == 1 -> 4 groups, tiles == 2 -> 2 (every wave one (tile, part))
  if (tiles <= 2) {
    gl = tiles == 1 ? 2 : 1;
    if (((1 << gl) - 1) * tiles * 256 > red_cap) gl = 0;
  }
  const int groups = 1 << gl;
  const int per = (ks + groups - 1) >> gl;
  const int acs4 = acs << 2, brs4 = brs << 2;
  if (groups == 1 && tiles > NWAVE) {
    // more tiles than waves: a wave runs tiles u and u + NWAVE together, their
    // two MFMA chains interleaved (a wave without a second tile repeats its
    // first and discards the copy)
    for (int u = w; u < tiles; u += 2 * NWAVE) {
      const bool two = u + NWAVE < tiles;
      const int tt[2] = {u, two ? u + NWAVE : u};
      const TT_LDS double *ap[2], *bp[2], *alast[2], *blast[2];
      int ti[2], jj[2];
      bool jv[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        ti[h] = (int)(((float)tt[h] + 0.5f) * itn);
        const int tj = tt[h] - ti[h] * tn;
        const int i = (ti[h] << 4) + li;
        jj[h] = (tj << 4) + li;
        jv[h] = jj[h] < N;
        ap[h] = pa + __mul24(i < M ? i : M - 1, ars);
        bp[h] = pb + __mul24(jv[h] ? jj[h] : N - 1, bcs);
        alast[h] = ap[h] + __mul24(K - 1, acs);
        blast[h] = bp[h] + __mul24(K - 1, brs);
        ap[h] += __mul24(lk, acs);
        bp[h] += __mul24(lk, brs);
      }
      int k = lk;
      dx4_t acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
      int s = 0;
      for (; s + 4 <= ks; s += 4) {
        double av[2][4], bv[2][4];
        bool kv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          kv[q] = k < K;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            av[h][q] = (kv[q] ? ap[h] : alast[h])[0];
            bv[h][q] = (kv[q] ? bp[h] : blast[h])[0];
            ap[h] += acs4;
            bp[h] += brs4;
          }
          k += 4;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int h = 0; h < 2; ++h)
            acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[h][q], kv[q] ? bv[h][q] : 0.0, acc[h], 0, 0, 0);
      }
      for (; s < ks; ++s) {
        const bool kq = k < K;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const double x = (kq ? ap[h] : alast[h])[0];
          const double y = (kq ? bp[h] : blast[h])[0];
          acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, kq ? y : 0.0, acc[h], 0, 0, 0);
          ap[h] += acs4;
          bp[h] += brs4;
        }
        k += 4;
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (h == 1 && !two) break;
        if (jv[h]) {
          const int cj = fmap_at(cc, jj[h]);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = (ti[h] << 4) + lk + 4 * e;
            if (row < M) pc[fmap_at(cr, row) + cj] = acc[h][e];
          }
        }
      }
    }
    __syncthreads();
    return;
  }
  for (int u = w; u < tiles * groups; u += NWAVE) {
    // groups > 1 only with tiles 1 or 2
    const int t = gl == 0 ? u : (tiles == 1 ? 0 : (u & 1));
    const int part = gl == 0 ? 0 : (tiles == 1 ? u : (u >> 1));
    const int ti = (int)(((float)t + 0.5f) * itn), tj = t - ti * tn;
    const int i = (ti << 4) + li, j = (tj << 4) + li;
    const bool iv = i < M, jv = j < N;
    const int s0 = part * per, s1 = min(ks, s0 + per);
    // Running operand pointers (no multiplies in the K loop).  Rows past M and
    // columns past N read a clamped address: their outputs are discarded.  A
    // K index past the end reads the last column/row instead and its B value
    // is zeroed (A finite there, so the product vanishes).
    const TT_LDS double* ap = pa + __mul24(iv ? i : M - 1, ars);
    const TT_LDS double* bp = pb + __mul24(jv ? j : N - 1, bcs);
    const TT_LDS double* alast = ap + __mul24(K - 1, acs);
    const TT_LDS double* blast = bp + __mul24(K - 1, brs);
    ap += __mul24((s0 << 2) + lk, acs);
    bp += __mul24((s0 << 2) + lk, brs);
    int k = (s0 << 2) + lk;
    dx4_t acc = {0.0, 0.0, 0.0, 0.0};
    int s = s0;
    for (; s + 4 <= s1; s += 4) {
      double av[4], bv[4];
      bool kv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        kv[q] = k < K;
        av[q] = (kv[q] ? ap : alast)[0];
        bv[q] = (kv[q] ? bp : blast)[0];
        ap += acs4;
        bp += brs4;
        k += 4;
      }
      // all eight loads in flight before the first MFMA waits on one
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < 4; ++q)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], kv[q] ? bv[q] : 0.0, acc, 0, 0, 0);
    }
    for (; s < s1; ++s) {
      const bool kq = k < K;
      const double x = (kq ? ap : alast)[0];
      const double y = (kq ? bp : blast)[0];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(x, kq ? y : 0.0, acc, 0, 0, 0);
      ap += acs4;
      bp += brs4;
      k += 4;
    }
    if (groups > 1) {
      // one work item per wave: partial parts park their sums, part 0 adds them
      if (part > 0) {
        TT_LDS double* r = red + ((part - 1) * tiles + t) * 256 + lane * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) r[e] = acc[e];
      }
      __syncthreads();
      if (part > 0) break;
      for (int p2 = 1; p2 < groups; ++p2) {
        const TT_LDS double* r = red + ((p2 - 1) * tiles + t) * 256 + lane * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[e] += r[e];
      }
    }
    if (jv) {
      const int cj = fmap_at(cc, j);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = (ti << 4) + lk + 4 * e;
        if (row < M) pc[fmap_at(cr, row) + cj] = acc[e];
      }
    }
  }
  __syncthreads();
}

// Stage 2 on the vector units when it is tiny (K = P*n and N = n*Q with
// K*N <= 64, e.g. n = 2, P = Q = 3): one output row (a', b) per thread, the
// row of T1 in registers, A' read as broadcasts.  A 16x16 matrix-core tile
// would pad N = 6 to 16 and spend its time on tile bookkeeping.
__device__ __forceinline__ void stage2_valu(int a, int P, int n, int Q, int b, const double* t1_, const double* Ap_,
                                            double* t2_) {
  const TT_LDS double* t1 = (const TT_LDS double*)t1_;
  const TT_LDS double* Ap = (const TT_LDS double*)Ap_;
  TT_LDS double* t2 = (TT_LDS double*)t2_;
  const int K = P * n, N = n * Q;
  const int s1 = odd_stride(K), s2 = odd_stride(b * Q);
  const float invb = 1.0f / (float)b, invQ = 1.0f / (float)Q;
  for (int r = threadIdx.x; r < a * b; r += NT) {
    double x[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = k < K ? t1[__mul24(r, s1) + k] : 0.0;
    const int ar = (int)(((float)r + 0.5f) * invb);
    const int base = __mul24(__mul24(ar, n), s2) + __mul24(r - __mul24(ar, b), Q);
    for (int c = 0; c < N; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 16; ++k)
        if (k < K) acc = fma(x[k], Ap[__mul24(k, N) + c], acc);
      const int mm = (int)(((float)c + 0.5f) * invQ);
      t2[base + __mul24(mm, s2) + (c - __mul24(mm, Q))] = acc;
    }
  }
  __syncthreads();
}

// Kept out of line: its MFMA accumulators would otherwise be allocated next
// to the Krylov solvers' registers and cost the second workgroup per CU.
template <class T>
__device__ __attribute__((noinline)) void local_matvec_mf(int a, int P, int n, int Q, int b, const T* L,
                                                          const T* Ap, const T* R, const T* v, T* y, T* t1, T* t2) {
  // the output maps are rebuilt here (four FP32 reciprocals) rather than
  // passed in: a plan passed by reference would be read from the stack
  const MvPlan pl = mv_plan(a, P, n, Q, b);
  static_assert(sizeof(T) == 8, "the on-chip matvec is FP64");
  const int s1 = odd_stride(P * n), s2 = odd_stride(b * Q);
  const FMap plain_r = fmap(1, b, 0), plain_c = fmap(1, 1, 0);
  // T1[(P,a'), (n,b)]: A = L viewed (P*a, a) with rows contiguous, B = v (a, n*b)
  mf_gemm_lin(P * a, n * b, a, L, 1, P * a, v, n * b, 1, t1, pl.c1r, pl.c1c, nullptr, 0);
  // T2[(a',b), (m,Q)]: A = T1 rows (stride s1), B = A' (P*n, n*Q)
  if (P * n <= 16 && P * n * n * Q <= 64)
    stage2_valu(a, P, n, Q, b, t1, Ap, t2);
  else
    mf_gemm_lin(a * b, n * Q, P * n, t1, s1, 1, Ap, n * Q, 1, t2, pl.c2r, pl.c2c, nullptr, 0);
  // y[(a',m), b']: A = T2 rows (stride s2), B = R (b*Q, b); T1 is free: K-split scratch
  mf_gemm_lin(a * n, b, b * Q, t2, s2, 1, R, b, 1, y, plain_r, plain_c, t1, (int)ds_t1_size(a, P, n, b));
}

// Interface stack update on-chip, left form:
//   out[(b,Q), b'] = sum L[a,P,a'] x[a,n,b] A[P,m,n,Q] x[a',m,b']
// L (a, P, a'), x (a, n, b), Ap = A'(P, n, m, Q) and out (b, Q, b') in LDS;
// t1 / t2 the three-stage matvec's intermediates.  Stages 1 and 2 are the
// matvec's with the core as the vector; stage 3 contracts T2[(a',m),(b,Q)]
// with the core.  A right stack is this update on the mirrored problem (core
// transposed to (b, n, a), operator core with P and Q swapped).
__device__ __attribute__((noinline)) void stack_left_mf(int a, int P, int n, int Q, int b, const double* L,
                                                        const double* Ap, const double* x, double* out, double* t1,
                                                        double* t2) {
  const MvPlan pl = mv_plan(a, P, n, Q, b);
  const int s1 = odd_stride(P * n), s2 = odd_stride(b * Q);
  mf_gemm_lin(P * a, n * b, a, L, 1, P * a, x, n * b, 1, t1, pl.c1r, pl.c1c, nullptr, 0);
  if (P * n <= 16 && P * n * n * Q <= 64)
    stage2_valu(a, P, n, Q, b, t1, Ap, t2);
  else
    mf_gemm_lin(a * b, n * Q, P * n, t1, s1, 1, Ap, n * Q, 1, t2, pl.c2r, pl.c2c, nullptr, 0);
  mf_gemm_lin(b * Q, b, a * n, t2, 1, s2, x, b, 1, out, fmap(1, b, 0), fmap(1, 1, 0), nullptr, 0);
}

// The operator core (P, m, n, Q) copied into LDS as A'(Q, n, m, P): the
// stage-2 layout of the mirrored problem (P and Q swapped).
__device__ __forceinline__ void wg_copy_opcore_qnmp(const double* src, double* dst, int P, int n, int Q) {
  const int tot = P * n * n * Q;
  const float iQ = 1.0f / (float)Q, in = 1.0f / (float)n;
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int r1 = fquot(e, iQ), q = e - r1 * Q;
    const int r2 = fquot(r1, in), nn = r1 - r2 * n;
    const int p = fquot(r2, in), mm = r2 - p * n;
    dst[((q * n + nn) * n + mm) * P + p] = src[e];
  }
}

// core (a, n, b) copied into LDS transposed to (b, n, a) (a n b < 2^22)
__device__ __forceinline__ void wg_copy_core_t(const double* src, double* dst, int a, int n, int b) {
  const int tot = a * n * b;
  const float ib = 1.0f / (float)b, in = 1.0f / (float)n;
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int r = fquot(e, ib), bb = e - r * b;
    const int aa = fquot(r, in), nn = r - aa * n;
    dst[(bb * n + nn) * a + aa] = src[e];
  }
}

// ---------------------------------------------------------------------------
// Two-stage on-chip matvec.  Per local solve the operator core and the right
// stack are merged once,
//   M2[(P,n,b), (m,b')] = sum_Q A[P,m,n,Q] R[b,Q,b'],
// and every Krylov matvec is then two matrix-core GEMMs:
//   T1'[a', (P,n,b)] = sum_a L[a,P,a'] x[a,(n,b)]                 (K = a)
//   y[a', (m,b')]    = sum_(P,n,b) T1'[a',(P,n,b)] M2[(P,n,b),(m,b')] (K = P n b)
// one stage, one barrier and the tiny vector-unit middle stage fewer than the
// three-stage form, for (P n b / (P n + b Q)) more flops in the second GEMM.
// T1' rows have the odd stride odd_stride(P n b) (conflict-free row reads).

// K-split scratch of the second GEMM (mf_gemm_lin's rule), in doubles.
__host__ __device__ __forceinline__ int m2_red_need(int a, int nb) {
  const int tiles = ((a + 15) >> 4) * ((nb + 15) >> 4);
  return (2 * tiles <= NWAVE && NWAVE % tiles == 0) ? (NWAVE / tiles - 1) * tiles * 256 : 0;
}

// M2 from the operator core A (P, m, n, Q) and the right stack R (b, Q, b') in
// global memory, both staged through `tmp` (b Q b + P n n Q doubles).  Ends
// with a barrier: M2 is complete and `tmp` free.
__device__ void build_m2(const double* A, const double* R, int P, int n, int Q, int b, double* M2, double* tmp) {
  const int bqb = b * Q * b, nb = n * b, k2 = P * n * b;
  TT_LDS double* Rs = (TT_LDS double*)tmp;
  TT_LDS double* As = Rs + bqb;
  for (int e = threadIdx.x; e < bqb; e += NT) Rs[e] = R[e];
  for (int e = threadIdx.x; e < P * n * n * Q; e += NT) As[e] = A[e];
  __syncthreads();
  const float inb = 1.0f / (float)nb, ib = 1.0f / (float)b, in = 1.0f / (float)n;
  TT_LDS double* out = (TT_LDS double*)M2;
  for (int e = threadIdx.x; e < k2 * nb; e += NT) {
    const int k = fquot(e, inb), j = e - k * nb;
    const int pn = fquot(k, ib), bb = k - pn * b;
    const int p = fquot(pn, in), nn = pn - p * n;
    const int mm = fquot(j, ib), bp = j - mm * b;
    const TT_LDS double* ap = As + ((p * n + mm) * n + nn) * Q;
    const TT_LDS double* rp = Rs + bb * Q * b + bp;
    double s = 0.0;
    for (int q = 0; q < Q; ++q) s = fma(ap[q], rp[q * b], s);
    out[e] = s;
  }
  __syncthreads();
}

// y = (L x M2) v on-chip; t1 holds T1' (a * odd_stride(P n b)), red the
// K-split scratch (m2_red_need).  Out of line for the same register reason
// as local_matvec_mf.
__device__ __attribute__((noinline)) void local_matvec_m2(int a, int P, int n, int b, const double* L,
                                                          const double* M2, const double* v, double* y, double* t1,
                                                          double* red) {
  const int nb = n * b, k2 = P * n * b, s = odd_stride(k2);
  const FMap plain = fmap(1, 1, 0);
  mf_gemm_lin(P * a, nb, a, L, 1, P * a, v, nb, 1, t1, fmap(a, nb, s), plain, nullptr, 0);
  mf_gemm_lin(a, nb, k2, t1, s, 1, M2, nb, 1, y, fmap(1, nb, 0), plain, red, m2_red_need(a, nb));
}

// Lout[b, Q, b'] = sum L[a,P,a'] x[a,n,b] A[P,m,n,Q] x[a',m,b']
template <class T, class ACC>
__device__ void stack_left_op(int a, int P, int n, int Q, int b, const T* L, const T* Ac, const T* x,
                              T* Lout, T* t1, T* t2, ACC* lds) {
  stage_left1<T, ACC>(a, P, n, b, L, x, t1, lds);
  // T2[b, Q, a', m] = sum_{P,n} T1[(a',b),(P,n)] A[P,m,n,Q]
  wg_gemm<T, ACC>(a * b, n * Q, P * n, crm<T>(t1, P * n),
                  cview<T>(Ac, split(n, n * n * Q, Q), split(Q, n * Q, 1)),
                  mview<T>(t2, split(b, n, Q * a * n), split(Q, 1, a * n)), lds);
  // Lout[(b,Q), b'] = T2[(b,Q),(a',m)] x[(a',m), b']
  wg_gemm<T, ACC>(b * Q, b, a * n, crm<T>(t2, a * n), crm<T>(x, b), mrm<T>(Lout, b), lds);
}

// Rout[a, P, a'] = sum x[a,n,b] A[P,m,n,Q] R[b,Q,b'] x[a',m,b']
template <class T, class ACC>
__device__ void stack_right_op(int a, int P, int n, int Q, int b, const T* R, const T* Ac, const T* x,
                               T* Rout, T* t1, T* t2, ACC* lds) {
  // U1[a', b, m, Q] = sum_b' x[(a',m), b'] R[(b,Q), b']
  wg_gemm<T, ACC>(a * n, b * Q, b, crm<T>(x, b), ctr<T>(R, b),
                  mview<T>(t1, split(n, b * n * Q, Q), split(Q, n * Q, 1)), lds);
  // U2[n, b, P, a'] = sum_{m,Q} U1[(a',b),(m,Q)] A[P,m,n,Q]
  wg_gemm<T, ACC>(a * b, P * n, n * Q, crm<T>(t1, n * Q),
                  cview<T>(Ac, split(Q, n * Q, 1), split(n, n * n * Q, Q)),
                  mview<T>(t2, split(b, 1, P * a), split(n, a, b * P * a)), lds);
  // Rout[a, (P,a')] = x[a,(n,b)] U2[(n,b),(P,a')]
  wg_gemm<T, ACC>(a, P * a, n * b, crm<T>(x, n * b), crm<T>(t2, P * a), mrm<T>(Rout, P * a), lds);
}

// Lbout[q, b'] = sum Lb[p,a'] B[p,m,q] x[a',m,b']
template <class T, class ACC>
__device__ void stack_left_rhs(int a, int p, int n, int q, int b, const T* Lb, const T* Bc, const T* x,
                               T* Lbout, T* t1, ACC* lds) {
  // W[q, a', m] = sum_p Lb[p,a'] B[p,(m,q)]
  wg_gemm<T, ACC>(a, n * q, p, ctr<T>(Lb, a), crm<T>(Bc, n * q),
                  mview<T>(t1, plain(n), split(q, 1, a * n)), lds);
  wg_gemm<T, ACC>(q, b, a * n, crm<T>(t1, a * n), crm<T>(x, b), mrm<T>(Lbout, b), lds);
}

// Rbout[p, a'] = sum B[p,m,q] Rb[q,b'] x[a',m,b']
template <class T, class ACC>
__device__ void stack_right_rhs(int a, int p, int n, int q, int b, const T* Rb, const T* Bc, const T* x,
                                T* Rbout, T* t1, ACC* lds) {
  // W[m, q, a'] = sum_b' x[(a',m), b'] Rb[q, b']
  wg_gemm<T, ACC>(a * n, q, b, crm<T>(x, b), ctr<T>(Rb, b),
                  mview<T>(t1, split(n, 1, q * a), plain(a)), lds);
  wg_gemm<T, ACC>(p, a, n * q, crm<T>(Bc, n * q), crm<T>(t1, a), mrm<T>(Rbout, a), lds);
}

// f[a', m, b'] = sum Lb[p,a'] B[p,m,q] Rb[q,b']
template <class T, class ACC>
__device__ void micro_rhs(int a, int p, int n, int q, int b, const T* Lb, const T* Bc, const T* Rb, T* f,
                          T* t1, ACC* lds) {
  wg_gemm<T, ACC>(a, n * q, p, ctr<T>(Lb, a), crm<T>(Bc, n * q), mrm<T>(t1, n * q), lds);
  wg_gemm<T, ACC>(a * n, b, q, crm<T>(t1, q), crm<T>(Rb, b), mrm<T>(f, b), lds);
}

// Dense local matrix M[(a',m,b'), (a,n,b)] = sum_{P,Q} L[a,P,a'] A[P,m,n,Q] R[b,Q,b']
// written with leading dimension ldm (used for small local systems).
template <class T>
__device__ void micro_dense(int a, int P, int n, int Q, int b, const T* L, const T* Ac, const T* R, T* Mout,
                            int ldm) {
  const int N = a * n * b;
  const long NN = (long)N * N;
  for (long e = threadIdx.x; e < NN; e += NT) {
    const int row = (int)(e / N), col = (int)(e - (long)row * N);
    const int a2 = row / (n * b), m = (row / b) % n, b2 = row % b;
    const int a1 = col / (n * b), nn = (col / b) % n, b1 = col % b;
    double s = 0.0;
    for (int pp = 0; pp < P; ++pp) {
      const double l = (double)L[(a1 * P + pp) * a + a2];
      if (l == 0.0) continue;
      for (int qq = 0; qq < Q; ++qq)
        s += l * (double)Ac[((pp * n + m) * n + nn) * Q + qq] * (double)R[(b1 * Q + qq) * b + b2];
    }
    Mout[(long)row * ldm + col] = (T)s;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Householder QR of a row-major m x c matrix W (ld = c), in place.
// On exit W holds R on/above the diagonal and the reflectors below it
// (unit leading entry implied); tau[0..k) with k = min(m, c).
// Applies H = I - t v v^T (v: rows j..m-1 of column 0 of `vcol`, leading
// dimension ldv, v_j = 1 implied) to columns c0..c1-1 of the row-major matrix
// X (leading dimension ldx), rows j..m-1.  X and vcol may be the same matrix
// (v's column lies outside c0..c1-1).  Waves own disjoint columns: no barrier
// inside.
//  - wide (>= 48 columns): each wave takes blocks of 16 consecutive columns
//    with its four 16-lane rows on interleaved matrix rows, so a load reads
//    four contiguous 128-byte row segments; four rows per lane are loaded
//    before any is used or stored (the stores may alias later loads, which
//    would otherwise serialise every row on its load latency); the four
//    partial dot products are combined by two lane exchanges, in the same
//    order in every lane;
//  - narrow: the 64 lanes of a wave walk the rows, four columns per wave at
//    a time, reduced over the wave; both passes load two rows per lane
//    before using either.
__device__ __forceinline__ void hh_apply(double* X, int ldx, int m, int j, const double* vcol, int ldv, int c0,
                                         int c1, double t) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (c1 - c0 >= 48) {
    const int cl = lane & 15, g = lane >> 4;
    const int nblk = (c1 - c0 + 15) >> 4;
    for (int blk = w; blk < nblk; blk += NWAVE) {
      const int col = c0 + (blk << 4) + cl;
      const bool on = col < c1;
      const int cc = on ? col : c0;  // in-bounds column for lanes past the end
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      int i = j + g;
      for (; i + 12 < m; i += 16) {
        const double v0 = (i == j) ? 1.0 : vcol[(long)i * ldv], v1 = vcol[(long)(i + 4) * ldv];
        const double v2 = vcol[(long)(i + 8) * ldv], v3 = vcol[(long)(i + 12) * ldv];
        const double x0 = X[(long)i * ldx + cc], x1 = X[(long)(i + 4) * ldx + cc];
        const double x2 = X[(long)(i + 8) * ldx + cc], x3 = X[(long)(i + 12) * ldx + cc];
        s0 += v0 * x0;
        s1 += v1 * x1;
        s2 += v2 * x2;
        s3 += v3 * x3;
      }
      for (; i < m; i += 4) s0 += ((i == j) ? 1.0 : vcol[(long)i * ldv]) * X[(long)i * ldx + cc];
      double sm = (s0 + s1) + (s2 + s3);
      sm += __shfl_xor(sm, 16, 64);
      sm += __shfl_xor(sm, 32, 64);
      sm *= t;
      if (on) {
        i = j + g;
        for (; i + 12 < m; i += 16) {
          const double v0 = (i == j) ? 1.0 : vcol[(long)i * ldv], v1 = vcol[(long)(i + 4) * ldv];
          const double v2 = vcol[(long)(i + 8) * ldv], v3 = vcol[(long)(i + 12) * ldv];
          const double x0 = X[(long)i * ldx + col], x1 = X[(long)(i + 4) * ldx + col];
          const double x2 = X[(long)(i + 8) * ldx + col], x3 = X[(long)(i + 12) * ldx + col];
          X[(long)i * ldx + col] = x0 - v0 * sm;
          X[(long)(i + 4) * ldx + col] = x1 - v1 * sm;
          X[(long)(i + 8) * ldx + col] = x2 - v2 * sm;
          X[(long)(i + 12) * ldx + col] = x3 - v3 * sm;
        }
        for (; i < m; i += 4) X[(long)i * ldx + col] -= ((i == j) ? 1.0 : vcol[(long)i * ldv]) * sm;
      }
    }
    return;
  }
  for (int q0 = c0 + w; q0 < c1; q0 += 4 * NWAVE) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    int i = j + lane;
    for (; i + 64 < m; i += 128) {
      const double v0 = (i == j) ? 1.0 : vcol[(long)i * ldv], v1 = vcol[(long)(i + 64) * ldv];
      double x0[4], x1[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int col = q0 + q * NWAVE;
        x0[q] = col < c1 ? X[(long)i * ldx + col] : 0.0;
        x1[q] = col < c1 ? X[(long)(i + 64) * ldx + col] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        s[q] += v0 * x0[q];  // row order kept: the same sums as one row at a time
        s[q] += v1 * x1[q];
      }
    }
    if (i < m) {
      const double v = (i == j) ? 1.0 : vcol[(long)i * ldv];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int col = q0 + q * NWAVE;
        if (col < c1) s[q] += v * X[(long)i * ldx + col];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) s[q] = wave_sum(s[q]) * t;
    i = j + lane;
    for (; i + 64 < m; i += 128) {
      const double v0 = (i == j) ? 1.0 : vcol[(long)i * ldv], v1 = vcol[(long)(i + 64) * ldv];
      double x0[4], x1[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int col = q0 + q * NWAVE;
        x0[q] = col < c1 ? X[(long)i * ldx + col] : 0.0;
        x1[q] = col < c1 ? X[(long)(i + 64) * ldx + col] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int col = q0 + q * NWAVE;
        if (col < c1) {
          X[(long)i * ldx + col] = x0[q] - v0 * s[q];
          X[(long)(i + 64) * ldx + col] = x1[q] - v1 * s[q];
        }
      }
    }
    if (i < m) {
      const double v = (i == j) ? 1.0 : vcol[(long)i * ldv];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int col = q0 + q * NWAVE;
        if (col < c1) X[(long)i * ldx + col] -= v * s[q];
      }
    }
  }
}

// The c x c part of one Cholesky QR pass: Rp holds the (shifted) Gram matrix
// on entry.  Cholesky Rp = R^T R (upper R in place, right-looking, lower part
// zeroed), then Rt <- R Rt (the accumulated factor; row by row in place, row
// p reads rows >= p of the old Rt), then Rp <- R^{-1} (column by column: for
// i < j, inv[i][j] = -(sum_{k=i}^{j-1} inv[i][k] R[k][j]) / R[j][j]).  tmp
// holds c doubles.  Returns false (nothing else changed after the failing
// column) when a pivot is not positive and finite: the Gram matrix is
// numerically singular.
__device__ __forceinline__ bool cholqr_small_step(double* Rp, double* Rt, double* tmp, int c) {
  for (int j = 0; j < c; ++j) {
    const double piv = Rp[(long)j * c + j];
    if (!(piv > 0.0) || !isfinite(piv)) return false;
    const double rjj = sqrt(piv), inv = 1.0 / rjj;
    __syncthreads();
    for (int q = j + 1 + threadIdx.x; q < c; q += NT) {
      Rp[(long)j * c + q] *= inv;
      Rp[(long)q * c + j] = 0.0;
    }
    if (threadIdx.x == 0) Rp[(long)j * c + j] = rjj;
    __syncthreads();
    const int rest = c - j - 1;
    for (int e = threadIdx.x; e < rest * rest; e += NT) {
      const int p = j + 1 + e / rest, q = j + 1 + e % rest;
      if (q >= p) Rp[(long)p * c + q] -= Rp[(long)j * c + p] * Rp[(long)j * c + q];
    }
    __syncthreads();
  }
  for (int p = 0; p < c; ++p) {
    for (int q = threadIdx.x; q < c; q += NT) {
      double t = 0.0;
      for (int k2 = p; k2 <= q; ++k2) t += Rp[(long)p * c + k2] * Rt[(long)k2 * c + q];
      tmp[q] = (q >= p) ? t : 0.0;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < c; q += NT) Rt[(long)p * c + q] = tmp[q];
    __syncthreads();
  }
  for (int j = 0; j < c; ++j) {
    const double rjj = Rp[(long)j * c + j];
    for (int i = threadIdx.x; i <= j; i += NT) {
      if (i == j) {
        tmp[i] = 1.0 / rjj;
      } else {
        double t = 0.0;
        for (int k2 = i; k2 < j; ++k2) t += Rp[(long)i * c + k2] * Rp[(long)k2 * c + j];
        tmp[i] = -t / rjj;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= j; i += NT) Rp[(long)i * c + j] = tmp[i];
    __syncthreads();
  }
  return true;
}

// The shift of the first shifted Cholesky QR pass for an (m x c) panel whose
// Gram matrix has trace tr (= ||W||_F^2): 11 (m c + c (c + 1)) u tr keeps the
// shifted Gram matrix positive definite in floating point for condition
// numbers up to ~1/u (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa,
// "Shifted Cholesky QR for computing the QR factorization of ill-conditioned
// matrices", SIAM J. Sci. Comput. 42 (2020)).
__device__ __forceinline__ double cholqr_shift(int m, int c, double tr) {
  return 11.0 * ((double)m * c + (double)c * (c + 1)) * 0x1p-53 * tr;
}

__device__ void wg_householder(double* W, int m, int c, double* tau, double* red) {
  const int k = m < c ? m : c;
  for (int j = 0; j < k; ++j) {
    double s2 = 0.0;
    for (int i = j + 1 + threadIdx.x; i < m; i += NT) {
      const double t = W[(long)i * c + j];
      s2 += t * t;
    }
    s2 = wg_sum(s2, red);
    const double alpha = W[(long)j * c + j];
    double t;
    if (s2 == 0.0) {
      t = 0.0;  // H_j = I (LAPACK convention for an already-reduced column)
    } else {
      const double nrm = sqrt(alpha * alpha + s2);
      const double beta = alpha >= 0.0 ? -nrm : nrm;
      t = (beta - alpha) / beta;
      const double sc = 1.0 / (alpha - beta);
      for (int i = j + 1 + threadIdx.x; i < m; i += NT) W[(long)i * c + j] *= sc;
      __syncthreads();
      if (threadIdx.x == 0) W[(long)j * c + j] = beta;
    }
    if (threadIdx.x == 0) tau[j] = t;
    __syncthreads();
    if (t != 0.0) hh_apply(W, c, m, j, W + j, c, j + 1, c, t);
    __syncthreads();
  }
}

__device__ __forceinline__ double rcp_nr(double x) {  // 1/x: hardware estimate + two Newton steps
  double r = __builtin_amdgcn_rcp(x);
  r = r * (2.0 - x * r);
  return r * (2.0 - x * r);
}

// Sums over each 16-lane row (DPP row) of the wave, C of them advanced stage
// by stage so their latencies overlap; every lane of a row gets its row's sum.
template <int C>
__device__ __forceinline__ void row16_sum_n(double (&p)[C]) {
#pragma unroll
  for (int u = 0; u < C; ++u) p[u] += dpp_mov<0xB1>(p[u]);
#pragma unroll
  for (int u = 0; u < C; ++u) p[u] += dpp_mov<0x4E>(p[u]);
#pragma unroll
  for (int u = 0; u < C; ++u) p[u] += dpp_mov<0x141>(p[u]);
#pragma unroll
  for (int u = 0; u < C; ++u) p[u] += dpp_mov<0x140>(p[u]);
}

// Householder QR of a small (m x c) row-major panel W, m <= 16 R, c <= 16 C,
// LAPACK convention (v_j = 1 at row j, H_j = I - tau_j v_j v_j^T, beta =
// -sign(alpha) |x|).  The workgroup's 16 DPP rows (4 waves x 4 rows of 16
// lanes) each own the columns r, r+16, ... (r = DPP row); lane l of a row
// holds the panel rows l, l+16, ... of those columns in registers, so a
// column's dot product is an in-register sum plus a 16-lane DPP reduction,
// with no cross-row traffic.  Step j: column j's owner row forms the
// reflector and publishes it in W's column j (finished R entries above the
// diagonal, beta on it, v below) with tau[j]; one barrier; every row applies
// it to its later columns.  Explicit Q (m x k, ld k) goes to Qm by backward
// accumulation in the same layout, the clean R (k x c) to Rm; W keeps the
// reflectors.
template <int R, int C>
__device__ void wg_qr_small_t(double* W, int m, int c, double* tau, double* Qm, double* Rm) {
  const int l16 = threadIdx.x & 15, r = threadIdx.x >> 4;  // lane in row, DPP row 0..15
  const int k = m < c ? m : c;
  double a[R][C];
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int u = 0; u < C; ++u) {
      const int i = l16 + 16 * q, jj = r + 16 * u;
      a[q][u] = (i < m && jj < c) ? W[i * c + jj] : 0.0;
    }
  __syncthreads();  // W is rewritten column by column below
  for (int j = 0; j < k; ++j) {
    if (r == (j & 15)) {
      const int uj = j >> 4;
      double x[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        x[q] = 0.0;
#pragma unroll
        for (int u = 0; u < C; ++u)
          if (u == uj) x[q] = a[q][u];
      }
      double sa[2] = {0.0, 0.0};  // sum of squares below the diagonal, alpha
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = l16 + 16 * q;
        if (i > j && i < m) sa[0] += x[q] * x[q];
        if (i == j) sa[1] = x[q];
      }
      row16_sum_n<2>(sa);  // alpha is the row's one nonzero term: exact
      const double alpha = sa[1];
      double t = 0.0, beta = alpha, sc = 0.0;
      if (sa[0] != 0.0) {
        const double nrm = sqrt(alpha * alpha + sa[0]);
        beta = alpha >= 0.0 ? -nrm : nrm;
        t = (beta - alpha) * rcp_nr(beta);
        sc = rcp_nr(alpha - beta);
      }
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = l16 + 16 * q;
        if (i < m) W[i * c + j] = i < j ? x[q] : (i == j ? beta : x[q] * sc);
      }
      if (l16 == 0) tau[j] = t;
    }
    __syncthreads();
    // tau and the reflector in one batch of LDS reads (rows past m read row m-1)
    const double t = tau[j];
    double v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = l16 + 16 * q;
      const double wv = W[(i < m ? i : m - 1) * c + j];
      v[q] = (i == j) ? 1.0 : ((i > j && i < m) ? wv : 0.0);
    }
    if (t != 0.0) {
      double p[C];
#pragma unroll
      for (int u = 0; u < C; ++u) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < R; ++q) s += v[q] * a[q][u];
        p[u] = (r + 16 * u > j) ? s : 0.0;
      }
      row16_sum_n<C>(p);
#pragma unroll
      for (int u = 0; u < C; ++u) {
        const double pt = p[u] * t;
#pragma unroll
        for (int q = 0; q < R; ++q) a[q][u] -= v[q] * pt;
      }
    }
  }
  // columns k..c-1 (wide panels) are R rows only: publish them
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int u = 0; u < C; ++u) {
      const int i = l16 + 16 * q, jj = r + 16 * u;
      if (jj >= k && jj < c && i < m) W[i * c + jj] = a[q][u];
    }
  __syncthreads();
  const float ic = 1.0f / (float)c;
  for (int e = threadIdx.x; e < k * c; e += NT) {
    const int i = fquot(e, ic), jj = e - i * c;
    Rm[e] = (jj >= i) ? W[i * c + jj] : 0.0;
  }
  // explicit Q = H_0 ... H_{k-1} I by backward accumulation; H_j leaves the
  // identity columns < j untouched (v_j vanishes above row j), and W, tau are
  // final, so no barrier is needed
  double qm[R][C];
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int u = 0; u < C; ++u) qm[q][u] = (l16 + 16 * q == r + 16 * u) ? 1.0 : 0.0;
  // reflector j-1 is read while H_j is applied
  double tn = 0.0, wn[R];
#pragma unroll
  for (int q = 0; q < R; ++q) wn[q] = 0.0;
  if (k > 0) {
    tn = tau[k - 1];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = l16 + 16 * q;
      wn[q] = W[(i < m ? i : m - 1) * c + (k - 1)];
    }
  }
  for (int j = k - 1; j >= 0; --j) {
    const double t = tn;
    double v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = l16 + 16 * q;
      v[q] = (i == j) ? 1.0 : ((i > j && i < m) ? wn[q] : 0.0);
    }
    if (j > 0) {
      tn = tau[j - 1];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = l16 + 16 * q;
        wn[q] = W[(i < m ? i : m - 1) * c + (j - 1)];
      }
    }
    if (t == 0.0) continue;
    double p[C];
#pragma unroll
    for (int u = 0; u < C; ++u) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < R; ++q) s += v[q] * qm[q][u];
      p[u] = s;
    }
    row16_sum_n<C>(p);
#pragma unroll
    for (int u = 0; u < C; ++u) {
      const double pt = p[u] * t;
#pragma unroll
      for (int q = 0; q < R; ++q) qm[q][u] -= v[q] * pt;
    }
  }
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int u = 0; u < C; ++u) {
      const int i = l16 + 16 * q, jj = r + 16 * u;
      if (i < m && jj < k) Qm[i * k + jj] = qm[q][u];
    }
  __syncthreads();
}

// m, c <= 64
__device__ void wg_qr_small(double* W, int m, int c, double* tau, double* Qm, double* Rm) {
  if (m <= 16 && c <= 16) wg_qr_small_t<1, 1>(W, m, c, tau, Qm, Rm);
  else if (m <= 32 && c <= 16) wg_qr_small_t<2, 1>(W, m, c, tau, Qm, Rm);
  else if (m <= 32 && c <= 32) wg_qr_small_t<2, 2>(W, m, c, tau, Qm, Rm);
  else wg_qr_small_t<4, 4>(W, m, c, tau, Qm, Rm);
}

// Explicit Q (m x k, row-major, ld = k) from the reflectors in W.
__device__ void wg_form_q(const double* W, int m, int c, const double* tau, double* Qm) {
  const int k = m < c ? m : c;
  for (long e = threadIdx.x; e < (long)m * k; e += NT) Qm[e] = 0.0;
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += NT) Qm[(long)i * k + i] = 1.0;
  __syncthreads();
  for (int j = k - 1; j >= 0; --j) {
    const double t = tau[j];
    if (t != 0.0) hh_apply(Qm, k, m, j, W + j, c, j, k, t);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Blocked Householder QR of a tall row-major (m x c) panel W in global memory
// (m >= c): the structure of LAPACK's dgeqrf / dlarft / dlarfb / dorgqr.
// Column blocks of NB are factored with the block in LDS (reflectors in
// wg_householder's dlarfg convention); each block's compact WY form
// I - V T V^T then updates the trailing columns, and backwards forms the
// explicit Q, as two streaming passes over the rows per block (Y = V^T X,
// X -= V (T' Y)) where the unblocked sweep makes two per column.  The Gram
// entries V^T V that T needs ride along with the column steps' reductions.

template <int NB>
struct HHBlockLDS {
  // rows x (NB + 1) block (odd stride: the rows of neighbouring threads fall
  // into different banks), then the reduction rows, the Y chunk, T and tau
  static __host__ __device__ long need(int m) { return (long)m * (NB + 1) + NWAVE * NB + NB * 64 + NB * NB + NB; }
};

// V(i, p) of the block's reflectors: 1 on the diagonal, 0 above, Pn below.
template <int NB>
__device__ __forceinline__ double hh_v(const double* Pn, int i, int p) {
  return i > p ? Pn[i * (NB + 1) + p] : (i == p ? 1.0 : 0.0);
}

// Factor the block Pn (rows x kb, ld NB + 1) in place (R on and above the
// diagonal, reflectors below), tau[0..kb) and the upper part of G = V^T V
// (G[l * NB + j], l < j).  dred: NWAVE * NB doubles.  One reduction per
// column: over the rows below the pivot, with x the column, |x|^2, x^T X_q
// (later columns) and x^T V_l (earlier reflectors) together; then with
// v = (1, sc x) the update's w_q = X_jq + sc x^T X_q and the Gram entries
// G[l][j] = V_jl + sc x^T V_l follow without a second pass.  Thread t owns
// the rows t, t + NT, ... in every step, so a row is written and read by one
// thread; the pivot row j, which every thread reads after the step's
// barrier, collects its new values in prow (NB * NB + NB doubles of LDS,
// with the step's w_q) until the block is done.  Two barriers per column.
template <int NB>
__device__ void hh_block_factor(double* Pn, int rows, int kb, double* tau, double* G, double* dred, double* prow) {
  constexpr int ld = NB + 1;
  double* wqs = prow + NB * NB;  // the step's w_q
  static_assert(NB < NT, "a block's pivot rows lie in the first row of every thread's set");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tid = threadIdx.x;
  for (int j = 0; j < kb; ++j) {
    const int i_gt = tid > j ? tid : tid + NT;  // first own row below the pivot
    double acc[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) acc[q] = 0.0;
    for (int i = i_gt; i < rows; i += NT) {
      const double x = Pn[i * ld + j];
#pragma unroll
      for (int q = 0; q < NB; ++q)
        if (q < kb) acc[q] += x * (q == j ? x : Pn[i * ld + q]);
    }
#pragma unroll
    for (int q = 0; q < NB; ++q)
      if (q < kb) acc[q] = wave_sum(acc[q]);
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < NB; ++q) dred[w * NB + q] = acc[q];
    }
    __syncthreads();
    double s2 = 0.0;
#pragma unroll
    for (int u = 0; u < NWAVE; ++u) s2 += dred[u * NB + j];
    const double alpha = Pn[j * ld + j];
    double t = 0.0, beta = alpha, sc = 0.0;
    if (s2 != 0.0) {
      const double nrm = sqrt(alpha * alpha + s2);
      beta = alpha >= 0.0 ? -nrm : nrm;
      t = (beta - alpha) / beta;
      sc = 1.0 / (alpha - beta);
    }
    // thread q < kb: w_q, the pivot row's new entry q and G[q][j] (q < j);
    // the pivot row itself is left alone (every thread reads it above)
    if (tid < kb) {
      double sm = 0.0;
#pragma unroll
      for (int u = 0; u < NWAVE; ++u) sm += dred[u * NB + tid];
      const double xq = Pn[j * ld + tid], wq = xq + sc * sm;
      wqs[tid] = wq;
      prow[j * NB + tid] = tid < j ? xq : (tid == j ? beta : xq - t * wq);
      if (tid < j) G[tid * NB + j] = wq;
    }
    if (tid == 0) tau[j] = t;
    __syncthreads();
    if (t != 0.0) {
      for (int i = i_gt; i < rows; i += NT) {
        const double v = Pn[i * ld + j] * sc;
        Pn[i * ld + j] = v;
#pragma unroll
        for (int q = 0; q < NB; ++q)
          if (q > j && q < kb) Pn[i * ld + q] -= t * v * wqs[q];
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < kb * kb; e += NT) {
    const int i = e / kb, q = e - i * kb;
    Pn[i * ld + q] = prow[i * NB + q];
  }
  __syncthreads();
}

// T of the block (dlarft, forward, columnwise): T[i][i] = tau_i,
// T[0:i, i] = -tau_i T[0:i, 0:i] G[0:i, i]; NB x NB, upper.
template <int NB>
__device__ void hh_block_t(const double* tau, const double* G, int kb, double* T) {
  for (int e = threadIdx.x; e < NB * NB; e += NT) T[e] = 0.0;
  __syncthreads();
  for (int i = 0; i < kb; ++i) {
    const int j = threadIdx.x;
    double v = 0.0;
    if (j < i) {
      for (int l = j; l < i; ++l) v += T[j * NB + l] * G[l * NB + i];
    }
    __syncthreads();
    if (j < i) T[j * NB + i] = -tau[i] * v;
    if (j == i) T[i * NB + i] = tau[i];
    __syncthreads();
  }
}

// X (rows x nc, global, ld ldx) <- (I - V T' V^T) X with T' = T^T (trans:
// H^T X, the factorisation's trailing update) or T (forming Q).  A wave's
// lanes cover ncp columns (nc rounded up to 8, 16, 32 or 64; wider X in
// chunks of 64) and 64 / ncp row groups; each lane streams its group's rows
// with HH_U loads issued before the first is used (the passes are bound by
// memory latency, not arithmetic).  Partial Y meet within the wave by lane
// exchanges and across the waves in Yb (NB x 64 doubles), in a fixed order.
// HH_U: 2 where the kernel must stay within 256 registers for two workgroups
// per CU (k_als), 8 in a translation unit that defines TTALS_HH_ROWS for a
// one-workgroup-per-CU kernel.
#ifndef TTALS_HH_ROWS
#define TTALS_HH_ROWS 2
#endif
constexpr int HH_U = TTALS_HH_ROWS;
template <int NB>
__device__ void hh_block_apply(double* X, int ldx, int rows, int nc, const double* Pn, int kb, const double* T,
                               bool trans, double* Yb) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ncp = nc > 32 ? 64 : (nc > 16 ? 32 : (nc > 8 ? 16 : 8));
  const int rg = 64 / ncp, cl = lane & (ncp - 1), sub = lane / ncp;
  const int stride = NWAVE * rg;  // rows between a lane's consecutive rows
  const int r0 = w * rg + sub;
  for (int c0 = 0; c0 < nc; c0 += ncp) {
    const int col = c0 + cl;
    const bool on = col < nc;
    const int cc = on ? col : c0;
    double y[NB];
#pragma unroll
    for (int p = 0; p < NB; ++p) y[p] = 0.0;
    int i = r0;
    if constexpr (HH_U >= 4) {
      // registers to spare: the next group's loads are issued before this
      // group is used, and the reflector reads are left to the scheduler
      double xn[HH_U];
      if (i + (HH_U - 1) * stride < rows) {
#pragma unroll
        for (int u = 0; u < HH_U; ++u) xn[u] = X[(long)(i + u * stride) * ldx + cc];
      }
      for (; i + (HH_U - 1) * stride < rows; i += HH_U * stride) {
        double x[HH_U];
#pragma unroll
        for (int u = 0; u < HH_U; ++u) x[u] = xn[u];
        const int inext = i + HH_U * stride;
        if (inext + (HH_U - 1) * stride < rows) {
#pragma unroll
          for (int u = 0; u < HH_U; ++u) xn[u] = X[(long)(inext + u * stride) * ldx + cc];
        }
#pragma unroll
        for (int u = 0; u < HH_U; ++u)
#pragma unroll
          for (int p = 0; p < NB; ++p)
            if (p < kb) y[p] += hh_v<NB>(Pn, i + u * stride, p) * x[u];
      }
    } else {
      for (; i + (HH_U - 1) * stride < rows; i += HH_U * stride) {
        double x[HH_U];
#pragma unroll
        for (int u = 0; u < HH_U; ++u) x[u] = X[(long)(i + u * stride) * ldx + cc];
#pragma unroll
        for (int u = 0; u < HH_U; ++u) {
          __builtin_amdgcn_sched_barrier(0);  // one row's reflector entries live at a time
#pragma unroll
          for (int p = 0; p < NB; ++p)
            if (p < kb) y[p] += hh_v<NB>(Pn, i + u * stride, p) * x[u];
        }
      }
    }
    for (; i < rows; i += stride) {
      const double x = X[(long)i * ldx + cc];
#pragma unroll
      for (int p = 0; p < NB; ++p)
        if (p < kb) y[p] += hh_v<NB>(Pn, i, p) * x;
    }
    // row groups of the wave: lanes cl, cl + ncp, ... (fixed butterfly order)
    for (int o = ncp; o < 64; o <<= 1) {
#pragma unroll
      for (int p = 0; p < NB; ++p) y[p] += __shfl_xor(y[p], o, 64);
    }
    for (int u = 0; u < NWAVE; ++u) {
      if (w == u && sub == 0) {
#pragma unroll
        for (int p = 0; p < NB; ++p) Yb[p * 64 + cl] = (u == 0 ? 0.0 : Yb[p * 64 + cl]) + y[p];
      }
      __syncthreads();
    }
    // Z = T' Y in place, one wave (a lane per column; Y's column in registers)
    if (w == 0 && sub == 0) {
#pragma unroll
      for (int p = 0; p < NB; ++p) y[p] = Yb[p * 64 + cl];
#pragma unroll 1
      for (int p = 0; p < NB; ++p) {
        double sm = 0.0;
#pragma unroll
        for (int l = 0; l < NB; ++l) {
          const double tv = trans ? (l <= p ? T[l * NB + p] : 0.0) : (l >= p ? T[p * NB + l] : 0.0);
          sm += tv * y[l];
        }
        Yb[p * 64 + cl] = sm;
      }
    }
    __syncthreads();
    double z[NB];
#pragma unroll
    for (int p = 0; p < NB; ++p) z[p] = Yb[p * 64 + cl];
    if (on) {
      i = r0;
      if constexpr (HH_U >= 4) {
        double xn[HH_U];
        if (i + (HH_U - 1) * stride < rows) {
#pragma unroll
          for (int u = 0; u < HH_U; ++u) xn[u] = X[(long)(i + u * stride) * ldx + col];
        }
        for (; i + (HH_U - 1) * stride < rows; i += HH_U * stride) {
          double x[HH_U];
#pragma unroll
          for (int u = 0; u < HH_U; ++u) x[u] = xn[u];
          // (rows of the next group are disjoint from this group's stores)
          const int inext = i + HH_U * stride;
          if (inext + (HH_U - 1) * stride < rows) {
#pragma unroll
            for (int u = 0; u < HH_U; ++u) xn[u] = X[(long)(inext + u * stride) * ldx + col];
          }
#pragma unroll
          for (int u = 0; u < HH_U; ++u) {
            double sm = 0.0;
#pragma unroll
            for (int p = 0; p < NB; ++p)
              if (p < kb) sm += hh_v<NB>(Pn, i + u * stride, p) * z[p];
            X[(long)(i + u * stride) * ldx + col] = x[u] - sm;
          }
        }
      } else {
        for (; i + (HH_U - 1) * stride < rows; i += HH_U * stride) {
          double x[HH_U];
#pragma unroll
          for (int u = 0; u < HH_U; ++u) x[u] = X[(long)(i + u * stride) * ldx + col];
#pragma unroll
          for (int u = 0; u < HH_U; ++u) {
            __builtin_amdgcn_sched_barrier(0);
            double sm = 0.0;
#pragma unroll
            for (int p = 0; p < NB; ++p)
              if (p < kb) sm += hh_v<NB>(Pn, i + u * stride, p) * z[p];
            X[(long)(i + u * stride) * ldx + col] = x[u] - sm;
          }
        }
      }
      for (; i < rows; i += stride) {
        double sm = 0.0;
#pragma unroll
        for (int p = 0; p < NB; ++p)
          if (p < kb) sm += hh_v<NB>(Pn, i, p) * z[p];
        X[(long)i * ldx + col] -= sm;
      }
    }
    __syncthreads();
  }
}

// Pn (rows x kb, ld NB + 1) <- the block W[jb.., jb..jb+kb) (row stride c):
// a thread per row, the row's kb loads issued together.
template <int NB>
__device__ __forceinline__ void hh_load_block(const double* W, int c, int jb, int rows, int kb, double* Pn) {
  for (int i = threadIdx.x; i < rows; i += NT) {
    const double* src = W + (long)(jb + i) * c + jb;
    double v[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) v[q] = q < kb ? src[q] : 0.0;
#pragma unroll
    for (int q = 0; q < NB; ++q)
      if (q < kb) Pn[i * (NB + 1) + q] = v[q];
  }
}

template <int NB>
__device__ __forceinline__ void hhqr_blocked_run(double* W, int m, int c, double* Qm, double* Rm, double* tws,
                                                 double* tau_all, double* lds, long lim, double* red) {
  constexpr int ld = NB + 1;
  double* Pn = lds;
  double* dred = Pn + (long)m * ld;
  double* Yb = dred + NWAVE * NB;
  double* T = Yb + NB * 64;
  double* tau = T + NB * NB;
  const int k = c;  // m >= c
  for (int jb = 0; jb < k; jb += NB) {
    const int kb = (k - jb) < NB ? (k - jb) : NB, rows = m - jb;
    hh_load_block<NB>(W, c, jb, rows, kb, Pn);
    __syncthreads();
    double* G = tws + (long)jb * NB;  // the block's Gram rows, then its T
    hh_block_factor<NB>(Pn, rows, kb, tau, G, dred, Yb);
    for (int i = threadIdx.x; i < rows; i += NT) {
      double* dst = W + (long)(jb + i) * c + jb;
#pragma unroll
      for (int q = 0; q < NB; ++q)
        if (q < kb) dst[q] = Pn[i * ld + q];
    }
    if (threadIdx.x < kb) tau_all[jb + threadIdx.x] = tau[threadIdx.x];
    hh_block_t<NB>(tau, G, kb, T);
    for (int e = threadIdx.x; e < NB * NB; e += NT) G[e] = T[e];
    if (jb + kb < c) hh_block_apply<NB>(W + (long)jb * c + jb + kb, c, rows, c - jb - kb, Pn, kb, T, true, Yb);
    __syncthreads();
  }
  // R
  for (long e = threadIdx.x; e < (long)k * c; e += NT) {
    const int i = (int)(e / c), j = (int)(e - (long)i * c);
    Rm[e] = (j >= i) ? W[e] : 0.0;
  }
  __syncthreads();
  if (GEMM_LDS_DOUBLES + 2L * k * k <= lim) {
    // Q = H_0 ... H_{k-1} [I; 0] = I - V T V1^T in one step, from the compact
    // WY form of all k reflectors (V: m x k, unit lower trapezoidal, made
    // explicit in W; V1 its top k x k; T: k x k upper, dlarft's recurrence
    // over the Gram matrix V^T V): Q = V (-T V1^T) + [I; 0], both products
    // on the matrix cores.
    double* Gs = lds + GEMM_LDS_DOUBLES;  // V^T V, then T in its upper part
    double* Ms = Gs + (long)k * k;
    for (long e = threadIdx.x; e < (long)k * c; e += NT) {
      const int i = (int)(e / c), j = (int)(e - (long)i * c);
      if (j >= i) W[e] = (j == i) ? 1.0 : 0.0;
    }
    __syncthreads();
    // three products through one GEMM call site (one inlined copy of its
    // registers): G = V^T V; Ms = T V1^T; Q = V (-Ms)
    for (int g = 0; g < 3; ++g) {
      const int gm = g == 2 ? m : k, gk = g == 0 ? m : k;
      const CView<double> A = g == 0 ? ctr<double>(W, c) : (g == 1 ? crm<double>(Gs, k) : crm<double>(W, c));
      const CView<double> Bv = g == 0 ? crm<double>(W, c) : (g == 1 ? ctr<double>(W, c) : crm<double>(Ms, k));
      const MView<double> Cv = g == 0 ? mrm<double>(Gs, k) : (g == 1 ? mrm<double>(Ms, k) : mrm<double>(Qm, k));
      wg_gemm_mf(gm, k, gk, A, Bv, Cv, lds);
      if (g == 0) {
        // T column by column over G's upper part: T[j][i] = -tau_i
        // sum_{l=j}^{i-1} T[j][l] G[l][i] (j < i), T[i][i] = tau_i
        for (int i = 0; i < k; ++i) {
          const int j = threadIdx.x;
          double v = 0.0;
          if (j < i)
            for (int l = j; l < i; ++l) v += Gs[j * k + l] * Gs[l * k + i];
          __syncthreads();
          if (j < i) Gs[j * k + i] = -tau_all[i] * v;
          if (j == i) Gs[i * k + i] = tau_all[i];
          __syncthreads();
        }
        for (int e = threadIdx.x; e < k * k; e += NT) {
          const int i = e / k, j = e - i * k;
          if (i > j) Gs[e] = 0.0;
        }
      } else if (g == 1) {
        for (int e = threadIdx.x; e < k * k; e += NT) Ms[e] = -Ms[e];
      } else {
        for (int j = threadIdx.x; j < k; j += NT) Qm[(long)j * k + j] += 1.0;
      }
      __syncthreads();
    }
    return;
  }
  // else Q = H_0 ... H_{k-1} I backwards over the blocks
  for (long e = threadIdx.x; e < (long)m * k; e += NT) {
    const int i = (int)(e / k), j = (int)(e - (long)i * k);
    Qm[e] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  const int nblk = (k + NB - 1) / NB;
  for (int b = nblk - 1; b >= 0; --b) {
    const int jb = b * NB, kb = (k - jb) < NB ? (k - jb) : NB, rows = m - jb;
    hh_load_block<NB>(W, c, jb, rows, kb, Pn);
    for (int e = threadIdx.x; e < NB * NB; e += NT) T[e] = tws[(long)jb * NB + e];
    __syncthreads();
    hh_block_apply<NB>(Qm + (long)jb * k + jb, k, rows, k - jb, Pn, kb, T, false, Yb);
  }
  __syncthreads();
}

// Blocked Householder QR (see above) of W (m x c, m >= c, global memory):
// Q (m x c, ld c) into Qm, R (c x c) into Rm; W is overwritten; tws (global)
// holds (c + 8) * 8 doubles, tau_all c.  The block (8 or 4 columns) is the
// larger whose LDS need fits min(cap, 61 KB) -- a bound independent of the
// engine's arena, so every engine factors a panel alike (and forms Q in one
// step when 2 c^2 doubles fit beside the GEMM tiles).  Returns false when
// neither block fits (the caller runs the unblocked sweep).
constexpr long HHQR_LDS_CAP = 61 * 1024 / 8;
__device__ bool wg_hhqr_blocked(double* W, int m, int c, double* Qm, double* Rm, double* tws, double* tau_all,
                                double* lds, long cap, double* red) {
  if (m < c || m >= (1 << 22) / 17) return false;
  const long lim = cap < HHQR_LDS_CAP ? cap : HHQR_LDS_CAP;
  if (HHBlockLDS<8>::need(m) <= lim)
    hhqr_blocked_run<8>(W, m, c, Qm, Rm, tws, tau_all, lds, lim, red);
  else if (HHBlockLDS<4>::need(m) <= lim)
    hhqr_blocked_run<4>(W, m, c, Qm, Rm, tws, tau_all, lds, lim, red);
  else
    return false;
  return true;
}

// ---------------------------------------------------------------------------
// One-sided (Hestenes) Jacobi SVD of a tall row-major c x k matrix B
// (c >= k), in place: on exit B = U * diag(sig) (columns orthogonal), V (k x
// k, row-major) accumulates the rotations, so B_in = U diag(sig) V^T.
// Round-robin pairing (k/2 disjoint column pairs per round, one wave per
// pair); sweeps until every pair satisfies |b_i.b_j| <= eps sqrt(|b_i|^2
// |b_j|^2) or 60 sweeps.  Columns are then sorted by decreasing sigma and
// the column norms returned in sig.  perm: int scratch of k.
__device__ void wg_jacobi_svd(double* B, int c, int k, double* V, double* sig, int* perm, double* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ int s_rot;
  for (long e = threadIdx.x; e < (long)k * k; e += NT) V[e] = ((e / k) == (e % k)) ? 1.0 : 0.0;
  __syncthreads();
  const int kk = (k & 1) ? k + 1 : k;  // padded for round-robin (index k = dummy)
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (threadIdx.x == 0) s_rot = 0;
    __syncthreads();
    for (int round = 0; round < kk - 1; ++round) {
      for (int pr = wv; pr < kk / 2; pr += NWAVE) {
        // round-robin: position 0 fixed, the others rotate
        int p0 = pr == 0 ? 0 : 1 + (pr - 1 + round) % (kk - 1);
        int p1 = 1 + (kk - 2 - pr + round) % (kk - 1);
        int i = p0 < p1 ? p0 : p1, j = p0 < p1 ? p1 : p0;
        if (j >= k) continue;  // dummy partner
        double aii = 0.0, ajj = 0.0, aij = 0.0;
        for (int r = lane; r < c; r += 64) {
          const double bi = B[(long)r * k + i], bj = B[(long)r * k + j];
          aii += bi * bi;
          ajj += bj * bj;
          aij += bi * bj;
        }
        aii = wave_sum(aii);
        ajj = wave_sum(ajj);
        aij = wave_sum(aij);
        if (fabs(aij) <= 1e-15 * sqrt(aii * ajj) || aij == 0.0) continue;
        const double zeta = (ajj - aii) / (2.0 * aij);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        for (int r = lane; r < c; r += 64) {
          const double bi = B[(long)r * k + i], bj = B[(long)r * k + j];
          B[(long)r * k + i] = cs * bi - sn * bj;
          B[(long)r * k + j] = sn * bi + cs * bj;
        }
        for (int r = lane; r < k; r += 64) {
          const double vi = V[(long)r * k + i], vj = V[(long)r * k + j];
          V[(long)r * k + i] = cs * vi - sn * vj;
          V[(long)r * k + j] = sn * vi + cs * vj;
        }
        if (lane == 0) s_rot = 1;
      }
      __syncthreads();
    }
    if (s_rot == 0) break;
    __syncthreads();
  }
  // column norms, then sort by decreasing norm (selection by thread 0)
  for (int j = wv; j < k; j += NWAVE) {
    double s2 = 0.0;
    for (int r = lane; r < c; r += 64) s2 += B[(long)r * k + j] * B[(long)r * k + j];
    s2 = wave_sum(s2);
    if (lane == 0) sig[j] = sqrt(s2);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < k; ++j) perm[j] = j;
    for (int j = 0; j < k; ++j) {
      int best = j;
      for (int q = j + 1; q < k; ++q)
        if (sig[perm[q]] > sig[perm[best]]) best = q;
      const int t = perm[j];
      perm[j] = perm[best];
      perm[best] = t;
    }
  }
  __syncthreads();
  (void)red;
}

// ---------------------------------------------------------------------------
// Dense LU with partial pivoting of an N x N row-major matrix in LDS/global
// memory (ld = N) and solve with right-hand side f (overwritten by the
// solution).  Returns false if an exactly zero pivot is met.
__device__ bool wg_lu_solve(double* Mx, int N, double* f, int* piv, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __shared__ int s_ip[NWAVE];
  __shared__ double s_iv[NWAVE];
  bool ok = true;
  for (int k = 0; k < N; ++k) {
    // pivot search: max |M[i][k]|, smallest index on ties
    double best = -1.0;
    int bi = k;
    for (int i = k + threadIdx.x; i < N; i += NT) {
      const double v = fabs(Mx[(long)i * N + k]);
      if (v > best) { best = v; bi = i; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    __syncthreads();
    if (lane == 0) { s_iv[w] = best; s_ip[w] = bi; }
    __syncthreads();
    best = s_iv[0];
    bi = s_ip[0];
    for (int q = 1; q < NWAVE; ++q)
      if (s_iv[q] > best || (s_iv[q] == best && s_ip[q] < bi)) { best = s_iv[q]; bi = s_ip[q]; }
    if (threadIdx.x == 0) piv[k] = bi;
    if (best == 0.0) { ok = false; }
    if (bi != k) {
      for (int j = threadIdx.x; j < N; j += NT) {
        const double t = Mx[(long)k * N + j];
        Mx[(long)k * N + j] = Mx[(long)bi * N + j];
        Mx[(long)bi * N + j] = t;
      }
      if (threadIdx.x == 0) { const double t = f[k]; f[k] = f[bi]; f[bi] = t; }
    }
    __syncthreads();
    if (best != 0.0) {
      const double inv = 1.0 / Mx[(long)k * N + k];
      for (int i = k + 1 + threadIdx.x; i < N; i += NT) Mx[(long)i * N + k] *= inv;
      __syncthreads();
      const int rem = N - k - 1;
      for (long e = threadIdx.x; e < (long)rem * rem; e += NT) {
        const int i = k + 1 + (int)(e / rem), j = k + 1 + (int)(e % rem);
        Mx[(long)i * N + j] -= Mx[(long)i * N + k] * Mx[(long)k * N + j];
      }
    }
    __syncthreads();
  }
  if (!ok) return false;
  // forward substitution (unit lower), then back substitution
  for (int k = 0; k < N; ++k) {
    const double fk = f[k];
    for (int i = k + 1 + threadIdx.x; i < N; i += NT) f[i] -= Mx[(long)i * N + k] * fk;
    __syncthreads();
  }
  for (int k = N - 1; k >= 0; --k) {
    if (threadIdx.x == 0) f[k] /= Mx[(long)k * N + k];
    __syncthreads();
    const double fk = f[k];
    for (int i = threadIdx.x; i < k; i += NT) f[i] -= Mx[(long)i * N + k] * fk;
    __syncthreads();
  }
  return true;
}

}  // namespace ttals
