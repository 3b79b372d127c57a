// Batched residual norms ||A x - b|| and ||b|| of TT systems, one workgroup
// per system, without forming A x - b: the cores of the residual train
//   Z_0 = [ (A x)_0  -b_0 ],  Z_i = [[ (A x)_i, 0 ], [ 0, b_i ]],
//   Z_{d-1} = [[ (A x)_{d-1} ], [ b_{d-1} ]]
// (the rank layout of scikit_tt/tensor_train.py:313-333 for x - y with the
// product ranks of tensor_train.py:471-479) are generated block by block,
// already multiplied by the R factor coming from the right, straight into the
// panel of a right-to-left Householder QR sweep.  Only R factors are kept:
// the norm is invariant under the orthogonal factors, and the last core's
// Frobenius norm is the residual norm (the orthogonalisation form of
// scikit_tt/tensor_train.py:1334-1430, here without ever storing Z).
#pragma once
#include "ttals_dev.hpp"

namespace ttals {

struct ResidDesc {
  int d;
  int pad;
  const int* n;              // [d] mode sizes
  const int* ra;             // [d+1] operator ranks
  const int* rx;             // [d+1] x ranks
  const int* rb;             // [d+1] rhs ranks
  const double* const* ac;   // operator cores (P, n, n, Q)
  const double* const* xc;   // x cores (r, n, 1, r')
  const double* const* bc;   // rhs cores (p, n, 1, q)
  double* W;                 // panel: max_i n_i k_{i+1} rho_i doubles
  double* U;                 // max_i r_i n_i R_{i+1} k_{i+1} doubles
  double* Rf;                // rho_max^2 doubles (R factor carried to the left)
  double* tau;               // rho_max doubles
  double* out;               // [2]: ||A x - b||, ||b||
};

// One right-to-left sweep.  with_ax: the residual train; else b alone.
__device__ double resid_sweep(const ResidDesc& D, bool with_ax, double* lds, double* red) {
  const int d = D.d;
  int k = 1;         // columns of the current right factor (k_{i+1})
  int ldr = 1;       // its leading dimension (rho_{i+1})
  if (threadIdx.x == 0) D.Rf[0] = 1.0;
  __syncthreads();
  for (int i = d - 1; i >= 0; --i) {
    const int n = D.n[i];
    const int P = with_ax ? D.ra[i] : 0, Q = D.ra[i + 1];
    const int a = D.rx[i], b = D.rx[i + 1];
    const int p = D.rb[i], q = D.rb[i + 1];
    // row block sizes of Z_i: (A x) block P a rows (one row on core 0), then the b rows
    const int top = (i == 0) ? (with_ax ? 1 : 0) : P * a;
    const int rho = (i == 0) ? 1 : top + p;
    const int boff_row = (i == 0) ? 0 : top;          // core 0: both blocks share row 0
    // column offset of the b block in Z_i (the last core stacks rows only)
    const int boff_col = (i == d - 1 || !with_ax) ? 0 : Q * b;
    const int mk = n * k;
    // W (mk x rho) = C_i^T, C_i[row, m, c] = sum_col Z_i[row, m, col] Rf[c, col]
    if (with_ax) {
      // U[a, n, Q, c] = sum_b x[a, n, b] Rf[c, Q b + b]
      wg_gemm<double, double>(a * n, Q * k, b, crm<double>(D.xc[i], b),
                              cview<double>(D.Rf, plain(1), split(k, b, ldr)), mrm<double>(D.U, Q * k), lds);
      // C_top[(P, a), m, c] = sum_{n', Q} A[P, m, n', Q] U[a, n', Q, c]; rows (P, m), cols (a, c)
      wg_gemm<double, double>(P * n, a * k, n * Q, crm<double>(D.ac[i], n * Q),
                              cview<double>(D.U, plain(k), split(k, n * Q * k, 1)),
                              mview<double>(D.W, split(n, a, k * rho), split(k, 1, rho)), lds);
    }
    // C_bot[j, m, c] = sum_j' b[j, m, j'] Rf[c, boff_col + j']; on core 0 it is subtracted
    // from the (A x) row, so it goes to U first
    double* dst = (i == 0 && with_ax) ? D.U : D.W;
    const int drho = (i == 0 && with_ax) ? 1 : rho;
    wg_gemm<double, double>(p * n, k, q, crm<double>(D.bc[i], q), cview<double>(D.Rf + boff_col, plain(1), plain(ldr)),
                            mview<double>(dst + (dst == D.W ? boff_row : 0), split(n, 1, k * drho), plain(drho)),
                            lds);
    if (i == 0) {
      // the last panel is one column: its norm is the residual norm
      double* v = D.W;
      if (with_ax)
        for (int e = threadIdx.x; e < mk; e += NT) v[e] -= D.U[e];
      __syncthreads();
      return wg_nrm2(v, mk, red);
    }
    wg_householder(D.W, mk, rho, D.tau, red);
    const int kk = mk < rho ? mk : rho;
    for (int e = threadIdx.x; e < kk * rho; e += NT) {
      const int r = e / rho, c = e - r * rho;
      D.Rf[e] = (c >= r) ? D.W[(long)r * rho + c] : 0.0;
    }
    __syncthreads();
    k = kk;
    ldr = rho;
  }
  return 0.0;
}

__global__ __launch_bounds__(NT) void k_residual_norm(const ResidDesc* descs) {
  __shared__ double lds[GEMM_LDS_DOUBLES];
  __shared__ double red[4 * NWAVE];
  const ResidDesc& D = descs[blockIdx.x];
  const double rn = resid_sweep(D, true, lds, red);
  const double bn = resid_sweep(D, false, lds, red);
  if (threadIdx.x == 0) {
    D.out[0] = rn;
    D.out[1] = bn;
  }
}

}  // namespace ttals
