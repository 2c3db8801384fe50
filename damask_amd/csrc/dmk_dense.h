// Register-resident dense solves for the per-cell Newton systems (row-major
// n x n, all indices compile-time so the matrix stays in VGPRs).
//
// The reference solves its 9x9 systems with LAPACK dgesv (LU with partial
// pivoting, src/math.f90:538-555 via src/phase_mechanical.f90:468-476,
// 1199-1208).  Row interchanges with a per-lane pivot index cannot address
// registers directly: every swap becomes a chain of selects, which costs
// several times the elimination itself on a 64-wide wavefront.  Householder
// QR needs no pivoting, is backward stable for any non-singular matrix and
// gives the same solution to rounding; that is what is used here.
//
// Singularity: a column whose remaining part is exactly zero makes R
// singular; the functions then report failure, as dgesv reports info > 0
// for an exactly zero pivot.
#pragma once

#include <hip/hip_runtime.h>

#define DMK_HD __host__ __device__ __forceinline__

// Reflector for column k: H = I - tau v v^T with v = (v0, x_{k+1..n-1}),
// H x = beta e_k.  Returns (beta, v0, tau); nrm == 0 flags a zero column.
struct DmkHouse {
  double beta, v0, tau;
  bool zero;
};

template <int n>
DMK_HD DmkHouse dmk_house(const double* J, int k) {
  double alpha = J[n * k + k];
  double sigma = 0.0;
#pragma unroll
  for (int i = k + 1; i < n; i++) sigma = fma(J[n * i + k], J[n * i + k], sigma);
  double nrm = sqrt(fma(alpha, alpha, sigma));
  DmkHouse h;
  h.beta = alpha >= 0.0 ? -nrm : nrm;
  h.v0 = alpha - h.beta;                 // |v0| = |alpha| + nrm: no cancellation
  h.tau = -1.0 / (h.beta * h.v0);        // 2 / (v^T v)
  h.zero = !(nrm > 0.0);
  return h;
}

// Solve J x = b in place (J destroyed, x returned in b).  Returns true when
// J is singular (the contents of b are then undefined).
template <int n>
DMK_HD bool dmk_qr_solve(double* J, double* b) {
  bool sing = false;
#pragma unroll
  for (int k = 0; k < n; k++) {
    DmkHouse h = dmk_house<n>(J, k);
    sing |= h.zero;
#pragma unroll
    for (int j = k + 1; j < n; j++) {
      double w = h.v0 * J[n * k + j];
#pragma unroll
      for (int i = k + 1; i < n; i++) w = fma(J[n * i + k], J[n * i + j], w);
      double s = h.tau * w;
      J[n * k + j] = fma(-s, h.v0, J[n * k + j]);
#pragma unroll
      for (int i = k + 1; i < n; i++) J[n * i + j] = fma(-s, J[n * i + k], J[n * i + j]);
    }
    {
      double w = h.v0 * b[k];
#pragma unroll
      for (int i = k + 1; i < n; i++) w = fma(J[n * i + k], b[i], w);
      double s = h.tau * w;
      b[k] = fma(-s, h.v0, b[k]);
#pragma unroll
      for (int i = k + 1; i < n; i++) b[i] = fma(-s, J[n * i + k], b[i]);
    }
    J[n * k + k] = h.beta;
  }
#pragma unroll
  for (int k = n - 1; k >= 0; k--) {
    double s = b[k];
#pragma unroll
    for (int j = k + 1; j < n; j++) s = fma(-J[n * k + j], b[j], s);
    b[k] = s / J[n * k + k];
  }
  return sing;
}

// Factorisation kept for several right-hand sides: R in the upper triangle
// (diagonal = beta), reflector tails below it, v0 and tau per column.
template <int n>
DMK_HD bool dmk_qr_factor(double* J, double* v0, double* tau) {
  bool sing = false;
#pragma unroll
  for (int k = 0; k < n; k++) {
    DmkHouse h = dmk_house<n>(J, k);
    sing |= h.zero;
#pragma unroll
    for (int j = k + 1; j < n; j++) {
      double w = h.v0 * J[n * k + j];
#pragma unroll
      for (int i = k + 1; i < n; i++) w = fma(J[n * i + k], J[n * i + j], w);
      double s = h.tau * w;
      J[n * k + j] = fma(-s, h.v0, J[n * k + j]);
#pragma unroll
      for (int i = k + 1; i < n; i++) J[n * i + j] = fma(-s, J[n * i + k], J[n * i + j]);
    }
    J[n * k + k] = h.beta;
    v0[k] = h.v0;
    tau[k] = h.tau;
  }
  return sing;
}

template <int n>
DMK_HD void dmk_qr_apply_solve(const double* J, const double* v0, const double* tau, double* b) {
#pragma unroll
  for (int k = 0; k < n; k++) {
    double w = v0[k] * b[k];
#pragma unroll
    for (int i = k + 1; i < n; i++) w = fma(J[n * i + k], b[i], w);
    double s = tau[k] * w;
    b[k] = fma(-s, v0[k], b[k]);
#pragma unroll
    for (int i = k + 1; i < n; i++) b[i] = fma(-s, J[n * i + k], b[i]);
  }
#pragma unroll
  for (int k = n - 1; k >= 0; k--) {
    double s = b[k];
#pragma unroll
    for (int j = k + 1; j < n; j++) s = fma(-J[n * k + j], b[j], s);
    b[k] = s / J[n * k + k];
  }
}
