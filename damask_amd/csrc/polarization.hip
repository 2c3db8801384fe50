// Per-cell kernels of the spectral-polarization scheme
// (src/grid/grid_mech_spectral_polarization.f90: forward :392-398,
// form_residual :582-634).  One thread per cell over component-major fields;
// every kernel is a single HBM pass (read the operands once, write once).
//
// invSym3333 (src/math.f90:511-532) goes through the Mandel 6x6 form
// (:826-904): rows/columns ordered 11, 22, 33, 23, 13, 12, shear rows and
// columns weighted sqrt(2), the result expanded with the inverse weights.
// Only the listed (i,j) of each pair are read, so for a tensor without minor
// symmetry (dPdF + C_scale) the result equals the reference's, which reads the
// same entries.  The residual needs the inverse applied to one tensor only, so
// the device solves that 6x6 system by Gaussian elimination with partial
// pivoting (first largest pivot, like dgetrf); a zero pivot, where the
// reference stops with 'matrix inversion error', makes the cell's residual NaN,
// which ends the nonlinear solve as diverged.
#include "dmk_common.h"
#include "dmk_linalg.h"
#include "dmk_polar.h"
#include "dmk_rot.h"

#include <cmath>
#include <utility>

namespace {

#define PB 256

// Mandel pairs and weights
__host__ __device__ constexpr int mp_i(int p) { return p < 3 ? p : (p == 3 ? 1 : 0); }
__host__ __device__ constexpr int mp_j(int p) { return p < 3 ? p : (p == 5 ? 1 : 2); }
__host__ __device__ constexpr double mp_w(int p) { return p < 3 ? 1.0 : 1.4142135623730951; }

__host__ __device__ inline int c4(int i, int j, int k, int l) { return i + 3 * j + 9 * k + 27 * l; }

// inverse of a 6x6 (row-major M[6 p + q]) by Gauss-Jordan elimination with
// partial pivoting, host side; false on a zero pivot
inline bool inv66(double* M) {
  double A[72];
  for (int p = 0; p < 6; p++)
    for (int q = 0; q < 6; q++) {
      A[12 * p + q] = M[6 * p + q];
      A[12 * p + 6 + q] = p == q ? 1.0 : 0.0;
    }
  for (int c = 0; c < 6; c++) {
    int piv = c;
    double best = std::fabs(A[12 * c + c]);
    for (int r = c + 1; r < 6; r++) {
      double v = std::fabs(A[12 * r + c]);
      if (v > best) { best = v; piv = r; }
    }
    if (best == 0.0) return false;
    if (piv != c)
      for (int q = 0; q < 12; q++) std::swap(A[12 * c + q], A[12 * piv + q]);
    double d = 1.0 / A[12 * c + c];
    for (int q = 0; q < 12; q++) A[12 * c + q] *= d;
    for (int r = 0; r < 6; r++) {
      if (r == c) continue;
      double f = A[12 * r + c];
      for (int q = 0; q < 12; q++) A[12 * r + q] = std::fma(-f, A[12 * c + q], A[12 * r + q]);
    }
  }
  for (int p = 0; p < 6; p++)
    for (int q = 0; q < 6; q++) M[6 * p + q] = A[12 * p + 6 + q];
  return true;
}

// M y = x in place (x <- y) by Gaussian elimination with partial pivoting
// (first largest pivot); every index is a compile-time constant after
// unrolling, the row exchange is a select, so M and x stay in registers
__device__ __forceinline__ bool solve66(double* M, double* x) {
  bool ok = true;
#pragma unroll
  for (int c = 0; c < 6; c++) {
    int piv = c;
    double best = fabs(M[6 * c + c]);
#pragma unroll
    for (int r = c + 1; r < 6; r++) {
      double v = fabs(M[6 * r + c]);
      if (v > best) { best = v; piv = r; }
    }
    ok = ok && best != 0.0;
#pragma unroll
    for (int r = c + 1; r < 6; r++) {
      const bool sw = r == piv;
#pragma unroll
      for (int q = c; q < 6; q++) {
        double a = M[6 * c + q], b = M[6 * r + q];
        M[6 * c + q] = sw ? b : a;
        M[6 * r + q] = sw ? a : b;
      }
      double a = x[c], b = x[r];
      x[c] = sw ? b : a;
      x[r] = sw ? a : b;
    }
    const double d = 1.0 / M[6 * c + c];
#pragma unroll
    for (int r = c + 1; r < 6; r++) {
      double f = M[6 * r + c] * d;
#pragma unroll
      for (int q = c + 1; q < 6; q++) M[6 * r + q] = fma(-f, M[6 * c + q], M[6 * r + q]);
      x[r] = fma(-f, x[c], x[r]);
    }
  }
#pragma unroll
  for (int c = 5; c >= 0; c--) {
    double s = x[c];
#pragma unroll
    for (int q = c + 1; q < 6; q++) s = fma(-M[6 * c + q], x[q], s);
    x[c] = s / M[6 * c + c];
  }
  return ok;
}

// (C:A)_ij = sum_kl C_ijkl A_kl, column-major 3x3
__device__ __forceinline__ void ddot3333(const double* __restrict__ C, const double* A, double* out) {
#pragma unroll
  for (int ij = 0; ij < 9; ij++) {
    double s = 0.0;
#pragma unroll
    for (int kl = 0; kl < 9; kl++) s = fma(C[ij + 9 * kl], A[kl], s);
    out[ij] = s;
  }
}

// column-major 3x3 product
__device__ __forceinline__ void mul33(const double* A, const double* B, double* out) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 3; k++) s = fma(A[i + 3 * k], B[k + 3 * j], s);
      out[i + 3 * j] = s;
    }
}

__device__ __forceinline__ void load9(const double* __restrict__ f, long N, long e, double* v) {
#pragma unroll
  for (int c = 0; c < 9; c++) v[c] = f[c * N + e];
}
__device__ __forceinline__ void store9(double* __restrict__ f, long N, long e, const double* v) {
#pragma unroll
  for (int c = 0; c < 9; c++) f[c * N + e] = v[c];
}

// W = F (C:(F_tau - F - I))
__device__ __forceinline__ void pol_W(const double* __restrict__ C, const double* F, const double* Ft, double* W) {
  double D[9], CD[9];
#pragma unroll
  for (int c = 0; c < 9; c++) D[c] = Ft[c] - F[c] - (c % 4 == 0 ? 1.0 : 0.0);
  ddot3333(C, D, CD);
  mul33(F, CD, W);
}

__global__ void __launch_bounds__(PB) pol_tau_kernel(long N, PolArgs a, const double* __restrict__ Fg,
                                                     const double* __restrict__ Ftg, double* __restrict__ T) {
  long e = (long)blockIdx.x * PB + threadIdx.x;
  if (e >= N) return;
  double F[9], Ft[9], W[9], D[9], CD[9], out[9];
  load9(Fg, N, e, F);
  load9(Ftg, N, e, Ft);
  pol_W(a.C, F, Ft, W);
#pragma unroll
  for (int c = 0; c < 9; c++) D[c] = F[c] - (c % 4 == 0 ? 1.0 : 0.0);
  ddot3333(a.C, D, CD);
#pragma unroll
  for (int c = 0; c < 9; c++) out[c] = a.beta * CD[c] - a.alpha * W[c];
  store9(T, N, e, out);
}

__global__ void __launch_bounds__(PB) pol_stage2_kernel(long n9, double beta, const double* __restrict__ F,
                                                        const double* __restrict__ G, double* __restrict__ r_tau,
                                                        double* __restrict__ F_resp) {
  long i = (long)blockIdx.x * PB + threadIdx.x;
  if (i >= n9) return;
  double f = F[i];
  double r = beta * f - G[i];
  r_tau[i] = r;
  F_resp[i] = f - r / beta;
}

__global__ void __launch_bounds__(PB) pol_rF_kernel(long N, PolArgs a, const double* __restrict__ Fg,
                                                    const double* __restrict__ Ftg, const double* __restrict__ Pg,
                                                    const double* __restrict__ dPdF,
                                                    const double* __restrict__ r_tau, double* __restrict__ r_F) {
  long e = (long)blockIdx.x * PB + threadIdx.x;
  if (e >= N) return;
  double x[6];
  {
    double F[9], Ft[9], W[9], P[9];
    load9(Fg, N, e, F);
    load9(Ftg, N, e, Ft);
    pol_W(a.C, F, Ft, W);
    load9(Pg, N, e, P);
#pragma unroll
    for (int c = 0; c < 9; c++) P[c] -= W[c];
    // x_q = X_kk (q < 3), (X_kl + X_lk) / sqrt(2) (q >= 3)
#pragma unroll
    for (int q = 0; q < 6; q++) {
      int k = mp_i(q), l = mp_j(q);
      x[q] = q < 3 ? P[k + 3 * k] : (P[k + 3 * l] + P[l + 3 * k]) / mp_w(q);
    }
  }
  double M[36];
#pragma unroll
  for (int p = 0; p < 6; p++)
#pragma unroll
    for (int q = 0; q < 6; q++) {
      int r = c4(mp_i(p), mp_j(p), mp_i(q), mp_j(q));
      M[6 * p + q] = (dPdF[(long)r * N + e] + a.C[r]) * (mp_w(p) * mp_w(q));
    }
  const bool ok = solve66(M, x);
  double y[6];
#pragma unroll
  for (int p = 0; p < 6; p++) y[p] = ok ? x[p] / mp_w(p) : NAN;
  // r_F_ij = y_p(ij) + r_F_tau_ij (the invSym3333 part is symmetric)
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      int p = i == j ? i : 6 - i - j;     // (1,2)->3, (0,2)->4, (0,1)->5
      long o = (long)(i + 3 * j) * N + e;
      r_F[o] = y[p] + r_tau[o];
    }
}

__global__ void __launch_bounds__(PB) pol_lambda_kernel(long N, PolArgs a, const double* __restrict__ Fg,
                                                        double* __restrict__ Ftg) {
  long e = (long)blockIdx.x * PB + threadIdx.x;
  if (e >= N) return;
  double F[9], L[9], A[9], B[9], H[9], D[9];
  load9(Fg, N, e, F);
  load9(Ftg, N, e, L);
#pragma unroll
  for (int c = 0; c < 9; c++) L[c] -= F[c];
  // A = L^T L - I
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 3; k++) s = fma(L[k + 3 * i], L[k + 3 * j], s);
      A[i + 3 * j] = s - (i == j ? 1.0 : 0.0);
    }
  ddot3333(a.C, A, B);
  mul33(L, B, H);
#pragma unroll
  for (int c = 0; c < 9; c++) H[c] *= 0.5;
  ddot3333(a.S, H, D);
#pragma unroll
  for (int c = 0; c < 9; c++) D[c] += (c % 4 == 0 ? 1.0 : 0.0) + F[c];
  store9(Ftg, N, e, D);
}

__global__ void pol_aim_pre_kernel(double q0, double q1, double q2, double q3, double beta,
                                   const double* __restrict__ faim, double* __restrict__ aim) {
  if (threadIdx.x != 0) return;
  double q[4] = {q0, q1, q2, q3}, A[9], R[9];
  for (int k = 0; k < 9; k++) A[k] = beta * faim[k];
  dmk_rotate33(q, A, true, R);
  for (int k = 0; k < 9; k++) aim[k] = R[k];
}

// end of form_residual (:627-632): P_av in the load frame, F_aim -= S:(P_av - P_aim),
// err_BC = max |C_scale:(F_aim - rotate(F_av))| on the deformation-controlled and
// |P_av - P_aim| on the stress-controlled components; report as the basic scheme's
__global__ void pol_residual_aim_kernel(AimArgs a, PolArgs pa, const double* __restrict__ pav_raw,
                                        const double* __restrict__ fav_raw, double* __restrict__ faim,
                                        const int* __restrict__ flag, double* __restrict__ rep) {
  if (threadIdx.x != 0) return;
  double P[9];
  for (int k = 0; k < 9; k++) P[k] = pav_raw[k];
  dmk_rotate33(a.q_R, P, false, P);
  for (int k = 0; k < 9; k++) rep[REP_PAV + k] = P[k];
  rep[REP_FLAG] = (double)flag[0];
  double dP[9], Fa[9];
  for (int k = 0; k < 9; k++) dP[k] = P[k] - a.P_aim[k];
  for (int ij = 0; ij < 9; ij++) {
    double s = 0.0;
    for (int kl = 0; kl < 9; kl++) s += a.S[ij + 9 * kl] * dP[kl];
    Fa[ij] = faim[ij] - s;
  }
  double Fav[9], D[9];
  for (int k = 0; k < 9; k++) Fav[k] = fav_raw[k];
  dmk_rotate33(a.q_R, Fav, false, Fav);
  for (int k = 0; k < 9; k++) D[k] = Fa[k] - Fav[k];
  double e = 0.0;
  for (int ij = 0; ij < 9; ij++) {
    double s = 0.0;
    for (int kl = 0; kl < 9; kl++) s += pa.C[ij + 9 * kl] * D[kl];
    e = fmax(e, fabs(a.stress_mask[ij] ? s : dP[ij]));
    faim[ij] = Fa[ij];
    rep[REP_FAIM + ij] = Fa[ij];
  }
  rep[REP_ERRBC] = e;
}

inline unsigned nblk(long n) { return (unsigned)((n + PB - 1) / PB); }

}  // namespace

int dmk_pol_tau(long N, const PolArgs* a, const double* F, const double* F_tau, double* T, hipStream_t s) {
  hipLaunchKernelGGL(pol_tau_kernel, dim3(nblk(N)), dim3(PB), 0, s, N, *a, F, F_tau, T);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_pol_stage2(long N, double beta, const double* F, const double* G, double* r_tau, double* F_resp,
                   hipStream_t s) {
  hipLaunchKernelGGL(pol_stage2_kernel, dim3(nblk(9 * N)), dim3(PB), 0, s, 9 * N, beta, F, G, r_tau, F_resp);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_pol_rF(long N, const PolArgs* a, const double* F, const double* F_tau, const double* P, const double* dPdF,
               const double* r_tau, double* r_F, hipStream_t s) {
  hipLaunchKernelGGL(pol_rF_kernel, dim3(nblk(N)), dim3(PB), 0, s, N, *a, F, F_tau, P, dPdF, r_tau, r_F);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_pol_lambda(long N, const PolArgs* a, const double* F, double* F_tau, hipStream_t s) {
  hipLaunchKernelGGL(pol_lambda_kernel, dim3(nblk(N)), dim3(PB), 0, s, N, *a, F, F_tau);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_pol_aim_pre(const double q_R[4], double beta, const double* faim, double* aim, hipStream_t s) {
  hipLaunchKernelGGL(pol_aim_pre_kernel, dim3(1), dim3(64), 0, s, q_R[0], q_R[1], q_R[2], q_R[3], beta, faim, aim);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_pol_residual_aim(const AimArgs* a, const PolArgs* pa, const double* pav_raw, const double* fav_raw,
                         double* faim, const int* flag, double* report, hipStream_t s) {
  hipLaunchKernelGGL(pol_residual_aim_kernel, dim3(1), dim3(64), 0, s, *a, *pa, pav_raw, fav_raw, faim, flag,
                     report);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

extern "C" int dmk_inv_sym3333(const double* A, double* out) {
  double M[36];
  for (int p = 0; p < 6; p++)
    for (int q = 0; q < 6; q++) M[6 * p + q] = A[c4(mp_i(p), mp_j(p), mp_i(q), mp_j(q))] * (mp_w(p) * mp_w(q));
  if (!inv66(M)) return 1;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) {
          int p = i == j ? i : 6 - i - j, q = k == l ? k : 6 - k - l;
          out[c4(i, j, k, l)] = M[6 * p + q] / (mp_w(p) * mp_w(q));
        }
  return 0;
}
