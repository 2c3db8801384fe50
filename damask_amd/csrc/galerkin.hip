// Kernels of the spectral-Galerkin scheme (src/grid/grid_mech_spectral_Galerkin.f90)
// besides the G convolution (spectral.hip): the tangent product K:dF of the
// matrix-free Jacobian GK_op (:607-660), the stress-BC bookkeeping of
// form_residual (:592-596), and the linear combination of the Krylov solver.
// One thread per cell (per element for lincomb), single HBM pass each.
#include "dmk_common.h"
#include "dmk_galerkin.h"
#include "dmk_linalg.h"
#include "dmk_rot.h"

namespace {

#define GB 256

__global__ void __launch_bounds__(GB) gk_apply_kernel(long N, const double* __restrict__ K,
                                                      const double* __restrict__ dF, double* __restrict__ out) {
  long e = (long)blockIdx.x * GB + threadIdx.x;
  if (e >= N) return;
  double d[9];
#pragma unroll
  for (int q = 0; q < 9; q++) d[q] = dF[q * N + e];
#pragma unroll
  for (int p = 0; p < 9; p++) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) s = fma(K[(long)(p + 9 * q) * N + e], d[q], s);
    out[p * N + e] = s;
  }
}

__global__ void __launch_bounds__(GB) lincomb_kernel(long n, double a0, const double* x0, CombineArgs c,
                                                     double* out) {
  long i = (long)blockIdx.x * GB + threadIdx.x;
  if (i >= n) return;
  double v = x0 ? a0 * x0[i] : 0.0;
  for (int k = 0; k < c.l; k++) v = fma(c.alpha[k], c.X[k][i], v);
  out[i] = v;
}

__global__ void gk_bc_kernel(AimArgs a, const double* __restrict__ pav_raw, const int* __restrict__ flag,
                             double* __restrict__ rep, double* __restrict__ dc) {
  if (threadIdx.x != 0) return;
  double P[9];
  for (int k = 0; k < 9; k++) P[k] = pav_raw[k];
  if (a.rotate) dmk_rotate33(a.q_R, P, false, P);
  for (int k = 0; k < 9; k++) rep[REP_PAV + k] = P[k];
  rep[REP_FLAG] = (double)flag[0];
  double e = 0.0;
  for (int k = 0; k < 9; k++) {
    double d = a.stress_mask[k] ? 0.0 : P[k] - a.P_aim[k];
    dc[k] = d;
    e = fmax(e, fabs(d));
  }
  rep[REP_ERRBC] = e;
}

}  // namespace

int dmk_gk_apply(long N, const double* dPdF, const double* dF, double* out, hipStream_t s) {
  hipLaunchKernelGGL(gk_apply_kernel, dim3((unsigned)((N + GB - 1) / GB)), dim3(GB), 0, s, N, dPdF, dF, out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_lincomb(long n, double a0, const double* x0, const CombineArgs* c, double* out, hipStream_t s) {
  hipLaunchKernelGGL(lincomb_kernel, dim3((unsigned)((n + GB - 1) / GB)), dim3(GB), 0, s, n, a0, x0, *c, out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_gk_bc(const AimArgs* a, const double* pav_raw, const int* flag, double* report, double* dc, hipStream_t s) {
  hipLaunchKernelGGL(gk_bc_kernel, dim3(1), dim3(64), 0, s, *a, pav_raw, flag, report, dc);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
