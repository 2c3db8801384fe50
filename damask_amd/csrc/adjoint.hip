// Adjoint kernels of the spectral-basic increment (one thread per cell):
// the local reverse step of the converged material-point increment
// (dmk_adjoint.h), the transposed tangent product of the adjoint equilibrium
// solve, and deterministic per-phase reductions of the parameter cotangents.
#include <float.h>

#include "dmk_adjoint.h"
#include "dmk_adjoint_k.h"
#include "../../include/damask_hip.h"

int dmk_build_phase(const dmk_phase_desc* d, DmkPhaseDev* p);

#define ADJ_WG 64

// field order (i + 3 j) <-> row-major
DHI void fld_to_rm(const double* f, double* T) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[3 * i + j] = f[i + 3 * j];
}
DHI void rm_to_fld(const double* T, double* f) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) f[i + 3 * j] = T[3 * i + j];
}

__device__ __forceinline__ void load_rm(const double* f, long N, int e, double* T) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[3 * i + j] = f[(size_t)(i + 3 * j) * N + e];
}
__device__ __forceinline__ void store_rm(double* f, long N, int e, const double* T) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) f[(size_t)(i + 3 * j) * N + e] = T[3 * i + j];
}

__global__ void __launch_bounds__(ADJ_WG) adj_local_kernel(AdjKArgs a, const DmkPhaseDev* __restrict__ php) {
  const int e = blockIdx.x * ADJ_WG + threadIdx.x;
  if (e >= a.N || a.phase[e] != a.phase_id) return;
  const DmkPhaseDev& ph = *php;
  const long N = a.N;
  AdjPoint pt;
  load_rm(a.F, N, e, pt.F);
  load_rm(a.Fp0, N, e, pt.Fp0);
  load_rm(a.Lp, N, e, pt.Lp);
  pt.dt = a.dt;
  AdjAcc acc{(size_t)N, (size_t)e, 9 + a.mmax, a.mmax, a.scr};
  AdjVec xi{(double*)a.xi, (size_t)N, (size_t)e}, xi0{(double*)a.xi0, (size_t)N, (size_t)e};
  AdjVec xib{(double*)a.xib_in, (size_t)N, (size_t)e};
  AdjFwd f;
  if (adj_forward(ph, pt, xi, f, acc)) {
    if (a.mode == 0) a.res[e] = HUGE_VAL;
    return;
  }
  double Pb[9], Fpb[9], Fb[9];
  for (int k = 0; k < 9; k++) Pb[k] = a.Pu[k];
  if (a.mode == 1 && a.mu)
    for (int k = 0; k < 9; k++) Pb[k] -= a.mu[(size_t)k * N + e];
  double Pr[9];
  fld_to_rm(Pb, Pr);
  load_rm(a.Fpb_in, N, e, Fpb);
  if (a.mode == 0) {
    a.res[e] = adj_residual(ph, pt, xi, xi0, acc);
    if (adj_factor(ph, pt, xi, acc)) {
      a.res[e] = HUGE_VAL;
      return;
    }
    double A[81];
    adj_tangent(ph, pt, f, acc, A);
    for (int k = 0; k < 81; k++) a.A[(size_t)k * N + e] = A[k];
    AdjVec none{nullptr, 0, 0};
    adj_vjp(ph, pt, f, xi, acc, Pr, Fpb, xib, Fb, false, nullptr, none, none, 0);
    store_rm(a.g, N, e, Fb);
  } else {
    double Fp0b[9];
    AdjVec xi0b{a.xib_out, (size_t)N, (size_t)e}, thb{a.thb, (size_t)N, (size_t)e};
    adj_vjp(ph, pt, f, xi, acc, Pr, Fpb, xib, Fb, true, Fp0b, xi0b, thb, a.mmax);
    store_rm(a.Fpb_out, N, e, Fp0b);
  }
}

// y_q = sum_p A[p + 9 q] mu_p per cell (the transposed tangent)
__global__ void adj_atmul_kernel(long N, const double* __restrict__ A, const double* __restrict__ mu,
                                 double* __restrict__ y) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N) return;
  double m[9];
  for (int p = 0; p < 9; p++) m[p] = mu[p * N + e];
  for (int q = 0; q < 9; q++) {
    double s = 0.0;
    for (int p = 0; p < 9; p++) s = fma(A[(p + 9 * q) * N + e], m[p], s);
    y[q * N + e] = s;
  }
}

// out[p * nslot + s] = sum over the cells of phase p of v[s * N + e]
#define ADJ_RB 256
__global__ void __launch_bounds__(ADJ_RB) adj_phase_sum_kernel(const double* __restrict__ v, long N,
                                                               const int32_t* __restrict__ phase, int n_phases,
                                                               int nslot, double* __restrict__ out) {
  __shared__ double sh[DMK_MAX_PHASES][ADJ_RB];
  const int s = blockIdx.x;
  double acc[DMK_MAX_PHASES];
  for (int p = 0; p < DMK_MAX_PHASES; p++) acc[p] = 0.0;
  for (long e = threadIdx.x; e < N; e += ADJ_RB) {
    int p = phase[e];
    double x = v[(size_t)s * N + e];
    for (int q = 0; q < DMK_MAX_PHASES; q++)
      if (q == p) acc[q] += x;
  }
  for (int p = 0; p < DMK_MAX_PHASES; p++) sh[p][threadIdx.x] = acc[p];
  __syncthreads();
  for (int k = ADJ_RB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k)
      for (int p = 0; p < DMK_MAX_PHASES; p++) sh[p][threadIdx.x] += sh[p][threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int p = 0; p < n_phases; p++) out[(size_t)p * nslot + s] = sh[p][0];
}

// out[0] = max over N (NaN if any entry is NaN)
__global__ void __launch_bounds__(ADJ_RB) adj_max_kernel(const double* __restrict__ v, long N,
                                                         double* __restrict__ out) {
  __shared__ double sh[ADJ_RB];
  __shared__ int nan_seen;
  if (threadIdx.x == 0) nan_seen = 0;
  __syncthreads();
  double m = 0.0;
  int bad = 0;
  for (long e = threadIdx.x; e < N; e += ADJ_RB) {
    double x = v[e];
    if (x != x) bad = 1;
    else m = fmax(m, x);
  }
  if (bad) nan_seen = 1;
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int k = ADJ_RB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = nan_seen ? NAN : sh[0];
}

// int32 -> double maximum (sub-step counts)
__global__ void __launch_bounds__(ADJ_RB) adj_imax_kernel(const int32_t* __restrict__ v, long N,
                                                          double* __restrict__ out) {
  __shared__ int sh[ADJ_RB];
  int m = 0;
  for (long e = threadIdx.x; e < N; e += ADJ_RB) m = max(m, v[e]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int k = ADJ_RB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = (double)sh[0];
}

// w -= sum_i c_i V_i (Gram-Schmidt update of the adjoint GMRES)
__global__ void adj_gs_kernel(long n, double* __restrict__ w, const double* __restrict__ V, long ldV, GsArgs c) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = w[i];
  for (int k = 0; k < c.l; k++) v = fma(-c.c[k], V[k * ldV + i], v);
  w[i] = v;
}

int dmk_adj_local(const AdjKArgs* a, const DmkPhaseDev* ph, hipStream_t s) {
  unsigned nb = (unsigned)((a->N + ADJ_WG - 1) / ADJ_WG);
  hipLaunchKernelGGL(adj_local_kernel, dim3(nb), dim3(ADJ_WG), 0, s, *a, ph);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_adj_atmul(long N, const double* A, const double* mu, double* y, hipStream_t s) {
  hipLaunchKernelGGL(adj_atmul_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, A, mu, y);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_adj_phase_sum(const double* v, long N, const int32_t* phase, int n_phases, int nslot, double* out,
                      hipStream_t s) {
  if (n_phases > DMK_MAX_PHASES) return 1;
  hipLaunchKernelGGL(adj_phase_sum_kernel, dim3(nslot), dim3(ADJ_RB), 0, s, v, N, phase, n_phases, nslot, out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_adj_max(const double* v, long N, double* out, hipStream_t s) {
  hipLaunchKernelGGL(adj_max_kernel, dim3(1), dim3(ADJ_RB), 0, s, v, N, out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_adj_imax(const int32_t* v, long N, double* out, hipStream_t s) {
  hipLaunchKernelGGL(adj_imax_kernel, dim3(1), dim3(ADJ_RB), 0, s, v, N, out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_adj_gs(long n, double* w, const double* V, long ldV, const GsArgs* c, hipStream_t s) {
  hipLaunchKernelGGL(adj_gs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, w, V, ldV, *c);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

// The local reverse step of one cell on the host (CPU checks of the
// derivatives; no GPU involved).  3x3 arguments in field order (i + 3 j),
// per-system arrays of length nsl.  thbar: per-system slots
// (slot * DMK_NSL_MAX + system) then interaction types, accumulated.
extern "C" int dmk_adjoint_point_host(const dmk_phase_desc* d, double dt, const double F[9], const double Fp0[9],
                                      const double* xi0, const double Lp[9], const double* xi,
                                      const double Pbar[9], const double Fpbar[9], const double* xibar,
                                      double Fbar[9], double Fp0bar[9], double* xi0bar, double* thbar,
                                      double A[81], double* residual) {
  static DmkPhaseDev ph;
  if (!d || dmk_build_phase(d, &ph)) return DMK_ERR_CONFIG;
  const int m = ph.nsl, mmax = DMK_NSL_MAX;
  double scr[(9 + DMK_NSL_MAX) * (9 + DMK_NSL_MAX) + 9 + DMK_NSL_MAX + AV_N * DMK_NSL_MAX];
  AdjAcc acc{1, 0, 9 + mmax, mmax, scr};
  double xiv[DMK_NSL_MAX], xi0v[DMK_NSL_MAX], xibv[DMK_NSL_MAX];
  for (int a = 0; a < m; a++) { xiv[a] = xi[a]; xi0v[a] = xi0[a]; xibv[a] = xibar ? xibar[a] : 0.0; }
  AdjVec X{xiv, 1, 0}, X0{xi0v, 1, 0}, XB{xibv, 1, 0};
  AdjPoint pt;
  fld_to_rm(F, pt.F);
  fld_to_rm(Fp0, pt.Fp0);
  fld_to_rm(Lp, pt.Lp);
  pt.dt = dt;
  AdjFwd f;
  if (adj_forward(ph, pt, X, f, acc)) return DMK_ERR_ARG;
  if (residual) *residual = adj_residual(ph, pt, X, X0, acc);
  if (adj_factor(ph, pt, X, acc)) return DMK_ERR_ARG;
  if (A) adj_tangent(ph, pt, f, acc, A);
  double Pr[9], Fpr[9], Fb[9], Fp0b[9], x0b[DMK_NSL_MAX];
  fld_to_rm(Pbar, Pr);
  if (Fpbar) fld_to_rm(Fpbar, Fpr);
  else for (int k = 0; k < 9; k++) Fpr[k] = 0.0;
  AdjVec X0B{x0b, 1, 0}, TH{thbar, 1, 0};
  adj_vjp(ph, pt, f, X, acc, Pr, Fpr, XB, Fb, true, Fp0b, X0B, TH, DMK_NSL_MAX);
  rm_to_fld(Fb, Fbar);
  rm_to_fld(Fp0b, Fp0bar);
  for (int a = 0; a < m; a++) xi0bar[a] = x0b[a];
  return 0;
}
