// Adjoint kernels of the spectral-basic increment (one thread per cell):
// the local reverse step of the converged material-point increment
// (dmk_adjoint.h), the transposed tangent product of the adjoint equilibrium
// solve, and deterministic per-phase reductions of the parameter cotangents.
#include <float.h>

#include "dmk_adjoint.h"
#include "dmk_adjoint_k.h"
#include "dmk_resid.h"
#include "dmk_vec.h"
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

// scratch regions, each entry-major over the cells ([entry][N]): per-system
// temporaries (AV_N mmax entries), the matrix block (nmax^2 + nmax: G_y row
// major, then the pivots) and the 9 tangent right-hand sides (9 nmax)
struct AdjScr {
  double *av, *g, *x;
  int nmax, blk;
};
__device__ __forceinline__ AdjScr adj_scr(const AdjKArgs& a) {
  AdjScr r;
  r.nmax = 9 + a.mmax;
  r.blk = r.nmax * r.nmax + r.nmax;
  r.av = a.scr;
  r.g = a.scr + (size_t)AV_N * a.mmax * a.N;
  r.x = r.g + (size_t)r.blk * a.N;
  return r;
}

__device__ __forceinline__ void load_point(const AdjKArgs& a, int e, AdjPoint& pt) {
  load_rm(a.F, a.N, e, pt.F);
  load_rm(a.Fp0, a.N, e, pt.Fp0);
  load_rm(a.Lp, a.N, e, pt.Lp);
  pt.dt = a.dt;
}

// K1 (per cell): forward intermediates, per-system temporaries, consistency
// residual, Jacobian G_y and the 9 tangent right-hand sides
__global__ void __launch_bounds__(ADJ_WG) adj_assemble_kernel(AdjKArgs a, const DmkPhaseDev* __restrict__ php) {
  const int e = blockIdx.x * ADJ_WG + threadIdx.x;
  if (e >= a.N || a.phase[e] != a.phase_id) return;
  const DmkPhaseDev& ph = *php;
  const AdjScr sc = adj_scr(a);
  AdjPoint pt;
  load_point(a, e, pt);
  AdjAcc acc{sc.av, (size_t)a.N, (size_t)e, sc.g, sc.nmax, a.mmax};
  AdjVec xi{(double*)a.xi, (size_t)a.N, (size_t)e}, xi0{(double*)a.xi0, (size_t)a.N, (size_t)e};
  AdjFwd f;
  if (adj_forward(ph, pt, xi, f, acc)) {
    a.res[e] = HUGE_VAL;
    return;
  }
  a.res[e] = adj_residual(ph, pt, xi, xi0, acc, a.eps_rel_Lp, a.eps_abs_Lp, a.eps_rel_state);
  adj_assemble(ph, pt, xi, acc);
  double* X = sc.x + e;      // entry r at X[r * N]
  for (int q = 0; q < 9; q++) {
    double xl[9], xs[DMK_NSL_MAX], w[DMK_NSL_MAX];
    adj_tangent_rhs(ph, pt, f, acc, q, xl, xs, w);
    for (int i = 0; i < 9; i++) X[(size_t)(q * sc.nmax + i) * a.N] = xl[i];
    for (int i = 0; i < ph.nsl; i++) X[(size_t)(q * sc.nmax + 9 + i) * a.N] = xs[i];
  }
}

// K2 (16 lanes per cell, 4 cells per wave): LU factorisation of G_y with
// partial pivoting in LDS, the 9 tangent right-hand sides carried along as
// extra columns, then back substitution.  Pivot choice, operation order and
// rounding per entry are those of adj_lu / adj_solve (first largest pivot).
// LDS per cell: n rows of n + 9 columns (+1 pad), sized at launch.
#define COOP_CELLS 4
#define COOP_LANES 16
__host__ __device__ constexpr int coop_ls(int n) { return n + 10; }
__global__ void __launch_bounds__(COOP_CELLS * COOP_LANES) adj_lu_kernel(AdjKArgs a) {
  extern __shared__ double lds[];
  const int g = threadIdx.x / COOP_LANES, l = threadIdx.x % COOP_LANES;
  const int e = blockIdx.x * COOP_CELLS + g;
  const bool act = e < a.N && a.phase[e] == a.phase_id;
  const int n = 9 + a.nsl, nc = n + 9, LS = coop_ls(n);
  double* G = lds + (size_t)g * n * LS;
  int* pv = (int*)(lds + (size_t)COOP_CELLS * n * LS) + g * n;
  const AdjScr sc = adj_scr(a);
  // entry r of this cell at Gg[r * N], Xg[r * N]
  double* Gg = sc.g + (act ? e : 0);
  double* Xg = sc.x + (act ? e : 0);
  const size_t N = a.N;
  if (act) {
    for (int i = 0; i < n; i++) {
      for (int j = l; j < n; j += COOP_LANES) G[i * LS + j] = Gg[(size_t)(i * sc.nmax + j) * N];
      if (l < 9) G[i * LS + n + l] = Xg[(size_t)(l * sc.nmax + i) * N];
    }
  }
  bool sing = false;
  __syncthreads();
  for (int j = 0; j < n; j++) {
    double best = -1.0;
    int bi = j;
    if (act)
      for (int i = j + l; i < n; i += COOP_LANES) {
        double v = fabs(G[i * LS + j]);
        if (v > best) { best = v; bi = i; }
      }
    for (int off = COOP_LANES / 2; off > 0; off >>= 1) {
      double ob = __shfl_xor(best, off, COOP_LANES);
      int oi = __shfl_xor(bi, off, COOP_LANES);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    const int p = bi;
    if (act && best == 0.0) sing = true;
    if (l == 0) pv[j] = p;
    if (act && !sing && p != j)
      for (int k = l; k < nc; k += COOP_LANES) {
        double t = G[j * LS + k];
        G[j * LS + k] = G[p * LS + k];
        G[p * LS + k] = t;
      }
    __syncthreads();
    if (act && !sing) {
      double r = 1.0 / G[j * LS + j];
      for (int i = j + 1 + l; i < n; i += COOP_LANES) G[i * LS + j] *= r;
    }
    __syncthreads();
    // trailing update; each lane owns columns k (matrix and right-hand sides),
    // rows four at a time so that their loads are in flight together
    if (act && !sing)
      for (int k = j + 1 + l; k < nc; k += COOP_LANES) {
        const double gjk = G[j * LS + k];
        int i = j + 1;
        for (; i + 4 <= n; i += 4) {
          double l0 = G[i * LS + j], l1 = G[(i + 1) * LS + j], l2 = G[(i + 2) * LS + j], l3 = G[(i + 3) * LS + j];
          double c0 = G[i * LS + k], c1 = G[(i + 1) * LS + k], c2 = G[(i + 2) * LS + k], c3 = G[(i + 3) * LS + k];
          G[i * LS + k] = fma(-l0, gjk, c0);
          G[(i + 1) * LS + k] = fma(-l1, gjk, c1);
          G[(i + 2) * LS + k] = fma(-l2, gjk, c2);
          G[(i + 3) * LS + k] = fma(-l3, gjk, c3);
        }
        for (; i < n; i++) G[i * LS + k] = fma(-G[i * LS + j], gjk, G[i * LS + k]);
      }
    __syncthreads();
  }
  // back substitution of the 9 right-hand sides (column n + d), each lane
  // owning right-hand side d for rows i = l / 9 + ..., four rows at a time
  for (int j = n - 1; j >= 0; j--) {
    if (act && !sing && l < 9) G[j * LS + n + l] /= G[j * LS + j];
    __syncthreads();
    if (act && !sing && l < 9) {
      const double xj = G[j * LS + n + l];
      int i = 0;
      for (; i + 4 <= j; i += 4) {
        double u0 = G[i * LS + j], u1 = G[(i + 1) * LS + j], u2 = G[(i + 2) * LS + j], u3 = G[(i + 3) * LS + j];
        double c0 = G[i * LS + n + l], c1 = G[(i + 1) * LS + n + l], c2 = G[(i + 2) * LS + n + l],
               c3 = G[(i + 3) * LS + n + l];
        G[i * LS + n + l] = fma(-u0, xj, c0);
        G[(i + 1) * LS + n + l] = fma(-u1, xj, c1);
        G[(i + 2) * LS + n + l] = fma(-u2, xj, c2);
        G[(i + 3) * LS + n + l] = fma(-u3, xj, c3);
      }
      for (; i < j; i++) G[i * LS + n + l] = fma(-G[i * LS + j], xj, G[i * LS + n + l]);
    }
    __syncthreads();
  }
  if (!act) return;
  if (sing) {
    if (l == 0) a.res[e] = HUGE_VAL;
    return;
  }
  for (int i = 0; i < n; i++) {
    for (int j = l; j < n; j += COOP_LANES) Gg[(size_t)(i * sc.nmax + j) * N] = G[i * LS + j];
    if (l < 9) Xg[(size_t)(l * sc.nmax + i) * N] = G[i * LS + n + l];
  }
  for (int j = l; j < n; j += COOP_LANES) Gg[(size_t)(sc.nmax * sc.nmax + j) * N] = (double)pv[j];
}

// K3a (per cell, with K3 mode 0): the tangent A from the 9 solutions,
// column by column
__global__ void __launch_bounds__(ADJ_WG) adj_tangent_kernel(AdjKArgs a, const DmkPhaseDev* __restrict__ php) {
  const int e = blockIdx.x * ADJ_WG + threadIdx.x;
  if (e >= a.N || a.phase[e] != a.phase_id) return;
  if (!(a.res[e] < HUGE_VAL)) return;
  const DmkPhaseDev& ph = *php;
  const long N = a.N;
  const AdjScr sc = adj_scr(a);
  AdjPoint pt;
  load_point(a, e, pt);
  AdjFwd f;
  adj_fwd33(ph, pt, f);
  const double* X = sc.x + e;      // entry r at X[r * N]
#pragma unroll 1
  for (int q = 0; q < 9; q++) {
    double x[9], c[9];
    for (int k = 0; k < 9; k++) x[k] = X[(size_t)(q * sc.nmax + k) * N];
    adj_tangent_dP(ph, pt, f, q, x, c);
    for (int k = 0; k < 9; k++) a.A[(size_t)(k + 9 * q) * N + e] = c[k];
  }
}

// K3 (per cell).  MODE 0: F cotangent g for the output cotangent (Pu,
// Fpb_in, xib_in).
// MODE 1: cotangents of F_p0, xi_0 and the parameters for (Pu - mu, Fpb_in,
// xib_in).  Both reuse the LU and the per-system temporaries of K1/K2 of the
// same increment; a kernel per mode keeps each one's registers to its own.
template <int MODE>
__global__ void __launch_bounds__(ADJ_WG) adj_vjp_kernel(AdjKArgs a, const DmkPhaseDev* __restrict__ php) {
  const int e = blockIdx.x * ADJ_WG + threadIdx.x;
  if (e >= a.N || a.phase[e] != a.phase_id) return;
  if (!(a.res[e] < HUGE_VAL)) return;
  const DmkPhaseDev& ph = *php;
  const long N = a.N;
  const AdjScr sc = adj_scr(a);
  AdjPoint pt;
  load_point(a, e, pt);
  AdjAcc acc{sc.av, (size_t)N, (size_t)e, sc.g, sc.nmax, a.mmax};
  AdjVec xi{(double*)a.xi, (size_t)N, (size_t)e};
  AdjVec xib{(double*)a.xib_in, (size_t)N, (size_t)e};
  AdjFwd f;
  adj_fwd33(ph, pt, f);
  double Pb[9], Pr[9], Fpb[9], Fb[9], Fp0ib[9], lam[9 + DMK_NSL_MAX];
  for (int k = 0; k < 9; k++) Pb[k] = a.Pu[k];
  if (MODE == 1 && a.mu)
    for (int k = 0; k < 9; k++) Pb[k] -= a.mu[(size_t)k * N + e];
  fld_to_rm(Pb, Pr);
  load_rm(a.Fpb_in, N, e, Fpb);
  adj_vjp_rhs(ph, pt, f, Pr, Fpb, xib, lam, Fb, Fp0ib);
  adj_solve_t(acc, 9 + ph.nsl, lam);
  if (MODE == 0) {
    AdjVec none{nullptr, 0, 0};
    adj_vjp_finish(ph, pt, f, xi, acc, lam, Fb, Fp0ib, false, nullptr, none, none, 0);
    store_rm(a.g, N, e, Fb);
  } else {
    double Fp0b[9];
    AdjVec xi0b{a.xib_out, (size_t)N, (size_t)e}, thb{a.thb, (size_t)N, (size_t)e};
    adj_vjp_finish(ph, pt, f, xi, acc, lam, Fb, Fp0ib, true, Fp0b, xi0b, thb, a.mmax);
    store_rm(a.Fpb_out, N, e, Fp0b);
  }
}

// y_q = sum_p A[p + 9 q] mu_p per cell (the transposed tangent)
__device__ __forceinline__ void adj_atmul_body(long N, const double* __restrict__ A, const double* __restrict__ mu,
                                               double* __restrict__ y, long e) {
  double m[9];
  for (int p = 0; p < 9; p++) m[p] = mu[p * N + e];
  for (int q = 0; q < 9; q++) {
    double s = 0.0;
    for (int p = 0; p < 9; p++) s = fma(A[(p + 9 * q) * N + e], m[p], s);
    y[q * N + e] = s;
  }
}

__global__ void adj_atmul_kernel(long N, const double* __restrict__ A, const double* __restrict__ mu,
                                 double* __restrict__ y) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N) return;
  adj_atmul_body(N, A, mu, y, e);
}

// the adjoint-operator requests of a lockstep round (blockIdx.y = request)
__global__ void adj_atmul_batch_kernel(const DmkResidDev* __restrict__ R) {
  const DmkResidDev& q = R[blockIdx.y];
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q.kind != DMK_RQ_ADJ_OP || !q.A || e >= q.N) return;   // no A: the operator's argument is Pnew itself
  adj_atmul_body(q.N, q.A, q.v, q.Pnew, e);
}

// out[p * nslot + s] = sum over the cells of phase p of v[s * N + e]
#define ADJ_RB 256
__global__ void __launch_bounds__(ADJ_RB) adj_phase_sum_kernel(const double* __restrict__ v, long N,
                                                               const int32_t* __restrict__ phase, int n_phases,
                                                               int nslot, int accumulate, double* __restrict__ out) {
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
    for (int p = 0; p < n_phases; p++) out[(size_t)p * nslot + s] = sh[p][0] + (accumulate ? out[(size_t)p * nslot + s] : 0.0);
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
  w[i] = dmk_gs_elem(w[i], V, ldV, i, c.l, c.c);
}

int dmk_adj_factor(const AdjKArgs* a, const DmkPhaseDev* ph, hipStream_t s) {
  if (a->nsl < 1 || a->nsl > a->mmax || a->mmax > DMK_NSL_MAX) return 1;
  unsigned nb = (unsigned)((a->N + ADJ_WG - 1) / ADJ_WG);
  hipLaunchKernelGGL(adj_assemble_kernel, dim3(nb), dim3(ADJ_WG), 0, s, *a, ph);
  unsigned nc = (unsigned)((a->N + COOP_CELLS - 1) / COOP_CELLS);
  const int n = 9 + a->nsl;
  size_t lds = sizeof(double) * COOP_CELLS * n * coop_ls(n) + sizeof(int) * COOP_CELLS * n;
  hipLaunchKernelGGL(adj_lu_kernel, dim3(nc), dim3(COOP_CELLS * COOP_LANES), lds, s, *a);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_adj_vjp(const AdjKArgs* a, const DmkPhaseDev* ph, hipStream_t s) {
  if (a->nsl < 1 || a->nsl > a->mmax || a->mmax > DMK_NSL_MAX) return 1;
  unsigned nb = (unsigned)((a->N + ADJ_WG - 1) / ADJ_WG);
  if (a->mode == 0) {
    hipLaunchKernelGGL(adj_tangent_kernel, dim3(nb), dim3(ADJ_WG), 0, s, *a, ph);
    hipLaunchKernelGGL(adj_vjp_kernel<0>, dim3(nb), dim3(ADJ_WG), 0, s, *a, ph);
  } else {
    hipLaunchKernelGGL(adj_vjp_kernel<1>, dim3(nb), dim3(ADJ_WG), 0, s, *a, ph);
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_adj_local(const AdjKArgs* a, const DmkPhaseDev* ph, hipStream_t s) {
  if (a->mode == 0) {
    int rc = dmk_adj_factor(a, ph, s);
    if (rc) return rc;
  }
  return dmk_adj_vjp(a, ph, s);
}

int dmk_adj_atmul(long N, const double* A, const double* mu, double* y, hipStream_t s) {
  hipLaunchKernelGGL(adj_atmul_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, A, mu, y);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_adj_atmul_batch(const DmkResidDev* R, int nr, long max_N, hipStream_t s) {
  if (nr < 1) return 0;
  hipLaunchKernelGGL(adj_atmul_batch_kernel, dim3((unsigned)((max_N + 255) / 256), nr), dim3(256), 0, s, R);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_adj_phase_sum(const double* v, long N, const int32_t* phase, int n_phases, int nslot, int accumulate,
                      double* out, hipStream_t s) {
  if (n_phases > DMK_MAX_PHASES) return 1;
  hipLaunchKernelGGL(adj_phase_sum_kernel, dim3(nslot), dim3(ADJ_RB), 0, s, v, N, phase, n_phases, nslot, accumulate,
                     out);
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
// derivatives; no GPU involved).  The consistency residual uses the default
// tolerances (eps_rel_Lp 1e-6, eps_abs_Lp 1e-8, eps_rel_state 1e-6).  3x3 arguments in field order (i + 3 j),
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
  double gblk[(9 + DMK_NSL_MAX) * (9 + DMK_NSL_MAX) + 9 + DMK_NSL_MAX], av[AV_N * DMK_NSL_MAX];
  AdjAcc acc{av, 1, 0, gblk, 9 + mmax, mmax};
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
  if (residual) *residual = adj_residual(ph, pt, X, X0, acc, 1.0e-6, 1.0e-8, 1.0e-6);
  if (adj_factor(ph, pt, X, acc)) return DMK_ERR_ARG;
  if (A) adj_tangent(ph, pt, f, acc, A);
  double Pr[9], Fpr[9], Fb[9], Fp0b[9], x0b[DMK_NSL_MAX];
  fld_to_rm(Pbar, Pr);
  if (Fpbar) fld_to_rm(Fpbar, Fpr);
  else for (int k = 0; k < 9; k++) Fpr[k] = 0.0;
  double th[ADJ_TH_SYS * DMK_NSL_MAX + DMK_NH_MAX] = {0};
  AdjVec X0B{x0b, 1, 0}, TH{th, 1, 0};
  adj_vjp(ph, pt, f, X, acc, Pr, Fpr, XB, Fb, true, Fp0b, X0B, TH, DMK_NSL_MAX);
  for (int k = 0; k < ADJ_TH_SYS * DMK_NSL_MAX + DMK_NH_MAX; k++) thbar[k] += th[k];
  rm_to_fld(Fb, Fbar);
  rm_to_fld(Fp0b, Fp0bar);
  for (int a = 0; a < m; a++) xi0bar[a] = x0b[a];
  return 0;
}
