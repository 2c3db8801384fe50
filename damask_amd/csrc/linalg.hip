// Field reductions and vector kernels of the solver (HBM-bound, FP64).
//
// Reductions use a fixed partition (2048 elements per block, tree reduction in
// LDS, then one block per result), so results are bitwise reproducible from
// run to run.
#include "dmk_common.h"
#include "dmk_linalg.h"
#include "dmk_resid.h"
#include "dmk_vec.h"
#include "dmk_rot.h"

#include <float.h>

#include <algorithm>

#define RB 256
#define PER_BLOCK 2048

namespace {

__device__ double block_sum(double v) {
  __shared__ double red[RB];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = RB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

// column sums: blockIdx.y = column (base + col * ld), blockIdx.x = chunk
__device__ __forceinline__ void colsum_partial_body(const double* __restrict__ base, long ld, long n,
                                                    double* __restrict__ partial, int nchunk, int bx, int col) {
  const double* x = base + (long)col * ld;
  long s = (long)bx * PER_BLOCK;
  long e = min(s + PER_BLOCK, n);
  double acc = 0.0;
  for (long i = s + threadIdx.x; i < e; i += RB) acc += x[i];
  double r = block_sum(acc);
  if (threadIdx.x == 0) partial[(long)col * nchunk + bx] = r;
}

__global__ void __launch_bounds__(RB) colsum_partial(const double* __restrict__ base, long ld, long n,
                                                     double* __restrict__ partial, int nchunk) {
  colsum_partial_body(base, ld, n, partial, nchunk, blockIdx.x, blockIdx.y);
}

__device__ __forceinline__ void finalize_sum_body(const double* __restrict__ partial, int nchunk, double scale,
                                                  double* __restrict__ out, int col) {
  const double* p = partial + (long)col * nchunk;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nchunk; i += RB) acc += p[i];
  double r = block_sum(acc);
  if (threadIdx.x == 0) out[col] = r * scale;
}

__global__ void __launch_bounds__(RB) finalize_sum(const double* __restrict__ partial, int nchunk,
                                                   double scale, double* __restrict__ out) {
  finalize_sum_body(partial, nchunk, scale, out, blockIdx.x);
}

__global__ void __launch_bounds__(RB) dot_partial(const double* __restrict__ x, const double* __restrict__ Y, long ldY,
                                                  int mode, long n, double* __restrict__ partial, int nchunk) {
  const double* y = Y + (long)blockIdx.y * ldY;
  long s = (long)blockIdx.x * PER_BLOCK;
  long e = min(s + PER_BLOCK, n);
  double acc = 0.0;
  if (mode == 0) {
    for (long i = s + threadIdx.x; i < e; i += RB) acc = fma(x[i], y[i], acc);
  } else {
    for (long i = s + threadIdx.x; i < e; i += RB) { double d = y[i] - x[i]; acc = fma(d, d, acc); }
  }
  double r = block_sum(acc);
  if (threadIdx.x == 0) partial[(long)blockIdx.y * nchunk + blockIdx.x] = r;
}

__device__ __forceinline__ void dot_jobs_partial_body(const DotJobs& J, long n, double* __restrict__ partial,
                                                      int nchunk, int bx, int j) {
  const double* __restrict__ x = J.x[j];
  const double* __restrict__ y = J.y[j];
  long s = (long)bx * PER_BLOCK;
  long e = min(s + PER_BLOCK, n);
  double acc = 0.0;
  if (J.mode[j] == 0) {
    for (long i = s + threadIdx.x; i < e; i += RB) acc = fma(x[i], y[i], acc);
  } else {
    for (long i = s + threadIdx.x; i < e; i += RB) { double d = y[i] - x[i]; acc = fma(d, d, acc); }
  }
  double r = block_sum(acc);
  if (threadIdx.x == 0) partial[(long)j * nchunk + bx] = r;
}

__global__ void __launch_bounds__(RB) dot_jobs_partial(DotJobs J, long n, double* __restrict__ partial, int nchunk) {
  dot_jobs_partial_body(J, n, partial, nchunk, blockIdx.x, blockIdx.y);
}

__device__ __forceinline__ void accept_P_body(const int* __restrict__ flag, const double* __restrict__ P_old,
                                              double* __restrict__ P_new, long n) {
  if (flag[0] == ST_OK) return;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    P_new[i] = P_old[i];
}

__global__ void accept_P_kernel(const int* __restrict__ flag, const double* __restrict__ P_old,
                                double* __restrict__ P_new, long n) {
  accept_P_body(flag, P_old, P_new, n);
}

// one thread: the order of operations is that of formResidual (:478-496)
__device__ void aim_update(const AimArgs& a, const double* __restrict__ pav_raw, double* __restrict__ faim,
                           int flagv, double* __restrict__ rep, double* __restrict__ aim) {
  double P[9];
  for (int k = 0; k < 9; k++) P[k] = pav_raw[k];
  if (a.rotate) dmk_rotate33(a.q_R, P, false, P);
  for (int k = 0; k < 9; k++) rep[REP_PAV + k] = P[k];
  rep[REP_FLAG] = (double)flagv;
  if (!a.update) return;
  double dP[9], dFa[9];
  for (int k = 0; k < 9; k++) dP[k] = P[k] - a.P_aim[k];
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) {
      double s = 0.0;
      for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++) s += a.S[i + 3 * j + 9 * m + 27 * n] * dP[m + 3 * n];
      dFa[i + 3 * j] = s;
    }
  double e = 0.0;
  for (int k = 0; k < 9; k++) {
    double f = faim[k] - dFa[k];
    faim[k] = f;
    rep[REP_FAIM + k] = f;
    if (!a.stress_mask[k]) e = fmax(e, fabs(dP[k]));
  }
  rep[REP_ERRBC] = e;
  dmk_rotate33(a.q_R, dFa, true, aim);
}

__global__ void residual_aim_kernel(AimArgs a, const double* __restrict__ pav_raw, double* __restrict__ faim,
                                    const int* __restrict__ flag, double* __restrict__ rep, double* __restrict__ aim) {
  if (threadIdx.x != 0) return;
  aim_update(a, pav_raw, faim, flag[0], rep, aim);
}

// the mean rule of the adjoint-operator requests of a round: column sums of
// A^T v (dmk_colsum), then dmk_residual_aim on them
__global__ void __launch_bounds__(RB) adj_colsum_batch(const DmkResidDev* __restrict__ R) {
  const DmkResidDev& q = R[blockIdx.z];
  if (q.kind != DMK_RQ_ADJ_OP || (int)blockIdx.x >= q.nc_resp) return;
  colsum_partial_body(q.Pnew, q.N, q.N, q.partial, q.nc_resp, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(RB) adj_colsum_finalize_batch(const DmkResidDev* __restrict__ R) {
  const DmkResidDev& q = R[blockIdx.y];
  if (q.kind != DMK_RQ_ADJ_OP) return;
  finalize_sum_body(q.partial, q.nc_resp, q.wgt, q.apav, blockIdx.x);
}

__global__ void adj_aim_batch(const DmkResidDev* __restrict__ R) {
  const DmkResidDev& q = R[blockIdx.x];
  if (q.kind != DMK_RQ_ADJ_OP || threadIdx.x != 0) return;
  aim_update(q.a, q.apav, q.afaim, q.flag[0], q.arep, q.aaim);
}

// The reductions of one constitutive response in two launches.  Partial
// (blockIdx.y 0: status, 1-9: the columns of P): per chunk of cells the
// any-failed flag, the Newton and state iteration counts and the nine column
// sums of the new P (each as status_partial /
// colsum_partial compute them).  Finalize (one block, a wavefront per reduced
// quantity): the convergence flag and counters; the raw volume average of P when
// every cell converged (otherwise P stays the previous response's, and so does
// its average); then, when asked, the stress-BC bookkeeping of aim_update.
__device__ __forceinline__ void response_partial_body(const int32_t* __restrict__ st, const int32_t* __restrict__ nn,
                                                      const int32_t* __restrict__ nf, const double* __restrict__ P,
                                                      long N, double* __restrict__ part, int nchunk, int bx, int by) {
  __shared__ double sb[RB], s1[RB], s2[RB];
  long s = (long)bx * PER_BLOCK, e = min(s + PER_BLOCK, N);
  if (by > 0) {
    // column c = by - 1 of P
    const int c = by - 1;
    const double* x = P + (long)c * N;
    double acc = 0.0;
    for (long i = s + threadIdx.x; i < e; i += RB) acc += x[i];
    double r = block_sum(acc);
    if (threadIdx.x == 0) part[(long)(3 + c) * nchunk + bx] = r;
    return;
  }
  double bad = 0.0, a = 0.0, b = 0.0;
  for (long i = s + threadIdx.x; i < e; i += RB) {
    if (st[i] != ST_OK) bad = 1.0;
    a += (double)nn[i];
    b += (double)nf[i];
  }
  sb[threadIdx.x] = bad; s1[threadIdx.x] = a; s2[threadIdx.x] = b;
  __syncthreads();
  for (int k = RB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + k]);
      s1[threadIdx.x] += s1[threadIdx.x + k];
      s2[threadIdx.x] += s2[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[bx] = sb[0];
    part[nchunk + bx] = s1[0];
    part[2 * nchunk + bx] = s2[0];
  }
}

__global__ void __launch_bounds__(RB) response_partial(const int32_t* __restrict__ st, const int32_t* __restrict__ nn,
                                                       const int32_t* __restrict__ nf, const double* __restrict__ P,
                                                       long N, double* __restrict__ part, int nchunk) {
  response_partial_body(st, nn, nf, P, N, part, nchunk, blockIdx.x, blockIdx.y);
}

// sum over the 64 lanes of a wavefront (fixed butterfly order)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// 12 wavefronts, one per reduced quantity: 0 the failure flag, 1-2 the
// iteration counts, 3-11 the columns of P; then one lane does the rest
__device__ __forceinline__ void response_finalize_body(const double* __restrict__ part, int nchunk, double wgt,
                                                       int* __restrict__ flag, unsigned long long* __restrict__ cnt,
                                                       double* __restrict__ pav, const AimArgs& a, int do_aim,
                                                       double* __restrict__ faim, double* __restrict__ rep,
                                                       double* __restrict__ aim) {
  __shared__ double res[12];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double* p = part + (long)w * nchunk;
  double v = 0.0;
  if (w == 0) {
    for (int i = lane; i < nchunk; i += 64) v = fmax(v, p[i]);
    v = wave_max(v);
  } else {
    for (int i = lane; i < nchunk; i += 64) v += p[i];
    v = wave_sum(v);
  }
  if (lane == 0) res[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const bool ok = !(res[0] > 0.0);
    flag[0] = ok ? ST_OK : ST_FAIL_MECH;
    cnt[0] += (unsigned long long)res[1];
    cnt[1] += (unsigned long long)res[2];
    if (ok)
      for (int c = 0; c < 9; c++) pav[c] = res[3 + c] * wgt;
    if (do_aim) aim_update(a, pav, faim, ok ? ST_OK : ST_FAIL_MECH, rep, aim);
  }
}

__global__ void __launch_bounds__(768) response_finalize(const double* __restrict__ part, int nchunk, double wgt,
                                                         int* __restrict__ flag, unsigned long long* __restrict__ cnt,
                                                         double* __restrict__ pav, AimArgs a, int do_aim,
                                                         double* __restrict__ faim, double* __restrict__ rep,
                                                         double* __restrict__ aim) {
  response_finalize_body(part, nchunk, wgt, flag, cnt, pav, a, do_aim, faim, rep, aim);
}

// several sums of partial results in one launch (blockIdx.x = job)
__device__ __forceinline__ void finalize_jobs_body(const SumJobs& J, int j) {
  const double* p = J.p[j];
  double acc = 0.0;
  for (int i = threadIdx.x; i < J.n[j]; i += RB) acc += p[i];
  double r = block_sum(acc);
  if (threadIdx.x == 0) J.out[j][0] = r * J.scale[j];
}

__global__ void __launch_bounds__(RB) finalize_jobs(SumJobs J) { finalize_jobs_body(J, blockIdx.x); }

// The reductions of a lockstep round: the kernels above over every RVE of
// the round (blockIdx.z or .y = RVE; blocks past an RVE's own count return)
__global__ void __launch_bounds__(RB) response_partial_batch(const DmkResidDev* __restrict__ R) {
  const DmkResidDev& q = R[blockIdx.z];
  if (q.kind != DMK_RQ_RESIDUAL || (int)blockIdx.x >= q.nc_resp) return;
  response_partial_body(q.st, q.nn, q.nf, q.Pnew, q.N, q.partial, q.nc_resp, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(768) response_finalize_batch(const DmkResidDev* __restrict__ R) {
  const DmkResidDev& q = R[blockIdx.x];
  if (q.kind != DMK_RQ_RESIDUAL) return;
  response_finalize_body(q.partial, q.nc_resp, q.wgt, q.flag, q.cnt, q.pav, q.a, 1, q.faim, q.rep, q.aim);
}

__global__ void accept_P_batch(const DmkResidDev* __restrict__ R) {
  const DmkResidDev& q = R[blockIdx.y];
  if (q.kind != DMK_RQ_RESIDUAL) return;
  accept_P_body(q.flag, q.Pold, q.Pnew, q.n9);
}

__global__ void __launch_bounds__(RB) dot_jobs_partial_batch(const DmkResidDev* __restrict__ R) {
  const DmkResidDev& q = R[blockIdx.z];
  if ((int)blockIdx.x >= q.nc_dots || (int)blockIdx.y >= q.J.n) return;
  dot_jobs_partial_body(q.J, q.nx, q.partial, q.nc_dots, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(RB) finalize_jobs_batch(const DmkResidDev* __restrict__ R) {
  const DmkResidDev& q = R[blockIdx.y];
  if ((int)blockIdx.x >= q.S.k) return;
  finalize_jobs_body(q.S, blockIdx.x);
}

// the report blocks of the round's RVEs, one after the other (REP_SIZE apart)
__global__ void rep_gather_batch(const DmkResidDev* __restrict__ R, double* __restrict__ out) {
  const DmkResidDev& q = R[blockIdx.x];
  for (int i = threadIdx.x; i < q.rep_n; i += blockDim.x) out[(long)blockIdx.x * REP_SIZE + i] = q.rep[i];
}

__global__ void axpby_kernel(long n, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                             double* __restrict__ z) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = dmk_axpby_elem(a, x[i], b, y[i]);
}

// the vector steps of a lockstep round's requests (blockIdx.y = request),
// each element through the steps in order
__global__ void pre_batch_kernel(const DmkResidDev* __restrict__ R) {
  const DmkResidDev& q = R[blockIdx.y];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = 0; k < q.npre; k++) {
    const DmkPreOp& o = q.pre[k];
    if (i >= o.n) continue;
    if (o.type == DMK_PRE_GS) o.w[i] = dmk_gs_elem(o.w[i], o.V, o.ldV, i, o.l, o.c);
    else o.z[i] = dmk_axpby_elem(o.a, o.x[i], o.b, o.y[i]);
  }
}

__global__ void post_batch_kernel(const DmkResidDev* __restrict__ R) {
  const DmkResidDev& q = R[blockIdx.y];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q.npost == 0) return;
  const DmkPreOp& o = q.post;
  if (i < o.n) o.z[i] = dmk_axpby_elem(o.a, o.x[i], o.b, o.y[i]);
}

__global__ void combine_kernel(long n, const double* __restrict__ XM, CombineArgs c, double* __restrict__ XA) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = XM[i] * (1.0 - c.atot);
  for (int k = 0; k < c.l; k++) v = fma(c.alpha[k], c.X[k][i], v);
  XA[i] = v;
}

__global__ void __launch_bounds__(RB) minmax_partial(const double* __restrict__ nrm, long N, double* __restrict__ pmax,
                                                     double* __restrict__ pmin) {
  __shared__ double smax[RB], smin[RB];
  long s = (long)blockIdx.x * PER_BLOCK, e = min(s + PER_BLOCK, N);
  double mx = 0.0, mn = DBL_MAX;
  for (long i = s + threadIdx.x; i < e; i += RB) {
    double v = nrm[i];
    if (v > mx) mx = v;
    if (v < mn) mn = v;
  }
  smax[threadIdx.x] = mx; smin[threadIdx.x] = mn;
  __syncthreads();
  for (int k = RB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + k]);
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { pmax[blockIdx.x] = smax[0]; pmin[blockIdx.x] = smin[0]; }
}

// global extrema of the per-cell norms from the chunk partials (one block)
__global__ void __launch_bounds__(RB) extrema_finalize(const double* __restrict__ pmax, const double* __restrict__ pmin,
                                                       int nchunk, double* __restrict__ ext) {
  __shared__ double smax[RB], smin[RB];
  double mx = 0.0, mn = DBL_MAX;
  for (int i = threadIdx.x; i < nchunk; i += RB) { mx = fmax(mx, pmax[i]); mn = fmin(mn, pmin[i]); }
  smax[threadIdx.x] = mx; smin[threadIdx.x] = mn;
  __syncthreads();
  for (int k = RB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + k]);
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { ext[0] = smax[0]; ext[1] = smin[0]; }
}

// first cell of each chunk that satisfies the (banded) extremum rule:
// norm (1 + band) >= max, norm (1 - band) <= min; N when none
__global__ void __launch_bounds__(RB) select_partial(const double* __restrict__ nrm, long N,
                                                     const double* __restrict__ ext, double band,
                                                     long* __restrict__ csel, int nchunk) {
  __shared__ long simax[RB], simin[RB];
  const double gmax = ext[0], gmin = ext[1];
  long s = (long)blockIdx.x * PER_BLOCK, e = min(s + PER_BLOCK, N);
  long imax = N, imin = N;
  for (long i = s + threadIdx.x; i < e; i += RB) {
    double v = nrm[i];
    if (imax == N && gmax > 0.0 && v * (1.0 + band) >= gmax) imax = i;
    if (imin == N && gmin < DBL_MAX && v * (1.0 - band) <= gmin) imin = i;
  }
  simax[threadIdx.x] = imax; simin[threadIdx.x] = imin;
  __syncthreads();
  for (int k = RB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      simax[threadIdx.x] = min(simax[threadIdx.x], simax[threadIdx.x + k]);
      simin[threadIdx.x] = min(simin[threadIdx.x], simin[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { csel[blockIdx.x] = simax[0]; csel[nchunk + blockIdx.x] = simin[0]; }
}

// first qualifying cell over all chunks (one block): sel[0] = imax, sel[1] = imin (-1: none)
__global__ void __launch_bounds__(RB) select_finalize(const long* __restrict__ csel, int nchunk, long N,
                                                      long* __restrict__ sel) {
  __shared__ long simax[RB], simin[RB];
  long imax = N, imin = N;
  for (int i = threadIdx.x; i < nchunk; i += RB) { imax = min(imax, csel[i]); imin = min(imin, csel[nchunk + i]); }
  simax[threadIdx.x] = imax; simin[threadIdx.x] = imin;
  __syncthreads();
  for (int k = RB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      simax[threadIdx.x] = min(simax[threadIdx.x], simax[threadIdx.x + k]);
      simin[threadIdx.x] = min(simin[threadIdx.x], simin[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sel[0] = simax[0] < N ? simax[0] : -1;
    sel[1] = simin[0] < N ? simin[0] : -1;
  }
}

__global__ void gather_cminmax(const double* __restrict__ dPdF, long N, const long* __restrict__ sel,
                               double* __restrict__ out) {
  int q = threadIdx.x;
  if (q >= 81) return;
  double mx = sel[0] >= 0 ? dPdF[q * N + sel[0]] : 0.0;
  double mn = sel[1] >= 0 ? dPdF[q * N + sel[1]] : DBL_MAX;
  out[q] = 0.5 * (mx + mn);
}

// per chunk: any cell not converged, sums of Newton and state iterations
// (integers, exact in double below 2^53)
__global__ void __launch_bounds__(RB) status_partial(const int32_t* __restrict__ st, const int32_t* __restrict__ nn,
                                                     const int32_t* __restrict__ nf, long N, double* __restrict__ part,
                                                     int nchunk) {
  __shared__ double sb[RB], s1[RB], s2[RB];
  long s = (long)blockIdx.x * PER_BLOCK, e = min(s + PER_BLOCK, N);
  double bad = 0.0, a = 0.0, b = 0.0;
  for (long i = s + threadIdx.x; i < e; i += RB) {
    if (st[i] != ST_OK) bad = 1.0;
    a += (double)nn[i];
    b += (double)nf[i];
  }
  sb[threadIdx.x] = bad; s1[threadIdx.x] = a; s2[threadIdx.x] = b;
  __syncthreads();
  for (int k = RB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + k]);
      s1[threadIdx.x] += s1[threadIdx.x + k];
      s2[threadIdx.x] += s2[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = sb[0];
    part[nchunk + blockIdx.x] = s1[0];
    part[2 * nchunk + blockIdx.x] = s2[0];
  }
}

// flag[0] = ST_OK if every cell converged, ST_FAIL_MECH otherwise; cnt[0..1]
// accumulate Newton and state iterations (one block)
__global__ void __launch_bounds__(RB) status_finalize(const double* __restrict__ part, int nchunk,
                                                      int* __restrict__ flag, unsigned long long* __restrict__ cnt) {
  __shared__ double sb[RB], s1[RB], s2[RB];
  double bad = 0.0, a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nchunk; i += RB) {
    bad = fmax(bad, part[i]);
    a += part[nchunk + i];
    b += part[2 * nchunk + i];
  }
  sb[threadIdx.x] = bad; s1[threadIdx.x] = a; s2[threadIdx.x] = b;
  __syncthreads();
  for (int k = RB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + k]);
      s1[threadIdx.x] += s1[threadIdx.x + k];
      s2[threadIdx.x] += s2[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    flag[0] = sb[0] > 0.0 ? ST_FAIL_MECH : ST_OK;
    cnt[0] += (unsigned long long)s1[0];
    cnt[1] += (unsigned long long)s2[0];
  }
}

// forwardTensorField / calculateRate (grid_utilities.f90:245-297), component-major
// F and Fnew may alias (element-wise update)
__global__ void forward_field_kernel(long N, int guess, double dt_old, double dt, ForwardArgs fa,
                                     const double* F, double* __restrict__ F_lastinc,
                                     double* __restrict__ Fdot, double* Fnew) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 9 * N) return;
  int c = (int)(i / N);
  if (!fa.cutback) {
    Fdot[i] = guess ? (F[i] - F_lastinc[i]) / dt_old : fa.avRate[c];
    F_lastinc[i] = F[i];
  }
  Fnew[i] = F_lastinc[i] + Fdot[i] * dt;
}

__global__ void shift_kernel(long N, ForwardArgs fa, double* __restrict__ F) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 9 * N) return;
  int c = (int)(i / N);
  F[i] -= fa.mean[c] - fa.aim[c];
}

__global__ void copy_jobs_kernel(CopyJobs j) {
  const int q = blockIdx.y;
  const double* __restrict__ src = j.src[q];
  double* __restrict__ dst = j.dst[q];
  const long n = j.n[q];
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) dst[i] = src[i];
}

// (9,N) cell-major <-> component-major
__global__ void aos_to_soa(const double* __restrict__ a, long N, int w, double* __restrict__ s) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)w * N) return;
  long e = i / w; int c = (int)(i % w);
  s[c * N + e] = a[i];
}
__global__ void soa_to_aos(const double* __restrict__ s, long N, int w, double* __restrict__ a) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)w * N) return;
  long e = i / w; int c = (int)(i % w);
  a[i] = s[c * N + e];
}

}  // namespace

// ---------------------------------------------------------------------------- host wrappers
int dmk_nchunk(long n) { return (int)((n + PER_BLOCK - 1) / PER_BLOCK); }

int dmk_colsum(const double* base, long ld, long n, int ncols, double* partial, double* out, double scale,
               hipStream_t s) {
  int nc = dmk_nchunk(n);
  hipLaunchKernelGGL(colsum_partial, dim3(nc, ncols), dim3(RB), 0, s, base, ld, n, partial, nc);
  hipLaunchKernelGGL(finalize_sum, dim3(ncols), dim3(RB), 0, s, partial, nc, scale, out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_mdot(const double* x, const double* Y, long ldY, int l, int mode, long n, double* partial, double* out,
             hipStream_t s) {
  int nc = dmk_nchunk(n);
  hipLaunchKernelGGL(dot_partial, dim3(nc, l), dim3(RB), 0, s, x, Y, ldY, mode, n, partial, nc);
  hipLaunchKernelGGL(finalize_sum, dim3(l), dim3(RB), 0, s, partial, nc, 1.0, out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_dots(const DotJobs* jobs, long n, double* partial, double* out, hipStream_t s) {
  if (jobs->n <= 0) return 0;
  int nc = dmk_nchunk(n);
  hipLaunchKernelGGL(dot_jobs_partial, dim3(nc, jobs->n), dim3(RB), 0, s, *jobs, n, partial, nc);
  hipLaunchKernelGGL(finalize_sum, dim3(jobs->n), dim3(RB), 0, s, partial, nc, 1.0, out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_response_reduce(const int32_t* st, const int32_t* nn, const int32_t* nf, const double* P, long N,
                        double wgt, double* partial, int* flag, unsigned long long* cnt, double* pav,
                        const AimArgs* a, double* faim, double* report, double* aim, hipStream_t s) {
  int nc = dmk_nchunk(N);
  hipLaunchKernelGGL(response_partial, dim3(nc, 10), dim3(RB), 0, s, st, nn, nf, P, N, partial, nc);
  AimArgs none{};
  hipLaunchKernelGGL(response_finalize, dim3(1), dim3(768), 0, s, partial, nc, wgt, flag, cnt, pav, a ? *a : none,
                     a ? 1 : 0, faim, report, aim);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_dots_sums(const DotJobs* jobs, long n, double* partial, double* out, const SumJobs* extra, hipStream_t s) {
  SumJobs J = extra ? *extra : SumJobs{};
  if (!extra) J.k = 0;
  int nc = dmk_nchunk(n);
  if (jobs && jobs->n > 0) {
    hipLaunchKernelGGL(dot_jobs_partial, dim3(nc, jobs->n), dim3(RB), 0, s, *jobs, n, partial, nc);
    for (int j = 0; j < jobs->n; j++) {
      if (J.k >= DMK_MAX_SUMS) return 1;
      J.p[J.k] = partial + (long)j * nc;
      J.n[J.k] = nc;
      J.scale[J.k] = 1.0;
      J.out[J.k] = out + j;
      J.k++;
    }
  }
  if (J.k == 0) return 0;
  hipLaunchKernelGGL(finalize_jobs, dim3(J.k), dim3(RB), 0, s, J);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_resid_reduce_batch(const DmkResidDev* R, int nr, int max_nc_resp, hipStream_t s) {
  if (nr < 1) return 0;
  hipLaunchKernelGGL(response_partial_batch, dim3(max_nc_resp, 10, nr), dim3(RB), 0, s, R);
  hipLaunchKernelGGL(response_finalize_batch, dim3(nr), dim3(768), 0, s, R);
  hipLaunchKernelGGL(accept_P_batch, dim3(512, nr), dim3(256), 0, s, R);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_pre_batch(const DmkResidDev* R, int nr, long max_n, hipStream_t s) {
  if (nr < 1 || max_n < 1) return 0;
  hipLaunchKernelGGL(pre_batch_kernel, dim3((unsigned)((max_n + 255) / 256), nr), dim3(256), 0, s, R);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_post_batch(const DmkResidDev* R, int nr, long max_n, hipStream_t s) {
  if (nr < 1 || max_n < 1) return 0;
  hipLaunchKernelGGL(post_batch_kernel, dim3((unsigned)((max_n + 255) / 256), nr), dim3(256), 0, s, R);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_adj_mean_batch(const DmkResidDev* R, int nr, int max_nc, hipStream_t s) {
  if (nr < 1) return 0;
  hipLaunchKernelGGL(adj_colsum_batch, dim3(max_nc, 9, nr), dim3(RB), 0, s, R);
  hipLaunchKernelGGL(adj_colsum_finalize_batch, dim3(9, nr), dim3(RB), 0, s, R);
  hipLaunchKernelGGL(adj_aim_batch, dim3(nr), dim3(64), 0, s, R);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_resid_dots_batch(const DmkResidDev* R, int nr, int max_nc_dots, int max_jobs, int max_sums, double* rep_out,
                         hipStream_t s) {
  if (nr < 1) return 0;
  if (max_jobs > 0) hipLaunchKernelGGL(dot_jobs_partial_batch, dim3(max_nc_dots, max_jobs, nr), dim3(RB), 0, s, R);
  if (max_sums > 0) hipLaunchKernelGGL(finalize_jobs_batch, dim3(max_sums, nr), dim3(RB), 0, s, R);
  hipLaunchKernelGGL(rep_gather_batch, dim3(nr), dim3(128), 0, s, R, rep_out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_residual_aim(const AimArgs* a, const double* pav_raw, double* faim, const int* flag, double* report,
                     double* aim, hipStream_t s) {
  hipLaunchKernelGGL(residual_aim_kernel, dim3(1), dim3(64), 0, s, *a, pav_raw, faim, flag, report, aim);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_accept_P(const int* flag, const double* P_old, double* P_new, long n, hipStream_t s) {
  hipLaunchKernelGGL(accept_P_kernel, dim3(512), dim3(256), 0, s, flag, P_old, P_new, n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_axpby(long n, double a, const double* x, double b, const double* y, double* z, hipStream_t s) {
  hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, x, b, y, z);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_combine(long n, const double* XM, const CombineArgs* c, double* XA, hipStream_t s) {
  hipLaunchKernelGGL(combine_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, XM, *c, XA);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_cminmax(const double* dPdF, long N, double band, const double* nrm, double* pmax, double* pmin,
                long* csel, long* sel, double* ext, double* out, hipStream_t s) {
  int nc = dmk_nchunk(N);
  hipLaunchKernelGGL(minmax_partial, dim3(nc), dim3(RB), 0, s, nrm, N, pmax, pmin);
  hipLaunchKernelGGL(extrema_finalize, dim3(1), dim3(RB), 0, s, pmax, pmin, nc, ext);
  hipLaunchKernelGGL(select_partial, dim3(nc), dim3(RB), 0, s, nrm, N, ext, band, csel, nc);
  hipLaunchKernelGGL(select_finalize, dim3(1), dim3(RB), 0, s, csel, nc, N, sel);
  hipLaunchKernelGGL(gather_cminmax, dim3(1), dim3(128), 0, s, dPdF, N, sel, out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_status(const int32_t* st, const int32_t* nn, const int32_t* nf, long N, double* partial, int* flag,
               unsigned long long* cnt, hipStream_t s) {
  int nc = dmk_nchunk(N);
  hipLaunchKernelGGL(status_partial, dim3(nc), dim3(RB), 0, s, st, nn, nf, N, partial, nc);
  hipLaunchKernelGGL(status_finalize, dim3(1), dim3(RB), 0, s, partial, nc, flag, cnt);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_forward_field(long N, int guess, double dt_old, double dt, const ForwardArgs* fa, const double* F,
                      double* F_lastinc, double* Fdot, double* Fnew, hipStream_t s) {
  hipLaunchKernelGGL(forward_field_kernel, dim3((unsigned)((9 * N + 255) / 256)), dim3(256), 0, s, N, guess, dt_old,
                     dt, *fa, F, F_lastinc, Fdot, Fnew);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_shift_field(long N, const ForwardArgs* fa, double* F, hipStream_t s) {
  hipLaunchKernelGGL(shift_kernel, dim3((unsigned)((9 * N + 255) / 256)), dim3(256), 0, s, N, *fa, F);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_copy_jobs(const CopyJobs* j, hipStream_t s) {
  if (j->k < 1 || j->k > 8) return 1;
  long mx = 0;
  for (int q = 0; q < j->k; q++) mx = j->n[q] > mx ? j->n[q] : mx;
  unsigned nb = (unsigned)std::min<long>((mx + 255) / 256, 4096);
  hipLaunchKernelGGL(copy_jobs_kernel, dim3(nb ? nb : 1, (unsigned)j->k), dim3(256), 0, s, *j);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_aos_to_soa(const double* a, long N, int w, double* soa, hipStream_t s) {
  hipLaunchKernelGGL(aos_to_soa, dim3((unsigned)((w * N + 255) / 256)), dim3(256), 0, s, a, N, w, soa);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_soa_to_aos(const double* soa, long N, int w, double* a, hipStream_t s) {
  hipLaunchKernelGGL(soa_to_aos, dim3((unsigned)((w * N + 255) / 256)), dim3(256), 0, s, soa, N, w, a);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
