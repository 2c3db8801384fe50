// Spectral operators of the basic scheme on MI355X (src/grid/spectral_utilities.f90).
//
// * Batched 3-D r2c / c2r FFTs of component-major fields, built from 1-D
//   mixed-radix (8, 4, 2, 3, 5, 7, generic) Stockham passes staged in LDS:
//     x pass: two real lines packed into one complex line (r2c/c2r split),
//     y and z passes: c2c on groups of lines adjacent in x so that every
//     global load and store is coalesced across the group.
//   Unnormalised, forward sign -1, like FFTW's r2c/c2r (:239-261,455,498,661).
// * gamma_div_kernel: one thread per frequency; the RMS divergence
//   (:650-691) and the memory-efficient Gamma convolution (:436-501) read the
//   same forward transform, which the reference computes twice.
//   With the continuous derivative xi = i a, Gamma_hat:P_hat collapses to
//   (A^-1 (P_hat a)) (x) a with the acoustic tensor A_lm = C_ref,ljmq a_j a_q.
// * curl_kernel: the RMS curl of the polarization scheme (:697-757).
// * galerkin_kernel: the G convolution of the Galerkin scheme (:507-589).
// * coords kernels: FFT integration of F for the node and cell positions
//   (:857-968).
#include "dmk_common.h"
#include "dmk_spectral.h"
#include "dmk_resid.h"

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// twiddle w^m, w = exp(sign 2 pi i / L); table holds the forward (sign -1) values
__device__ __forceinline__ double2 twid(const double2* __restrict__ tw, int m, int sign) {
  double2 t = tw[m];
  return sign < 0 ? t : cconj(t);
}

// Work-groups of a launch are dispatched round-robin over the eight XCDs of
// the MI355X, each with its own L2.  Kernels whose neighbouring blocks read
// neighbouring columns (sharing 128-byte lines) number their blocks through
// this map, which hands each XCD a contiguous range of logical blocks, so the
// blocks that share a line share an L2.  A bijection of [0, nb).
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned q = nb >> 3, r = nb & 7u, x = b & 7u, s = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + s;
}

template <int R>
__device__ __forceinline__ void butterfly(double2* v, const double2* __restrict__ tw, int L, int sign) {
  if constexpr (R == 2) {
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    double2 a = v[0], b = v[1], c = v[2], d = v[3];
    double2 s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d), s3 = csub(b, d);
    // multiply s3 by -i (forward) or +i (inverse)
    double2 r3 = sign < 0 ? make_double2(s3.y, -s3.x) : make_double2(-s3.y, s3.x);
    v[0] = cadd(s0, s2);
    v[2] = csub(s0, s2);
    v[1] = cadd(s1, r3);
    v[3] = csub(s1, r3);
  } else if constexpr (R == 8) {
    // radix 2 x 4: a_j = x_j + x_{j+4} gives the even outputs, (x_j - x_{j+4}) w8^j the odd ones
    const double c = 0.70710678118654752440;
    double2 a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; j++) { a[j] = cadd(v[j], v[j + 4]); b[j] = csub(v[j], v[j + 4]); }
    if (sign < 0) {
      b[1] = make_double2((b[1].x + b[1].y) * c, (b[1].y - b[1].x) * c);
      b[2] = make_double2(b[2].y, -b[2].x);
      b[3] = make_double2((b[3].y - b[3].x) * c, -(b[3].x + b[3].y) * c);
    } else {
      b[1] = make_double2((b[1].x - b[1].y) * c, (b[1].y + b[1].x) * c);
      b[2] = make_double2(-b[2].y, b[2].x);
      b[3] = make_double2(-(b[3].x + b[3].y) * c, (b[3].x - b[3].y) * c);
    }
    butterfly<4>(a, tw, L, sign);
    butterfly<4>(b, tw, L, sign);
#pragma unroll
    for (int j = 0; j < 4; j++) { v[2 * j] = a[j]; v[2 * j + 1] = b[j]; }
  } else {
    double2 w[R];
    const int st = L / R;
#pragma unroll
    for (int m = 0; m < R; m++) w[m] = twid(tw, m * st, sign);
    double2 o[R];
#pragma unroll
    for (int p = 0; p < R; p++) {
      double2 acc = v[0];
#pragma unroll
      for (int q = 1; q < R; q++) acc = cadd(acc, cmul(v[q], w[(p * q) % R]));
      o[p] = acc;
    }
#pragma unroll
    for (int p = 0; p < R; p++) v[p] = o[p];
  }
}

// one Stockham stage of radix R over `nlines` lines of length L held in LDS
template <int R>
__device__ void stockham_stage(const double2* in, double2* out, int nlines, int L, int ldl, int Ns,
                               const double2* __restrict__ tw, int sign) {
  const int nb = L / R;
  const int tot = nlines * nb;
  const int tstep = L / (Ns * R);
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    int line = idx / nb, j = idx - line * nb;
    const double2* src = in + line * ldl;
    double2* dst = out + line * ldl;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = src[j + r * nb];
    int k = j % Ns;
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) v[r] = cmul(v[r], twid(tw, (r * k * tstep) % L, sign));
    }
    butterfly<R>(v, tw, L, sign);
    int o = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) dst[o + r * Ns] = v[r];
  }
}

// generic prime radix (rare sizes): O(R^2) DFT
__device__ void stockham_stage_generic(const double2* in, double2* out, int nlines, int L, int ldl, int Ns, int R,
                                       const double2* __restrict__ tw, int sign) {
  const int nb = L / R;
  const int tot = nlines * nb;
  const int tstep = L / (Ns * R);
  const int st = L / R;
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    int line = idx / nb, j = idx - line * nb;
    const double2* src = in + line * ldl;
    double2* dst = out + line * ldl;
    int k = j % Ns;
    int o = (j / Ns) * Ns * R + k;
    for (int p = 0; p < R; p++) {
      double2 acc = make_double2(0.0, 0.0);
      for (int q = 0; q < R; q++) {
        double2 x = src[j + q * nb];
        if (Ns > 1 && q > 0) x = cmul(x, twid(tw, (q * k * tstep) % L, sign));
        acc = cadd(acc, q == 0 ? x : cmul(x, twid(tw, ((p * q) % R) * st, sign)));
      }
      dst[o + p * Ns] = acc;
    }
  }
}

// full 1-D FFT of nlines LDS lines; returns the buffer holding the result
__device__ double2* lds_fft(double2* b0, double2* b1, int nlines, int L, int ldl, const FftPlan& pl,
                            const double2* __restrict__ tw, int sign) {
  double2 *in = b0, *out = b1;
  int Ns = 1;
  for (int s = 0; s < pl.nstages; s++) {
    int R = pl.radix[s];
    switch (R) {
      case 2: stockham_stage<2>(in, out, nlines, L, ldl, Ns, tw, sign); break;
      case 3: stockham_stage<3>(in, out, nlines, L, ldl, Ns, tw, sign); break;
      case 4: stockham_stage<4>(in, out, nlines, L, ldl, Ns, tw, sign); break;
      case 5: stockham_stage<5>(in, out, nlines, L, ldl, Ns, tw, sign); break;
      case 7: stockham_stage<7>(in, out, nlines, L, ldl, Ns, tw, sign); break;
      case 8: stockham_stage<8>(in, out, nlines, L, ldl, Ns, tw, sign); break;
      default: stockham_stage_generic(in, out, nlines, L, ldl, Ns, R, tw, sign); break;
    }
    __syncthreads();
    double2* t = in; in = out; out = t;
    Ns *= R;
  }
  return in;
}

// x pass, forward: real lines (nl lines of length L, contiguous) -> half spectra
// (nl lines of length h, contiguous); pairs of lines share one complex FFT.
__global__ void __launch_bounds__(256) fft_x_r2c(const double* __restrict__ in, double2* __restrict__ out,
                                                 int nl, FftPlan pl, const double2* __restrict__ tw, int pairs_per_block) {
  extern __shared__ double2 lds[];
  const int L = pl.L, h = L / 2 + 1, ldl = L + 1;
  double2* b0 = lds;
  double2* b1 = lds + pairs_per_block * ldl;
  const int p0 = blockIdx.x * pairs_per_block;
  const long l0 = 2L * p0;
  const int nlines_here = min(2 * pairs_per_block, nl - (int)l0);
  const int npairs = (nlines_here + 1) / 2;
  for (int i = threadIdx.x; i < 2 * pairs_per_block * L; i += blockDim.x) {
    int line = i / L, pos = i - line * L;
    double v = line < nlines_here ? in[(l0 + line) * L + pos] : 0.0;
    double2* c = &b0[(line >> 1) * ldl + pos];
    if (line & 1) c->y = v; else c->x = v;
  }
  __syncthreads();
  double2* r = lds_fft(b0, b1, npairs, L, ldl, pl, tw, -1);
  for (int i = threadIdx.x; i < 2 * pairs_per_block * h; i += blockDim.x) {
    int line = i / h, k = i - line * h;
    if (line >= nlines_here) continue;
    const double2* z = r + (line >> 1) * ldl;
    double2 a = z[k], b = cconj(z[(L - k) % L]);
    double2 v;
    if (line & 1) {            // (Z[k] - conj Z[L-k]) / (2i)
      double2 d = csub(a, b);
      v = make_double2(0.5 * d.y, -0.5 * d.x);
    } else {
      double2 s = cadd(a, b);
      v = make_double2(0.5 * s.x, 0.5 * s.y);
    }
    out[(l0 + line) * h + k] = v;
  }
}

// x pass, inverse: half spectra -> real lines (imaginary parts of the DC and
// Nyquist bins are ignored, as a c2r transform does)
__global__ void __launch_bounds__(256) fft_x_c2r(const double2* __restrict__ in, double* __restrict__ out,
                                                 int nl, FftPlan pl, const double2* __restrict__ tw, int pairs_per_block) {
  extern __shared__ double2 lds[];
  const int L = pl.L, h = L / 2 + 1, ldl = L + 1;
  double2* b0 = lds;
  double2* b1 = lds + pairs_per_block * ldl;
  const int p0 = blockIdx.x * pairs_per_block;
  const long l0 = 2L * p0;
  const int nlines_here = min(2 * pairs_per_block, nl - (int)l0);
  const int npairs = (nlines_here + 1) / 2;
  // Z[k] = A[k] + i B[k] with A, B extended by Hermitian symmetry; the
  // imaginary parts of the DC and Nyquist bins do not enter
  for (int i = threadIdx.x; i < pairs_per_block * L; i += blockDim.x) {
    int p = i / L, k = i - p * L;
    double2 z = make_double2(0.0, 0.0);
    if (p < npairs) {
      int kk = (k < h) ? k : L - k;
      bool cj = k >= h;
      long la = l0 + 2 * p, lb = la + 1;
      double2 a = in[la * h + kk];
      if (kk == 0 || 2 * kk == L) a.y = 0.0;
      if (cj) a.y = -a.y;
      double2 b = make_double2(0.0, 0.0);
      if (2 * p + 1 < nlines_here) {
        b = in[lb * h + kk];
        if (kk == 0 || 2 * kk == L) b.y = 0.0;
        if (cj) b.y = -b.y;
      }
      z = make_double2(a.x - b.y, a.y + b.x);
    }
    b0[p * ldl + k] = z;
  }
  __syncthreads();
  double2* r = lds_fft(b0, b1, npairs, L, ldl, pl, tw, +1);
  for (int i = threadIdx.x; i < 2 * pairs_per_block * L; i += blockDim.x) {
    int line = i / L, pos = i - line * L;
    if (line >= nlines_here) continue;
    double2 z = r[(line >> 1) * ldl + pos];
    out[(l0 + line) * L + pos] = (line & 1) ? z.y : z.x;
  }
}

// In-place variants for radix 2, 4 and 8 plans: each thread keeps its
// butterflies (at most KMAX) in registers across a barrier and writes them
// back into the same LDS buffer, so one buffer suffices and more blocks fit on
// a CU.  Same arithmetic as stockham_stage / lds_fft.
constexpr int IP_KMAX = 2;     // butterflies per thread and stage
constexpr int IP_KU = 12;      // unpacked values per thread

template <int R>
__device__ void stockham_stage_ip(double2* buf, int nlines, int L, int ldl, int Ns, const double2* __restrict__ tw,
                                  int sign) {
  const int nb = L / R;
  const int tot = nlines * nb;
  const int tstep = L / (Ns * R);
  double2 v[IP_KMAX][R];
  int dst[IP_KMAX];
#pragma unroll
  for (int q = 0; q < IP_KMAX; q++) {
    const int idx = threadIdx.x + q * blockDim.x;
    dst[q] = -1;
    if (idx < tot) {
      const int line = idx / nb, j = idx - line * nb;
      const double2* src = buf + line * ldl;
#pragma unroll
      for (int r = 0; r < R; r++) v[q][r] = src[j + r * nb];
      const int k = j % Ns;
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) v[q][r] = cmul(v[q][r], twid(tw, (r * k * tstep) % L, sign));
      }
      butterfly<R>(v[q], tw, L, sign);
      dst[q] = line * ldl + (j / Ns) * Ns * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < IP_KMAX; q++)
    if (dst[q] >= 0) {
#pragma unroll
      for (int r = 0; r < R; r++) buf[dst[q] + r * Ns] = v[q][r];
    }
  __syncthreads();
}

__device__ void lds_fft_ip(double2* buf, int nlines, int L, int ldl, const FftPlan& pl, const double2* __restrict__ tw,
                           int sign) {
  int Ns = 1;
  for (int s = 0; s < pl.nstages; s++) {
    const int R = pl.radix[s];
    if (R == 8) stockham_stage_ip<8>(buf, nlines, L, ldl, Ns, tw, sign);
    else if (R == 4) stockham_stage_ip<4>(buf, nlines, L, ldl, Ns, tw, sign);
    else stockham_stage_ip<2>(buf, nlines, L, ldl, Ns, tw, sign);
    Ns *= R;
  }
}

__device__ __forceinline__ void xy_r2c_ip_body(const double* __restrict__ in, double2* __restrict__ out, int N1,
                                               int N2, int N3, const FftPlan& px, const double2* __restrict__ twx,
                                               const FftPlan& py, const double2* __restrict__ twy, unsigned plane) {
  extern __shared__ double2 lds[];
  const int h = N1 / 2 + 1, np = (N2 + 1) / 2, lx = N1 + 1, ly = N2 + 1;
  double2* A = lds;
  const double* src = in + (long)plane * N1 * N2;
  for (int i = threadIdx.x; i < 2 * np * N1; i += blockDim.x) {
    int line = i / N1, pos = i - line * N1;
    double v = line < N2 ? src[(long)line * N1 + pos] : 0.0;
    double2* c = &A[(line >> 1) * lx + pos];
    if (line & 1) c->y = v; else c->x = v;
  }
  __syncthreads();
  lds_fft_ip(A, np, N1, lx, px, twx, -1);
  // unpack the pairs into y lines (in registers across the barrier)
  double2 u[IP_KU];
#pragma unroll
  for (int q = 0; q < IP_KU; q++) {
    const int i = threadIdx.x + q * blockDim.x;
    if (i < N2 * h) {
      int line = i / h, k = i - line * h;
      const double2* z = A + (line >> 1) * lx;
      double2 a = z[k], b = cconj(z[(N1 - k) % N1]);
      if (line & 1) {
        double2 d = csub(a, b);
        u[q] = make_double2(0.5 * d.y, -0.5 * d.x);
      } else {
        double2 s = cadd(a, b);
        u[q] = make_double2(0.5 * s.x, 0.5 * s.y);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < IP_KU; q++) {
    const int i = threadIdx.x + q * blockDim.x;
    if (i < N2 * h) {
      int line = i / h, k = i - line * h;
      A[k * ly + line] = u[q];
    }
  }
  __syncthreads();
  lds_fft_ip(A, h, N2, ly, py, twy, -1);
  double2* dst = out + (plane / N3) * ((long)h * N2 * N3) + (long)(plane % N3) * h * N2;
  for (int i = threadIdx.x; i < N2 * h; i += blockDim.x) {
    int ky = i / h, kx = i - ky * h;
    dst[i] = A[kx * ly + ky];
  }
}

__global__ __launch_bounds__(512) void fft_xy_r2c_ip(const double* __restrict__ in, double2* __restrict__ out, int N1, int N2, int N3,
                              FftPlan px, const double2* __restrict__ twx, FftPlan py,
                              const double2* __restrict__ twy) {
  xy_r2c_ip_body(in, out, N1, N2, N3, px, twx, py, twy, blockIdx.x);
}

__device__ __forceinline__ void yx_c2r_ip_body(const double2* __restrict__ in, double* __restrict__ out, int N1,
                                               int N2, int N3, const FftPlan& px, const double2* __restrict__ twx,
                                               const FftPlan& py, const double2* __restrict__ twy, unsigned plane) {
  extern __shared__ double2 lds[];
  const int h = N1 / 2 + 1, np = (N2 + 1) / 2, lx = N1 + 1, ly = N2 + 1;
  double2* A = lds;
  const double2* src = in + (plane / N3) * ((long)h * N2 * N3) + (long)(plane % N3) * h * N2;
  for (int i = threadIdx.x; i < N2 * h; i += blockDim.x) {
    int ky = i / h, kx = i - ky * h;
    A[kx * ly + ky] = src[i];
  }
  __syncthreads();
  lds_fft_ip(A, h, N2, ly, py, twy, +1);
  // pack pairs of y lines into complex x lines (registers across the barrier)
  double2 u[IP_KU];
#pragma unroll
  for (int q = 0; q < IP_KU; q++) {
    const int i = threadIdx.x + q * blockDim.x;
    if (i < np * N1) {
      int p = i / N1, k = i - p * N1;
      int kk = (k < h) ? k : N1 - k;
      bool cj = k >= h;
      double2 a = A[kk * ly + 2 * p];
      if (kk == 0 || 2 * kk == N1) a.y = 0.0;
      if (cj) a.y = -a.y;
      double2 b = make_double2(0.0, 0.0);
      if (2 * p + 1 < N2) {
        b = A[kk * ly + 2 * p + 1];
        if (kk == 0 || 2 * kk == N1) b.y = 0.0;
        if (cj) b.y = -b.y;
      }
      u[q] = make_double2(a.x - b.y, a.y + b.x);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < IP_KU; q++) {
    const int i = threadIdx.x + q * blockDim.x;
    if (i < np * N1) {
      int p = i / N1, k = i - p * N1;
      A[p * lx + k] = u[q];
    }
  }
  __syncthreads();
  lds_fft_ip(A, np, N1, lx, px, twx, +1);
  double* dst = out + (long)plane * N1 * N2;
  for (int i = threadIdx.x; i < N1 * N2; i += blockDim.x) {
    int line = i / N1, pos = i - line * N1;
    double2 z = A[(line >> 1) * lx + pos];
    dst[i] = (line & 1) ? z.y : z.x;
  }
}

__global__ __launch_bounds__(512) void fft_yx_c2r_ip(const double2* __restrict__ in, double* __restrict__ out, int N1, int N2, int N3,
                              FftPlan px, const double2* __restrict__ twx, FftPlan py,
                              const double2* __restrict__ twy) {
  yx_c2r_ip_body(in, out, N1, N2, N3, px, twx, py, twy, blockIdx.x);
}

// Fused x and y passes, one block per (field, z) plane: the real plane and its
// half-spectrum plane are each contiguous in memory, so every global access
// is coalesced and the plane stays in LDS between the passes.  The arithmetic
// is that of fft_x_r2c / fft_x_c2r followed or preceded by fft_strided along
// y (same line packing, same Stockham stages), so for even N2 results are
// bitwise those of the separate passes (for odd N2 the x pass pairs the last
// line of a plane with zeros instead of the next plane's first line: equal to
// rounding).  LDS: two buffers of max(ceil(N2/2) (N1+1), h (N2+1)) complex
// values.
__device__ __forceinline__ int xy_buf(int N1, int N2) {
  const int h = N1 / 2 + 1, np = (N2 + 1) / 2;
  int a = np * (N1 + 1), b = h * (N2 + 1);
  return a > b ? a : b;
}

__device__ __forceinline__ void xy_r2c_body(const double* __restrict__ in, double2* __restrict__ out, int N1, int N2,
                                            int N3, const FftPlan& px, const double2* __restrict__ twx,
                                            const FftPlan& py, const double2* __restrict__ twy, unsigned plane) {
  extern __shared__ double2 lds[];
  const int h = N1 / 2 + 1, np = (N2 + 1) / 2, lx = N1 + 1, ly = N2 + 1;
  double2* A = lds;
  double2* B = lds + xy_buf(N1, N2);
  const double* src = in + (long)plane * N1 * N2;   // plane = field * N3 + z
  for (int i = threadIdx.x; i < 2 * np * N1; i += blockDim.x) {
    int line = i / N1, pos = i - line * N1;
    double v = line < N2 ? src[(long)line * N1 + pos] : 0.0;
    double2* c = &A[(line >> 1) * lx + pos];
    if (line & 1) c->y = v; else c->x = v;
  }
  __syncthreads();
  double2* r = lds_fft(A, B, np, N1, lx, px, twx, -1);
  double2* t = r == A ? B : A;
  for (int i = threadIdx.x; i < N2 * h; i += blockDim.x) {
    int line = i / h, k = i - line * h;
    const double2* z = r + (line >> 1) * lx;
    double2 a = z[k], b = cconj(z[(N1 - k) % N1]);
    double2 v;
    if (line & 1) {
      double2 d = csub(a, b);
      v = make_double2(0.5 * d.y, -0.5 * d.x);
    } else {
      double2 s = cadd(a, b);
      v = make_double2(0.5 * s.x, 0.5 * s.y);
    }
    t[k * ly + line] = v;
  }
  __syncthreads();
  double2* q = lds_fft(t, r, h, N2, ly, py, twy, -1);
  double2* dst = out + (plane / N3) * ((long)h * N2 * N3) + (long)(plane % N3) * h * N2;
  for (int i = threadIdx.x; i < N2 * h; i += blockDim.x) {
    int ky = i / h, kx = i - ky * h;
    dst[i] = q[kx * ly + ky];
  }
}

__global__ void fft_xy_r2c(const double* __restrict__ in, double2* __restrict__ out, int N1, int N2, int N3,
                           FftPlan px, const double2* __restrict__ twx, FftPlan py,
                           const double2* __restrict__ twy) {
  xy_r2c_body(in, out, N1, N2, N3, px, twx, py, twy, blockIdx.x);
}

__device__ __forceinline__ void yx_c2r_body(const double2* __restrict__ in, double* __restrict__ out, int N1, int N2,
                                            int N3, const FftPlan& px, const double2* __restrict__ twx,
                                            const FftPlan& py, const double2* __restrict__ twy, unsigned plane) {
  extern __shared__ double2 lds[];
  const int h = N1 / 2 + 1, np = (N2 + 1) / 2, lx = N1 + 1, ly = N2 + 1;
  double2* A = lds;
  double2* B = lds + xy_buf(N1, N2);
  const double2* src = in + (plane / N3) * ((long)h * N2 * N3) + (long)(plane % N3) * h * N2;
  for (int i = threadIdx.x; i < N2 * h; i += blockDim.x) {
    int ky = i / h, kx = i - ky * h;
    A[kx * ly + ky] = src[i];
  }
  __syncthreads();
  double2* r = lds_fft(A, B, h, N2, ly, py, twy, +1);
  double2* t = r == A ? B : A;
  // pairs of y lines into one complex line, Hermitian-extended in x; the
  // imaginary parts of the DC and Nyquist bins do not enter
  for (int i = threadIdx.x; i < np * N1; i += blockDim.x) {
    int p = i / N1, k = i - p * N1;
    int kk = (k < h) ? k : N1 - k;
    bool cj = k >= h;
    double2 a = r[kk * ly + 2 * p];
    if (kk == 0 || 2 * kk == N1) a.y = 0.0;
    if (cj) a.y = -a.y;
    double2 b = make_double2(0.0, 0.0);
    if (2 * p + 1 < N2) {
      b = r[kk * ly + 2 * p + 1];
      if (kk == 0 || 2 * kk == N1) b.y = 0.0;
      if (cj) b.y = -b.y;
    }
    t[p * lx + k] = make_double2(a.x - b.y, a.y + b.x);
  }
  __syncthreads();
  double2* q = lds_fft(t, r, np, N1, lx, px, twx, +1);
  double* dst = out + (long)plane * N1 * N2;
  for (int i = threadIdx.x; i < N1 * N2; i += blockDim.x) {
    int line = i / N1, pos = i - line * N1;
    double2 z = q[(line >> 1) * lx + pos];
    dst[i] = (line & 1) ? z.y : z.x;
  }
}

__global__ void fft_yx_c2r(const double2* __restrict__ in, double* __restrict__ out, int N1, int N2, int N3,
                           FftPlan px, const double2* __restrict__ twx, FftPlan py,
                           const double2* __restrict__ twy) {
  yx_c2r_body(in, out, N1, N2, N3, px, twx, py, twy, blockIdx.x);
}

// c2c pass along a strided axis.  A line group is `gx` lines adjacent in x
// (unit stride) starting at x0, of length L with element stride `stride`;
// groups are enumerated over (outer index, x chunk); outer o has base offset
// (o / M) * comp_stride + (o % M) * outer_stride.
__global__ void __launch_bounds__(256) fft_strided(double2* __restrict__ data, int nx, int gx, int M, long comp_stride,
                                                   long outer_stride, long stride, FftPlan pl,
                                                   const double2* __restrict__ tw, int sign) {
  extern __shared__ double2 lds[];
  const int L = pl.L, ldl = L + 1;
  double2* b0 = lds;
  double2* b1 = lds + gx * ldl;
  const int nchunk = (nx + gx - 1) / gx;
  const int blk = (int)xcd_block(blockIdx.x, gridDim.x);
  const int o = blk / nchunk, ch = blk - o * nchunk;
  const int x0 = ch * gx;
  const int nlx = min(gx, nx - x0);
  double2* base = data + (long)(o / M) * comp_stride + (long)(o % M) * outer_stride + x0;
  for (int i = threadIdx.x; i < gx * L; i += blockDim.x) {
    int j = i / gx, x = i - j * gx;
    b0[x * ldl + j] = x < nlx ? base[(long)j * stride + x] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  double2* r = lds_fft(b0, b1, gx, L, ldl, pl, tw, sign);
  for (int i = threadIdx.x; i < gx * L; i += blockDim.x) {
    int j = i / gx, x = i - j * gx;
    if (x < nlx) base[(long)j * stride + x] = r[x * ldl + j];
  }
}

// ---------------------------------------------------------------- Gamma + div
__device__ __forceinline__ double wrap_freq(int k, int n) { return (double)(k > n / 2 ? k - n : k); }

// divergence contribution and (with do_gamma) the Gamma convolution of one
// frequency (kx, ky, kz), P in place; dc: the frequency is the DC term
__device__ __forceinline__ double gamma_div_point(double2* P, int kx, int ky, int kz, bool dc, const GammaArgs& g) {
  const double TAU = 6.283185307179586;
  double a[3];
  a[0] = TAU * (double)kx / g.size[0];
  a[1] = TAU * wrap_freq(ky, g.n[1]) / g.size[1];
  a[2] = TAU * wrap_freq(kz, g.n[2]) / g.size[2];
  // first-derivative frequencies: Nyquist set to zero on even grids
  if (g.n[0] % 2 == 0 && 2 * kx == g.n[0]) a[0] = 0.0;
  if (g.n[1] % 2 == 0 && 2 * ky == g.n[1]) a[1] = 0.0;
  if (g.n[2] % 2 == 0 && 2 * kz == g.n[2]) a[2] = 0.0;
  double contrib = 0.0;
  if (g.do_div) {
    // |P_hat (conj(-xi) s)|^2 with conj(-xi) = i a
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < 3; l++) {
      double2 u = make_double2(0.0, 0.0);
#pragma unroll
      for (int m = 0; m < 3; m++) {
        double w = a[m] * g.rg[m];
        u.x = fma(P[l + 3 * m].x, w, u.x);
        u.y = fma(P[l + 3 * m].y, w, u.y);
      }
      s = fma(u.x, u.x, fma(u.y, u.y, s));
    }
    double wgt;
    if (g.h == 1) wgt = 2.0;
    else wgt = (kx >= 1 && kx <= g.h - 2) ? 2.0 : 1.0;
    contrib = wgt * s;
  }
  if (g.do_gamma) {
    if (dc) {
#pragma unroll
      for (int c = 0; c < 9; c++) P[c] = make_double2(g.aim_dev ? g.aim_dev[c] : g.aim[c], 0.0);
    } else {
      double aa[6] = {a[0] * a[0], a[1] * a[1], a[2] * a[2], a[1] * a[2], a[0] * a[2], a[0] * a[1]};
      double A[9];
#pragma unroll
      for (int q = 0; q < 9; q++) {
        double s = 0.0;
#pragma unroll
        for (int p = 0; p < 6; p++) s = fma(g.Ccoef[q][p], aa[p], s);
        A[q] = s;                      // A_lm, q = 3 l + m
      }
      double detA = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6])
                  + A[2] * (A[3] * A[7] - A[4] * A[6]);
      if (fabs(detA) > 1.0e-16) {
        double inv[9];
        double r = 1.0 / detA;
        inv[0] = (A[4] * A[8] - A[5] * A[7]) * r;
        inv[1] = (A[2] * A[7] - A[1] * A[8]) * r;
        inv[2] = (A[1] * A[5] - A[2] * A[4]) * r;
        inv[3] = (A[5] * A[6] - A[3] * A[8]) * r;
        inv[4] = (A[0] * A[8] - A[2] * A[6]) * r;
        inv[5] = (A[2] * A[3] - A[0] * A[5]) * r;
        inv[6] = (A[3] * A[7] - A[4] * A[6]) * r;
        inv[7] = (A[1] * A[6] - A[0] * A[7]) * r;
        inv[8] = (A[0] * A[4] - A[1] * A[3]) * r;
        double2 v[3], w[3];
#pragma unroll
        for (int n = 0; n < 3; n++) {
          double2 s = make_double2(0.0, 0.0);
#pragma unroll
          for (int o2 = 0; o2 < 3; o2++) {
            s.x = fma(P[n + 3 * o2].x, a[o2], s.x);
            s.y = fma(P[n + 3 * o2].y, a[o2], s.y);
          }
          v[n] = s;
        }
#pragma unroll
        for (int l = 0; l < 3; l++) {
          double2 s = make_double2(0.0, 0.0);
#pragma unroll
          for (int n = 0; n < 3; n++) {
            s.x = fma(inv[3 * l + n], v[n].x, s.x);
            s.y = fma(inv[3 * l + n], v[n].y, s.y);
          }
          w[l] = s;
        }
#pragma unroll
        for (int l = 0; l < 3; l++)
#pragma unroll
          for (int m = 0; m < 3; m++) P[l + 3 * m] = make_double2(w[l].x * a[m], w[l].y * a[m]);
      } else {
#pragma unroll
        for (int c = 0; c < 9; c++) P[c] = make_double2(0.0, 0.0);
      }
    }
  }
  return contrib;
}

__global__ void __launch_bounds__(256) gamma_div_kernel(double2* __restrict__ Z, GammaArgs g, double* __restrict__ div_partial) {
  const long Nk = (long)g.h * g.n[1] * g.n[2];
  const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  if (f < Nk) {
    int kx = (int)(f % g.h);
    long r = f / g.h;
    int ky = (int)(r % g.n[1]);
    int kz = (int)(r / g.n[1]);
    double2 P[9];
#pragma unroll
    for (int c = 0; c < 9; c++) P[c] = Z[c * Nk + f];
    contrib = gamma_div_point(P, kx, ky, kz, f == 0, g);
    if (g.do_gamma) {
#pragma unroll
      for (int c = 0; c < 9; c++) Z[c * Nk + f] = P[c];
    }
  }
  if (g.do_div) {
    __shared__ double red[256];
    red[threadIdx.x] = contrib;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) div_partial[blockIdx.x] = red[0];
  }
}

// z pass, Gamma convolution with the divergence RMS, and inverse z pass of
// the nine fields in one kernel.  A block holds gz consecutive columns
// m = kx + h ky (all z) of the nine fields in LDS, as lines (field, column)
// of length N3: each frequency then has its nine components at hand, and the
// spectrum makes one round trip through HBM instead of three.
__device__ __forceinline__ void zgz_body(double2* __restrict__ Z, const GammaArgs& g, const FftPlan& pz,
                                         const double2* __restrict__ twz, int gz, double* __restrict__ div_partial,
                                         unsigned blk) {
  extern __shared__ double2 lds[];
  __shared__ double red[1024];
  const int h = g.h, N3 = g.n[2], ldl = N3 + 1;
  const long M = (long)h * g.n[1], Nk = M * N3;
  const long m0 = (long)blk * gz;
  const int nl = (int)min((long)gz, M - m0);
  const int nlines = 9 * gz;
  double2* b0 = lds;
  double2* b1 = lds + nlines * ldl;
  for (int i = threadIdx.x; i < 9 * N3 * gz; i += blockDim.x) {
    int j = i % gz, t = i / gz, kz = t % N3, c = t / N3;
    b0[(c * gz + j) * ldl + kz] = j < nl ? Z[c * Nk + (long)kz * M + m0 + j] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  double2* r = lds_fft(b0, b1, nlines, N3, ldl, pz, twz, -1);
  double2* o = r == b0 ? b1 : b0;
  double contrib = 0.0;
  for (int i = threadIdx.x; i < gz * N3; i += blockDim.x) {
    int j = i % gz, kz = i / gz;
    if (j >= nl) continue;
    const long m = m0 + j;
    double2 P[9];
#pragma unroll
    for (int c = 0; c < 9; c++) P[c] = r[(c * gz + j) * ldl + kz];
    contrib += gamma_div_point(P, (int)(m % h), (int)(m / h), kz, m == 0 && kz == 0, g);
    if (g.do_gamma) {
#pragma unroll
      for (int c = 0; c < 9; c++) r[(c * gz + j) * ldl + kz] = P[c];
    }
  }
  if (g.do_div) {
    red[threadIdx.x] = contrib;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
      if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) div_partial[blk] = red[0];
  }
  __syncthreads();
  double2* q = lds_fft(r, o, nlines, N3, ldl, pz, twz, +1);
  for (int i = threadIdx.x; i < 9 * N3 * gz; i += blockDim.x) {
    int j = i % gz, t = i / gz, kz = t % N3, c = t / N3;
    if (j < nl) Z[c * Nk + (long)kz * M + m0 + j] = q[(c * gz + j) * ldl + kz];
  }
}

__global__ void fft_zgz(double2* __restrict__ Z, GammaArgs g, FftPlan pz, const double2* __restrict__ twz, int gz,
                        double* __restrict__ div_partial) {
  zgz_body(Z, g, pz, twz, gz, div_partial, xcd_block(blockIdx.x, gridDim.x));
}

// fft_zgz with one LDS buffer (z plans of radix 2, 4 and 8: the in-place
// Stockham stages of the fused x-y kernels), so twice the columns fit per block
// and a block reads 16 gz contiguous bytes per (field, z) row.  Same arithmetic
// as fft_zgz.
__device__ __forceinline__ void zgz_ip_body(double2* __restrict__ Z, const GammaArgs& g, const FftPlan& pz,
                                            const double2* __restrict__ twz, int gz, double* __restrict__ div_partial,
                                            unsigned blk) {
  extern __shared__ double2 lds[];
  __shared__ double red[512];
  const int h = g.h, N3 = g.n[2], ldl = N3 + 1;
  const long M = (long)h * g.n[1], Nk = M * N3;
  const long m0 = (long)blk * gz;
  const int nl = (int)min((long)gz, M - m0);
  const int nlines = 9 * gz;
  double2* A = lds;
  for (int i = threadIdx.x; i < 9 * N3 * gz; i += blockDim.x) {
    int j = i % gz, t = i / gz, kz = t % N3, c = t / N3;
    A[(c * gz + j) * ldl + kz] = j < nl ? Z[c * Nk + (long)kz * M + m0 + j] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  lds_fft_ip(A, nlines, N3, ldl, pz, twz, -1);
  double contrib = 0.0;
  for (int i = threadIdx.x; i < gz * N3; i += blockDim.x) {
    int j = i % gz, kz = i / gz;
    if (j >= nl) continue;
    const long m = m0 + j;
    double2 P[9];
#pragma unroll
    for (int c = 0; c < 9; c++) P[c] = A[(c * gz + j) * ldl + kz];
    contrib += gamma_div_point(P, (int)(m % h), (int)(m / h), kz, m == 0 && kz == 0, g);
    if (g.do_gamma) {
#pragma unroll
      for (int c = 0; c < 9; c++) A[(c * gz + j) * ldl + kz] = P[c];
    }
  }
  if (g.do_div) {
    red[threadIdx.x] = contrib;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
      if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) div_partial[blk] = red[0];
  }
  __syncthreads();
  lds_fft_ip(A, nlines, N3, ldl, pz, twz, +1);
  for (int i = threadIdx.x; i < 9 * N3 * gz; i += blockDim.x) {
    int j = i % gz, t = i / gz, kz = t % N3, c = t / N3;
    if (j < nl) Z[c * Nk + (long)kz * M + m0 + j] = A[(c * gz + j) * ldl + kz];
  }
}

__global__ void __launch_bounds__(512) fft_zgz_ip(double2* __restrict__ Z, GammaArgs g, FftPlan pz,
                                                  const double2* __restrict__ twz, int gz,
                                                  double* __restrict__ div_partial) {
  zgz_ip_body(Z, g, pz, twz, gz, div_partial, xcd_block(blockIdx.x, gridDim.x));
}

// The transforms of a lockstep round: the fused kernels above over the RVEs
// of one FFT configuration (blockIdx.y = RVE), each with its own fields and
// Gamma arguments
__global__ __launch_bounds__(512) void fft_xy_r2c_ip_batch(const DmkResidDev* __restrict__ R, int N1, int N2, int N3,
                                                            FftPlan px, const double2* __restrict__ twx, FftPlan py,
                                                            const double2* __restrict__ twy) {
  const DmkResidDev& q = R[blockIdx.y];
  xy_r2c_ip_body(q.Pnew, q.Z, N1, N2, N3, px, twx, py, twy, blockIdx.x);
}

__global__ void fft_xy_r2c_batch(const DmkResidDev* __restrict__ R, int N1, int N2, int N3, FftPlan px,
                                 const double2* __restrict__ twx, FftPlan py, const double2* __restrict__ twy) {
  const DmkResidDev& q = R[blockIdx.y];
  xy_r2c_body(q.Pnew, q.Z, N1, N2, N3, px, twx, py, twy, blockIdx.x);
}

__global__ __launch_bounds__(512) void fft_yx_c2r_ip_batch(const DmkResidDev* __restrict__ R, int N1, int N2, int N3,
                                                            FftPlan px, const double2* __restrict__ twx, FftPlan py,
                                                            const double2* __restrict__ twy) {
  const DmkResidDev& q = R[blockIdx.y];
  yx_c2r_ip_body(q.Z, q.r, N1, N2, N3, px, twx, py, twy, blockIdx.x);
}

__global__ void fft_yx_c2r_batch(const DmkResidDev* __restrict__ R, int N1, int N2, int N3, FftPlan px,
                                 const double2* __restrict__ twx, FftPlan py, const double2* __restrict__ twy) {
  const DmkResidDev& q = R[blockIdx.y];
  yx_c2r_body(q.Z, q.r, N1, N2, N3, px, twx, py, twy, blockIdx.x);
}

__global__ void __launch_bounds__(512) fft_zgz_ip_batch(const DmkResidDev* __restrict__ R, FftPlan pz,
                                                        const double2* __restrict__ twz, int gz) {
  const DmkResidDev& q = R[blockIdx.y];
  zgz_ip_body(q.Z, q.g, pz, twz, gz, q.divp, xcd_block(blockIdx.x, gridDim.x));
}

__global__ void fft_zgz_batch(const DmkResidDev* __restrict__ R, FftPlan pz, const double2* __restrict__ twz,
                              int gz) {
  const DmkResidDev& q = R[blockIdx.y];
  zgz_body(q.Z, q.g, pz, twz, gz, q.divp, xcd_block(blockIdx.x, gridDim.x));
}

// utilities_curlRMS (spectral_utilities.f90:697-757): per frequency the
// squared curl of F_hat with xi = i a (rescaled like the divergence), interior
// x planes counted twice.  On the x = 0 and x = N1/2 planes the reference
// assigns each curl expression to the whole 3x3 array, so what it sums there
// is 9 |F_hat(3,2) xi_1 - F_hat(3,1) xi_2|^2 (the last assignment); this kernel
// sums the same.  For N1 = 1 that plane is visited twice.
__global__ void __launch_bounds__(256) curl_kernel(const double2* __restrict__ Z, GammaArgs g,
                                                   double* __restrict__ partial) {
  const long Nk = (long)g.h * g.n[1] * g.n[2];
  const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  if (f < Nk) {
    int kx = (int)(f % g.h);
    long r = f / g.h;
    int ky = (int)(r % g.n[1]);
    int kz = (int)(r / g.n[1]);
    const double TAU = 6.283185307179586;
    double w[3];
    w[0] = TAU * (double)kx / g.size[0];
    w[1] = TAU * wrap_freq(ky, g.n[1]) / g.size[1];
    w[2] = TAU * wrap_freq(kz, g.n[2]) / g.size[2];
    if (g.n[0] % 2 == 0 && 2 * kx == g.n[0]) w[0] = 0.0;
    if (g.n[1] % 2 == 0 && 2 * ky == g.n[1]) w[1] = 0.0;
    if (g.n[2] % 2 == 0 && 2 * kz == g.n[2]) w[2] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) w[d] *= g.rg[d];
    auto cdiff = [](double2 p, double wp, double2 q, double wq) {
      double re = p.x * wp - q.x * wq, im = p.y * wp - q.y * wq;
      return re * re + im * im;
    };
    if (kx >= 1 && kx <= g.h - 2) {
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < 3; l++) {
        double2 F0 = Z[(l + 0) * Nk + f], F1 = Z[(l + 3) * Nk + f], F2 = Z[(l + 6) * Nk + f];
        s += cdiff(F2, w[1], F1, w[2]) + cdiff(F0, w[2], F2, w[0]) + cdiff(F1, w[0], F0, w[1]);
      }
      contrib = 2.0 * s;
    } else {
      double2 F20 = Z[2 * Nk + f], F21 = Z[5 * Nk + f];
      double s = 9.0 * cdiff(F21, w[0], F20, w[1]);
      contrib = g.h == 1 ? 2.0 * s : s;
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = contrib;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// utilities_G_Convolution (spectral_utilities.f90:544-589) with G_hat of
// G_hat_init (:507-533): G_lmno = delta_ln conj(-xi_m) xi_o / |xi|^2, i.e. with
// xi = i a, (G:F_hat)_lm = -a_m (F_hat a)_l / |a|^2 (zero where |a|^2 <= 1e-16).
// DC term: residual mode sets it to the aim (dc_dev * dc_scale, the reference's
// fieldAim / wgt), Jacobian mode keeps the transform's own; both then zero the
// components whose stress is not prescribed.  The optional RMS divergence reads
// the transform before it is overwritten (same sum as gamma_div_kernel).
__global__ void __launch_bounds__(256) galerkin_kernel(double2* __restrict__ Z, GammaArgs g, GalerkinArgs ga,
                                                       double* __restrict__ div_partial) {
  const long Nk = (long)g.h * g.n[1] * g.n[2];
  const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  if (f < Nk) {
    int kx = (int)(f % g.h);
    long r = f / g.h;
    int ky = (int)(r % g.n[1]);
    int kz = (int)(r / g.n[1]);
    const double TAU = 6.283185307179586;
    double a[3];
    a[0] = TAU * (double)kx / g.size[0];
    a[1] = TAU * wrap_freq(ky, g.n[1]) / g.size[1];
    a[2] = TAU * wrap_freq(kz, g.n[2]) / g.size[2];
    if (g.n[0] % 2 == 0 && 2 * kx == g.n[0]) a[0] = 0.0;
    if (g.n[1] % 2 == 0 && 2 * ky == g.n[1]) a[1] = 0.0;
    if (g.n[2] % 2 == 0 && 2 * kz == g.n[2]) a[2] = 0.0;
    double2 P[9];
#pragma unroll
    for (int c = 0; c < 9; c++) P[c] = Z[c * Nk + f];
    if (ga.do_div) {
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < 3; l++) {
        double2 u = make_double2(0.0, 0.0);
#pragma unroll
        for (int m = 0; m < 3; m++) {
          double w = a[m] * g.rg[m];
          u.x = fma(P[l + 3 * m].x, w, u.x);
          u.y = fma(P[l + 3 * m].y, w, u.y);
        }
        s = fma(u.x, u.x, fma(u.y, u.y, s));
      }
      double wgt;
      if (g.h == 1) wgt = 2.0;
      else wgt = (kx >= 1 && kx <= g.h - 2) ? 2.0 : 1.0;
      contrib = wgt * s;
    }
    if (f == 0) {
#pragma unroll
      for (int c = 0; c < 9; c++) {
        double2 v = ga.residual ? make_double2(ga.dc_dev[c] * ga.dc_scale, 0.0) : P[c];
        Z[c * Nk] = ga.stress_mask[c] ? make_double2(0.0, 0.0) : v;
      }
    } else {
      const double n2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
      if (n2 > 1.0e-16) {
#pragma unroll
        for (int l = 0; l < 3; l++) {
          double2 v = make_double2(0.0, 0.0);
#pragma unroll
          for (int o = 0; o < 3; o++) {
            v.x = fma(P[l + 3 * o].x, a[o], v.x);
            v.y = fma(P[l + 3 * o].y, a[o], v.y);
          }
#pragma unroll
          for (int m = 0; m < 3; m++) {
            double w = -a[m] / n2;
            Z[(l + 3 * m) * Nk + f] = make_double2(v.x * w, v.y * w);
          }
        }
      } else {
#pragma unroll
        for (int c = 0; c < 9; c++) Z[c * Nk + f] = make_double2(0.0, 0.0);
      }
    }
  }
  if (ga.do_div) {
    __shared__ double red[256];
    red[threadIdx.x] = contrib;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) div_partial[blockIdx.x] = red[0];
  }
}

// updateCoords: per frequency u_hat = F_hat xi / (conj(-xi) . xi), xi second-derivative set
__global__ void coords_freq_kernel(const double2* __restrict__ Fh, double2* __restrict__ U, GammaArgs g) {
  const long Nk = (long)g.h * g.n[1] * g.n[2];
  const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= Nk) return;
  int kx = (int)(f % g.h);
  long r = f / g.h;
  int ky = (int)(r % g.n[1]);
  int kz = (int)(r / g.n[1]);
  const double TAU = 6.283185307179586;
  double a[3] = {TAU * (double)kx / g.size[0], TAU * wrap_freq(ky, g.n[1]) / g.size[1],
                 TAU * wrap_freq(kz, g.n[2]) / g.size[2]};
  if (f == 0) {
    for (int l = 0; l < 3; l++) U[l * Nk] = make_double2(0.0, 0.0);
    return;
  }
  // xi = i a: F_hat xi = i (F_hat a);  conj(-xi).xi = -|a|^2
  double den = -(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  for (int l = 0; l < 3; l++) {
    double2 s = make_double2(0.0, 0.0);
    for (int m = 0; m < 3; m++) {
      double2 p = Fh[(l + 3 * m) * Nk + f];
      s.x = fma(p.x, a[m], s.x);
      s.y = fma(p.y, a[m], s.y);
    }
    // i s / den
    U[l * Nk + f] = make_double2(-s.y / den, s.x / den);
  }
}

__global__ void coords_nodes_kernel(const double* __restrict__ u, CoordsArgs c, double* __restrict__ x_n,
                                    double* __restrict__ x_p) {
  const int N1 = c.n[0], N2 = c.n[1], N3 = c.n[2];
  const long N = (long)N1 * N2 * N3;
  const long nn = (long)(N1 + 1) * (N2 + 1) * (N3 + 1);
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  // mean F from the DC terms of the forward transform (component c at Fdc[c * Nk])
  double Favg[9];
  for (int q = 0; q < 9; q++) Favg[q] = c.Fdc[q * c.Nk].x * c.wgt;
  if (t < nn) {
    int i = (int)(t % (N1 + 1));
    long r = t / (N1 + 1);
    int j = (int)(r % (N2 + 1));
    int k = (int)(r / (N2 + 1));
    double pos[3] = {c.step[0] * i, c.step[1] * j, c.step[2] * k};
    double x[3];
    for (int a = 0; a < 3; a++) x[a] = Favg[a] * pos[0] + Favg[a + 3] * pos[1] + Favg[a + 6] * pos[2];
    for (int nb = 0; nb < 8; nb++) {
      int dx = (nb == 1 || nb == 2 || nb == 5 || nb == 6), dy = (nb == 2 || nb == 3 || nb == 6 || nb == 7), dz = nb >= 4;
      int mx = (i + dx - 1 + N1) % N1, my = (j + dy - 1 + N2) % N2, mz = (k + dz - 1 + N3) % N3;
      long ci = mx + (long)N1 * (my + (long)N2 * mz);
      for (int a = 0; a < 3; a++) x[a] = x[a] + u[a * N + ci] * c.wgt * 0.125;
    }
    for (int a = 0; a < 3; a++) x_n[3 * t + a] = x[a];
  }
  if (t < N) {
    int i = (int)(t % N1);
    long r = t / N1;
    int j = (int)(r % N2);
    int k = (int)(r / N2);
    double pos[3] = {c.step[0] * (i + 1 - 0.5), c.step[1] * (j + 1 - 0.5), c.step[2] * (k + 1 - 0.5)};
    for (int a = 0; a < 3; a++)
      x_p[3 * t + a] = u[a * N + t] * c.wgt + (Favg[a] * pos[0] + Favg[a + 3] * pos[1] + Favg[a + 6] * pos[2]);
  }
}

}  // namespace

// ---------------------------------------------------------------------------- host
static FftPlan make_plan(int L) {
  FftPlan p{};
  p.L = L;
  int n = L;
  while (n % 8 == 0 && n > 1) { p.radix[p.nstages++] = 8; n /= 8; }
  while (n % 4 == 0 && n > 1) { p.radix[p.nstages++] = 4; n /= 4; }
  while (n % 2 == 0) { p.radix[p.nstages++] = 2; n /= 2; }
  for (int f = 3; n > 1; f += 2)
    while (n % f == 0) { p.radix[p.nstages++] = f; n /= f; }
  return p;
}

static int x_pairs_per_block(int L) { int p = 1024 / L; return p < 1 ? 1 : p; }
// lines per block of a strided pass covering h lines: the chunks are as even
// as possible (h = 33 in groups of 16 would leave a block with one line)
static int gx_even(int h, int gmax) {
  int nch = (h + gmax - 1) / gmax;
  return (h + nch - 1) / nch;
}
static int gx_for(int L) { int g = 1024 / L; if (g > 32) g = 32; return g < 1 ? 1 : g; }

// geomSize / scaledGeomSize of the divergence measure for the
// divergence_correction option dc (0 none, 1 grid+size, 2 size),
// spectral_utilities.f90:181-197: the median dimension (by size, or by size per
// cell for grid+size) scales the others; equal extremes pick the last candidate
void dmk_div_scale(const int cells[3], const double size[3], int dc, double rg[3]) {
  double scaled[3] = {size[0], size[1], size[2]};
  double r[3];
  for (int d = 0; d < 3; d++) r[d] = size[d] / cells[d];
  if (dc == 0 || dc == 1) {
    const double* v = dc == 1 ? r : size;
    int imin = 0, imax = 0;
    for (int d = 1; d < 3; d++) { if (v[d] < v[imin]) imin = d; if (v[d] > v[imax]) imax = d; }
    for (int j = 0; j < 3; j++)
      if (j != imin && j != imax)
        for (int d = 0; d < 3; d++) scaled[d] = dc == 1 ? size[d] / size[j] * (double)cells[j] : size[d] / size[j];
  }
  for (int d = 0; d < 3; d++) rg[d] = size[d] / scaled[d];
}

int dmk_fft_setup(DmkFFT* fft, const int cells[3]) {
  for (int d = 0; d < 3; d++) {
    fft->n[d] = cells[d];
    fft->plan[d] = make_plan(cells[d]);
    if (fft->plan[d].nstages > 16) return 1;
    int L = cells[d];
    double2* host = (double2*)malloc(sizeof(double2) * L);
    for (int k = 0; k < L; k++) {
      long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)L;
      host[k] = make_double2((double)cosl(ang), (double)sinl(ang));
    }
    if (hipMalloc(&fft->tw[d], sizeof(double2) * L) != hipSuccess) { free(host); return 2; }
    if (hipMemcpy(fft->tw[d], host, sizeof(double2) * L, hipMemcpyHostToDevice) != hipSuccess) { free(host); return 2; }
    free(host);
  }
  fft->h = cells[0] / 2 + 1;
  fft->gx[0] = x_pairs_per_block(cells[0]);
  for (int d = 1; d < 3; d++) fft->gx[d] = gx_even(fft->h, gx_for(cells[d]));
  return 0;
}

// twiddles exp(-2 pi i k / L), k < L, of the three axes into caller memory
__global__ void twiddle_kernel(double2* __restrict__ tw, int L0, int L1, int L2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int L, k;
  if (i < L0) { L = L0; k = i; }
  else if (i < L0 + L1) { L = L1; k = i - L0; }
  else if (i < L0 + L1 + L2) { L = L2; k = i - L0 - L1; }
  else return;
  double sn, cs;
  sincospi(-2.0 * (double)k / (double)L, &sn, &cs);
  tw[i] = make_double2(cs, sn);
}

int dmk_fft_setup_into(DmkFFT* fft, const int cells[3], double2* tw, hipStream_t s) {
  for (int d = 0; d < 3; d++) {
    fft->n[d] = cells[d];
    fft->plan[d] = make_plan(cells[d]);
    if (fft->plan[d].nstages > 16) return 1;
  }
  fft->tw[0] = tw;
  fft->tw[1] = tw + cells[0];
  fft->tw[2] = tw + cells[0] + cells[1];
  const int tot = cells[0] + cells[1] + cells[2];
  hipLaunchKernelGGL(twiddle_kernel, dim3((tot + 255) / 256), dim3(256), 0, s, tw, cells[0], cells[1], cells[2]);
  fft->h = cells[0] / 2 + 1;
  fft->gx[0] = x_pairs_per_block(cells[0]);
  for (int d = 1; d < 3; d++) fft->gx[d] = gx_even(fft->h, gx_for(cells[d]));
  fft->xy_threads = 0;
  fft->xy_ip = 0;
  fft->zgz_g = 0;
  fft->zgz_threads = 0;
  fft->zgz_ip = 0;
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

void dmk_fft_free(DmkFFT* fft) {
  for (int d = 0; d < 3; d++) if (fft->tw[d]) hipFree(fft->tw[d]);
}

static const size_t FFT_LDS_MAX = 64 * 1024;
// the fused x-y kernels may take more (the MI355X allows 160 KB per block)
static const size_t FFT_XY_LDS_MAX = 160 * 1024;

static size_t xy_lds(const DmkFFT* fft) {
  const int N1 = fft->n[0], N2 = fft->n[1], h = fft->h, np = (N2 + 1) / 2;
  size_t a = (size_t)np * (N1 + 1), b = (size_t)h * (N2 + 1);
  return 2 * sizeof(double2) * (a > b ? a : b);
}

static void launch_xy_r2c(const DmkFFT* fft, int howmany, const double* in, double2* out, int threads, int ip,
                          hipStream_t s) {
  if (ip)
    hipLaunchKernelGGL(fft_xy_r2c_ip, dim3(howmany * fft->n[2]), dim3(threads), xy_lds(fft) / 2, s, in, out,
                       fft->n[0], fft->n[1], fft->n[2], fft->plan[0], fft->tw[0], fft->plan[1], fft->tw[1]);
  else
    hipLaunchKernelGGL(fft_xy_r2c, dim3(howmany * fft->n[2]), dim3(threads), xy_lds(fft), s, in, out, fft->n[0],
                       fft->n[1], fft->n[2], fft->plan[0], fft->tw[0], fft->plan[1], fft->tw[1]);
}

static void launch_yx_c2r(const DmkFFT* fft, int howmany, const double2* in, double* out, int threads, int ip,
                          hipStream_t s) {
  if (ip)
    hipLaunchKernelGGL(fft_yx_c2r_ip, dim3(howmany * fft->n[2]), dim3(threads), xy_lds(fft) / 2, s, in, out,
                       fft->n[0], fft->n[1], fft->n[2], fft->plan[0], fft->tw[0], fft->plan[1], fft->tw[1]);
  else
    hipLaunchKernelGGL(fft_yx_c2r, dim3(howmany * fft->n[2]), dim3(threads), xy_lds(fft), s, in, out, fft->n[0],
                       fft->n[1], fft->n[2], fft->plan[0], fft->tw[0], fft->plan[1], fft->tw[1]);
}

// the in-place x-y kernels need radix 2/4/8 plans and every stage, unpack and
// pack to fit the per-thread register slots at this block size
static bool xy_ip_fits(const DmkFFT* fft, int threads) {
  const int N1 = fft->n[0], N2 = fft->n[1], h = fft->h, np = (N2 + 1) / 2;
  for (int d = 0; d < 2; d++) {
    const FftPlan& p = fft->plan[d];
    const int lines = d == 0 ? np : h, L = fft->n[d];
    for (int st = 0; st < p.nstages; st++) {
      const int R = p.radix[st];
      if (R != 2 && R != 4 && R != 8) return false;
      if ((long)lines * (L / R) > (long)IP_KMAX * threads) return false;
    }
  }
  return (long)N2 * h <= (long)IP_KU * threads && (long)np * N1 <= (long)IP_KU * threads;
}

static void launch_x_r2c(const DmkFFT* fft, int howmany, const double* in, double2* out, int ppb, hipStream_t s) {
  const int N1 = fft->n[0];
  int nl = howmany * fft->n[1] * fft->n[2];
  int nblk = (nl + 2 * ppb - 1) / (2 * ppb);
  size_t lds = sizeof(double2) * 2 * ppb * (N1 + 1);
  hipLaunchKernelGGL(fft_x_r2c, dim3(nblk), dim3(256), lds, s, in, out, nl, fft->plan[0], fft->tw[0], ppb);
}

static void launch_x_c2r(const DmkFFT* fft, int howmany, const double2* in, double* out, int ppb, hipStream_t s) {
  const int N1 = fft->n[0];
  int nl = howmany * fft->n[1] * fft->n[2];
  int nblk = (nl + 2 * ppb - 1) / (2 * ppb);
  size_t lds = sizeof(double2) * 2 * ppb * (N1 + 1);
  hipLaunchKernelGGL(fft_x_c2r, dim3(nblk), dim3(256), lds, s, in, out, nl, fft->plan[0], fft->tw[0], ppb);
}

// y (axis 1) or z (axis 2) pass over the half spectra, in place
static void launch_strided(const DmkFFT* fft, int howmany, double2* data, int axis, int gx, int sign, hipStream_t s) {
  const int N2 = fft->n[1], N3 = fft->n[2], h = fft->h;
  const long Nk = (long)h * N2 * N3;
  const int L = fft->n[axis];
  int nch = (h + gx - 1) / gx;
  size_t lds = sizeof(double2) * 2 * gx * (L + 1);
  if (axis == 1)
    hipLaunchKernelGGL(fft_strided, dim3(howmany * N3 * nch), dim3(256), lds, s, data, h, gx, N3, Nk, (long)h * N2,
                       (long)h, fft->plan[1], fft->tw[1], sign);
  else
    hipLaunchKernelGGL(fft_strided, dim3(howmany * N2 * nch), dim3(256), lds, s, data, h, gx, N2, Nk, (long)h,
                       (long)h * N2, fft->plan[2], fft->tw[2], sign);
}

static size_t zgz_lds(const DmkFFT* fft, int gz, int ip) {
  return (ip ? 1 : 2) * sizeof(double2) * 9 * gz * (fft->n[2] + 1);
}

// the single-buffer z kernel needs a radix 2/4/8 z plan whose every stage fits
// the per-thread butterfly slots at this block size
static bool zgz_ip_fits(const DmkFFT* fft, int gz, int threads) {
  const FftPlan& p = fft->plan[2];
  for (int st = 0; st < p.nstages; st++) {
    const int R = p.radix[st];
    if (R != 2 && R != 4 && R != 8) return false;
    if ((long)9 * gz * (fft->n[2] / R) > (long)IP_KMAX * threads) return false;
  }
  return true;
}

static void launch_zgz(const DmkFFT* fft, double2* Z, const GammaArgs* g, int gz, int threads, int ip,
                       double* div_partial, hipStream_t s) {
  const long M = (long)fft->h * fft->n[1];
  unsigned nb = (unsigned)((M + gz - 1) / gz);
  if (ip)
    hipLaunchKernelGGL(fft_zgz_ip, dim3(nb), dim3(threads), zgz_lds(fft, gz, 1), s, Z, *g, fft->plan[2],
                       fft->tw[2], gz, div_partial);
  else
    hipLaunchKernelGGL(fft_zgz, dim3(nb), dim3(threads), zgz_lds(fft, gz, 0), s, Z, *g, fft->plan[2], fft->tw[2],
                       gz, div_partial);
}

int dmk_fft_tune(DmkFFT* fft, double2* cscratch, double* rscratch, hipStream_t s) {
  const int howmany = 9;
  const long Nk = (long)fft->h * fft->n[1] * fft->n[2], N = (long)fft->n[0] * fft->n[1] * fft->n[2];
  if (hipMemsetAsync(cscratch, 0, sizeof(double2) * howmany * Nk, s) != hipSuccess) return 2;
  if (hipMemsetAsync(rscratch, 0, sizeof(double) * howmany * N, s) != hipSuccess) return 2;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 2;
  auto timed = [&](auto&& launch) {
    launch();                                    // warm-up
    hipEventRecord(e0, s);
    for (int r = 0; r < 3; r++) launch();
    hipEventRecord(e1, s);
    hipEventSynchronize(e1);
    float ms = 0.0f;
    hipEventElapsedTime(&ms, e0, e1);
    return ms;
  };
  // x pass: line pairs per block
  {
    const int L = fft->n[0];
    float best = 1e30f;
    for (int ppb : {1024 / L, 512 / L, 256 / L, 2048 / L}) {
      if (ppb < 1 || sizeof(double2) * 2 * ppb * (L + 1) > FFT_LDS_MAX) continue;
      float t = timed([&] { launch_x_r2c(fft, howmany, rscratch, cscratch, ppb, s); });
      if (t < best) { best = t; fft->gx[0] = ppb; }
    }
  }
  // strided passes: even chunkings of the h lines
  float t_y = 0.0f;
  for (int axis = 1; axis <= 2; axis++) {
    const int L = fft->n[axis];
    if (L <= 1) continue;
    float best = 1e30f;
    int last = -1;
    for (int nch = 1; nch <= fft->h && nch <= 16; nch++) {
      int gx = (fft->h + nch - 1) / nch;
      if (gx == last || sizeof(double2) * 2 * gx * (L + 1) > FFT_LDS_MAX) continue;
      last = gx;
      float t = timed([&] { launch_strided(fft, howmany, cscratch, axis, gx, -1, s); });
      if (t < best) { best = t; fft->gx[axis] = gx; }
    }
    if (axis == 1) t_y = best;
  }
  // fused x-y passes against the separate ones (forward direction);
  // DMK_FFT_XY=0 / 1 in the environment forces the separate / fused passes
  fft->xy_threads = 0;
  fft->xy_ip = 0;
  const char* force = getenv("DMK_FFT_XY");
  if (fft->n[0] >= 2 && fft->n[1] >= 2 && xy_lds(fft) <= FFT_XY_LDS_MAX && !(force && force[0] == '0')) {
    float sep = timed([&] { launch_x_r2c(fft, howmany, rscratch, cscratch, fft->gx[0], s); }) + t_y;
    if (force && (force[0] == '1' || force[0] == '2')) sep = 1e30f;
    for (int ip = 0; ip <= 1; ip++)
      for (int th : {256, 512, 1024}) {
        if (ip && (th > 512 || !xy_ip_fits(fft, th))) continue;   // launch bounds of the in-place kernels
        if (force && force[0] == '1' && ip) continue;       // DMK_FFT_XY=1: the two-buffer variant
        if (force && force[0] == '2' && !ip) continue;      // DMK_FFT_XY=2: the in-place variant
        float t = timed([&] { launch_xy_r2c(fft, howmany, rscratch, cscratch, th, ip, s); });
        if (t < sep) { sep = t; fft->xy_threads = th; fft->xy_ip = ip; }
      }
  }
  // fused z passes and Gamma against the separate kernels (a reference medium
  // with an invertible acoustic tensor; DMK_FFT_ZGZ=0 / 1 / 2 forces separate /
  // fused / fused single-buffer where the z plan allows it)
  fft->zgz_g = 0;
  fft->zgz_ip = 0;
  const char* fz = getenv("DMK_FFT_ZGZ");
  if (fft->n[2] >= 2 && !(fz && fz[0] == '0')) {
    GammaArgs g{};
    for (int d = 0; d < 3; d++) { g.n[d] = fft->n[d]; g.size[d] = 1.0; g.rg[d] = 1.0; }
    g.h = fft->h;
    for (int q = 0; q < 9; q++) g.Ccoef[q][0] = g.Ccoef[q][1] = g.Ccoef[q][2] = (q % 4 == 0) ? 1.0 : 0.0;
    g.do_div = 1;
    g.do_gamma = 1;
    double* part = rscratch;      // per-block sums: rscratch holds 9 N >= h N2 doubles
    float sep = timed([&] {
      launch_strided(fft, howmany, cscratch, 2, fft->gx[2], -1, s);
      hipLaunchKernelGGL(gamma_div_kernel, dim3((unsigned)((Nk + 255) / 256)), dim3(256), 0, s, cscratch, g, part);
      launch_strided(fft, howmany, cscratch, 2, fft->gx[2], +1, s);
    });
    if (fz && (fz[0] == '1' || fz[0] == '2')) sep = 1e30f;
    bool any_ip = false;
    for (int gz : {2, 4, 8, 16})
      for (int th : {256, 512}) any_ip |= zgz_lds(fft, gz, 1) <= FFT_XY_LDS_MAX && zgz_ip_fits(fft, gz, th);
    for (int ip = 0; ip <= 1; ip++)
      for (int gz : {1, 2, 4, 8, 16}) {
        if (zgz_lds(fft, gz, ip) > FFT_XY_LDS_MAX) continue;
        if (fz && fz[0] == '1' && ip) continue;                 // DMK_FFT_ZGZ=1: the two-buffer kernel
        if (fz && fz[0] == '2' && any_ip && !ip) continue;      // DMK_FFT_ZGZ=2: the single-buffer kernel
        for (int th : {256, 512}) {
          if (ip && !zgz_ip_fits(fft, gz, th)) continue;
          float t = timed([&] { launch_zgz(fft, cscratch, &g, gz, th, ip, part, s); });
          if (t < sep) { sep = t; fft->zgz_g = gz; fft->zgz_threads = th; fft->zgz_ip = ip; }
        }
      }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_fft_r2c_xy(const DmkFFT* fft, int howmany, const double* in, double2* out, hipStream_t s) {
  if (fft->xy_threads) {
    launch_xy_r2c(fft, howmany, in, out, fft->xy_threads, fft->xy_ip, s);
  } else {
    launch_x_r2c(fft, howmany, in, out, fft->gx[0], s);
    if (fft->n[1] > 1) launch_strided(fft, howmany, out, 1, fft->gx[1], -1, s);
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_fft_c2r_xy(const DmkFFT* fft, int howmany, double2* in, double* out, hipStream_t s) {
  if (fft->xy_threads) {
    launch_yx_c2r(fft, howmany, in, out, fft->xy_threads, fft->xy_ip, s);
  } else {
    if (fft->n[1] > 1) launch_strided(fft, howmany, in, 1, fft->gx[1], +1, s);
    launch_x_c2r(fft, howmany, in, out, fft->gx[0], s);
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_transform_batch(const DmkFFT* fft, const DmkResidDev* R, int nr, hipStream_t s) {
  if (nr < 1) return 0;
  if (!fft->xy_threads || !fft->zgz_g) return 1;
  const int N1 = fft->n[0], N2 = fft->n[1], N3 = fft->n[2];
  const dim3 gxy(9 * N3, nr);
  const long M = (long)fft->h * N2;
  const dim3 gzz((unsigned)((M + fft->zgz_g - 1) / fft->zgz_g), nr);
  if (fft->xy_ip)
    hipLaunchKernelGGL(fft_xy_r2c_ip_batch, gxy, dim3(fft->xy_threads), xy_lds(fft) / 2, s, R, N1, N2, N3,
                       fft->plan[0], fft->tw[0], fft->plan[1], fft->tw[1]);
  else
    hipLaunchKernelGGL(fft_xy_r2c_batch, gxy, dim3(fft->xy_threads), xy_lds(fft), s, R, N1, N2, N3, fft->plan[0],
                       fft->tw[0], fft->plan[1], fft->tw[1]);
  if (fft->zgz_ip)
    hipLaunchKernelGGL(fft_zgz_ip_batch, gzz, dim3(fft->zgz_threads), zgz_lds(fft, fft->zgz_g, 1), s, R,
                       fft->plan[2], fft->tw[2], fft->zgz_g);
  else
    hipLaunchKernelGGL(fft_zgz_batch, gzz, dim3(fft->zgz_threads), zgz_lds(fft, fft->zgz_g, 0), s, R, fft->plan[2],
                       fft->tw[2], fft->zgz_g);
  if (fft->xy_ip)
    hipLaunchKernelGGL(fft_yx_c2r_ip_batch, gxy, dim3(fft->xy_threads), xy_lds(fft) / 2, s, R, N1, N2, N3,
                       fft->plan[0], fft->tw[0], fft->plan[1], fft->tw[1]);
  else
    hipLaunchKernelGGL(fft_yx_c2r_batch, gxy, dim3(fft->xy_threads), xy_lds(fft), s, R, N1, N2, N3, fft->plan[0],
                       fft->tw[0], fft->plan[1], fft->tw[1]);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_gamma_conv_nb(const DmkFFT* fft) {
  const long M = (long)fft->h * fft->n[1], Nk = M * fft->n[2];
  return (int)(fft->zgz_g ? (M + fft->zgz_g - 1) / fft->zgz_g : (Nk + 255) / 256);
}

long dmk_gamma_conv_partials(const DmkFFT* fft) {
  const long M = (long)fft->h * fft->n[1], Nk = M * fft->n[2];
  long a = (Nk + 255) / 256;
  return a > M ? a : M;
}

int dmk_gamma_conv(const DmkFFT* fft, double2* Z, const GammaArgs* g, double* div_partial, int* nb, hipStream_t s) {
  const long M = (long)fft->h * fft->n[1], Nk = M * fft->n[2];
  if (fft->zgz_g) {
    launch_zgz(fft, Z, g, fft->zgz_g, fft->zgz_threads, fft->zgz_ip, div_partial, s);
    *nb = (int)((M + fft->zgz_g - 1) / fft->zgz_g);
  } else {
    if (fft->n[2] > 1) launch_strided(fft, 9, Z, 2, fft->gx[2], -1, s);
    hipLaunchKernelGGL(gamma_div_kernel, dim3((unsigned)((Nk + 255) / 256)), dim3(256), 0, s, Z, *g, div_partial);
    if (fft->n[2] > 1) launch_strided(fft, 9, Z, 2, fft->gx[2], +1, s);
    *nb = (int)((Nk + 255) / 256);
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_fft_r2c(const DmkFFT* fft, int howmany, const double* in, double2* out, hipStream_t s) {
  if (fft->xy_threads) {
    launch_xy_r2c(fft, howmany, in, out, fft->xy_threads, fft->xy_ip, s);
  } else {
    launch_x_r2c(fft, howmany, in, out, fft->gx[0], s);
    if (fft->n[1] > 1) launch_strided(fft, howmany, out, 1, fft->gx[1], -1, s);
  }
  if (fft->n[2] > 1) launch_strided(fft, howmany, out, 2, fft->gx[2], -1, s);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

// in is overwritten (the y and z passes run in place)
int dmk_fft_c2r(const DmkFFT* fft, int howmany, double2* in, double* out, hipStream_t s) {
  if (fft->n[2] > 1) launch_strided(fft, howmany, in, 2, fft->gx[2], +1, s);
  if (fft->xy_threads) {
    launch_yx_c2r(fft, howmany, in, out, fft->xy_threads, fft->xy_ip, s);
  } else {
    if (fft->n[1] > 1) launch_strided(fft, howmany, in, 1, fft->gx[1], +1, s);
    launch_x_c2r(fft, howmany, in, out, fft->gx[0], s);
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_gamma_div(double2* Z, const GammaArgs* g, double* div_partial, hipStream_t s) {
  const long Nk = (long)g->h * g->n[1] * g->n[2];
  int nblk = (int)((Nk + 255) / 256);
  hipLaunchKernelGGL(gamma_div_kernel, dim3(nblk), dim3(256), 0, s, Z, *g, div_partial);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_gamma_div_blocks(const GammaArgs* g) {
  const long Nk = (long)g->h * g->n[1] * g->n[2];
  return (int)((Nk + 255) / 256);
}

int dmk_curl_partial(const double2* Z, const GammaArgs* g, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(curl_kernel, dim3(dmk_gamma_div_blocks(g)), dim3(256), 0, s, Z, *g, partial);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_galerkin(double2* Z, const GammaArgs* g, const GalerkinArgs* ga, double* div_partial, hipStream_t s) {
  hipLaunchKernelGGL(galerkin_kernel, dim3(dmk_gamma_div_blocks(g)), dim3(256), 0, s, Z, *g, *ga, div_partial);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_coords_freq(const double2* Fh, double2* U, const GammaArgs* g, hipStream_t s) {
  const long Nk = (long)g->h * g->n[1] * g->n[2];
  hipLaunchKernelGGL(coords_freq_kernel, dim3((unsigned)((Nk + 255) / 256)), dim3(256), 0, s, Fh, U, *g);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int dmk_coords_nodes(const double* u, const CoordsArgs* c, double* x_n, double* x_p, hipStream_t s) {
  long nn = (long)(c->n[0] + 1) * (c->n[1] + 1) * (c->n[2] + 1);
  hipLaunchKernelGGL(coords_nodes_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, u, *c, x_n, x_p);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
