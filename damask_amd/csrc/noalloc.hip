// utilities_divergenceRMS (src/grid/spectral_utilities.f90:650-691) of a
// device-resident field with no device allocation: dmk_divergence_rms_noalloc.
//
// Every buffer the transform and the reduction need (the component-major copy
// of the field, its half spectrum, the twiddle tables and the per-block sums)
// lives in one file-scope device array, `scratch`, of a fixed size sized for
// grids up to DMK_NOALLOC_MAX_CELLS cells.  Each GPU holds its own copy of the
// array from the first use of the library until the process ends; its device
// address is looked up once per device (hipGetSymbolAddress) and only that
// address is handed to kernels and copies.  Calls that overlap on one GPU share
// the array and race; the caller serialises them (one stream, or events).
#include <mutex>

#include "dmk_common.h"
#include "dmk_linalg.h"
#include "dmk_spectral.h"
#include "../../include/damask_hip.h"

namespace {

constexpr long MAXC = DMK_NOALLOC_MAX_CELLS;
constexpr int MAX_EDGE = 1024;                  // one axis (LDS of the 1-D passes)
// layout in doubles: field, component-major (9 N) | half spectrum (9 Nk complex,
// Nk <= N) | twiddles (N1 + N2 + N3 <= 3 MAX_EDGE complex) | per-block sums
constexpr long OFF_SOA = 0;
constexpr long OFF_Z = OFF_SOA + 9 * MAXC;
constexpr long OFF_TW = OFF_Z + 18 * MAXC;
constexpr long OFF_PART = OFF_TW + 2 * 3 * MAX_EDGE;
constexpr long N_PART = (MAXC + 255) / 256;
constexpr long SCRATCH_DOUBLES = OFF_PART + N_PART;

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_ptr[MAX_DEVICES] = {};

// device address of the current device's copy of `scratch`, looked up once per device
int scratch_address(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    dmk_set_error("dmk_divergence_rms_noalloc: no current HIP device");
    return DMK_ERR_HIP;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      dmk_set_error("dmk_divergence_rms_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
      return DMK_ERR_HIP;
    }
    g_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_ptr[dev];
  return 0;
}

// sum of the per-block divergence sums -> sqrt(sum) * scale (one block)
__global__ void __launch_bounds__(256) rms_finalize(const double* __restrict__ part, int nb, double scale,
                                                    double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sqrt(red[0]) * scale;
}

}  // namespace

extern "C" int dmk_divergence_rms_noalloc(const int32_t cells[3], const double size[3], int32_t divergence_correction,
                                          const double* field_dev, double* rms_dev, void* stream) {
  if (!cells || !size || !field_dev || !rms_dev) {
    dmk_set_error("dmk_divergence_rms_noalloc: null argument");
    return DMK_ERR_ARG;
  }
  long N = 1;
  for (int d = 0; d < 3; d++) {
    if (cells[d] < 1 || !(size[d] > 0.0)) {
      dmk_set_error("dmk_divergence_rms_noalloc: invalid grid");
      return DMK_ERR_ARG;
    }
    if (cells[d] > MAX_EDGE) {
      dmk_set_error("dmk_divergence_rms_noalloc: %d cells along axis %d exceed the maximum of %d", cells[d], d,
                    MAX_EDGE);
      return DMK_ERR_ARG;
    }
    N *= cells[d];
  }
  if (N > MAXC) {
    dmk_set_error("dmk_divergence_rms_noalloc: grid of %ld cells exceeds the maximum of %ld cells", N, MAXC);
    return DMK_ERR_ARG;
  }
  if (divergence_correction < 0 || divergence_correction > 2) {
    dmk_set_error("dmk_divergence_rms_noalloc: divergence_correction must be 0, 1 or 2");
    return DMK_ERR_ARG;
  }
  double* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* soa = base + OFF_SOA;
  double2* Z = reinterpret_cast<double2*>(base + OFF_Z);
  double2* tw = reinterpret_cast<double2*>(base + OFF_TW);
  double* part = base + OFF_PART;

  DmkFFT fft;
  const int cl[3] = {cells[0], cells[1], cells[2]};
  if (dmk_fft_setup_into(&fft, cl, tw, s)) {
    dmk_set_error("dmk_divergence_rms_noalloc: FFT set-up failed: %s", hipGetErrorString(hipGetLastError()));
    return DMK_ERR_HIP;
  }
  if (dmk_aos_to_soa(field_dev, N, 9, soa, s) || dmk_fft_r2c(&fft, 9, soa, Z, s)) {
    dmk_set_error("dmk_divergence_rms_noalloc: transform launch failed: %s", hipGetErrorString(hipGetLastError()));
    return DMK_ERR_HIP;
  }
  GammaArgs g{};
  for (int d = 0; d < 3; d++) {
    g.n[d] = cells[d];
    g.size[d] = size[d];
  }
  g.h = fft.h;
  dmk_div_scale(cl, size, divergence_correction, g.rg);
  g.do_div = 1;
  g.do_gamma = 0;
  const int nb = dmk_gamma_div_blocks(&g);
  double scale = 1.0 / (double)N;
  if (cells[0] == 1) scale *= 0.5;
  if (dmk_gamma_div(Z, &g, part, s)) {
    dmk_set_error("dmk_divergence_rms_noalloc: divergence launch failed: %s", hipGetErrorString(hipGetLastError()));
    return DMK_ERR_HIP;
  }
  hipLaunchKernelGGL(rms_finalize, dim3(1), dim3(256), 0, s, part, nb, scale, rms_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    dmk_set_error("dmk_divergence_rms_noalloc: launch failed: %s", hipGetErrorString(e));
    return DMK_ERR_HIP;
  }
  return 0;
}
