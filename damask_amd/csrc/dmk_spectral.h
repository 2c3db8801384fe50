// Spectral-operator interfaces (FFT plans and kernel arguments).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct FftPlan {
  int L;
  int nstages;
  int radix[16];
};

// tile shapes: gx[0] = line pairs per block of the x pass, gx[1], gx[2] =
// lines per block of the y and z passes; xy_threads > 0: the x and y passes
// run fused, one block of that many threads per (field, z) plane (all picked
// by dmk_fft_tune)
struct DmkFFT {
  int n[3];
  int h;
  FftPlan plan[3];
  double2* tw[3] = {nullptr, nullptr, nullptr};
  int gx[3] = {0, 0, 0};
  int xy_threads = 0;
  int xy_ip = 0;                    // 1: the fused x-y kernels' single-buffer (in-place) variant
  int zgz_g = 0, zgz_threads = 0;   // > 0: z passes and Gamma fused, zgz_g columns per block
  int zgz_ip = 0;                   // 1: the fused z kernel's single-buffer variant
};

struct GammaArgs {
  int n[3];
  int h;
  double size[3];
  double rg[3];            // geomSize / scaledGeomSize
  double Ccoef[9][6];      // acoustic-tensor coefficients of C_ref
  double aim[9];           // DC value, column-major
  const double* aim_dev;   // when set, the DC value is read from device memory instead
  int do_div, do_gamma;
};

// utilities_G_Convolution options: residual mode sets the DC term to
// dc_dev[c] * dc_scale, Jacobian mode keeps it; components with stress_mask
// set (stress not prescribed) get a zero DC term either way
struct GalerkinArgs {
  int do_div;
  int residual;
  int32_t stress_mask[9];
  const double* dc_dev;
  double dc_scale;
};

struct CoordsArgs {
  int n[3];
  double step[3];
  const double2* Fdc;      // forward transform of F: mean F_c = Fdc[c * Nk].x * wgt
  long Nk;
  double wgt;
};

// geomSize / scaledGeomSize for divergence_correction dc (0 none, 1 grid+size, 2 size)
void dmk_div_scale(const int cells[3], const double size[3], int dc, double rg[3]);
int dmk_fft_setup(DmkFFT* fft, const int cells[3]);
// time the candidate tile shapes on the device (scratch: 9 Nk complex and 9 N
// real doubles, overwritten) and keep the fastest per pass
int dmk_fft_tune(DmkFFT* fft, double2* cscratch, double* rscratch, hipStream_t s);
void dmk_fft_free(DmkFFT* fft);
// plans with the default (untuned) tile shapes and the twiddles of the three
// axes written by a kernel into tw (N1 + N2 + N3 complex values, caller memory,
// never freed by dmk_fft_free)
int dmk_fft_setup_into(DmkFFT* fft, const int cells[3], double2* tw, hipStream_t s);
int dmk_fft_r2c(const DmkFFT* fft, int howmany, const double* in, double2* out, hipStream_t s);
// the x and y passes of dmk_fft_r2c / dmk_fft_c2r only
int dmk_fft_r2c_xy(const DmkFFT* fft, int howmany, const double* in, double2* out, hipStream_t s);
int dmk_fft_c2r_xy(const DmkFFT* fft, int howmany, double2* in, double* out, hipStream_t s);
// on the x-y transformed nine fields Z: forward z pass, Gamma convolution /
// divergence per g, inverse z pass (fused or separate as tuned); *nb = number
// of per-block divergence sums written to div_partial (at most
// dmk_gamma_conv_partials)
int dmk_gamma_conv(const DmkFFT* fft, double2* Z, const GammaArgs* g, double* div_partial, int* nb, hipStream_t s);
long dmk_gamma_conv_partials(const DmkFFT* fft);
int dmk_fft_c2r(const DmkFFT* fft, int howmany, double2* in, double* out, hipStream_t s);
int dmk_gamma_div(double2* Z, const GammaArgs* g, double* div_partial, hipStream_t s);
int dmk_gamma_div_blocks(const GammaArgs* g);
// sum of squared curl amplitudes of a forward transform, per block of 256
// frequencies (dmk_gamma_div_blocks partial sums)
int dmk_galerkin(double2* Z, const GammaArgs* g, const GalerkinArgs* ga, double* div_partial, hipStream_t s);
int dmk_curl_partial(const double2* Z, const GammaArgs* g, double* partial, hipStream_t s);
int dmk_coords_freq(const double2* Fh, double2* U, const GammaArgs* g, hipStream_t s);
int dmk_coords_nodes(const double* u, const CoordsArgs* c, double* x_n, double* x_p, hipStream_t s);
