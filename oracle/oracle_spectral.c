/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restatement of src/grid/spectral_utilities.f90 for the spectral-basic path:
 * unnormalised multi-field r2c/c2r FFTs (own mixed-radix FFT in place of
 * FFTW), the continuous-derivative frequency tables, the memory-efficient
 * Gamma convolution, the RMS divergence and curl and the FFT coordinate update.
 * Fourier data are kept un-transposed, (howmany, N1/2+1, N2, N3), component
 * fastest; FFTW's transposed-out layout only changes memory order.
 */
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* the tangent-linear build (oracle_fwd.cpp) supplies its own complex type */
#ifndef ORC_CPLX_DEFINED
#include <complex.h>
typedef double complex cplx;
#endif

struct orc_grid {
  int n[3], h;            /* cells, N1/2+1 */
  double size[3];
  double wgt;
  double scaled[3];       /* scaledGeomSize */
  orc_num_grid num;
  double C_ref[81];
  cplx *tw[3][2];         /* twiddles per axis, sign -1 / +1 */
  cplx *work;             /* 9 * h * N2 * N3 complex */
  double *tmpr;           /* 9 * N real */
};

void orc_default_num_grid(orc_num_grid *n) {
  n->N_iter_min = 1;
  n->N_iter_max = 100;
  n->eps_abs_div = 1.0e-4;
  n->eps_rel_div = 5.0e-4;
  n->eps_abs_P = 1.0e3;
  n->eps_rel_P = 1.0e-3;
  n->update_gamma = 0;
  n->divergence_correction = 1;
  n->solver = 0;
  n->ngmres_m = 30;
  n->snes_max_it = 50;
  n->N_cutback_max = 3;
}

/* ------------------------------------------------------------------ 1-D FFT */
static int smallest_factor(int n) {
  for (int p = 2; p * p <= n; p++)
    if (n % p == 0) return p;
  return n;
}

/* recursive mixed-radix decimation in time: y[k] = sum_j x[j*xs] w^(jk),
 * w = exp(sign*2*pi*i/n) taken from tw with stride tws */
static void fft_rec(int n, const cplx *x, int xs, cplx *y, const cplx *tw, int tws, cplx *tmp) {
  if (n == 1) { y[0] = x[0]; return; }
  int p = smallest_factor(n), m = n / p;
  for (int r = 0; r < p; r++) fft_rec(m, x + (size_t)r * xs, xs * p, y + (size_t)r * m, tw, tws * p, tmp);
  for (int k = 0; k < n; k++) {
    cplx s = 0;
    for (int r = 0; r < p; r++) s += tw[(size_t)(((long)r * k) % n) * tws] * y[(size_t)r * m + (k % m)];
    tmp[k] = s;
  }
  memcpy(y, tmp, sizeof(cplx) * n);
}

static cplx *make_tw(int n, int sign) {
  cplx *t = (cplx *)malloc(sizeof(cplx) * n);
  const double TAU = 2.0 * acos(-1.0);
  for (int k = 0; k < n; k++) t[k] = cos(TAU * k / n) + I * (sign * sin(TAU * k / n));
  return t;
}

/* ------------------------------------------------------------------ grid */
orc_grid *orc_grid_create(const int cells[3], const double size[3], const orc_num_grid *num) {
  orc_grid *g = (orc_grid *)calloc(1, sizeof *g);
  for (int d = 0; d < 3; d++) { g->n[d] = cells[d]; g->size[d] = size[d]; }
  g->h = cells[0] / 2 + 1;
  g->wgt = 1.0 / (double)((long)cells[0] * cells[1] * cells[2]);
  if (num) g->num = *num; else orc_default_num_grid(&g->num);
  /* spectral_utilities.f90:181-197 divergence_correction */
  double r[3];
  for (int d = 0; d < 3; d++) r[d] = size[d] / (double)cells[d];
  if (g->num.divergence_correction == 2) {
    for (int d = 0; d < 3; d++) g->scaled[d] = size[d];
  } else {
    const double *v = g->num.divergence_correction == 1 ? r : size;
    int imin = 0, imax = 0;
    for (int d = 1; d < 3; d++) { if (v[d] < v[imin]) imin = d; if (v[d] > v[imax]) imax = d; }
    for (int j = 0; j < 3; j++)
      if (j != imin && j != imax)
        for (int d = 0; d < 3; d++)
          g->scaled[d] = g->num.divergence_correction == 1 ? size[d] / size[j] * (double)cells[j]
                                                            : size[d] / size[j];
  }
  for (int d = 0; d < 3; d++) { g->tw[d][0] = make_tw(cells[d], -1); g->tw[d][1] = make_tw(cells[d], +1); }
  size_t Nk = (size_t)g->h * cells[1] * cells[2], N = (size_t)cells[0] * cells[1] * cells[2];
  g->work = (cplx *)calloc(9 * Nk, sizeof(cplx));
  g->tmpr = (double *)calloc(9 * N, sizeof(double));
  return g;
}

void orc_grid_destroy(orc_grid *g) {
  if (!g) return;
  for (int d = 0; d < 3; d++) { free(g->tw[d][0]); free(g->tw[d][1]); }
  free(g->work); free(g->tmpr); free(g);
}

/* c2c along y and z of a (hm, h, N2, N3) complex field, in place */
static void fft_yz(const orc_grid *g, int hm, cplx *c, int sign) {
  int h = g->h, N2 = g->n[1], N3 = g->n[2];
  int si = sign < 0 ? 0 : 1;
  /* along y */
#pragma omp parallel for collapse(2)
  for (int z = 0; z < N3; z++)
    for (int kx = 0; kx < h; kx++) {
      cplx in[1024], out[1024], tmp[1024];
      for (int q = 0; q < hm; q++) {
        for (int y = 0; y < N2; y++) in[y] = c[q + (size_t)hm * (kx + (size_t)h * (y + (size_t)N2 * z))];
        fft_rec(N2, in, 1, out, g->tw[1][si], 1, tmp);
        for (int y = 0; y < N2; y++) c[q + (size_t)hm * (kx + (size_t)h * (y + (size_t)N2 * z))] = out[y];
      }
    }
  /* along z */
#pragma omp parallel for collapse(2)
  for (int y = 0; y < N2; y++)
    for (int kx = 0; kx < h; kx++) {
      cplx in[1024], out[1024], tmp[1024];
      for (int q = 0; q < hm; q++) {
        for (int z = 0; z < N3; z++) in[z] = c[q + (size_t)hm * (kx + (size_t)h * (y + (size_t)N2 * z))];
        fft_rec(N3, in, 1, out, g->tw[2][si], 1, tmp);
        for (int z = 0; z < N3; z++) c[q + (size_t)hm * (kx + (size_t)h * (y + (size_t)N2 * z))] = out[z];
      }
    }
}

static void r2c(const orc_grid *g, int hm, const double *in, cplx *c) {
  int N1 = g->n[0], N2 = g->n[1], N3 = g->n[2], h = g->h;
#pragma omp parallel for collapse(2)
  for (int z = 0; z < N3; z++)
    for (int y = 0; y < N2; y++) {
      cplx a[1024], o[1024], tmp[1024];
      for (int q = 0; q < hm; q++) {
        for (int x = 0; x < N1; x++) a[x] = in[q + (size_t)hm * (x + (size_t)N1 * (y + (size_t)N2 * z))];
        fft_rec(N1, a, 1, o, g->tw[0][0], 1, tmp);
        for (int kx = 0; kx < h; kx++) c[q + (size_t)hm * (kx + (size_t)h * (y + (size_t)N2 * z))] = o[kx];
      }
    }
  fft_yz(g, hm, c, -1);
}

/* c2r: the input is consumed (overwritten) */
static void c2r(const orc_grid *g, int hm, cplx *c, double *out) {
  int N1 = g->n[0], N2 = g->n[1], N3 = g->n[2], h = g->h;
  fft_yz(g, hm, c, +1);
#pragma omp parallel for collapse(2)
  for (int z = 0; z < N3; z++)
    for (int y = 0; y < N2; y++) {
      cplx a[1024], o[1024], tmp[1024];
      for (int q = 0; q < hm; q++) {
        for (int kx = 0; kx < h; kx++) a[kx] = c[q + (size_t)hm * (kx + (size_t)h * (y + (size_t)N2 * z))];
        a[0] = creal(a[0]);
        if (N1 % 2 == 0) a[N1 / 2] = creal(a[N1 / 2]);
        for (int kx = h; kx < N1; kx++) a[kx] = conj(a[N1 - kx]);
        fft_rec(N1, a, 1, o, g->tw[0][1], 1, tmp);
        for (int x = 0; x < N1; x++) out[q + (size_t)hm * (x + (size_t)N1 * (y + (size_t)N2 * z))] = creal(o[x]);
      }
    }
}

void orc_fft_r2c(const orc_grid *g, int howmany, const double *in, double *out) {
  r2c(g, howmany, in, (cplx *)out);
}

void orc_fft_c2r(const orc_grid *g, int howmany, const double *in, double *out) {
  size_t Nk = (size_t)g->h * g->n[1] * g->n[2];
  cplx *c = (cplx *)malloc(sizeof(cplx) * howmany * Nk);
  memcpy(c, in, sizeof(cplx) * howmany * Nk);
  c2r(g, howmany, c, out);
  free(c);
}

/* spectral_utilities.f90:315-333 + getFreqDerivative (continuous): xi1st */
static void xi_at(const orc_grid *g, int kx, int ky, int kz, int first, cplx xi[3]) {
  const double TAU = 2.0 * acos(-1.0);
  int idx[3] = {kx, ky, kz};
  for (int d = 0; d < 3; d++) {
    int k = idx[d];
    int ks = k;
    if (d > 0 && k > g->n[d] / 2) ks = k - g->n[d];
    xi[d] = I * (TAU * (double)ks / g->size[d]);
    if (first && g->n[d] % 2 == 0 && k == g->n[d] / 2) xi[d] = 0.0;
  }
}

/* spectral_utilities.f90:356-368 utilities_updateGamma (memory-efficient) */
void orc_update_gamma(orc_grid *g, const double C[81]) {
  for (int k = 0; k < 81; k++) g->C_ref[k] = C[k] / g->wgt;
}

/* spectral_utilities.f90:436-501 utilities_GammaConvolution */
void orc_gamma_convolution(orc_grid *g, const double *field, const double aim[9], double *out) {
  int h = g->h, N2 = g->n[1], N3 = g->n[2];
  cplx *c = g->work;
  r2c(g, 9, field, c);
  const double *Cr = g->C_ref;
#pragma omp parallel for collapse(2)
  for (int z = 0; z < N3; z++)
    for (int y = 0; y < N2; y++)
      for (int kx = 0; kx < h; kx++) {
        if (kx == 0 && y == 0 && z == 0) continue;
        cplx xi[3], D[9], T[9];
        xi_at(g, kx, y, z, 1, xi);
        for (int m = 0; m < 3; m++)
          for (int l = 0; l < 3; l++) D[I2(l, m)] = conj(-xi[l]) * xi[m];
        for (int m = 0; m < 3; m++)
          for (int l = 0; l < 3; l++) {
            cplx s = 0;
            for (int q = 0; q < 3; q++)
              for (int j = 0; j < 3; j++) s += Cr[I4(l, j, m, q)] * D[I2(j, q)];
            T[I2(l, m)] = s;
          }
        double A[36], Ainv[36], Re[9];
        for (int m = 0; m < 3; m++)
          for (int l = 0; l < 3; l++) {
            A[l + 6 * m] = creal(T[I2(l, m)]);
            A[(l + 3) + 6 * (m + 3)] = creal(T[I2(l, m)]);
            A[l + 6 * (m + 3)] = cimag(T[I2(l, m)]);
            A[(l + 3) + 6 * m] = -cimag(T[I2(l, m)]);
            Re[I2(l, m)] = creal(T[I2(l, m)]);
          }
        cplx *Ph = c + (size_t)9 * (kx + (size_t)h * (y + (size_t)N2 * z));
        if (fabs(m33_det(Re)) > 1.e-16) {
          orc_invert(6, A, Ainv);
          cplx K[9], G[81], R[9];
          for (int n = 0; n < 3; n++)
            for (int l = 0; l < 3; l++) K[I2(l, n)] = Ainv[l + 6 * n] + I * Ainv[l + 6 * (n + 3)];
          for (int o = 0; o < 3; o++)
            for (int n = 0; n < 3; n++)
              for (int m = 0; m < 3; m++)
                for (int l = 0; l < 3; l++) G[I4(l, m, n, o)] = K[I2(l, n)] * D[I2(o, m)];
          for (int m = 0; m < 3; m++)
            for (int l = 0; l < 3; l++) {
              cplx s = 0;
              for (int o = 0; o < 3; o++)
                for (int n = 0; n < 3; n++) s += G[I4(l, m, n, o)] * Ph[I2(n, o)];
              R[I2(l, m)] = s;
            }
          for (int k = 0; k < 9; k++) Ph[k] = R[k];
        } else {
          for (int k = 0; k < 9; k++) Ph[k] = 0.0;
        }
      }
  for (int k = 0; k < 9; k++) c[k] = aim[k];
  c2r(g, 9, c, out);
}

/* spectral_utilities.f90:650-691 utilities_divergenceRMS */
double orc_divergence_rms(orc_grid *g, const double *field) {
  int h = g->h, N2 = g->n[1], N3 = g->n[2];
  cplx *c = g->work;
  r2c(g, 9, field, c);
  double rg[3];
  for (int d = 0; d < 3; d++) rg[d] = g->size[d] / g->scaled[d];
  double total = 0.0;
  for (int y = 0; y < N2; y++)
    for (int z = 0; z < N3; z++) {
      double s_inner = 0.0;
      for (int kx = 0; kx < h; kx++) {
        cplx xi[3], v[3];
        xi_at(g, kx, y, z, 1, xi);
        const cplx *Ph = c + (size_t)9 * (kx + (size_t)h * (y + (size_t)N2 * z));
        for (int l = 0; l < 3; l++) {
          cplx s = 0;
          for (int m = 0; m < 3; m++) s += Ph[I2(l, m)] * (conj(-xi[m]) * rg[m]);
          v[l] = s;
        }
        double a = 0.0;
        for (int l = 0; l < 3; l++) a += creal(v[l]) * creal(v[l]);
        double b = 0.0;
        for (int l = 0; l < 3; l++) b += cimag(v[l]) * cimag(v[l]);
        if (kx >= 1 && kx <= h - 2) s_inner += 2.0 * (a + b);
        else if (kx == 0 || kx == h - 1) s_inner += a + b;
      }
      if (h == 1) {
        /* the DC layer doubles as the "Nyquist" layer: counted twice */
        cplx xi[3], v[3];
        xi_at(g, 0, y, z, 1, xi);
        const cplx *Ph = c + (size_t)9 * ((size_t)h * (y + (size_t)N2 * z));
        for (int l = 0; l < 3; l++) {
          cplx s = 0;
          for (int m = 0; m < 3; m++) s += Ph[I2(l, m)] * (conj(-xi[m]) * rg[m]);
          v[l] = s;
        }
        for (int l = 0; l < 3; l++) s_inner += creal(v[l]) * creal(v[l]) + cimag(v[l]) * cimag(v[l]);
      }
      total += s_inner;
    }
  double r = sqrt(total) * g->wgt;
  if (g->n[0] == 1) r *= 0.5;
  return r;
}

/* spectral_utilities.f90:697-757 utilities_curlRMS.  On the DC and last x
 * planes the reference assigns each of the three curl expressions to the whole
 * 3x3 curl array inside the loop over l, so the sum it takes there is nine
 * times |F(3,2) xi_1 - F(3,1) xi_2|^2 of l = 3; that is restated as written. */
double orc_curl_rms(orc_grid *g, const double *field) {
  int h = g->h, N2 = g->n[1], N3 = g->n[2];
  cplx *c = g->work;
  r2c(g, 9, field, c);
  double rg[3];
  for (int d = 0; d < 3; d++) rg[d] = g->size[d] / g->scaled[d];
  double total = 0.0;
  for (int y = 0; y < N2; y++)
    for (int z = 0; z < N3; z++) {
      for (int kx = 1; kx <= h - 2; kx++) {
        cplx xi[3], cu[3][3];
        xi_at(g, kx, y, z, 1, xi);
        const cplx *Fh = c + (size_t)9 * (kx + (size_t)h * (y + (size_t)N2 * z));
        for (int l = 0; l < 3; l++) {
          cu[l][0] = Fh[I2(l, 2)] * xi[1] * rg[1] - Fh[I2(l, 1)] * xi[2] * rg[2];
          cu[l][1] = Fh[I2(l, 0)] * xi[2] * rg[2] - Fh[I2(l, 2)] * xi[0] * rg[0];
          cu[l][2] = Fh[I2(l, 1)] * xi[0] * rg[0] - Fh[I2(l, 0)] * xi[1] * rg[1];
        }
        double s = 0.0;
        for (int l = 0; l < 3; l++)
          for (int m = 0; m < 3; m++) s += creal(cu[l][m]) * creal(cu[l][m]) + cimag(cu[l][m]) * cimag(cu[l][m]);
        total += 2.0 * s;
      }
      /* the DC plane, then the plane kx = h - 1 (the same plane when h == 1) */
      int planes[2] = {0, h - 1};
      for (int pi = 0; pi < 2; pi++) {
        int kx = planes[pi];
        cplx xi[3], v = 0;
        xi_at(g, kx, y, z, 1, xi);
        const cplx *Fh = c + (size_t)9 * (kx + (size_t)h * (y + (size_t)N2 * z));
        for (int l = 0; l < 3; l++) {
          v = Fh[I2(l, 2)] * xi[1] * rg[1] - Fh[I2(l, 1)] * xi[2] * rg[2];
          v = Fh[I2(l, 0)] * xi[2] * rg[2] - Fh[I2(l, 2)] * xi[0] * rg[0];
          v = Fh[I2(l, 1)] * xi[0] * rg[0] - Fh[I2(l, 0)] * xi[1] * rg[1];
        }
        total += 9.0 * (creal(v) * creal(v) + cimag(v) * cimag(v));
      }
    }
  double r = sqrt(total) * g->wgt;
  if (g->n[0] == 1) r *= 0.5;
  return r;
}

/* spectral_utilities.f90:857-968 utilities_updateCoords (one process) */
void orc_update_coords(orc_grid *g, const double *F, double *x_p, double *x_n) {
  int N1 = g->n[0], N2 = g->n[1], N3 = g->n[2], h = g->h;
  size_t Nk = (size_t)h * N2 * N3, N = (size_t)N1 * N2 * N3;
  cplx *c = g->work;
  r2c(g, 9, F, c);
  double step[3], Favg[9];
  for (int d = 0; d < 3; d++) step[d] = g->size[d] / (double)g->n[d];
  for (int k = 0; k < 9; k++) Favg[k] = creal(c[k]) * g->wgt;
  cplx *v = (cplx *)calloc(3 * Nk, sizeof(cplx));
  for (int z = 0; z < N3; z++)
    for (int y = 0; y < N2; y++)
      for (int kx = 0; kx < h; kx++) {
        size_t f = kx + (size_t)h * (y + (size_t)N2 * z);
        if (kx == 0 && y == 0 && z == 0) continue;
        cplx xi[3];
        xi_at(g, kx, y, z, 0, xi);
        cplx den = 0;
        for (int d = 0; d < 3; d++) den += conj(-xi[d]) * xi[d];
        const cplx *Fh = c + 9 * f;
        for (int l = 0; l < 3; l++) {
          cplx s = 0;
          for (int m = 0; m < 3; m++) s += Fh[I2(l, m)] * xi[m];
          v[3 * f + l] = s / den;
        }
      }
  double *u = (double *)malloc(sizeof(double) * 3 * N);
  c2r(g, 3, v, u);
  for (size_t i = 0; i < 3 * N; i++) u[i] *= g->wgt;
  if (x_n) {
    for (int k = 0; k <= N3; k++)
      for (int j = 0; j <= N2; j++)
        for (int i = 0; i <= N1; i++) {
          size_t nidx = i + (size_t)(N1 + 1) * (j + (size_t)(N2 + 1) * k);
          double pos[3] = {step[0] * i, step[1] * j, step[2] * k};
          double x[3];
          for (int a = 0; a < 3; a++)
            x[a] = Favg[I2(a, 0)] * pos[0] + Favg[I2(a, 1)] * pos[1] + Favg[I2(a, 2)] * pos[2];
          for (int nb = 0; nb < 8; nb++) {
            static const int off[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1}};
            int mx = ((i + off[nb][0] - 1) % N1 + N1) % N1;
            int my = ((j + off[nb][1] - 1) % N2 + N2) % N2;
            int mz = ((k + off[nb][2] - 1) % N3 + N3) % N3;
            size_t ci = mx + (size_t)N1 * (my + (size_t)N2 * mz);
            for (int a = 0; a < 3; a++) x[a] = x[a] + u[3 * ci + a] * 0.125;
          }
          for (int a = 0; a < 3; a++) x_n[3 * nidx + a] = x[a];
        }
  }
  if (x_p) {
    for (int k = 0; k < N3; k++)
      for (int j = 0; j < N2; j++)
        for (int i = 0; i < N1; i++) {
          size_t ci = i + (size_t)N1 * (j + (size_t)N2 * k);
          double pos[3] = {step[0] * (i + 1 - 0.5), step[1] * (j + 1 - 0.5), step[2] * (k + 1 - 0.5)};
          for (int a = 0; a < 3; a++)
            x_p[3 * ci + a] = u[3 * ci + a]
                            + (Favg[I2(a, 0)] * pos[0] + Favg[I2(a, 1)] * pos[1] + Favg[I2(a, 2)] * pos[2]);
        }
  }
  free(u);
  free(v);
}

double orc_grid_wgt(const orc_grid *g) { return g->wgt; }
const int *orc_grid_cells(const orc_grid *g) { return g->n; }
