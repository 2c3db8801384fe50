// Host-side crystallography for the per-phase device tables:
//   slip systems and Schmid / non-Schmid projections (src/crystal.f90:1348-1454,
//   buildCoordinateSystem :1937-2001), slip-slip interaction matrices
//   (:591-902 with buildInteraction :1893-1930), cubic stiffness in Mandel
//   notation (:1711-1759) and its contraction with the projections.
// The slip-system directions/normals and the interaction-type tables are
// crystallographic data and are necessarily the reference's values.
#include <cmath>
#include <cstring>
#include <vector>

#include "dmk_common.h"
#include "../../include/damask_hip.h"

namespace {

const double CF_SLIP[18][6] = {
    {0, 1, -1, 1, 1, 1},   {-1, 0, 1, 1, 1, 1},   {1, -1, 0, 1, 1, 1},   {0, -1, -1, -1, -1, 1},
    {1, 0, 1, -1, -1, 1},  {-1, 1, 0, -1, -1, 1}, {0, -1, 1, 1, -1, -1}, {-1, 0, -1, 1, -1, -1},
    {1, 1, 0, 1, -1, -1},  {0, 1, 1, -1, 1, -1},  {1, 0, -1, -1, 1, -1}, {-1, -1, 0, -1, 1, -1},
    {1, 1, 0, 1, -1, 0},   {1, -1, 0, 1, 1, 0},   {1, 0, 1, 1, 0, -1},   {1, 0, -1, 1, 0, 1},
    {0, 1, 1, 0, 1, -1},   {0, 1, -1, 0, 1, 1}};

const double CI_SLIP[24][6] = {
    {1, -1, 1, 0, 1, 1},   {1, -1, 1, 1, 0, -1},  {1, -1, 1, -1, -1, 0}, {-1, -1, 1, 0, -1, -1},
    {-1, -1, 1, 1, 0, 1},  {-1, -1, 1, -1, 1, 0}, {1, 1, 1, 0, 1, -1},   {1, 1, 1, -1, 0, 1},
    {1, 1, 1, 1, -1, 0},   {-1, 1, 1, 0, -1, 1},  {-1, 1, 1, -1, 0, -1}, {-1, 1, 1, 1, 1, 0},
    {1, -1, 1, 2, 1, -1},  {1, -1, 1, -1, 1, 2},  {1, -1, 1, 1, 2, 1},   {-1, -1, 1, 2, -1, 1},
    {-1, -1, 1, 1, 1, 2},  {-1, -1, 1, -1, 2, 1}, {1, 1, 1, 1, 1, -2},   {1, 1, 1, 1, -2, 1},
    {1, 1, 1, -2, 1, 1},   {-1, 1, 1, 1, -1, 2},  {-1, 1, 1, 1, 2, -1},  {-1, 1, 1, 2, 1, 1}};

const unsigned char CF_INT[18][18] = {
    {1, 2, 2, 4, 7, 6, 3, 5, 5, 4, 6, 7, 10, 11, 10, 11, 12, 13},
    {2, 1, 2, 7, 4, 6, 6, 4, 7, 5, 3, 5, 10, 11, 12, 13, 10, 11},
    {2, 2, 1, 5, 5, 3, 6, 7, 4, 7, 6, 4, 12, 13, 10, 11, 10, 11},
    {4, 7, 6, 1, 2, 2, 4, 6, 7, 3, 5, 5, 10, 11, 11, 10, 13, 12},
    {7, 4, 6, 2, 1, 2, 5, 3, 5, 6, 4, 7, 10, 11, 13, 12, 11, 10},
    {5, 5, 3, 2, 2, 1, 7, 6, 4, 6, 7, 4, 12, 13, 11, 10, 11, 10},
    {3, 5, 5, 4, 6, 7, 1, 2, 2, 4, 7, 6, 11, 10, 11, 10, 12, 13},
    {6, 4, 7, 5, 3, 5, 2, 1, 2, 7, 4, 6, 11, 10, 13, 12, 10, 11},
    {6, 7, 4, 7, 6, 4, 2, 2, 1, 5, 5, 3, 13, 12, 11, 10, 10, 11},
    {4, 6, 7, 3, 5, 5, 4, 7, 6, 1, 2, 2, 11, 10, 10, 11, 13, 12},
    {5, 3, 5, 6, 4, 7, 7, 4, 6, 2, 1, 2, 11, 10, 12, 13, 11, 10},
    {7, 6, 4, 6, 7, 4, 5, 5, 3, 2, 2, 1, 13, 12, 10, 11, 11, 10},
    {10, 10, 12, 10, 10, 12, 11, 11, 13, 11, 11, 13, 1, 8, 9, 9, 9, 9},
    {11, 11, 13, 11, 11, 13, 10, 10, 12, 10, 10, 12, 8, 1, 9, 9, 9, 9},
    {10, 12, 10, 11, 13, 11, 11, 13, 11, 10, 12, 10, 9, 9, 1, 8, 9, 9},
    {11, 13, 11, 10, 12, 10, 10, 12, 10, 11, 13, 11, 9, 9, 8, 1, 9, 9},
    {12, 10, 10, 13, 11, 11, 12, 10, 10, 13, 11, 11, 9, 9, 9, 9, 1, 8},
    {13, 11, 11, 12, 10, 10, 13, 11, 11, 12, 10, 10, 9, 9, 9, 9, 8, 1}};

const unsigned char CI_INT[24][24] = {
    {1, 2, 2, 3, 5, 5, 6, 4, 7, 6, 7, 4, 8, 9, 9, 11, 13, 13, 14, 17, 18, 17, 14, 18},
    {2, 1, 2, 4, 6, 7, 5, 3, 5, 7, 6, 4, 9, 9, 8, 17, 14, 18, 13, 11, 13, 17, 18, 14},
    {2, 2, 1, 4, 7, 6, 7, 4, 6, 5, 5, 3, 9, 8, 9, 17, 18, 14, 18, 17, 14, 11, 13, 13},
    {3, 5, 5, 1, 2, 2, 6, 7, 4, 6, 4, 7, 11, 13, 13, 8, 9, 9, 17, 14, 18, 14, 17, 18},
    {4, 6, 7, 2, 1, 2, 7, 6, 4, 5, 3, 5, 17, 14, 18, 9, 9, 8, 17, 18, 14, 13, 11, 13},
    {4, 7, 6, 2, 2, 1, 5, 5, 3, 7, 4, 6, 17, 18, 14, 9, 8, 9, 11, 13, 13, 18, 17, 14},
    {6, 4, 7, 6, 7, 4, 1, 2, 2, 3, 5, 5, 18, 14, 17, 18, 17, 14, 9, 9, 8, 13, 13, 11},
    {5, 3, 5, 7, 6, 4, 2, 1, 2, 4, 6, 7, 13, 13, 11, 14, 17, 18, 9, 8, 9, 14, 18, 17},
    {7, 4, 6, 5, 5, 3, 2, 2, 1, 4, 7, 6, 14, 18, 17, 13, 11, 13, 8, 9, 9, 18, 14, 17},
    {6, 7, 4, 6, 4, 7, 3, 5, 5, 1, 2, 2, 18, 17, 14, 18, 14, 17, 13, 13, 11, 9, 9, 8},
    {7, 6, 4, 5, 3, 5, 4, 6, 7, 2, 1, 2, 14, 17, 18, 13, 13, 11, 14, 18, 17, 9, 8, 9},
    {5, 5, 3, 7, 4, 6, 4, 7, 6, 2, 2, 1, 13, 11, 13, 14, 18, 17, 18, 14, 17, 8, 9, 9},
    {8, 9, 9, 10, 16, 16, 19, 12, 15, 19, 15, 12, 1, 2, 2, 20, 21, 21, 23, 22, 24, 22, 23, 24},
    {9, 9, 8, 12, 15, 19, 15, 12, 19, 16, 16, 10, 2, 1, 2, 22, 24, 23, 24, 22, 23, 20, 21, 21},
    {9, 8, 9, 12, 19, 15, 16, 10, 16, 15, 19, 12, 2, 2, 1, 22, 23, 24, 21, 20, 21, 22, 24, 23},
    {10, 16, 16, 8, 9, 9, 19, 15, 12, 19, 12, 15, 20, 21, 21, 1, 2, 2, 22, 23, 24, 23, 22, 24},
    {12, 15, 19, 9, 9, 8, 16, 16, 10, 15, 12, 19, 22, 24, 23, 2, 1, 2, 20, 21, 21, 24, 22, 23},
    {12, 19, 15, 9, 8, 9, 15, 19, 12, 16, 10, 16, 21, 23, 24, 2, 2, 1, 22, 24, 23, 21, 20, 21},
    {15, 12, 19, 16, 16, 10, 9, 9, 8, 12, 15, 19, 23, 24, 22, 21, 20, 21, 1, 2, 2, 24, 23, 22},
    {16, 10, 16, 15, 19, 12, 9, 8, 9, 12, 19, 15, 22, 21, 20, 23, 22, 24, 2, 1, 2, 23, 24, 22},
    {19, 12, 15, 19, 15, 12, 8, 9, 9, 10, 16, 16, 24, 23, 22, 24, 22, 23, 2, 2, 1, 21, 21, 20},
    {16, 16, 10, 15, 12, 19, 12, 15, 19, 9, 9, 8, 21, 20, 21, 23, 24, 22, 24, 23, 22, 1, 2, 2},
    {15, 19, 12, 16, 10, 16, 12, 19, 15, 9, 8, 9, 23, 22, 24, 21, 21, 20, 23, 24, 22, 2, 1, 2},
    {19, 15, 12, 19, 12, 15, 10, 16, 16, 8, 9, 9, 24, 22, 23, 24, 23, 22, 21, 21, 20, 2, 2, 1}};

const int CF_MAX[2] = {12, 6};
const int CI_MAX[2] = {12, 12};

// passive Rodrigues rotation of a vector, P = -1 (src/rotations.f90:776-791)
void rot_vec(const double* q, const double* v, double* o) {
  double A = q[0] * q[0] - (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  double B = 2.0 * (q[1] * v[0] + q[2] * v[1] + q[3] * v[2]);
  double C = -2.0 * q[0];
  double c[3] = {q[2] * v[2] - q[3] * v[1], q[3] * v[0] - q[1] * v[2], q[1] * v[1] - q[2] * v[0]};
  for (int k = 0; k < 3; k++) o[k] = A * v[k] + B * q[k + 1] + C * c[k];
}

void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

void add_outer(double* M, double s, const double* a, const double* b) {   // M (row-major) += s a (x) b
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) M[3 * i + j] += s * a[i] * b[j];
}

}  // namespace

// Mandel cubic stiffness expanded to the full tensor, C[i][j][k][l] row-major
static void cubic_3333(double C11, double C12, double C44, double* C) {
  for (int q = 0; q < 81; q++) C[q] = 0.0;
  auto at = [&](int i, int j, int k, int l) -> double& { return C[27 * i + 9 * j + 3 * k + l]; };
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) at(i, i, k, k) = (i == k) ? C11 : C12;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (i != j) { at(i, j, i, j) = C44; at(i, j, j, i) = C44; }
}

// returns 0 on success, otherwise a message index
int dmk_build_phase(const dmk_phase_desc* d, DmkPhaseDev* p) {
  std::memset(p, 0, sizeof(*p));
  if (d->lattice != 0 && d->lattice != 1) return 1;
  if (d->n_fam < 1 || d->n_fam > 2) return 2;
  const int* maxs = d->lattice == 0 ? CF_MAX : CI_MAX;
  int nsl = 0;
  for (int f = 0; f < d->n_fam; f++) {
    if (d->N_sl[f] < 0 || d->N_sl[f] > maxs[f]) return 3;
    nsl += d->N_sl[f];
  }
  if (nsl != 12 && nsl != 24) return 4;   // kernels are instantiated for 12 and 24 systems
  p->nsl = nsl;
  p->C11 = d->C_11;
  p->C12 = d->C_12;
  p->C44 = d->C_44;
  double C[81];
  cubic_3333(d->C_11, d->C_12, d->C_44, C);
  std::vector<int> used;
  std::vector<int> fam;
  for (int f = 0; f < d->n_fam; f++) {
    int off = f == 0 ? 0 : maxs[0];
    for (int s = 0; s < d->N_sl[f]; s++) { used.push_back(off + s); fam.push_back(f); }
  }
  p->nonschmid = (d->lattice == 1 && d->n_nS > 0) ? 1 : 0;
  for (int a = 0; a < nsl; a++) {
    const double* row = d->lattice == 0 ? CF_SLIP[used[a]] : CI_SLIP[used[a]];
    double dv[3] = {row[0], row[1], row[2]}, nv[3] = {row[3], row[4], row[5]};
    double nd = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    double nn = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    for (int k = 0; k < 3; k++) { dv[k] /= nd; nv[k] /= nn; }
    std::memset(p->P[a], 0, sizeof(p->P[a]));
    add_outer(p->P[a], 1.0, dv, nv);
    for (int sense = -1; sense <= 1; sense += 2) {
      double* M = sense > 0 ? p->Pp[a] : p->Pm[a];
      double ds[3] = {dv[0] * sense, dv[1] * sense, dv[2] * sense};
      std::memset(M, 0, 9 * sizeof(double));
      add_outer(M, 1.0, ds, nv);
      if (p->nonschmid && fam[a] == 0) {
        double c[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < d->n_nS && k < 6; k++) c[k] = d->a_nS[k];
        // R = fromAxisAngle([ds, 60 deg], P=+1): axis negated, then ax2qu
        double ang = 60.0 * (2.0 * std::acos(-1.0) / 360.0);
        double q[4] = {std::cos(ang * 0.5), -ds[0] * std::sin(ang * 0.5), -ds[1] * std::sin(ang * 0.5),
                       -ds[2] * std::sin(ang * 0.5)};
        double np[3], nxd[3], npxd[3];
        rot_vec(q, nv, np);
        cross(nv, ds, nxd);
        cross(np, ds, npxd);
        add_outer(M, c[0], ds, np);
        add_outer(M, c[1], nxd, nv);
        add_outer(M, c[2], npxd, np);
        add_outer(M, c[3], nv, nv);
        add_outer(M, c[4], nxd, nxd);
        add_outer(M, c[5], ds, ds);
      }
      // Q = C : Pn  (Q_od = sum_mn Pn_mn C_mnod)
      double* Q = sense > 0 ? p->Qp[a] : p->Qm[a];
      for (int o = 0; o < 3; o++)
        for (int dd = 0; dd < 3; dd++) {
          double s = 0.0;
          for (int m = 0; m < 3; m++)
            for (int n = 0; n < 3; n++) s += M[3 * m + n] * C[27 * m + 9 * n + 3 * o + dd];
          Q[3 * o + dd] = s;
        }
    }
  }
  // interaction (reacting r, acting a)
  for (int r = 0; r < nsl; r++)
    for (int a = 0; a < nsl; a++) {
      int t = d->lattice == 0 ? CF_INT[used[r]][used[a]] : CI_INT[used[r]][used[a]];
      if (t > d->n_h) return 5;
      p->H[r][a] = d->h_sl_sl[t - 1];
      p->Ht[r][a] = t - 1;
    }
  p->n_h = d->n_h;
  int a = 0;
  for (int f = 0; f < d->n_fam; f++)
    for (int s = 0; s < d->N_sl[f]; s++, a++) {
      p->g0[a] = d->dot_gamma_0[f];
      p->n[a] = d->n[f];
      p->a[a] = d->a[f];
      p->h0[a] = d->h_0[f];
      p->xi_inf[a] = d->xi_inf[f];
      p->c1[a] = d->c_1[f];
      p->c2[a] = d->c_2[f];
      p->fsat[a] = d->f_sat[f];
      p->xi_sat[a] = d->xi_inf[f] + d->f_sat[f] * std::sqrt(0.0);
      p->hf[a] = d->h_0[f] * (1.0 + d->c_1[f] * std::pow(0.0, d->c_2[f]));
    }
  p->quarter_pow = 1;
  for (int b = 0; b < nsl; b++)
    for (double y : {p->n[b], p->a[b]})
      if (!(y > 0.0 && y <= 64.0 && y * 4.0 == std::rint(y * 4.0))) p->quarter_pow = 0;
  p->atol_xi = d->atol_xi;
  p->atol_gamma = d->atol_gamma;
  return 0;
}

// qu2om with P = -1 and det-normalisation (src/rotations.f90:399-423), row-major
void dmk_qu2om(const double* qu, double* om) {
  double qq = qu[0] * qu[0] - (qu[1] * qu[1] + qu[2] * qu[2] + qu[3] * qu[3]);
  double o[9];
  o[0] = qq + 2.0 * qu[1] * qu[1];
  o[4] = qq + 2.0 * qu[2] * qu[2];
  o[8] = qq + 2.0 * qu[3] * qu[3];
  o[1] = 2.0 * (qu[1] * qu[2] - qu[0] * qu[3]);
  o[5] = 2.0 * (qu[2] * qu[3] - qu[0] * qu[1]);
  o[6] = 2.0 * (qu[3] * qu[1] - qu[0] * qu[2]);
  o[3] = 2.0 * (qu[2] * qu[1] + qu[0] * qu[3]);
  o[7] = 2.0 * (qu[3] * qu[2] + qu[0] * qu[1]);
  o[2] = 2.0 * (qu[1] * qu[3] + qu[0] * qu[2]);
  double t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[3 * i + j] = o[3 * j + i];
  double det = t[0] * (t[4] * t[8] - t[5] * t[7]) - t[1] * (t[3] * t[8] - t[5] * t[6])
             + t[2] * (t[3] * t[7] - t[4] * t[6]);
  double dd = std::pow(det, 1.0 / 3.0);
  for (int k = 0; k < 9; k++) om[k] = t[k] / dd;
}

// Miller indices of the a-th active slip system (direction, normal) for the
// family selection of dmk_build_phase: the labels of the xi_sl / gamma_sl
// results (crystal.f90:1642-1676).  Returns 0, or 1 for an invalid request.
extern "C" int dmk_slip_system(int32_t lattice, int32_t n_fam, const int32_t* N_sl, int32_t a, double out[6]) {
  if ((lattice != 0 && lattice != 1) || n_fam < 1 || n_fam > 2 || !N_sl || !out || a < 0) return 1;
  const int* maxs = lattice == 0 ? CF_MAX : CI_MAX;
  int k = a;
  for (int f = 0; f < n_fam; f++) {
    if (N_sl[f] < 0 || N_sl[f] > maxs[f]) return 1;
    if (k < N_sl[f]) {
      const int off = f == 0 ? 0 : maxs[0];
      const double* row = lattice == 0 ? CF_SLIP[off + k] : CI_SLIP[off + k];
      for (int c = 0; c < 6; c++) out[c] = row[c];
      return 0;
    }
    k -= N_sl[f];
  }
  return 1;
}
