/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restatement of the tensor algebra, rotation and crystallography routines the
 * spectral-basic hot path uses.  Each function cites the reference routine it
 * follows.  Indexing follows the Fortran arrays: T(i,j) at [i + 3j].
 */
#include "oracle_internal.h"

#include <float.h>
#include <math.h>
#include <string.h>

/* prec.f90:131-143 dEq0 without tolerance */
int orc_dEq0(double a) { return fabs(a) <= DBL_MIN * 10.0; }
/* prec.f90:93-105 dEq */
int orc_dEq(double a, double b, double tol) { return fabs(a - b) <= tol; }

void m33_mul(const double *A, const double *B, double *C) {
  double T[9];
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += A[I2(i, k)] * B[I2(k, j)];
      T[I2(i, j)] = s;
    }
  memcpy(C, T, sizeof T);
}

void m33_transpose(const double *A, double *B) {
  double T[9];
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) T[I2(i, j)] = A[I2(j, i)];
  memcpy(B, T, sizeof T);
}

void m33_eye(double *A) {
  for (int k = 0; k < 9; k++) A[k] = 0.0;
  A[0] = A[4] = A[8] = 1.0;
}

/* norm2 of a 3x3 (Frobenius) */
double m33_norm2(const double *A) {
  double s = 0.0;
  for (int k = 0; k < 9; k++) s += A[k] * A[k];
  return sqrt(s);
}

/* math.f90:630-638 math_det33 */
double m33_det(const double *m) {
  return m[I2(0, 0)] * (m[I2(1, 1)] * m[I2(2, 2)] - m[I2(1, 2)] * m[I2(2, 1)])
       - m[I2(0, 1)] * (m[I2(1, 0)] * m[I2(2, 2)] - m[I2(1, 2)] * m[I2(2, 0)])
       + m[I2(0, 2)] * (m[I2(1, 0)] * m[I2(2, 1)] - m[I2(1, 1)] * m[I2(2, 0)]);
}

/* math.f90:471-505 math_invert33 (Cramer).  Returns 1 on error. */
int m33_invert(const double *A, double *InvA, double *DetA) {
  double I[9];
  I[I2(0, 0)] = A[I2(1, 1)] * A[I2(2, 2)] - A[I2(1, 2)] * A[I2(2, 1)];
  I[I2(1, 0)] = -A[I2(1, 0)] * A[I2(2, 2)] + A[I2(1, 2)] * A[I2(2, 0)];
  I[I2(2, 0)] = A[I2(1, 0)] * A[I2(2, 1)] - A[I2(1, 1)] * A[I2(2, 0)];
  double det = A[I2(0, 0)] * I[I2(0, 0)] + A[I2(0, 1)] * I[I2(1, 0)] + A[I2(0, 2)] * I[I2(2, 0)];
  if (orc_dEq0(det)) {
    for (int k = 0; k < 9; k++) InvA[k] = 0.0;
    if (DetA) *DetA = 0.0;
    return 1;
  }
  I[I2(0, 1)] = -A[I2(0, 1)] * A[I2(2, 2)] + A[I2(0, 2)] * A[I2(2, 1)];
  I[I2(1, 1)] = A[I2(0, 0)] * A[I2(2, 2)] - A[I2(0, 2)] * A[I2(2, 0)];
  I[I2(2, 1)] = -A[I2(0, 0)] * A[I2(2, 1)] + A[I2(0, 1)] * A[I2(2, 0)];
  I[I2(0, 2)] = A[I2(0, 1)] * A[I2(1, 2)] - A[I2(0, 2)] * A[I2(1, 1)];
  I[I2(1, 2)] = -A[I2(0, 0)] * A[I2(1, 2)] + A[I2(0, 2)] * A[I2(1, 0)];
  I[I2(2, 2)] = A[I2(0, 0)] * A[I2(1, 1)] - A[I2(0, 1)] * A[I2(1, 0)];
  for (int k = 0; k < 9; k++) InvA[k] = I[k] / det;
  if (DetA) *DetA = det;
  return 0;
}

/* math.f90:452-465 math_inv33: zeros on failure */
void m33_inv(const double *A, double *B) {
  if (m33_invert(A, B, NULL))
    for (int k = 0; k < 9; k++) B[k] = 0.0;
}

/* LAPACK dgetrf: LU with partial pivoting (first maximum wins), column-major,
 * in place.  Returns info > 0 if U(info,info) is exactly zero. */
int lu_factor(int n, double *A, int *ipiv) {
  int info = 0;
  for (int j = 0; j < n; j++) {
    int p = j;
    double amax = fabs(A[j + n * j]);
    for (int i = j + 1; i < n; i++)
      if (fabs(A[i + n * j]) > amax) { amax = fabs(A[i + n * j]); p = i; }
    ipiv[j] = p;
    if (A[p + n * j] != 0.0) {
      if (p != j)
        for (int k = 0; k < n; k++) { double t = A[j + n * k]; A[j + n * k] = A[p + n * k]; A[p + n * k] = t; }
      double r = 1.0 / A[j + n * j];
      for (int i = j + 1; i < n; i++) A[i + n * j] *= r;
    } else if (info == 0) {
      info = j + 1;
    }
    for (int k = j + 1; k < n; k++) {
      double a = A[j + n * k];
      for (int i = j + 1; i < n; i++) A[i + n * k] -= A[i + n * j] * a;
    }
  }
  return info;
}

/* solve with an LU factorisation (dgetrs, no transpose), b overwritten */
void lu_solve(int n, const double *LU, const int *ipiv, double *b) {
  for (int i = 0; i < n; i++) {
    int p = ipiv[i];
    if (p != i) { double t = b[i]; b[i] = b[p]; b[p] = t; }
  }
  for (int j = 0; j < n; j++)
    for (int i = j + 1; i < n; i++) b[i] -= LU[i + n * j] * b[j];
  for (int j = n - 1; j >= 0; j--) {
    b[j] /= LU[j + n * j];
    for (int i = 0; i < j; i++) b[i] -= LU[i + n * j] * b[j];
  }
}

/* LAPACK dgesv restated: returns info */
int dgesv_n(int n, double *A, double *b) {
  int ipiv[64];
  int info = lu_factor(n, A, ipiv);
  if (info) return info;
  lu_solve(n, A, ipiv, b);
  return 0;
}

/* math.f90:538-555 math_invert (dgetrf + dgetri).  Returns 1 on error. */
int orc_invert(int n, const double *A, double *invA) {
  double LU[81 * 4];
  int ipiv[64];
  double *lu = n * n <= 81 * 4 ? LU : NULL;
  memcpy(lu, A, sizeof(double) * n * n);
  int info = lu_factor(n, lu, ipiv);
  if (info) {
    memcpy(invA, lu, sizeof(double) * n * n);
    return 1;
  }
  for (int c = 0; c < n; c++) {
    double col[64];
    for (int i = 0; i < n; i++) col[i] = (i == c) ? 1.0 : 0.0;
    lu_solve(n, lu, ipiv, col);
    for (int i = 0; i < n; i++) invA[i + n * c] = col[i];
  }
  return 0;
}

/* math.f90:826-845 math_sym3333to66: Mandel 6x6 (column-major) of a 3x3x3x3;
 * Voigt order 11 22 33 23 13 12, weights 1 and sqrt(2) (constants.f90:19) */
static const int MV_I[6] = {0, 1, 2, 1, 0, 0}, MV_J[6] = {0, 1, 2, 2, 2, 1};
static void sym3333to66(const double *T, double *m) {
  const double w[6] = {1.0, 1.0, 1.0, sqrt(2.0), sqrt(2.0), sqrt(2.0)};
  for (int q = 0; q < 6; q++)
    for (int p = 0; p < 6; p++) m[p + 6 * q] = T[I4(MV_I[p], MV_J[p], MV_I[q], MV_J[q])] * w[p] * w[q];
}

/* math.f90:851-904 math_66toSym3333 */
static void m66toSym3333(const double *m, double *T) {
  const double w[6] = {1.0, 1.0, 1.0, sqrt(2.0), sqrt(2.0), sqrt(2.0)};
  for (int q = 0; q < 6; q++)
    for (int p = 0; p < 6; p++) {
      double v = m[p + 6 * q] / (w[p] * w[q]);
      int i = MV_I[p], j = MV_J[p], k = MV_I[q], l = MV_J[q];
      T[I4(i, j, k, l)] = v;
      T[I4(j, i, k, l)] = v;
      T[I4(i, j, l, k)] = v;
      T[I4(j, i, l, k)] = v;
    }
}

/* math.f90:511-532 math_invSym3333 (dgetrf + dgetri on the Mandel form).
 * Returns 1 on a singular matrix. */
int orc_inv_sym3333(const double *A, double *invA) {
  double m[36], mi[36];
  sym3333to66(A, m);
  if (orc_invert(6, m, mi)) return 1;
  m66toSym3333(mi, invA);
  return 0;
}

/* math.f90:405-418 math_mul3333xx3333 */
void t3333_mul(const double *A, const double *B, double *C) {
  double T[81];
  for (int l = 0; l < 3; l++)
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          double s = 0.0;
          for (int n = 0; n < 3; n++)
            for (int m = 0; m < 3; m++) s += A[I4(i, j, m, n)] * B[I4(m, n, k, l)];
          T[I4(i, j, k, l)] = s;
        }
  memcpy(C, T, sizeof T);
}

/* math.f90:386-398 math_mul3333xx33 */
void t3333_mul33(const double *A, const double *B, double *C) {
  double T[9];
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) {
      double s = 0.0;
      for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++) s += A[I4(i, j, m, n)] * B[I2(m, n)];
      T[I2(i, j)] = s;
    }
  memcpy(C, T, sizeof T);
}

/* math.f90:736-744 math_3333to99: M99(r,c) with r = 3a+b, c = 3c'+d (0-based),
 * stored column-major as a 9x9 LAPACK matrix */
void t3333_to99(const double *T, double *M) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) M[(3 * a + b) + 9 * (3 * c + d)] = T[I4(a, b, c, d)];
}

/* math.f90:764-772 math_99to3333 */
void t99_to3333(const double *M, double *T) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) T[I4(a, b, c, d)] = M[(3 * a + b) + 9 * (3 * c + d)];
}

/* math.f90:680-688 math_33to9 (row-major flatten) and 708-716 math_9to33 */
void m33_to9(const double *A, double *v) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) v[3 * a + b] = A[I2(a, b)];
}
void m9_to33(const double *v, double *A) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) A[I2(a, b)] = v[3 * a + b];
}

/* math.f90:792-802 math_sym33to6 (Mandel) */
void m33_sym_to6(const double *m, double *t) {
  const double W = sqrt(2.0);
  t[0] = m[I2(0, 0)];
  t[1] = m[I2(1, 1)];
  t[2] = m[I2(2, 2)];
  t[3] = m[I2(2, 1)] * W;
  t[4] = m[I2(2, 0)] * W;
  t[5] = m[I2(1, 0)] * W;
}

/* math.f90:808-820 math_6toSym33 */
void m6_to_sym33(const double *t, double *m) {
  const double IW = 1.0 / sqrt(2.0);
  m[I2(0, 0)] = t[0];
  m[I2(1, 0)] = t[5] * IW;
  m[I2(2, 0)] = t[4] * IW;
  m[I2(0, 1)] = t[5] * IW;
  m[I2(1, 1)] = t[1];
  m[I2(2, 1)] = t[3] * IW;
  m[I2(0, 2)] = t[4] * IW;
  m[I2(1, 2)] = t[3] * IW;
  m[I2(2, 2)] = t[2];
}

/* math.f90:851-904 math_66toSym3333 */
void m66_to_sym3333(const double C[6][6], double *T) {
  const double IW1 = 1.0 / sqrt(2.0), IW2 = 1.0 / 2.0;
  /* for Mandel column c, the 3x3 block built from rows 1..6 of column c */
  for (int c = 0; c < 6; c++) {
    double w_d = c < 3 ? 1.0 : IW1;   /* weight of (i,i) rows */
    double w_s = c < 3 ? IW1 : IW2;   /* weight of (i,j) rows, i /= j */
    double B[9];
    B[I2(0, 0)] = C[0][c] * w_d;
    B[I2(1, 0)] = C[5][c] * w_s;
    B[I2(2, 0)] = C[4][c] * w_s;
    B[I2(0, 1)] = C[5][c] * w_s;
    B[I2(1, 1)] = C[1][c] * w_d;
    B[I2(2, 1)] = C[3][c] * w_s;
    B[I2(0, 2)] = C[4][c] * w_s;
    B[I2(1, 2)] = C[3][c] * w_s;
    B[I2(2, 2)] = C[2][c] * w_d;
    int kl[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
    int k = kl[c][0], l = kl[c][1];
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        T[I4(i, j, k, l)] = B[I2(i, j)];
        T[I4(i, j, l, k)] = B[I2(i, j)];
      }
  }
}

/* ------------------------------------------------------------------ rotations */
static const double ROT_P = -1.0;   /* rotations.f90:57 */

/* rotations.f90:776-791 rotate_vector */
static void rotate_vector(const double *v, const double *q, double *out) {
  double A = q[0] * q[0] - (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  double B = 2.0 * (q[1] * v[0] + q[2] * v[1] + q[3] * v[2]);
  double C = 2.0 * ROT_P * q[0];
  double cx = q[2] * v[2] - q[3] * v[1];
  double cy = q[3] * v[0] - q[1] * v[2];
  double cz = q[1] * v[1] - q[2] * v[0];
  out[0] = A * v[0] + B * q[1] + C * cx;
  out[1] = A * v[1] + B * q[2] + C * cy;
  out[2] = A * v[2] + B * q[3] + C * cz;
}

void orc_rotate_vector(const double q[4], const double v[3], int active, double out[3]) {
  double qq[4] = {q[0], active ? -q[1] : q[1], active ? -q[2] : q[2], active ? -q[3] : q[3]};
  rotate_vector(v, qq, out);
}

/* rotations.f90:282-301 rotate3x3 */
void orc_rotate33(const double q[4], const double T[9], int active, double out[9]) {
  double qq[4] = {q[0], active ? -q[1] : q[1], active ? -q[2] : q[2], active ? -q[3] : q[3]};
  double R[9];
  for (int i = 0; i < 3; i++) {
    double row[3] = {T[I2(i, 0)], T[I2(i, 1)], T[I2(i, 2)]}, o[3];
    rotate_vector(row, qq, o);
    R[I2(i, 0)] = o[0]; R[I2(i, 1)] = o[1]; R[I2(i, 2)] = o[2];
  }
  for (int j = 0; j < 3; j++) {
    double col[3] = {R[I2(0, j)], R[I2(1, j)], R[I2(2, j)]}, o[3];
    rotate_vector(col, qq, o);
    R[I2(0, j)] = o[0]; R[I2(1, j)] = o[1]; R[I2(2, j)] = o[2];
  }
  memcpy(out, R, sizeof R);
}

/* rotations.f90:310-339 rotate3x3x3x3 */
void orc_rotate3333(const double q[4], const double T[81], int active, double out[81]) {
  double qq[4] = {q[0], active ? -q[1] : q[1], active ? -q[2] : q[2], active ? -q[3] : q[3]};
  double R[81];
  memcpy(R, T, sizeof R);
  double v[3], o[3];
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) for (int k = 0; k < 3; k++) {
    for (int x = 0; x < 3; x++) v[x] = R[I4(i, j, k, x)];
    rotate_vector(v, qq, o);
    for (int x = 0; x < 3; x++) R[I4(i, j, k, x)] = o[x];
  }
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) for (int l = 0; l < 3; l++) {
    for (int x = 0; x < 3; x++) v[x] = R[I4(i, j, x, l)];
    rotate_vector(v, qq, o);
    for (int x = 0; x < 3; x++) R[I4(i, j, x, l)] = o[x];
  }
  for (int i = 0; i < 3; i++) for (int k = 0; k < 3; k++) for (int l = 0; l < 3; l++) {
    for (int x = 0; x < 3; x++) v[x] = R[I4(i, x, k, l)];
    rotate_vector(v, qq, o);
    for (int x = 0; x < 3; x++) R[I4(i, x, k, l)] = o[x];
  }
  for (int j = 0; j < 3; j++) for (int k = 0; k < 3; k++) for (int l = 0; l < 3; l++) {
    for (int x = 0; x < 3; x++) v[x] = R[I4(x, j, k, l)];
    rotate_vector(v, qq, o);
    for (int x = 0; x < 3; x++) R[I4(x, j, k, l)] = o[x];
  }
  memcpy(out, R, sizeof R);
}

/* rotations.f90:399-423 qu2om */
void orc_qu2om(const double qu[4], double om[9]) {
  double qq = qu[0] * qu[0] - (qu[1] * qu[1] + qu[2] * qu[2] + qu[3] * qu[3]);
  double o[9];
  o[I2(0, 0)] = qq + 2.0 * qu[1] * qu[1];
  o[I2(1, 1)] = qq + 2.0 * qu[2] * qu[2];
  o[I2(2, 2)] = qq + 2.0 * qu[3] * qu[3];
  o[I2(0, 1)] = 2.0 * (qu[1] * qu[2] - qu[0] * qu[3]);
  o[I2(1, 2)] = 2.0 * (qu[2] * qu[3] - qu[0] * qu[1]);
  o[I2(2, 0)] = 2.0 * (qu[3] * qu[1] - qu[0] * qu[2]);
  o[I2(1, 0)] = 2.0 * (qu[2] * qu[1] + qu[0] * qu[3]);
  o[I2(2, 1)] = 2.0 * (qu[3] * qu[2] + qu[0] * qu[1]);
  o[I2(0, 2)] = 2.0 * (qu[1] * qu[3] + qu[0] * qu[2]);
  if (ROT_P < 0.0) m33_transpose(o, o);
  double d = cbrt(m33_det(o));
  d = pow(m33_det(o), 1.0 / 3.0);
  for (int k = 0; k < 9; k++) om[k] = o[k] / d;
}

/* rotations.f90:486-514 om2qu */
void orc_om2qu(const double om[9], double qu[4]) {
  double trace = om[I2(0, 0)] + om[I2(1, 1)] + om[I2(2, 2)];
  double s;
  if (trace > 0.0) {
    s = 0.5 / sqrt(trace + 1.0);
    qu[0] = 0.25 / s;
    qu[1] = (om[I2(2, 1)] - om[I2(1, 2)]) * s;
    qu[2] = (om[I2(0, 2)] - om[I2(2, 0)]) * s;
    qu[3] = (om[I2(1, 0)] - om[I2(0, 1)]) * s;
  } else if (om[I2(0, 0)] > om[I2(1, 1)] && om[I2(0, 0)] > om[I2(2, 2)]) {
    s = 2.0 * sqrt(1.0 + om[I2(0, 0)] - om[I2(1, 1)] - om[I2(2, 2)]);
    qu[0] = (om[I2(2, 1)] - om[I2(1, 2)]) / s;
    qu[1] = 0.25 * s;
    qu[2] = (om[I2(0, 1)] + om[I2(1, 0)]) / s;
    qu[3] = (om[I2(0, 2)] + om[I2(2, 0)]) / s;
  } else if (om[I2(1, 1)] > om[I2(2, 2)]) {
    s = 2.0 * sqrt(1.0 + om[I2(1, 1)] - om[I2(0, 0)] - om[I2(2, 2)]);
    qu[0] = (om[I2(0, 2)] - om[I2(2, 0)]) / s;
    qu[1] = (om[I2(0, 1)] + om[I2(1, 0)]) / s;
    qu[2] = 0.25 * s;
    qu[3] = (om[I2(1, 2)] + om[I2(2, 1)]) / s;
  } else {
    s = 2.0 * sqrt(1.0 + om[I2(2, 2)] - om[I2(0, 0)] - om[I2(1, 1)]);
    qu[0] = (om[I2(1, 0)] - om[I2(0, 1)]) / s;
    qu[1] = (om[I2(0, 2)] + om[I2(2, 0)]) / s;
    qu[2] = (om[I2(1, 2)] + om[I2(2, 1)]) / s;
    qu[3] = 0.25 * s;
  }
  if (copysign(1.0, qu[0]) < 0.0)
    for (int k = 0; k < 4; k++) qu[k] = -qu[k];
  for (int k = 1; k < 4; k++)
    if (!orc_dEq0(qu[k])) qu[k] *= ROT_P;
  double n = sqrt(qu[0] * qu[0] + qu[1] * qu[1] + qu[2] * qu[2] + qu[3] * qu[3]);
  for (int k = 0; k < 4; k++) qu[k] /= n;
}

/* rotations.f90:670-686 ax2qu, with fromAxisAngle's P handling (:218-236) */
void orc_ax2qu(const double ax[4], int degrees, int P, double qu[4]) {
  double angle = degrees ? ax[3] * (2.0 * acos(-1.0) / 360.0) : ax[3];
  double sgn = (P == 1) ? -1.0 : 1.0;
  if (orc_dEq0(angle)) {
    qu[0] = 1.0; qu[1] = qu[2] = qu[3] = 0.0;
    return;
  }
  double c = cos(angle * 0.5), s = sin(angle * 0.5);
  qu[0] = c;
  qu[1] = ax[0] * sgn * s;
  qu[2] = ax[1] * sgn * s;
  qu[3] = ax[2] * sgn * s;
}

/* math.f90:1108-1119 math_invariantsSym33 (with math_detSym33) */
static void invariants_sym33(const double *m, double *I) {
  I[0] = m[I2(0, 0)] + m[I2(1, 1)] + m[I2(2, 2)];
  I[1] = m[I2(0, 0)] * m[I2(1, 1)] + m[I2(0, 0)] * m[I2(2, 2)] + m[I2(1, 1)] * m[I2(2, 2)]
       - (m[I2(0, 1)] * m[I2(0, 1)] + m[I2(0, 2)] * m[I2(0, 2)] + m[I2(1, 2)] * m[I2(1, 2)]);
  I[2] = -(m[I2(0, 0)] * m[I2(1, 2)] * m[I2(1, 2)] + m[I2(1, 1)] * m[I2(0, 2)] * m[I2(0, 2)]
           + m[I2(2, 2)] * m[I2(0, 1)] * m[I2(0, 1)])
       + m[I2(0, 0)] * m[I2(1, 1)] * m[I2(2, 2)] + 2.0 * m[I2(0, 1)] * m[I2(0, 2)] * m[I2(1, 2)];
}

static double clipl(double a, double l) { return a < l ? l : a; }

/* math.f90:1004-1043 math_rotationalPart */
void orc_rotational_part(const double F[9], double R[9]) {
  const double TAU = 2.0 * acos(-1.0);
  double Ft[9], C[9], FF[9], FC[9];
  m33_transpose(F, Ft);
  m33_mul(Ft, F, C);
  double I_C[3];
  invariants_sym33(C, I_C);
  m33_mul(F, F, FF);
  double trF = F[0] + F[4] + F[8];
  double trFF = FF[0] + FF[4] + FF[8];
  double I_F[2] = {trF, 0.5 * (trF * trF - trFF)};
  double lambda[3];
  double x = pow(clipl(I_C[0] * I_C[0] - 3.0 * I_C[1], 0.0), 3.0 / 2.0);
  if (!orc_dEq0(x)) {
    double arg = (I_C[0] * I_C[0] * I_C[0] - 4.5 * I_C[0] * I_C[1] + 13.5 * I_C[2]) / x;
    if (arg < -1.0) arg = -1.0;
    if (arg > 1.0) arg = 1.0;
    double Phi = acos(arg);
    for (int k = 0; k < 3; k++) {
      lambda[k] = I_C[0] + (2.0 * sqrt(clipl(I_C[0] * I_C[0] - 3.0 * I_C[1], 0.0)))
                  * cos((Phi - TAU * (double)(k + 1)) / 3.0);
      lambda[k] = sqrt(clipl(lambda[k], 0.0) / 3.0);
    }
  } else {
    for (int k = 0; k < 3; k++) lambda[k] = sqrt(I_C[0] / 3.0);
  }
  double I_U[3] = {lambda[0] + lambda[1] + lambda[2],
                   lambda[0] * lambda[1] + lambda[1] * lambda[2] + lambda[2] * lambda[0],
                   lambda[0] * lambda[1] * lambda[2]};
  double FFt[9];
  m33_transpose(FF, FFt);
  m33_mul(F, C, FC);
  double Rt[9];
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++)
      Rt[I2(i, j)] = I_U[0] * I_F[1] * (i == j ? 1.0 : 0.0) + (I_U[0] * I_U[0] - I_U[1]) * F[I2(i, j)]
                   - I_U[0] * I_F[0] * Ft[I2(i, j)] + I_U[0] * FFt[I2(i, j)] - FC[I2(i, j)];
  double d = pow(m33_det(Rt), -1.0 / 3.0);
  for (int k = 0; k < 9; k++) R[k] = Rt[k] * d;
}

/* ------------------------------------------------------------------ crystal */
/* slip systems, direction then plane normal, in the order of crystal.f90:42-60 (cF)
 * and crystal.f90:122-175 (cI).  Crystallographic data: identical by necessity. */
static const double CF_SLIP[18][6] = {
  { 0, 1,-1,  1, 1, 1}, {-1, 0, 1,  1, 1, 1}, { 1,-1, 0,  1, 1, 1},
  { 0,-1,-1, -1,-1, 1}, { 1, 0, 1, -1,-1, 1}, {-1, 1, 0, -1,-1, 1},
  { 0,-1, 1,  1,-1,-1}, {-1, 0,-1,  1,-1,-1}, { 1, 1, 0,  1,-1,-1},
  { 0, 1, 1, -1, 1,-1}, { 1, 0,-1, -1, 1,-1}, {-1,-1, 0, -1, 1,-1},
  { 1, 1, 0,  1,-1, 0}, { 1,-1, 0,  1, 1, 0}, { 1, 0, 1,  1, 0,-1},
  { 1, 0,-1,  1, 0, 1}, { 0, 1, 1,  0, 1,-1}, { 0, 1,-1,  0, 1, 1}};
static const double CI_SLIP[48][6] = {
  { 1,-1, 1,  0, 1, 1}, { 1,-1, 1,  1, 0,-1}, { 1,-1, 1, -1,-1, 0},
  {-1,-1, 1,  0,-1,-1}, {-1,-1, 1,  1, 0, 1}, {-1,-1, 1, -1, 1, 0},
  { 1, 1, 1,  0, 1,-1}, { 1, 1, 1, -1, 0, 1}, { 1, 1, 1,  1,-1, 0},
  {-1, 1, 1,  0,-1, 1}, {-1, 1, 1, -1, 0,-1}, {-1, 1, 1,  1, 1, 0},
  { 1,-1, 1,  2, 1,-1}, { 1,-1, 1, -1, 1, 2}, { 1,-1, 1,  1, 2, 1},
  {-1,-1, 1,  2,-1, 1}, {-1,-1, 1,  1, 1, 2}, {-1,-1, 1, -1, 2, 1},
  { 1, 1, 1,  1, 1,-2}, { 1, 1, 1,  1,-2, 1}, { 1, 1, 1, -2, 1, 1},
  {-1, 1, 1,  1,-1, 2}, {-1, 1, 1,  1, 2,-1}, {-1, 1, 1,  2, 1, 1},
  { 1,-1, 1, -1, 2, 3}, { 1,-1, 1,  1, 3, 2}, { 1,-1, 1, -2, 1, 3},
  { 1,-1, 1,  2, 3, 1}, { 1,-1, 1,  3, 1,-2}, { 1,-1, 1,  3, 2,-1},
  {-1,-1, 1,  1, 2, 3}, {-1,-1, 1, -1, 3, 2}, {-1,-1, 1,  2, 1, 3},
  {-1,-1, 1, -2, 3, 1}, {-1,-1, 1,  3,-1, 2}, {-1,-1, 1,  3,-2, 1},
  { 1, 1, 1,  1, 2,-3}, { 1, 1, 1,  1,-3, 2}, { 1, 1, 1,  2, 1,-3},
  { 1, 1, 1,  2,-3, 1}, { 1, 1, 1, -3, 1, 2}, { 1, 1, 1, -3, 2, 1},
  {-1, 1, 1,  1,-2, 3}, {-1, 1, 1,  1, 3,-2}, {-1, 1, 1,  2,-1, 3},
  {-1, 1, 1,  2, 3,-1}, {-1, 1, 1,  3, 1, 2}, {-1, 1, 1,  3, 2, 1}};
static const int CF_NSLIPMAX[2] = {12, 6};
static const int CI_NSLIPMAX[3] = {12, 12, 24};

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static double norm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static void outer3(const double *a, const double *b, double *M) {
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) M[I2(i, j)] = a[i] * b[j];
}

/* crystal.f90:1348-1454 crystal_SchmidMatrix_slip with buildCoordinateSystem
 * (crystal.f90:1937-2001).  P is [nsl][9] column-major 3x3. */
void orc_schmid_slip(int lattice, int n_fam, const int *N_sl, int sense,
                     int n_nS, const double *a_nS, double *P) {
  const double (*sys)[6] = lattice == ORC_LATTICE_CF ? CF_SLIP : CI_SLIP;
  const int *maxs = lattice == ORC_LATTICE_CF ? CF_NSLIPMAX : CI_NSLIPMAX;
  int a = 0;
  for (int f = 0; f < n_fam; f++) {
    int off = 0;
    for (int g = 0; g < f; g++) off += maxs[g];
    for (int s = 0; s < N_sl[f]; s++, a++) {
      const double *row = sys[off + s];
      double d[3] = {row[0], row[1], row[2]}, n[3] = {row[3], row[4], row[5]};
      double nd = norm3(d), nn = norm3(n);
      for (int k = 0; k < 3; k++) { d[k] /= nd; n[k] /= nn; }
      for (int k = 0; k < 3; k++) d[k] *= (double)sense;
      double *M = P + 9 * a;
      outer3(d, n, M);
      if (n_nS > 0 && lattice == ORC_LATTICE_CI && f == 0) {
        double c[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < n_nS && k < 6; k++) c[k] = a_nS[k];
        double ax[4] = {d[0], d[1], d[2], 60.0}, q[4], np[3];
        orc_ax2qu(ax, 1, 1, q);
        rotate_vector(n, q, np);
        double nxd[3], npxd[3], T[9];
        cross3(n, d, nxd);
        cross3(np, d, npxd);
        outer3(d, np, T);    for (int k = 0; k < 9; k++) M[k] += c[0] * T[k];
        outer3(nxd, n, T);   for (int k = 0; k < 9; k++) M[k] += c[1] * T[k];
        outer3(npxd, np, T); for (int k = 0; k < 9; k++) M[k] += c[2] * T[k];
        outer3(n, n, T);     for (int k = 0; k < 9; k++) M[k] += c[3] * T[k];
        outer3(nxd, nxd, T); for (int k = 0; k < 9; k++) M[k] += c[4] * T[k];
        outer3(d, d, T);     for (int k = 0; k < 9; k++) M[k] += c[5] * T[k];
      }
    }
  }
}

/* slip-slip interaction types, rows = reacting system, columns = acting system,
 * values as tabulated in crystal.f90:603-620 (cF) and crystal.f90:639-690 (cI). */
static const unsigned char CF_INT[18][18] = {
  { 1, 2, 2, 4, 7, 6, 3, 5, 5, 4, 6, 7, 10,11,10,11,12,13},
  { 2, 1, 2, 7, 4, 6, 6, 4, 7, 5, 3, 5, 10,11,12,13,10,11},
  { 2, 2, 1, 5, 5, 3, 6, 7, 4, 7, 6, 4, 12,13,10,11,10,11},
  { 4, 7, 6, 1, 2, 2, 4, 6, 7, 3, 5, 5, 10,11,11,10,13,12},
  { 7, 4, 6, 2, 1, 2, 5, 3, 5, 6, 4, 7, 10,11,13,12,11,10},
  { 5, 5, 3, 2, 2, 1, 7, 6, 4, 6, 7, 4, 12,13,11,10,11,10},
  { 3, 5, 5, 4, 6, 7, 1, 2, 2, 4, 7, 6, 11,10,11,10,12,13},
  { 6, 4, 7, 5, 3, 5, 2, 1, 2, 7, 4, 6, 11,10,13,12,10,11},
  { 6, 7, 4, 7, 6, 4, 2, 2, 1, 5, 5, 3, 13,12,11,10,10,11},
  { 4, 6, 7, 3, 5, 5, 4, 7, 6, 1, 2, 2, 11,10,10,11,13,12},
  { 5, 3, 5, 6, 4, 7, 7, 4, 6, 2, 1, 2, 11,10,12,13,11,10},
  { 7, 6, 4, 6, 7, 4, 5, 5, 3, 2, 2, 1, 13,12,10,11,11,10},
  {10,10,12,10,10,12,11,11,13,11,11,13,  1, 8, 9, 9, 9, 9},
  {11,11,13,11,11,13,10,10,12,10,10,12,  8, 1, 9, 9, 9, 9},
  {10,12,10,11,13,11,11,13,11,10,12,10,  9, 9, 1, 8, 9, 9},
  {11,13,11,10,12,10,10,12,10,11,13,11,  9, 9, 8, 1, 9, 9},
  {12,10,10,13,11,11,12,10,10,13,11,11,  9, 9, 9, 9, 1, 8},
  {13,11,11,12,10,10,13,11,11,12,10,10,  9, 9, 9, 9, 8, 1}};

/* cI, first 24 systems ({110} and {112} families) fully tabulated; the {123}
 * family interacts through types 25..28 only (generated below). */
static const unsigned char CI_INT24[24][24] = {
  { 1, 2, 2, 3, 5, 5, 6, 4, 7, 6, 7, 4,   8, 9, 9,11,13,13,14,17,18,17,14,18},
  { 2, 1, 2, 4, 6, 7, 5, 3, 5, 7, 6, 4,   9, 9, 8,17,14,18,13,11,13,17,18,14},
  { 2, 2, 1, 4, 7, 6, 7, 4, 6, 5, 5, 3,   9, 8, 9,17,18,14,18,17,14,11,13,13},
  { 3, 5, 5, 1, 2, 2, 6, 7, 4, 6, 4, 7,  11,13,13, 8, 9, 9,17,14,18,14,17,18},
  { 4, 6, 7, 2, 1, 2, 7, 6, 4, 5, 3, 5,  17,14,18, 9, 9, 8,17,18,14,13,11,13},
  { 4, 7, 6, 2, 2, 1, 5, 5, 3, 7, 4, 6,  17,18,14, 9, 8, 9,11,13,13,18,17,14},
  { 6, 4, 7, 6, 7, 4, 1, 2, 2, 3, 5, 5,  18,14,17,18,17,14, 9, 9, 8,13,13,11},
  { 5, 3, 5, 7, 6, 4, 2, 1, 2, 4, 6, 7,  13,13,11,14,17,18, 9, 8, 9,14,18,17},
  { 7, 4, 6, 5, 5, 3, 2, 2, 1, 4, 7, 6,  14,18,17,13,11,13, 8, 9, 9,18,14,17},
  { 6, 7, 4, 6, 4, 7, 3, 5, 5, 1, 2, 2,  18,17,14,18,14,17,13,13,11, 9, 9, 8},
  { 7, 6, 4, 5, 3, 5, 4, 6, 7, 2, 1, 2,  14,17,18,13,13,11,14,18,17, 9, 8, 9},
  { 5, 5, 3, 7, 4, 6, 4, 7, 6, 2, 2, 1,  13,11,13,14,18,17,18,14,17, 8, 9, 9},
  { 8, 9, 9,10,16,16,19,12,15,19,15,12,   1, 2, 2,20,21,21,23,22,24,22,23,24},
  { 9, 9, 8,12,15,19,15,12,19,16,16,10,   2, 1, 2,22,24,23,24,22,23,20,21,21},
  { 9, 8, 9,12,19,15,16,10,16,15,19,12,   2, 2, 1,22,23,24,21,20,21,22,24,23},
  {10,16,16, 8, 9, 9,19,15,12,19,12,15,  20,21,21, 1, 2, 2,22,23,24,23,22,24},
  {12,15,19, 9, 9, 8,16,16,10,15,12,19,  22,24,23, 2, 1, 2,20,21,21,24,22,23},
  {12,19,15, 9, 8, 9,15,19,12,16,10,16,  21,23,24, 2, 2, 1,22,24,23,21,20,21},
  {15,12,19,16,16,10, 9, 9, 8,12,15,19,  23,24,22,21,20,21, 1, 2, 2,24,23,22},
  {16,10,16,15,19,12, 9, 8, 9,12,19,15,  22,21,20,23,22,24, 2, 1, 2,23,24,22},
  {19,12,15,19,15,12, 8, 9, 9,10,16,16,  24,23,22,24,22,23, 2, 2, 1,21,21,20},
  {16,16,10,15,12,19,12,15,19, 9, 9, 8,  21,20,21,23,24,22,24,23,22, 1, 2, 2},
  {15,19,12,16,10,16,12,19,15, 9, 8, 9,  23,22,24,21,21,20,23,24,22, 2, 1, 2},
  {19,15,12,19,12,15,10,16,16, 8, 9, 9,  24,22,23,24,23,22,21,21,20, 2, 2, 1}};

/* interaction type between two cI systems (0-based), crystal.f90:639-690 */
static int ci_type(int r, int a) {
  if (r < 24 && a < 24) return CI_INT24[r][a];
  if (r >= 24 && a >= 24) {
    if (r == a) return 1;
    return ((r - 24) / 6 == (a - 24) / 6) ? 27 : 28;
  }
  int s123 = r >= 24 ? r : a, o = r >= 24 ? a : r;   /* o: system of {110} or {112} */
  int grp = (s123 - 24) / 6;                          /* 4 groups of 6, by <111> direction */
  int og = (o % 12) / 3;                              /* 4 groups of 3 in each family */
  if (grp != og) return 28;
  return o < 12 ? 25 : 26;
}

/* crystal.f90:591-902 crystal_interaction_SlipBySlip via buildInteraction
 * (crystal.f90:1893-1930).  H[r*nsl + a] = values(type(r,a)). */
void orc_interaction_slipslip(int lattice, int n_fam, const int *N_sl,
                              int n_val, const double *val, double *H) {
  const int *maxs = lattice == ORC_LATTICE_CF ? CF_NSLIPMAX : CI_NSLIPMAX;
  int nsl = 0;
  for (int f = 0; f < n_fam; f++) nsl += N_sl[f];
  int idx_used[ORC_NSL_MAX];
  int a = 0;
  for (int f = 0; f < n_fam; f++) {
    int off = 0;
    for (int g = 0; g < f; g++) off += maxs[g];
    for (int s = 0; s < N_sl[f]; s++) idx_used[a++] = off + s;
  }
  for (int r = 0; r < nsl; r++)
    for (int c = 0; c < nsl; c++) {
      int t = lattice == ORC_LATTICE_CF ? CF_INT[idx_used[r]][idx_used[c]]
                                        : ci_type(idx_used[r], idx_used[c]);
      H[r * nsl + c] = (t <= n_val) ? val[t - 1] : NAN;
    }
}

/* crystal.f90:1711-1759 crystal_assembleStiffness (cubic) */
void orc_cubic_stiffness(double C11, double C12, double C44, double C[6][6]) {
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) C[i][j] = 0.0;
  C[0][0] = C11;
  C[1][1] = C11;
  C[0][1] = C12;
  C[3][3] = C44 * 2.0;
  C[4][4] = C44 * 2.0;
  C[2][2] = C11;
  C[0][2] = C12;
  C[1][2] = C12;
  C[5][5] = C44 * 2.0;
  for (int i = 0; i < 3; i++)
    for (int j = i + 1; j < 3; j++) C[j][i] = C[i][j];
}
