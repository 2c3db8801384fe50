/* CPU oracle internals -- TEST INFRASTRUCTURE ONLY (see oracle.h). */
#ifndef DAMASK_ORACLE_INTERNAL_H
#define DAMASK_ORACLE_INTERNAL_H

#include "oracle.h"

#define I2(i, j) ((i) + 3 * (j))
#define I4(i, j, k, l) ((i) + 3 * (j) + 9 * (k) + 27 * (l))

int orc_dEq0(double a);
int orc_dEq(double a, double b, double tol);
void m33_mul(const double *A, const double *B, double *C);
void m33_transpose(const double *A, double *B);
void m33_eye(double *A);
double m33_norm2(const double *A);
double m33_det(const double *m);
int m33_invert(const double *A, double *InvA, double *DetA);
void m33_inv(const double *A, double *B);
int lu_factor(int n, double *A, int *ipiv);
void lu_solve(int n, const double *LU, const int *ipiv, double *b);
int dgesv_n(int n, double *A, double *b);
void t3333_mul(const double *A, const double *B, double *C);
void t3333_mul33(const double *A, const double *B, double *C);
void t3333_to99(const double *T, double *M);
void t99_to3333(const double *M, double *T);
void m33_to9(const double *A, double *v);
void m9_to33(const double *v, double *A);
void m33_sym_to6(const double *m, double *t);
void m6_to_sym33(const double *t, double *m);
void m66_to_sym3333(const double C[6][6], double *T);
void orc_rotate_vector(const double q[4], const double v[3], int active, double out[3]);
void orc_ax2qu(const double ax[4], int degrees, int P, double qu[4]);
void orc_cubic_stiffness(double C11, double C12, double C44, double C[6][6]);

#endif
