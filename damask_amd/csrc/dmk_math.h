// Register-resident 3x3 algebra for the per-cell kernels (row-major: T[3 i + j] = T_ij).
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#define DI __device__ __forceinline__
// pure register algebra, also compiled for the host (adjoint point checks)
#define DHI __host__ __device__ __forceinline__

// Per-cell field access.  Fields are component-major ([c][N]); a lane reaches
// its cell through a 32-bit byte offset (cell index * 8, so N < 2^29) added to
// a wave-uniform component base.  That maps to global loads/stores with a
// scalar base register: no 64-bit per-lane address has to stay live in VGPRs.
struct Lane {
  size_t sN;        // field stride (cells)
  uint32_t off;     // byte offset of the lane's cell
};

DI Lane make_lane(int N, int e) { return Lane{(size_t)N, (uint32_t)e * 8u}; }

// The offset is passed through an empty asm per access so the optimiser cannot
// re-associate (base_c + off) into (base + off) + c N, a per-lane 64-bit
// pointer shared by all components.
DI uint32_t lane_off(const Lane& l) {
  uint32_t o = l.off;
  asm volatile("" : "+v"(o));
  return o;
}

DI double ld(const double* f, const Lane& l, int c) {
  return *(const double*)((const char*)(f + (size_t)c * l.sN) + lane_off(l));
}

DI void st(double* f, const Lane& l, int c, double v) {
  *(double*)((char*)(f + (size_t)c * l.sN) + lane_off(l)) = v;
}

// T_ij (row-major in registers) <-> component i + 3 j of the field
DI void m_load(const double* f, const Lane& l, double* T) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) T[3 * i + j] = ld(f, l, i + 3 * j);
}

DI void m_store(double* f, const Lane& l, const double* T) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) st(f, l, i + 3 * j, T[3 * i + j]);
}

// Make a value opaque to the optimiser (no code emitted): address arithmetic
// derived from it cannot be hoisted out of the enclosing loop.
template <class T>
DI T* opaque(T* p) {
  asm volatile("" : "+v"(p));
  return p;
}

DHI void m_mul(const double* A, const double* B, double* C) {
  double T[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double s = A[3 * i] * B[j];
      s = fma(A[3 * i + 1], B[3 + j], s);
      s = fma(A[3 * i + 2], B[6 + j], s);
      T[3 * i + j] = s;
    }
#pragma unroll
  for (int k = 0; k < 9; k++) C[k] = T[k];
}

// C = A^T B
DHI void m_mul_tn(const double* A, const double* B, double* C) {
  double T[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double s = A[i] * B[j];
      s = fma(A[3 + i], B[3 + j], s);
      s = fma(A[6 + i], B[6 + j], s);
      T[3 * i + j] = s;
    }
#pragma unroll
  for (int k = 0; k < 9; k++) C[k] = T[k];
}

// C = A B^T
DHI void m_mul_nt(const double* A, const double* B, double* C) {
  double T[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double s = A[3 * i] * B[3 * j];
      s = fma(A[3 * i + 1], B[3 * j + 1], s);
      s = fma(A[3 * i + 2], B[3 * j + 2], s);
      T[3 * i + j] = s;
    }
#pragma unroll
  for (int k = 0; k < 9; k++) C[k] = T[k];
}

DHI void m_transpose(const double* A, double* B) {
  double T[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) T[3 * j + i] = A[3 * i + j];
#pragma unroll
  for (int k = 0; k < 9; k++) B[k] = T[k];
}

// math_det33 (src/math.f90:630-638)
DHI double m_det(const double* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6])
       + m[2] * (m[3] * m[7] - m[4] * m[6]);
}

DHI bool dEq0(double a) { return fabs(a) <= DBL_MIN * 10.0; }

// math_invert33 (src/math.f90:471-505): Cramer; returns true on error (zeros out)
DHI bool m_invert(const double* A, double* inv) {
  double c00 = A[4] * A[8] - A[5] * A[7];
  double c10 = -A[3] * A[8] + A[5] * A[6];
  double c20 = A[3] * A[7] - A[4] * A[6];
  double det = A[0] * c00 + A[1] * c10 + A[2] * c20;
  if (dEq0(det)) {
#pragma unroll
    for (int k = 0; k < 9; k++) inv[k] = 0.0;
    return true;
  }
  double c01 = -A[1] * A[8] + A[2] * A[7];
  double c11 = A[0] * A[8] - A[2] * A[6];
  double c21 = -A[0] * A[7] + A[1] * A[6];
  double c02 = A[1] * A[5] - A[2] * A[4];
  double c12 = -A[0] * A[5] + A[2] * A[3];
  double c22 = A[0] * A[4] - A[1] * A[3];
  double r = 1.0 / det;
  inv[0] = c00 * r; inv[1] = c01 * r; inv[2] = c02 * r;
  inv[3] = c10 * r; inv[4] = c11 * r; inv[5] = c12 * r;
  inv[6] = c20 * r; inv[7] = c21 * r; inv[8] = c22 * r;
  return false;
}

DHI double m_norm(const double* A) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) s = fma(A[k], A[k], s);
  return sqrt(s);
}

// C : X for cubic stiffness in the lattice frame, X symmetric, through the
// Mandel vector (src/math.f90:792-820 and src/crystal.f90:1711-1759)
DHI void hooke_sym(double C11, double C12, double C44, const double* X, double* S) {
  const double W = 1.4142135623730951;          // sqrt(2)
  const double IW = 0.7071067811865475;         // 1/sqrt(2)
  double v1 = X[0], v2 = X[4], v3 = X[8];
  double v4 = X[7] * W, v5 = X[6] * W, v6 = X[3] * W;
  double s1 = C11 * v1 + C12 * v2 + C12 * v3;
  double s2 = C12 * v1 + C11 * v2 + C12 * v3;
  double s3 = C12 * v1 + C12 * v2 + C11 * v3;
  double G = C44 * 2.0;
  double s4 = G * v4, s5 = G * v5, s6 = G * v6;
  S[0] = s1;      S[1] = s6 * IW; S[2] = s5 * IW;
  S[3] = s6 * IW; S[4] = s2;      S[5] = s4 * IW;
  S[6] = s5 * IW; S[7] = s4 * IW; S[8] = s3;
}
