// Rotation of 3x3 tensors by a unit quaternion, usable on host and device
// (rotations.f90:148-172 rotVector, :282-301 rotate3x3; P = -1 convention).
// Tensors are column-major, T_ij at [i + 3 j].
#pragma once

#include <hip/hip_runtime.h>

#define DMK_RHD __host__ __device__ inline

DMK_RHD void dmk_rot_vec(const double* q, const double* v, double* o) {
  double A = q[0] * q[0] - (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  double B = 2.0 * (q[1] * v[0] + q[2] * v[1] + q[3] * v[2]);
  double C = -2.0 * q[0];
  double c[3] = {q[2] * v[2] - q[3] * v[1], q[3] * v[0] - q[1] * v[2], q[1] * v[1] - q[2] * v[0]};
  for (int k = 0; k < 3; k++) o[k] = A * v[k] + B * q[k + 1] + C * c[k];
}

// active rotation conjugates q
DMK_RHD void dmk_rotate33(const double* q0, const double* T, bool active, double* out) {
  double q[4] = {q0[0], active ? -q0[1] : q0[1], active ? -q0[2] : q0[2], active ? -q0[3] : q0[3]};
  double R[9], v[3], o[3];
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) v[j] = T[i + 3 * j];
    dmk_rot_vec(q, v, o);
    for (int j = 0; j < 3; j++) R[i + 3 * j] = o[j];
  }
  for (int j = 0; j < 3; j++) {
    for (int i = 0; i < 3; i++) v[i] = R[i + 3 * j];
    dmk_rot_vec(q, v, o);
    for (int i = 0; i < 3; i++) R[i + 3 * j] = o[i];
  }
  for (int k = 0; k < 9; k++) out[k] = R[k];
}
