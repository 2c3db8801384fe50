// Element steps of the solver's vector kernels, shared by the single-RVE
// kernels and the batched ones of a lockstep round so that both round alike.
#pragma once

#include <hip/hip_runtime.h>

// z_i = a x_i + b y_i
__device__ __forceinline__ double dmk_axpby_elem(double a, double x, double b, double y) { return a * x + b * y; }

// w_i - sum_k c_k V_k,i over l basis vectors (stride ldV)
__device__ __forceinline__ double dmk_gs_elem(double w, const double* __restrict__ V, long ldV, long i, int l,
                                              const double* __restrict__ c) {
  for (int k = 0; k < l; k++) w = fma(-c[k], V[k * ldV + i], w);
  return w;
}
