"""Algorithmic FLOP counts of the per-cell kernels, derived from the operations
the stress-integration kernel performs (a fused multiply-add counts 2, a
division, square root or pow counts 1).  The counts multiply device-side
iteration counters (Newton iterations, state-integrator iterations) to give
the FLOPs of one response.

Newton iteration of integrateStress (N_sl slip systems), per cell:
  B = I - dt Lp                                  18
  F_e = A B                                      54
  Hooke: F_e^T F_e, E, Mandel C:E                94
  kinetics: 2 per system tau (9 FMA), ratio, pow, sign, dgamma        26 N_sl
  L_p = sum gamma_dot P                          18 N_sl
  residual, norms, tolerance                     72
  G = F_e^T A                                    54
  Jacobian rank-N_sl update: G^T Q (27 FMA), scale, 9x9 outer (81 FMA)   226 N_sl
  9x9 LU with partial pivoting + solve           650
  step update                                    36
State-integrator iteration (beyond its Newton iterations):
  dotState: kinetics without tangent (23 N_sl), hardening h0 |1-xi/xi_inf|^a (H |gamma_dot|)
  (2 N_sl + 8 per system), damper and state update (9 per state entry, 2 N_sl entries)
  + the integrateStress prologue/epilogue (inverse F_p0, A, F_p, P, det regularisation) 320
Tangent kernel (once per cell and response): kinetics 26 N_sl; K assembly
  (G2 P, symmetrise, C:, 9x9 outer) 279 N_sl; LU 650; 9 columns of
  (rhs 80, solve 162, 37 N_sl, 405); invariants / polar decomposition 300.
"""


def newton_flops(nsl: int) -> int:
    return 18 + 54 + 94 + 26 * nsl + 18 * nsl + 72 + 54 + 226 * nsl + 650 + 36


def state_flops(nsl: int) -> int:
    return 23 * nsl + nsl * (2 * nsl + 8) + 9 * 2 * nsl + 10 + 320


def tangent_flops(nsl: int) -> int:
    return 26 * nsl + 80 + 108 + 279 * nsl + 650 + 216 + 9 * (80 + 162 + 37 * nsl + 405) + 300


def stress_kernel_flops(nsl: int, newton_iterations: int, state_iterations: int) -> int:
    return newton_flops(nsl) * newton_iterations + state_flops(nsl) * state_iterations


# Algorithmic HBM bytes per cell of the stress kernel: read F, F0, F_p0, L_p0,
# S0 (5 x 72), xi0, gamma0 (2 N_sl x 8); write F_phase, P, S, L_p, F_p, F_e
# (6 x 72), xi, gamma (2 N_sl x 8), status/counters (12)
def stress_kernel_bytes(nsl: int) -> int:
    return 5 * 72 + 16 * nsl + 6 * 72 + 16 * nsl + 12


# Per residual evaluation of the spectral part (N cells, Nk = (N1/2+1) N2 N3):
def spectral_bytes(N: int, Nk: int) -> dict:
    return {"fft_fwd": 9 * N * 8 + 9 * Nk * 16, "gamma_div": 2 * 9 * Nk * 16, "fft_inv": 9 * Nk * 16 + 9 * N * 8}
