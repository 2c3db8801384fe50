// Per-material-point constitutive update on MI355X: one lane per cell, FP64.
//
// stress_kernel: homogenization_mechanical_response's first loop
//   (src/homogenization.f90:309-348): restore, isostrain partition,
//   phase_mechanical_constitutive substepping (src/phase_mechanical.f90:1008-1090),
//   integrateStateFPI (:591-676) and integrateStress (:379-584) with
//   phenopowerlaw kinetics (src/phase_mechanical_plastic_phenopowerlaw.f90:303-399,449-488)
//   and Hooke's law (src/phase_mechanical_elastic.f90:172-205).
// tangent_kernel: the second loop (:352-360): crystallite orientation
//   (src/phase.f90:659-673) and phase_mechanical_dPdF (src/phase_mechanical.f90:1121-1242).
//
// The material point carries no eigen kinematics and V_e = I, so F_i = I and
// L_i = 0 throughout; the 3x3x3x3 chains of the reference then collapse to
// rank-N_sl updates of the 9x9 identity:
//   dR_Lp/dL_p = I9 + dt * sum_a dgamma_a vec(P_a) vec(G^T Q_a)^T,  G = F_e^T A,
//   Q_a = C : P_nS,a (tabulated per phase), vec row-major (math_33to9),
// which is assembled in registers and solved by Householder QR (dmk_dense.h).
//
// Register budget.  The 8x8 Newton system (64 doubles) is the only large
// register object.  Slip-system loops are rolled (the kinetics loop by two) so
// the phase tables are read through scalar loads inside the loop instead of
// being hoisted into VGPRs; the per-system vectors of a lane (xi, rates, rate
// derivatives) and its sub-step bookkeeping live in LDS, [system][lane], one
// wavefront per work-group.  Sub-step start values, the state-rate history of
// the fixed-point iteration and the line search's fallback live in global
// scratch; which start values apply is a per-lane flag, not a per-lane
// pointer, so no 64-bit addresses stay live across the Newton loop.
#include <cmath>

#include "dmk_common.h"
#include "dmk_math.h"
#include "dmk_dense.h"

namespace {

constexpr int SWG = DMK_STRESS_WG;   // work-group of the per-cell kernels: one wavefront

// The stress kernel runs two waves per SIMD: measured on the 64^3 tension
// workload this is 20 % faster than one wave per SIMD without spills, the
// second wave hiding the scalar-load and dependency latency of the slip loops
// (at the cost of spilling a few loop-carried tensors around the QR solve).
#ifndef DMK_STRESS_WAVES
#define DMK_STRESS_WAVES 2
#endif

// Library transcendental functions are kept out of line: inlined, their
// polynomial constants are hoisted out of every enclosing loop and stay live
// across the 9x9 Newton system.
__device__ __noinline__ double pow_ool(double x, double y) { return pow(x, y); }
__device__ __noinline__ double tanh_ool(double x) { return tanh(x); }

// x^y for the power laws (phenopowerlaw n_sl, a_sl).  With QP (every
// exponent of the phase is a positive multiple of 1/4 up to 64, checked on the
// host) y = k/4 is evaluated by binary powering of x (k a multiple of 4),
// sqrt(x) (k even) or sqrt(sqrt(x)) (k odd); y is a per-phase constant, so the
// powering loop is uniform across the wavefront.  Its error is that of the
// root, amplified by the power taken of it, plus the products' roundings,
// each amplified by the powering after it; against a long-double reference
// over the power laws' base ranges: n = 20 within 12 ulp, a = 2.25 within 11
// ulp, a = 2 within 0.5 ulp, and up to 376 ulp (8e-14 relative) for the
// largest odd multiples of 1/4 (y near 64, taken as (x^(1/4))^255).  It
// follows pow's rules for zero, negative (NaN unless y is an integer, sign of
// odd powers), infinite and NaN bases.  Phases with other exponents use the
// library pow (within 2 ulp, 23 % slower per stress launch at 64^3).
template <bool QP>
DI double pow_law(double x, double y) {
  if (!QP) return pow_ool(x, y);
  int k = (int)(y * 4.0);
  double b = x;
  if (k & 1) {
    b = sqrt(sqrt(x));
  } else if (k & 2) {
    b = sqrt(x);
    k >>= 1;
  } else {
    k >>= 2;
  }
  double r = 1.0;
  bool first = true;
  while (k) {
    if (k & 1) { r = first ? b : r * b; first = false; }
    k >>= 1;
    if (k) b = b * b;
  }
  return r;
}

// the power laws' x^y on the device, for checking their arithmetic
template <bool QP>
__global__ void pow_law_kernel(const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ out,
                               int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = pow_law<QP>(x[i], y[i]);
}

// per-lane slip-system vectors, entry a of a lane at [a][lane]
template <int NSL>
struct SlipLds {
  double xi[NSL][SWG];    // critical resolved shear stress of the current state
  double g[NSL][SWG];     // slip rates (signed), |rate| once the state rate is formed
  double w[NSL][SWG];     // d rate / d tau in the stress Newton loop, xi rate in the state loop
  // sub-step bookkeeping of a lane, kept out of the registers the Newton needs
  double frac[SWG], step[SWG], dt_done[SWG];
  int nsub[SWG], fpi[SWG];
};

// kinetics_sl (src/phase_mechanical_plastic_phenopowerlaw.f90:449-488) for all
// systems of one lane: rates -> gs, derivatives -> ws (DERIV), and
// L_p = sum_a rate_a P_a (src/phase_mechanical_plastic_phenopowerlaw.f90:325-338) when LP.
template <int NSL, bool QP, bool DERIV, bool LP>
DI void kinetics(const DmkPhaseDev* __restrict__ ph, const double* Mp, const double* xs, double* gs,
                 double* ws, double* Lp) {
  if (LP) {
#pragma unroll
    for (int k = 0; k < 9; k++) Lp[k] = 0.0;
  }
  const bool nonschmid = ph->nonschmid;
#pragma unroll 2
  for (int a = 0; a < NSL; a++) {
    double tp = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) tp = fma(Mp[k], ph->Pp[a][k], tp);
    double tn;
    if (nonschmid) {
      tn = 0.0;
#pragma unroll
      for (int k = 0; k < 9; k++) tn = fma(Mp[k], ph->Pm[a][k], tn);
    } else {
      tn = -tp;
    }
    double sgn = tp > tn ? 1.0 : -1.0;
    double tmax = fmax(tp, tn);
    double gg = sgn * ph->g0[a] * pow_law<QP>(tmax / xs[a * SWG], ph->n[a]);
    gs[a * SWG] = gg;
    if (DERIV) ws[a * SWG] = !dEq0(gg) ? gg * ph->n[a] / tmax : 0.0;
    if (LP) {
#pragma unroll
      for (int k = 0; k < 9; k++) Lp[k] = fma(gg, ph->P[a][k], Lp[k]);
    }
  }
}

// phenopowerlaw_dotState without twins (src/phase_mechanical_plastic_phenopowerlaw.f90:356-397)
// from the slip rates in gs: gs <- |rate| (gamma rate), ws <- xi rate.
// xi_sat = xi_inf + f_sat sqrt(sum F_tw) and hf = h_0 (1 + c_1 (sum F_tw)^c_2) at sum F_tw = 0
// are tabulated per phase.  Returns true when a rate is NaN.
template <int NSL, bool QP>
DI bool dot_state(const DmkPhaseDev* __restrict__ ph, const double* xs, double* gs, double* ws) {
  double hg[NSL];
#pragma unroll
  for (int r = 0; r < NSL; r++) hg[r] = 0.0;
  bool nan = false;
#pragma unroll 1
  for (int a = 0; a < NSL; a++) {
    double ga = fabs(gs[a * SWG]);
    gs[a * SWG] = ga;
    nan |= (ga != ga);
#pragma unroll
    for (int r = 0; r < NSL; r++) hg[r] = fma(ph->H[r][a], ga, hg[r]);
  }
#pragma unroll
  for (int r = 0; r < NSL; r++) ws[r * SWG] = hg[r];
#pragma unroll 1
  for (int r = 0; r < NSL; r++) {
    double x = 1.0 - xs[r * SWG] / ph->xi_sat[r];
    double left = copysign(pow_law<QP>(fabs(x), ph->a[r]), x);
    double d = ph->hf[r] * left * ws[r * SWG];
    ws[r * SWG] = d;
    nan |= (d != d);
  }
  return nan;
}

// E = (Fe^T Fe - I)/2 and S = C:E (lattice frame)
DI void hooke(const DmkPhaseDev* __restrict__ ph, const double* Fe, double* S) {
  double E[9];
  m_mul_tn(Fe, Fe, E);
  E[0] -= 1.0; E[4] -= 1.0; E[8] -= 1.0;
#pragma unroll
  for (int k = 0; k < 9; k++) E[k] *= 0.5;
  hooke_sym(ph->C11, ph->C12, ph->C44, E, S);
}

// a per-lane flag re-read in every loop iteration, so that pointers selected
// by it are formed where they are used instead of being kept live across loops
DI bool live_flag(bool f) {
  int v = f ? 1 : 0;
  asm volatile("" : "+v"(v));
  return v != 0;
}

DI void copy9(double* dst, const double* src, const Lane& l) {
#pragma unroll
  for (int c = 0; c < 9; c++) st(dst, l, c, ld(src, l, c));
}

// per-cell scratch of the stress integration ([DMK_SUB_ROWS][N]): sub-step
// F_0, sub-step F, F_p0 and L_p0 (the last two once a sub-step has completed),
// and the Newton line search's last accepted L_p and direction
enum { SUB_F0 = 0, SUB_F = 9, SUB_FP0 = 18, SUB_LP0 = 27, SUB_LSOLD = 36, SUB_LSDIR = 45 };
static_assert(SUB_LSDIR + 9 <= DMK_SUB_ROWS, "sub-step scratch rows");

// integrateStress (src/phase_mechanical.f90:379-584) with F_i = I, L_i = 0, for
// the sub-step F (scratch row SUB_F) and sub-step F_p0 (Fp0_p), starting from
// the L_p in mp.Lp (the previous call's result: phase_mechanical_Lp is not
// restored between calls).  On success the converged L_p goes back to mp.Lp
// and S is returned in registers.  P, F_p and F_e follow from (F, F_p0, L_p, S)
// alone, so they are formed once, by finish_point, after the state integration
// has converged instead of after every call.
template <int NSL, bool QP>
DI int integrate_stress(const DmkPhaseDev* __restrict__ ph, const DmkNumMech& num, const DmkMPDev& mp,
                        const Lane& ln, bool started, double dt, SlipLds<NSL>& sh, int lane, int& newton,
                        double* S) {
  const double* Fsub_p = mp.sub + SUB_F * ln.sN;
  const double* xs = &sh.xi[0][lane];
  double* gs = &sh.g[0][lane];
  double* ws = &sh.w[0][lane];
  double A[9];
  {
    const double* Fp0_p = live_flag(started) ? mp.sub + SUB_FP0 * ln.sN : mp.Fp0;
    double Fp0[9], invFp0[9], F[9];
    m_load(Fp0_p, ln, Fp0);
    if (m_invert(Fp0, invFp0)) return ST_FAIL_STRESS;
    m_load(Fsub_p, ln, F);
    m_mul(F, invFp0, A);
  }
  double Lpg[9];
  m_load(mp.Lp, ln, Lpg);
  // the line search's last accepted iterate and direction live in the
  // scratch (read back only when a step is rejected, and to form the next
  // iterate), out of the registers the Newton system needs; the first
  // iteration is always accepted, so both are written before they are read
  double* ls_old = mp.sub + SUB_LSOLD * ln.sN;
  double* ls_dir = mp.sub + SUB_LSDIR * ln.sN;
  double step = 1.0, res_old = DBL_MAX;
  int it = 0;
  for (;;) {
    it++;
    if (it > num.N_iter_Lp_max) return ST_FAIL_LP_MAXIT;
    newton++;
    {
      double B[9], Fe[9];
#pragma unroll
      for (int k = 0; k < 9; k++) B[k] = ((k % 4) == 0 ? 1.0 : 0.0) - dt * Lpg[k];
      m_mul(A, B, Fe);
      hooke(ph, Fe, S);
    }
    double Lpc[9];
    kinetics<NSL, QP, true, true>(ph, S, xs, gs, ws, Lpc);
    double ng = m_norm(Lpg), nc = m_norm(Lpc);
    double atol = fmax(num.eps_rel_Lp * fmax(ng, nc), num.eps_abs_Lp);
    double res[9];
    bool nan = false;
#pragma unroll
    for (int k = 0; k < 9; k++) { res[k] = Lpg[k] - Lpc[k]; nan |= (res[k] != res[k]); }
    double nres = m_norm(res);
    if (nan) return ST_FAIL_LP_NAN;
    if (nres < atol) break;
    if (nres < res_old) {
      res_old = nres;
#pragma unroll
      for (int k = 0; k < 9; k++) st(ls_old, ln, k, Lpg[k]);
      step = 1.0;
    } else {
      step = num.r_linesearch_Lp * step;
#pragma unroll
      for (int k = 0; k < 9; k++) Lpg[k] = fma(ld(ls_dir, ln, k), step, ld(ls_old, ln, k));
      continue;
    }
    // Jacobian, rebuilt every iteration (f_update_jacobi_Lp = 1):
    // J = I9 + sum_a u_a vec(P_a) vec(Y_a)^T.  The Schmid tensors P_a are
    // traceless, so tr(J x) = tr(x) and the solution of J x = res has
    // x_22 = tr(res) - x_00 - x_11.  Substituting it into rows 0..7 leaves an
    // 8x8 system (columns 0 and 4 reduced by column 8), which is what is
    // assembled: J8 = rows/columns 0..7, j8 = column 8 of rows 0..7.
    double G[9];
    {
      // F_e again (cheaper than keeping it live across the kinetics loop)
      double B[9], Fe[9];
#pragma unroll
      for (int k = 0; k < 9; k++) B[k] = ((k % 4) == 0 ? 1.0 : 0.0) - dt * Lpg[k];
      m_mul(A, B, Fe);
      m_mul_tn(Fe, A, G);
    }
    // The trace substitution x_22 = tr(res) - x_00 - x_11 is applied per slip
    // system: column 8 of each rank-1 term is subtracted from columns 0 and 4
    // and its product with tr(res) from the right-hand side, so no ninth column
    // is kept.
    double J[64];
#pragma unroll
    for (int r = 0; r < 64; r++) J[r] = (r % 9 == 0) ? 1.0 : 0.0;
    const double tr = res[0] + res[4] + res[8];
    const bool nonschmid = ph->nonschmid;
#pragma unroll 1
    for (int a = 0; a < NSL; a++) {
      double u = dt * ws[a * SWG];
      const bool pos = gs[a * SWG] > 0.0;
      double Y[9];
      if (nonschmid) {
        double Q[9];
#pragma unroll
        for (int k = 0; k < 9; k++) Q[k] = pos ? ph->Qp[a][k] : ph->Qm[a][k];
        m_mul_tn(G, Q, Y);
      } else {
        // Schmid law: Q_neg = -Q_pos exactly, folded into the sign of u
        m_mul_tn(G, ph->Qp[a], Y);
        u = pos ? u : -u;
      }
      Y[0] -= Y[8];
      Y[4] -= Y[8];
      const double y8tr = Y[8] * tr;
#pragma unroll
      for (int r = 0; r < 8; r++) {
        double pr = u * ph->P[a][r];
#pragma unroll
        for (int c = 0; c < 8; c++) J[8 * r + c] = fma(pr, Y[c], J[8 * r + c]);
        res[r] = fma(-pr, y8tr, res[r]);
      }
    }
    if (dmk_qr_solve<8>(J, res)) return ST_FAIL_LP_DGESV;
    res[8] = tr - res[0] - res[4];
    // the accepted iterate was just stored: L_p is read back instead of being
    // kept live across the solve
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double d = -res[k];
      st(ls_dir, ln, k, d);
      Lpg[k] = fma(d, step, ld(ls_old, ln, k));
    }
  }
  // the reference's error exit when the new plastic deformation is singular
  // (math_invert33 of F_p0^-1 (I - dt L_p))
  {
    const double* Fp0_p = live_flag(started) ? mp.sub + SUB_FP0 * ln.sN : mp.Fp0;
    double Fp0[9], invFp0[9], B[9], invFp_new[9], Fp_new[9];
    m_load(Fp0_p, ln, Fp0);
    m_invert(Fp0, invFp0);
#pragma unroll
    for (int k = 0; k < 9; k++) B[k] = ((k % 4) == 0 ? 1.0 : 0.0) - dt * Lpg[k];
    m_mul(invFp0, B, invFp_new);
    if (m_invert(invFp_new, Fp_new)) return ST_FAIL_STRESS;
  }
  m_store(mp.Lp, ln, Lpg);
  return ST_OK;
}

// The end of integrateStress (src/phase_mechanical.f90:571-581) for the
// converged L_p (mp.Lp) and S (mp.S) of a sub-step with F (scratch row SUB_F),
// F_p0 (Fp0_p) and time step dt: the regularised F_p = F_p,new / det(F_p,new)^(1/3)
// goes to Fp_dst; with OUT also P = F F_p,new^-1 S F_p,new^-T and F_e = F F_p,new^-1.
template <bool OUT>
DI void finish_point(const DmkMPDev& mp, const Lane& ln, const double* Fp0_p, double dt, double* Fp_dst) {
  double invFp_new[9], Fp_new[9];
  {
    double Fp0[9], invFp0[9], B[9], Lp[9];
    m_load(Fp0_p, ln, Fp0);
    m_invert(Fp0, invFp0);
    m_load(mp.Lp, ln, Lp);
#pragma unroll
    for (int k = 0; k < 9; k++) B[k] = ((k % 4) == 0 ? 1.0 : 0.0) - dt * Lp[k];
    m_mul(invFp0, B, invFp_new);
  }
  m_invert(invFp_new, Fp_new);
  if (OUT) {
    double T1[9], P[9];
    {
      double F[9], T2[9], S[9];
      m_load(mp.sub + SUB_F * ln.sN, ln, F);
      m_load(mp.S, ln, S);
      m_mul(F, invFp_new, T1);
      m_mul_nt(S, invFp_new, T2);
      m_mul(T1, T2, P);
    }
    m_store(mp.Pph, ln, P);
    m_store(mp.Fe, ln, T1);
  }
  double d = pow_ool(m_det(Fp_new), 1.0 / 3.0);
#pragma unroll
  for (int k = 0; k < 9; k++) Fp_new[k] = Fp_new[k] / d;
  m_store(Fp_dst, ln, Fp_new);
}

// integrateStateFPI (src/phase_mechanical.f90:591-676).  State = (xi, gamma):
// xi in LDS, gamma in mp.gam; its start values xi0_p / gam0_p (the converged
// increment, or st0s after a completed sub-step); the two previous state
// rates (dotState_last) in the ping-pong buffers mp.last1 / mp.last2.  The
// initial rate is taken at the stress S_p (S_0 after a restore, else the
// previous sub-step's S); on success the S of the converged iterate is stored
// in mp.S.
template <int NSL, bool QP>
DI int integrate_state_fpi(const DmkPhaseDev* __restrict__ ph, const DmkNumMech& num, const DmkMPDev& mp,
                           const Lane& ln, bool started, bool s_cur, double dt, SlipLds<NSL>& sh, int lane,
                           int& newton) {
  double* xs = &sh.xi[0][lane];
  double* gs = &sh.g[0][lane];
  double* ws = &sh.w[0][lane];
  {
    double S0[9];
    m_load(s_cur ? mp.S : mp.S0, ln, S0);
    kinetics<NSL, QP, false, false>(ph, S0, xs, gs, nullptr, nullptr);
  }
  if (dot_state<NSL, QP>(ph, xs, gs, ws)) return ST_FAIL_STATE;
#pragma unroll 1
  for (int r = 0; r < NSL; r++) {
    double dx = ws[r * SWG], dgm = gs[r * SWG];
    const bool stf = live_flag(started);
    xs[r * SWG] = fma(dx, dt, ld(stf ? mp.st0s : mp.xi0, ln, r));
    st(mp.gam, ln, r, fma(dgm, dt, ld(stf ? mp.st0s + NSL * ln.sN : mp.gam0, ln, r)));
    st(mp.last1, ln, r, dx);
    st(mp.last1, ln, NSL + r, dgm);
  }
  int status = ST_FAIL_STATE;
  bool cur = false;                   // last1 holds dotState_last(:,1) when !cur
  for (int it = 1; it <= num.N_iter_state_max; it++) {
    sh.fpi[lane]++;
    double S[9];
    status = integrate_stress<NSL, QP>(ph, num, mp, ln, started, dt, sh, lane, newton, S);
    if (status != ST_OK) break;
    m_store(mp.S, ln, S);
    // the converged Newton iterate evaluated kinetics_sl at this (S, xi): rates
    // are in gs.  A NaN rate ends the iteration with the status of the stress
    // integration, as the reference's exit does (phase_mechanical.f90:638)
    if (dot_state<NSL, QP>(ph, xs, gs, ws)) break;
    const double* w1p = opaque(cur ? mp.last2 : mp.last1);
    double* w2p = opaque(cur ? mp.last1 : mp.last2);
    const bool stf = live_flag(started);
    const double* x0 = opaque(stf ? mp.st0s : mp.xi0);
    const double* g0 = opaque(stf ? mp.st0s + NSL * ln.sN : mp.gam0);
    // damper (src/phase_mechanical.f90:653-674)
    double w1v[2 * NSL];
    double d12 = 0.0, d22 = 0.0, d01 = 0.0;
#pragma unroll
    for (int r = 0; r < 2 * NSL; r++) {
      double w1 = ld(w1p, ln, r);
      double w2 = it > 1 ? ld(w2p, ln, r) : 0.0;
      double d0 = r < NSL ? ws[r * SWG] : gs[(r - NSL) * SWG];
      w1v[r] = w1;
      d12 = fma(d0 - w1, w1 - w2, d12);
      d22 = fma(w1 - w2, w1 - w2, d22);
      d01 = fma(d0, w1, d01);
    }
    double zeta = (fmin(d01, d12) < 0.0 && d22 > 0.0) ? 0.75 + 0.25 * tanh_ool(2.0 + 4.0 * d12 / d22) : 1.0;
    bool conv = true;
#pragma unroll
    for (int r = 0; r < 2 * NSL; r++) {
      double w1 = w1v[r];
      double d0 = r < NSL ? ws[r * SWG] : gs[(r - NSL) * SWG];
      double d = d0 * zeta + w1 * (1.0 - zeta);
      st(w2p, ln, r, d);             // dotState_last(:,1) of the next iteration
      double s0 = r < NSL ? ld(x0, ln, r) : ld(g0, ln, r - NSL);
      double s = r < NSL ? xs[r * SWG] : ld(mp.gam, ln, r - NSL);
      double res = s - s0 - d * dt;
      s = s - res;
      double atol = r < NSL ? ph->atol_xi : ph->atol_gamma;
      conv = conv && (fabs(res) <= fmax(atol, num.eps_rel_state * fabs(s)));
      if (r < NSL) xs[r * SWG] = s; else st(mp.gam, ln, r - NSL, s);
    }
    cur = !cur;
    if (conv) { status = ST_OK; break; }
  }
  return status;
}

// one material point update per lane: block `blk` of the cells of mp
template <int NSL, bool QP>
DI void stress_body(const DmkMPDev& mp, const double* __restrict__ Ftarget, const DmkPhaseDev* __restrict__ ph,
                    int phase_id, const DmkNumMech& num, double dt, int blk, SlipLds<NSL>& sh) {
  const int lane = threadIdx.x;
  const int e = blk * SWG + lane;
  if (e >= mp.N) return;
  if (mp.phase[e] != phase_id) return;
  const Lane ln = make_lane(mp.N, e);
  double* xs = &sh.xi[0][lane];
  double* sub = mp.sub;

  // isostrain partition (src/homogenization_mechanical_isostrain.f90:60-75)
  copy9(mp.Fphase, Ftarget, ln);
  // start values of the sub-step: those of the converged increment until a
  // sub-step completes (started), then copies in the sub-step scratch / st0s;
  // the initial state rate is taken at S_0 until a sub-step completes, then at
  // the last completed sub-step's S (s_cur)
  bool started = false, s_cur = false;

  int newton = 0;
  sh.fpi[lane] = 0;
  sh.nsub[lane] = 0;
  sh.dt_done[lane] = 0.0;            // time step of the last completed sub-step
  // phase_mechanical_constitutive substepping; the first pass takes the
  // cut-back branch, which is phase_restore (src/phase_mechanical.f90:1096-1115)
  sh.frac[lane] = 0.0;
  sh.step[lane] = 1.0 / num.r_cutback;
  int status = ST_ITERATING;
  bool todo = true;
  while (todo) {
    if (status == ST_OK) {
      const double stepFrac = sh.frac[lane] + sh.step[lane];
      const double step = fmin(1.0 - stepFrac, num.r_increase * sh.step[lane]);
      sh.frac[lane] = stepFrac;
      sh.step[lane] = step;
      todo = step > 0.0;
      if (todo) {
        // the completed sub-step's F_p becomes the next one's F_p0
        finish_point<false>(mp, ln, started ? sub + SUB_FP0 * ln.sN : mp.Fp0, sh.dt_done[lane], sub + SUB_FP0 * ln.sN);
        copy9(sub + SUB_F0 * ln.sN, sub + SUB_F * ln.sN, ln);
        copy9(sub + SUB_LP0 * ln.sN, mp.Lp, ln);
#pragma unroll 1
        for (int r = 0; r < NSL; r++) {
          st(mp.st0s, ln, r, xs[r * SWG]);
          st(mp.st0s, ln, NSL + r, ld(mp.gam, ln, r));
        }
        started = true;
      }
    } else {
      // cut back: restore the sub-step start (F_p, S = S_0, L_p for an actual
      // cut-back, state).  F_p, S and gamma are outputs only until they are
      // read again, so they are restored where they are read: F_p0 and S come
      // from the sub-step start / S_0, and gamma is overwritten by the state
      // integration
      const double step = num.r_cutback * sh.step[lane];
      sh.step[lane] = step;
      s_cur = false;
      if (step < 1.0) copy9(mp.Lp, started ? sub + SUB_LP0 * ln.sN : mp.Lp0, ln);
#pragma unroll 1
      for (int r = 0; r < NSL; r++) xs[r * SWG] = ld(started ? mp.st0s : mp.xi0, ln, r);
      todo = step > num.r_cutback_min;
    }
    if (todo) {
      const double* pF0 = started ? sub + SUB_F0 * ln.sN : mp.F0;
#pragma unroll
      for (int c = 0; c < 9; c++)
        st(sub + SUB_F * ln.sN, ln, c, fma(sh.step[lane], ld(Ftarget, ln, c) - ld(mp.F0, ln, c), ld(pF0, ln, c)));
      status = integrate_state_fpi<NSL, QP>(ph, num, mp, ln, started, s_cur, sh.step[lane] * dt, sh, lane, newton);
      sh.nsub[lane]++;
      if (status == ST_OK) {
        s_cur = true;
        sh.dt_done[lane] = sh.step[lane] * dt;
      }
    }
  }
  if (status == ST_OK) {
    finish_point<true>(mp, ln, started ? sub + SUB_FP0 * ln.sN : mp.Fp0, sh.dt_done[lane], mp.Fp);
  } else {
    // the restored sub-step start, as the last cut-back left it
    copy9(mp.Fp, started ? sub + SUB_FP0 * ln.sN : mp.Fp0, ln);
    copy9(mp.S, mp.S0, ln);
#pragma unroll 1
    for (int r = 0; r < NSL; r++) st(mp.gam, ln, r, ld(started ? mp.st0s + NSL * ln.sN : mp.gam0, ln, r));
  }
#pragma unroll 1
  for (int r = 0; r < NSL; r++) st(mp.xi, ln, r, xs[r * SWG]);
  mp.status[e] = status;
  mp.n_newton[e] = newton;
  mp.n_fpi[e] = sh.fpi[lane];
  mp.n_sub[e] = sh.nsub[lane];
}

template <int NSL, bool QP>
__global__ void __launch_bounds__(SWG) __attribute__((amdgpu_waves_per_eu(DMK_STRESS_WAVES)))
stress_kernel(DmkMPDev mp, const double* __restrict__ Ftarget, const DmkPhaseDev* __restrict__ ph, int phase_id,
              DmkNumMech num, double dt) {
  __shared__ SlipLds<NSL> sh;
  stress_body<NSL, QP>(mp, Ftarget, ph, phase_id, num, dt, blockIdx.x, sh);
}

// The stress integration of several RVEs (and phases) in one launch: item i
// owns blocks [block0_i, block0_{i+1}) and runs them on its own material
// point, phase tables, load and time step, exactly as stress_kernel would.
template <int NSL, bool QP>
__global__ void __launch_bounds__(SWG) __attribute__((amdgpu_waves_per_eu(DMK_STRESS_WAVES)))
stress_kernel_batch(const DmkStressItem* __restrict__ items, int n_items) {
  __shared__ SlipLds<NSL> sh;
  const int b = blockIdx.x;
  int lo = 0, hi = n_items - 1;                     // last item with block0 <= b (uniform)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].block0 <= b) lo = mid; else hi = mid - 1;
  }
  const DmkStressItem* it = items + lo;
  const DmkMPDev mp = it->mp;
  const DmkNumMech num = it->num;
  stress_body<NSL, QP>(mp, it->F, it->ph, it->phase_id, num, it->dt, b - it->block0, sh);
}

// math_rotationalPart (src/math.f90:1004-1043)
DI void rotational_part(const double* F, double* R) {
  const double TAU = 6.283185307179586;
  double C[9], FF[9], FC[9];
  m_mul_tn(F, F, C);
  double I1 = C[0] + C[4] + C[8];
  double I2 = C[0] * C[4] + C[0] * C[8] + C[4] * C[8] - (C[1] * C[1] + C[2] * C[2] + C[5] * C[5]);
  double I3 = -(C[0] * C[5] * C[5] + C[4] * C[2] * C[2] + C[8] * C[1] * C[1]) + C[0] * C[4] * C[8]
            + 2.0 * C[1] * C[2] * C[5];
  m_mul(F, F, FF);
  double trF = F[0] + F[4] + F[8];
  double IF2 = 0.5 * (trF * trF - (FF[0] + FF[4] + FF[8]));
  double l[3];
  double q = fmax(I1 * I1 - 3.0 * I2, 0.0);
  double x = pow(q, 1.5);
  if (!dEq0(x)) {
    double arg = fmin(fmax((I1 * I1 * I1 - 4.5 * I1 * I2 + 13.5 * I3) / x, -1.0), 1.0);
    double Phi = acos(arg);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double lam = I1 + (2.0 * sqrt(q)) * cos((Phi - TAU * (double)(k + 1)) / 3.0);
      l[k] = sqrt(fmax(lam, 0.0) / 3.0);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) l[k] = sqrt(I1 / 3.0);
  }
  double U1 = l[0] + l[1] + l[2], U2 = l[0] * l[1] + l[1] * l[2] + l[2] * l[0];
  m_mul(F, C, FC);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      R[3 * i + j] = U1 * IF2 * (i == j ? 1.0 : 0.0) + (U1 * U1 - U2) * F[3 * i + j] - U1 * trF * F[3 * j + i]
                   + U1 * FF[3 * j + i] - FC[3 * i + j];
  double d = pow(m_det(R), -1.0 / 3.0);
#pragma unroll
  for (int k = 0; k < 9; k++) R[k] *= d;
}

// om2qu (src/rotations.f90:486-514), P = -1
DI void om2qu(const double* om, double* qu) {
  double tr = om[0] + om[4] + om[8];
  double s;
  if (tr > 0.0) {
    s = 0.5 / sqrt(tr + 1.0);
    qu[0] = 0.25 / s; qu[1] = (om[7] - om[5]) * s; qu[2] = (om[2] - om[6]) * s; qu[3] = (om[3] - om[1]) * s;
  } else if (om[0] > om[4] && om[0] > om[8]) {
    s = 2.0 * sqrt(1.0 + om[0] - om[4] - om[8]);
    qu[0] = (om[7] - om[5]) / s; qu[1] = 0.25 * s; qu[2] = (om[1] + om[3]) / s; qu[3] = (om[2] + om[6]) / s;
  } else if (om[4] > om[8]) {
    s = 2.0 * sqrt(1.0 + om[4] - om[0] - om[8]);
    qu[0] = (om[2] - om[6]) / s; qu[1] = (om[1] + om[3]) / s; qu[2] = 0.25 * s; qu[3] = (om[5] + om[7]) / s;
  } else {
    s = 2.0 * sqrt(1.0 + om[8] - om[0] - om[4]);
    qu[0] = (om[3] - om[1]) / s; qu[1] = (om[2] + om[6]) / s; qu[2] = (om[5] + om[7]) / s; qu[3] = 0.25 * s;
  }
  if (signbit(qu[0])) for (int k = 0; k < 4; k++) qu[k] = -qu[k];
  for (int k = 1; k < 4; k++) if (!dEq0(qu[k])) qu[k] = -qu[k];
  double n = sqrt(qu[0] * qu[0] + qu[1] * qu[1] + qu[2] * qu[2] + qu[3] * qu[3]);
  for (int k = 0; k < 4; k++) qu[k] /= n;
}


// row r (runtime, uniform) of a register 3x3 without dynamic register indexing
DI void row3(const double* M, int r, double* v) {
#pragma unroll
  for (int j = 0; j < 3; j++) v[j] = r == 0 ? M[j] : (r == 1 ? M[3 + j] : M[6 + j]);
}

// Orientation and consistent tangent of the homogenization (one constituent,
// v = 1) for a response in which every cell converged (flag[0] == ST_OK).
// Also writes the squared Frobenius norm of each cell's dPdF, summed in
// Fortran element order (the C_minMaxAvg selection works on it).
template <int NSL, bool QP>
__global__ void __launch_bounds__(SWG) tangent_kernel(DmkMPDev mp, const DmkPhaseDev* __restrict__ ph,
                                                      int phase_id, double dt, const int* __restrict__ flag,
                                                      double* __restrict__ dPdF, double* __restrict__ nrm) {
  __shared__ SlipLds<NSL> sh;
  const int lane = threadIdx.x;
  const int e = blockIdx.x * SWG + lane;
  if (e >= mp.N || flag[0] != ST_OK) return;
  if (mp.phase[e] != phase_id) return;
  const Lane ln = make_lane(mp.N, e);
  double* xs = &sh.xi[0][lane];
  double* gs = &sh.g[0][lane];
  double* ws = &sh.w[0][lane];
  double Fe[9];
  m_load(mp.Fe, ln, Fe);
  {
    double R[9], Rt[9], q[4];
    rotational_part(Fe, R);
    m_transpose(R, Rt);
    om2qu(Rt, q);
#pragma unroll
    for (int k = 0; k < 4; k++) st(mp.O, ln, k, q[k]);
  }
#pragma unroll 1
  for (int r = 0; r < NSL; r++) xs[r * SWG] = ld(mp.xi, ln, r);
  {
    double S[9];
    m_load(mp.S, ln, S);
    kinetics<NSL, QP, true, false>(ph, S, xs, gs, ws, nullptr);
  }
  const bool nonschmid = ph->nonschmid;

  double invFp[9], invSubFp0[9], A2[9];
  {
    double Fp[9], Fp0[9], F[9];
    m_load(mp.Fp, ln, Fp);
    m_load(mp.Fp0, ln, Fp0);
    m_load(mp.Fphase, ln, F);
    m_invert(Fp, invFp);
    m_invert(Fp0, invSubFp0);
    m_mul(F, invSubFp0, A2);
  }
  // K = I9 + dt sum_a dg_a vec(C:sym(G2 P_a)) vec(Pn_a)^T,  G2 = F_e^T A2
  double K[81];
  {
    double G2[9];
    m_mul_tn(Fe, A2, G2);
#pragma unroll
    for (int r = 0; r < 81; r++) K[r] = (r % 10 == 0) ? 1.0 : 0.0;
#pragma unroll 1
    for (int a = 0; a < NSL; a++) {
      double X[9], Xs[9], W[9];
      m_mul(G2, ph->P[a], X);
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Xs[3 * i + j] = 0.5 * (X[3 * i + j] + X[3 * j + i]);
      hooke_sym(ph->C11, ph->C12, ph->C44, Xs, W);
      const bool pos = gs[a * SWG] > 0.0;
      double u = dt * ws[a * SWG];
      if (nonschmid) {
        double Pn[9];
#pragma unroll
        for (int c = 0; c < 9; c++) Pn[c] = pos ? ph->Pp[a][c] : ph->Pm[a][c];
#pragma unroll
        for (int r = 0; r < 9; r++) {
          double wr = u * W[r];
#pragma unroll
          for (int c = 0; c < 9; c++) K[9 * r + c] = fma(wr, Pn[c], K[9 * r + c]);
        }
      } else {
        u = pos ? u : -u;            // Schmid law: P_nS,neg = -P_nS,pos exactly
#pragma unroll
        for (int r = 0; r < 9; r++) {
          double wr = u * W[r];
#pragma unroll
          for (int c = 0; c < 9; c++) K[9 * r + c] = fma(wr, ph->Pp[a][c], K[9 * r + c]);
        }
      }
    }
  }
  double v0[9], tau[9];
  const bool sing = dmk_qr_factor<9>(K, v0, tau);

  double M1[9], M2[9], M3[9], Mid[9];
  {
    double S[9], F[9];
    m_load(mp.S, ln, S);
    m_load(mp.Fphase, ln, F);
    m_mul_nt(S, invFp, M1);        // S invFp^T
    m_mul(F, invFp, M2);           // F invFp
    m_mul(M2, S, M3);              // F invFp S
    m_mul(invFp, M1, Mid);         // invFp S invFp^T
  }
  double n2 = 0.0;
#pragma unroll 1
  for (int l = 0; l < 3; l++) {
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
      // rhs column: C : sym(Fe(k,:) (x) invFp(l,:))
      double fk[3], pl[3], ml[3];
      row3(Fe, k, fk);
      row3(invFp, l, pl);
      row3(Mid, l, ml);
      double X[9], y[9];
#pragma unroll
      for (int n = 0; n < 3; n++)
#pragma unroll
        for (int m = 0; m < 3; m++) X[3 * n + m] = 0.5 * (fk[n] * pl[m] + fk[m] * pl[n]);
      hooke_sym(ph->C11, ph->C12, ph->C44, X, y);
      if (!sing) dmk_qr_apply_solve<9>(K, v0, tau, y);
      // t = sum_a dg_a P_a (Pn_a : dSdF)
      double t[9];
#pragma unroll
      for (int q = 0; q < 9; q++) t[q] = 0.0;
#pragma unroll 1
      for (int a = 0; a < NSL; a++) {
        const bool pos = gs[a * SWG] > 0.0;
        double c = 0.0;
        if (nonschmid) {
#pragma unroll
          for (int q = 0; q < 9; q++) c = fma(pos ? ph->Pp[a][q] : ph->Pm[a][q], y[q], c);
        } else {
#pragma unroll
          for (int q = 0; q < 9; q++) c = fma(ph->Pp[a][q], y[q], c);
          c = pos ? c : -c;
        }
        c *= ws[a * SWG];
#pragma unroll
        for (int q = 0; q < 9; q++) t[q] = fma(c, ph->P[a][q], t[q]);
      }
      // dFpinv = -dt invSubFp0 t;  F dFpinv = -dt A2 t
      double D[9], AD[9];
      m_mul(invSubFp0, t, D);
      m_mul(A2, t, AD);
#pragma unroll
      for (int q = 0; q < 9; q++) { D[q] = -D[q] * dt; AD[q] = -AD[q] * dt; }
      double R2[9], R3[9], R4[9], R1[9], col[9];
      m_mul(AD, M1, R2);            // F dFpinv S invFp^T
      m_mul(M2, y, R3);
      m_mul_nt(R3, invFp, R4);      // F invFp dSdF invFp^T
      m_mul_nt(M3, D, R1);          // F invFp S dFpinv^T
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          double base = (i == k) ? ml[j] : 0.0;
          col[3 * i + j] = base + R2[3 * i + j] + R4[3 * i + j] + R1[3 * i + j];
        }
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          double v = col[3 * i + j];
          st(dPdF, ln, i + 3 * j + 9 * k + 27 * l, v);
          n2 = fma(v, v, n2);
        }
    }
  }
  nrm[e] = n2;
}

}  // namespace

// ---------------------------------------------------------------------------- launchers
template <int NSL, bool QP>
static void launch_stress(const DmkMPDev* mp, const double* Ftarget, const DmkPhaseDev* ph, int phase_id,
                          const DmkNumMech* num, double dt, hipStream_t s) {
  dim3 blk(SWG), grd((mp->N + SWG - 1) / SWG);
  hipLaunchKernelGGL((stress_kernel<NSL, QP>), grd, blk, 0, s, *mp, Ftarget, ph, phase_id, *num, dt);
}

template <int NSL, bool QP>
static void launch_tangent(const DmkMPDev* mp, const DmkPhaseDev* ph, int phase_id, double dt, const int* flag,
                           double* dPdF, double* nrm, hipStream_t s) {
  dim3 blk(SWG), grd((mp->N + SWG - 1) / SWG);
  hipLaunchKernelGGL((tangent_kernel<NSL, QP>), grd, blk, 0, s, *mp, ph, phase_id, dt, flag, dPdF, nrm);
}

extern "C" int dmk_launch_stress(const DmkMPDev* mp, const double* Ftarget, const DmkPhaseDev* ph,
                                 const DmkPhaseDev* ph_host, int phase_id, const DmkNumMech* num, double dt,
                                 hipStream_t s) {
  const bool qp = ph_host->quarter_pow != 0;
  switch (ph_host->nsl) {
    case 12: qp ? launch_stress<12, true>(mp, Ftarget, ph, phase_id, num, dt, s)
                : launch_stress<12, false>(mp, Ftarget, ph, phase_id, num, dt, s); break;
    case 24: qp ? launch_stress<24, true>(mp, Ftarget, ph, phase_id, num, dt, s)
                : launch_stress<24, false>(mp, Ftarget, ph, phase_id, num, dt, s); break;
    default: return 1;
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

extern "C" int dmk_launch_stress_batch(const DmkStressItem* items_dev, int n_items, int n_blocks, int nsl,
                                       int quarter_pow, hipStream_t s) {
  if (n_items < 1 || n_blocks < 1) return 0;
  dim3 blk(SWG), grd(n_blocks);
  if (nsl == 12 && quarter_pow)
    hipLaunchKernelGGL((stress_kernel_batch<12, true>), grd, blk, 0, s, items_dev, n_items);
  else if (nsl == 12)
    hipLaunchKernelGGL((stress_kernel_batch<12, false>), grd, blk, 0, s, items_dev, n_items);
  else if (nsl == 24 && quarter_pow)
    hipLaunchKernelGGL((stress_kernel_batch<24, true>), grd, blk, 0, s, items_dev, n_items);
  else if (nsl == 24)
    hipLaunchKernelGGL((stress_kernel_batch<24, false>), grd, blk, 0, s, items_dev, n_items);
  else
    return 1;
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

extern "C" int dmk_launch_tangent(const DmkMPDev* mp, const DmkPhaseDev* ph, const DmkPhaseDev* ph_host,
                                  int phase_id, double dt, const int* flag, double* dPdF, double* nrm,
                                  hipStream_t s) {
  const bool qp = ph_host->quarter_pow != 0;
  switch (ph_host->nsl) {
    case 12: qp ? launch_tangent<12, true>(mp, ph, phase_id, dt, flag, dPdF, nrm, s)
                : launch_tangent<12, false>(mp, ph, phase_id, dt, flag, dPdF, nrm, s); break;
    case 24: qp ? launch_tangent<24, true>(mp, ph, phase_id, dt, flag, dPdF, nrm, s)
                : launch_tangent<24, false>(mp, ph, phase_id, dt, flag, dPdF, nrm, s); break;
    default: return 1;
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

extern "C" int dmk_pow_law(const double* x, const double* y, double* out, int32_t n, int32_t quarter) {
  if (!x || !y || !out || n < 1) { dmk_set_error("dmk_pow_law: invalid argument"); return DMK_ERR_ARG; }
  if (quarter)
    for (int i = 0; i < n; i++)
      if (!(y[i] > 0.0 && y[i] <= 64.0 && y[i] * 4.0 == std::rint(y[i] * 4.0))) {
        dmk_set_error("dmk_pow_law: quarter-power exponents are positive multiples of 1/4 up to 64");
        return DMK_ERR_ARG;
      }
  double* d = nullptr;
  const size_t b = sizeof(double) * (size_t)n;
  if (hipMalloc((void**)&d, 3 * b) != hipSuccess) { hipGetLastError(); dmk_set_error("dmk_pow_law: device memory"); return DMK_ERR_HIP; }
  int rc = 0;
  if (hipMemcpy(d, x, b, hipMemcpyHostToDevice) != hipSuccess || hipMemcpy(d + n, y, b, hipMemcpyHostToDevice) != hipSuccess)
    rc = DMK_ERR_HIP;
  if (!rc) {
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (quarter) hipLaunchKernelGGL(pow_law_kernel<true>, dim3(nb), dim3(256), 0, 0, d, d + n, d + 2 * n, n);
    else hipLaunchKernelGGL(pow_law_kernel<false>, dim3(nb), dim3(256), 0, 0, d, d + n, d + 2 * n, n);
    if (hipGetLastError() != hipSuccess || hipMemcpy(out, d + 2 * n, b, hipMemcpyDeviceToHost) != hipSuccess)
      rc = DMK_ERR_HIP;
  }
  hipFree(d);
  if (rc) dmk_set_error("dmk_pow_law: launch or copy failed");
  return rc;
}
