// Tangent-linear (forward-mode) build of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  The oracle's C sources are
// compiled a second time as C++ with every `double` replaced by a dual number
// (oracle_dual.hpp), so one load case of DAMASK_grid (DAMASK_grid.f90:288-460
// as restated in oracle_solver.c) carries, next to each value, its derivative
// along one direction in the phenopowerlaw / elastic parameters.  The values
// are the plain oracle's, operation for operation; the derivative of the
// volume-averaged stress history is the FP64 forward-mode reference gradient
// the adjoint is checked against.  The derivative rides along the solver's
// own iterations (the Lp Newton, the state fixed point and the spectral fixed
// point, NGMRES or Richardson as the numerics choose), and each spectral
// solve stops only once the derivative part of its residual has also fallen
// below tangent_rtol of the derivative field: the derivative of a convergent
// fixed-point iteration converges to the derivative of its fixed point.  The
// material point's own iterations stop on their values, so the derivative
// residual levels off at a small fraction of the field, larger for parameters
// of little influence (1e-8 for an h_sl-sl entry of the 24-system BCC phase);
// 1e-7 is the default, ample for checks at 1e-5.
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "oracle_dual.hpp"

namespace orc_fwd {
#pragma omp declare reduction(+ : Dual : omp_out += omp_in) initializer(omp_priv = Dual(0.0))

double tangent_rtol = 1e-7;

// max |r'| <= tangent_rtol max |x'| over the solver's field and residual
inline int tangent_converged(const Dual* r, const Dual* x, size_t n) {
  double rm = 0.0, xm = 0.0;
  for (size_t i = 0; i < n; i++) {
    rm = std::fmax(rm, std::fabs(r[i].d));
    xm = std::fmax(xm, std::fabs(x[i].d));
  }
  return rm <= tangent_rtol * xm;
}

// Tangent of the truncated least squares solution a = H^+ b of the NGMRES
// step (H symmetric, H = V diag(lam) V^T in values, eigenvalues at or below
// `cut` dropped): the derivative of the pseudo-inverse (Golub & Pereyra 1973)
//   a' = H^+ (b' - H' a) + (I - P) H' H^+ a + H^+ H^+ H' (I - P) b,
// P = H H^+ the projector on the kept eigenvectors.  Replaces the derivative
// carried through the Jacobi sweeps, whose rotation angles of vanishing
// pivots have unbounded derivatives.
inline void lsq_tangent(int m, const Dual* H, const Dual* b, const Dual* lam, const Dual* V, Dual cut, Dual* a) {
  double Hp[30 * 30] = {0}, P[30 * 30] = {0};
  for (int k = 0; k < m; k++) {
    const double l = lam[k + m * k].v;
    if (std::fabs(l) <= cut.v) continue;
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++) {
        const double vv = V[i + m * k].v * V[j + m * k].v;
        Hp[i + m * j] += vv / l;
        P[i + m * j] += vv;
      }
  }
  double r[30], u[30], w[30], Hpa[30], t1[30], res[30];
  for (int i = 0; i < m; i++) {       // r = b' - H' a,  w = (I - P) b
    double s = b[i].d, q = b[i].v;
    for (int j = 0; j < m; j++) {
      s -= H[i + m * j].d * a[j].v;
      q -= P[i + m * j] * b[j].v;
    }
    r[i] = s;
    w[i] = q;
  }
  for (int i = 0; i < m; i++) {       // H^+ a and H' (I - P) b
    double s = 0.0, q = 0.0;
    for (int j = 0; j < m; j++) {
      s += Hp[i + m * j] * a[j].v;
      q += H[i + m * j].d * w[j];
    }
    Hpa[i] = s;
    u[i] = q;
  }
  for (int i = 0; i < m; i++) {       // H' H^+ a
    double s = 0.0;
    for (int j = 0; j < m; j++) s += H[i + m * j].d * Hpa[j];
    t1[i] = s;
  }
  for (int i = 0; i < m; i++) {       // H^+ r + (I - P) t1, and H^+ u
    double s = 0.0, q = t1[i], z = 0.0;
    for (int j = 0; j < m; j++) {
      s += Hp[i + m * j] * r[j];
      q -= P[i + m * j] * t1[j];
      z += Hp[i + m * j] * u[j];
    }
    res[i] = s + q;
    w[i] = z;
  }
  for (int i = 0; i < m; i++) {       // + H^+ (H^+ u)
    double s = 0.0;
    for (int j = 0; j < m; j++) s += Hp[i + m * j] * w[j];
    a[i].d = res[i] + s;
  }
}
}  // namespace orc_fwd

#define ORC_NO_EXTERN_C
#define ORC_CPLX_DEFINED
#define ORC_TANGENT_CONVERGED(r, x, n) tangent_converged(r, x, n)
#define ORC_LSQ_TANGENT(m, H, b, lam, V, cut, a) lsq_tangent(m, H, b, lam, V, cut, a)

namespace orc_fwd {
typedef CDual cplx;
#define double Dual
#include "oracle_math.c"
#include "oracle_phase.c"
#define I CDual(Dual(0.0), Dual(1.0))
#include "oracle_spectral.c"
#undef I
#include "oracle_solver.c"
#undef double
}  // namespace orc_fwd

// the plain-double structures of the C interface
#undef DAMASK_ORACLE_H
namespace orc_dbl {
#include "oracle.h"
}

namespace {

using orc_fwd::Dual;

Dual dual(double v, double d) { return Dual(v, d); }

void convert(const orc_dbl::orc_phase_desc& a, const orc_dbl::orc_phase_desc& s, orc_fwd::orc_phase_desc& o) {
  o.lattice = a.lattice;
  o.n_fam = a.n_fam;
  for (int i = 0; i < 4; i++) o.N_sl[i] = a.N_sl[i];
  o.C_11 = dual(a.C_11, s.C_11);
  o.C_12 = dual(a.C_12, s.C_12);
  o.C_44 = dual(a.C_44, s.C_44);
  for (int i = 0; i < 4; i++) {
    o.dot_gamma_0[i] = dual(a.dot_gamma_0[i], s.dot_gamma_0[i]);
    o.n[i] = dual(a.n[i], s.n[i]);
    o.a[i] = dual(a.a[i], s.a[i]);
    o.h_0[i] = dual(a.h_0[i], s.h_0[i]);
    o.xi_0[i] = dual(a.xi_0[i], s.xi_0[i]);
    o.xi_inf[i] = dual(a.xi_inf[i], s.xi_inf[i]);
    o.c_1[i] = dual(a.c_1[i], s.c_1[i]);
    o.c_2[i] = dual(a.c_2[i], s.c_2[i]);
    o.f_sat[i] = dual(a.f_sat[i], s.f_sat[i]);
  }
  o.n_h = a.n_h;
  for (int i = 0; i < 32; i++) o.h_sl_sl[i] = dual(a.h_sl_sl[i], s.h_sl_sl[i]);
  o.n_nS = a.n_nS;
  for (int i = 0; i < 6; i++) o.a_nS[i] = dual(a.a_nS[i], s.a_nS[i]);
  o.atol_xi = a.atol_xi;
  o.atol_gamma = a.atol_gamma;
}

void convert(const orc_dbl::orc_num_mech& a, orc_fwd::orc_num_mech& o) {
  o.r_cutback_min = a.r_cutback_min;
  o.r_cutback = a.r_cutback;
  o.r_increase = a.r_increase;
  o.eps_rel_state = a.eps_rel_state;
  o.N_iter_state_max = a.N_iter_state_max;
  o.N_iter_Lp_max = a.N_iter_Lp_max;
  o.r_linesearch_Lp = a.r_linesearch_Lp;
  o.eps_rel_Lp = a.eps_rel_Lp;
  o.eps_abs_Lp = a.eps_abs_Lp;
  o.f_update_jacobi_Lp = a.f_update_jacobi_Lp;
  o.N_iter_Li_max = a.N_iter_Li_max;
  o.r_linesearch_Li = a.r_linesearch_Li;
  o.eps_rel_Li = a.eps_rel_Li;
  o.eps_abs_Li = a.eps_abs_Li;
  o.f_update_jacobi_Li = a.f_update_jacobi_Li;
}

void convert(const orc_dbl::orc_num_grid& a, orc_fwd::orc_num_grid& o) {
  o.N_iter_min = a.N_iter_min;
  o.N_iter_max = a.N_iter_max;
  o.eps_abs_div = a.eps_abs_div;
  o.eps_rel_div = a.eps_rel_div;
  o.eps_abs_P = a.eps_abs_P;
  o.eps_rel_P = a.eps_rel_P;
  o.update_gamma = a.update_gamma;
  o.divergence_correction = a.divergence_correction;
  o.solver = a.solver;
  o.ngmres_m = a.ngmres_m;
  o.snes_max_it = a.snes_max_it;
  o.N_cutback_max = a.N_cutback_max;
}

void convert(const orc_dbl::orc_loadcase& a, orc_fwd::orc_loadcase& o) {
  o.def_type = a.def_type;
  o.stress_type = a.stress_type;
  for (int i = 0; i < 9; i++) {
    o.def_values[i] = a.def_values[i];
    o.def_mask[i] = a.def_mask[i];
    o.stress_values[i] = a.stress_values[i];
    o.stress_mask[i] = a.stress_mask[i];
  }
  for (int i = 0; i < 4; i++) o.q_R[i] = a.q_R[i];
  o.t = a.t;
  o.r = a.r;
  o.N = a.N;
  o.f_out = a.f_out;
  o.estimate_rate = a.estimate_rate;
}

}  // namespace

// One load case from a fresh state (the plain oracle's orc_grid_create,
// orc_mp_create, orc_solver_create, orc_solver_set_cref_rule and
// orc_run_loadcase), with the phase
// parameters seeded by the tangent `seed` (same layout as the descriptors;
// zero where a parameter is held fixed).
// Per increment: P_av (9, column-major) and its directional derivative dP_av,
// stats as orc_run_loadcase's (4 values).  Returns orc_run_loadcase's code, or
// -1 when a phase descriptor is rejected.
extern "C" __attribute__((visibility("default"))) int orc_fwd_run_loadcase(
    const int cells[3], const double size[3], int n_phases, const orc_dbl::orc_phase_desc* desc,
    const orc_dbl::orc_phase_desc* seed, const int32_t* phase_of, const double* q0,
    const orc_dbl::orc_num_mech* nm, const orc_dbl::orc_num_grid* ng, int cref_rule,
    const orc_dbl::orc_loadcase* lc, double tangent_rtol, double* P_av, double* dP_av, double* stats,
    int64_t* n_updates) {
  orc_fwd::tangent_rtol = tangent_rtol;
  std::vector<orc_fwd::orc_phase_desc> dd(n_phases);
  std::vector<orc_fwd::orc_phase> ph(n_phases);
  for (int p = 0; p < n_phases; p++) {
    convert(desc[p], seed[p], dd[p]);
    if (orc_fwd::orc_phase_build(&dd[p], &ph[p])) return -1;
  }
  const long N = (long)cells[0] * cells[1] * cells[2];
  std::vector<Dual> q(4 * N);
  for (long i = 0; i < 4 * N; i++) q[i] = q0[i];
  orc_fwd::orc_num_mech nmd;
  orc_fwd::orc_num_grid ngd;
  orc_fwd::orc_loadcase lcd;
  convert(*nm, nmd);
  convert(*ng, ngd);
  convert(*lc, lcd);
  Dual sz[3] = {size[0], size[1], size[2]};
  orc_fwd::orc_mp* m = orc_fwd::orc_mp_create((int)N, n_phases, ph.data(), phase_of, q.data(), &nmd);
  orc_fwd::orc_grid* g = orc_fwd::orc_grid_create(cells, sz, &ngd);
  orc_fwd::orc_solver* s = orc_fwd::orc_solver_create(g, m, &ngd);
  orc_fwd::orc_solver_set_cref_rule(s, cref_rule);
  const int n = lc->N;
  std::vector<Dual> st(4 * n), P(9 * n), Fa(9 * n);
  Dual t = 0.0;
  int64_t nu = 0;
  int rc = orc_fwd::orc_run_loadcase(s, &lcd, &t, st.data(), P.data(), Fa.data(), &nu);
  for (int i = 0; i < 9 * n; i++) {
    P_av[i] = P[i].v;
    dP_av[i] = P[i].d;
  }
  if (stats)
    for (int i = 0; i < 4 * n; i++) stats[i] = st[i].v;
  if (n_updates) *n_updates = nu;
  orc_fwd::orc_solver_destroy(s);
  orc_fwd::orc_grid_destroy(g);
  orc_fwd::orc_mp_destroy(m);
  return rc;
}
