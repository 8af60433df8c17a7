/* CPU restatement of CCAKE's BSQ diffusion for D = 2 (test infrastructure
 * only, like the rest of oracle/):
 *   - diffusion transport coefficients kappa_q, tau_q
 *     (src/transport_coefficients.cpp:571-620, 917-1025);
 *   - the diffusion regulator (SPHWorkstation::regulate_diffusion,
 *     src/sph_workstation.cpp:108-218; PhysRevC.98.034916 Eqs. C6-C8);
 *   - EoM_default::calculate_MRF_diffusion (src/eom_default.cpp:1416-1822);
 *   - SPHWorkstation::set_diffusion (src/sph_workstation.cpp:1436-1459).
 * The reconstruction of q^mu from the extensive diffusion is part of
 * orc_reset_pi_tensor, the gradient-pass sums part of orc_smooth_gradients,
 * the charge and diffusion derivatives part of the EoM.
 */
#include "orc_internal.h"

#define TINY_TC 1e-25   /* include/eos_header.h:16 */

/* kappa (571-590) and tauq (592-620) for one particle; th = thermo_info order */
void orc_diffusion_transport(const orc_params* P, const double* th, double kappa[3][3], double tauq[3][3])
{
  const double T = th[0];
  memset(kappa, 0, 9*sizeof(double));
  memset(tauq, 0, 9*sizeof(double));
  if (P->diffusion_mode == 2) {                      /* kappaB_coth (963-995) */
    const double muB = th[1], nB = th[6], w = th[11];
    const double muB_over_T_floor = 1.0e-2;
    const double x = (T > TINY_TC) ? muB/T : 0.0;
    if (T > TINY_TC && fabs(w) > TINY_TC && fabs(x) > muB_over_T_floor) {
      const double kB = (P->C_B/T)*nB*((1.0/3.0)/tanh(x) - nB*T/w);
      if (isfinite(kB) && kB > 0.0) kappa[0][0] = kB;
    }
    if (P->diffusion_relax_mode == 1) {              /* tauqB_only (1003-1012) */
      if (T > TINY_TC) for (int i = 0; i < 3; ++i) tauq[i][i] = P->C_B/T;
      return;
    }
  } else {                                           /* default_kappa (917-951) */
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) kappa[i][j] = 1.0*P->kappa_over_T2[3*i + j]*(T*T);
  }
  for (int i = 0; i < 3; ++i) tauq[i][i] += 0.2/T;   /* default_tauq (955-971) */
}

/* regulate_diffusion (133-218) for particle i after q^mu has been rebuilt */
void orc_regulate_diffusion(orc_event* ev)
{
  const orc_params* P = &ev->p;
  const double hbarc_GeVfm = 0.1973269804;
  const double chi0 = P->regulator_chi0;
  const double e0 = P->regulator_e0/hbarc_GeVfm;
  const double xi0 = P->regulator_xi0/hbarc_GeVfm;
  const double rq_max = P->regulator_rq_max;
  const double fermi_offset = 1.0/(exp(e0/xi0) + 1.0);
  const double t2 = ev->t*ev->t;
  #pragma omp parallel for schedule(static) num_threads(orc_nthreads())
  for (int i = 0; i < ev->n; ++i) {
    const double e = FLD(ev, ORC_F_E)[i];
    const double ferm = 1.0/(exp(-(e - e0)/xi0) + 1.0);
    const double f_str = chi0*(ferm - fermi_offset);
    const double sigma = FLD(ev, ORC_F_SIGMA)[i];
    const double rho[3] = { FLD(ev, ORC_F_RHOB)[i], FLD(ev, ORC_F_RHOS)[i], FLD(ev, ORC_F_RHOQ)[i] };
    for (int c = 0; c < 3; ++c) {
      const double n_abs = fabs(rho[c]);
      if (n_abs < 1.0e-15) continue;
      double* q[4];
      for (int m = 0; m < 4; ++m) q[m] = &FLD(ev, ORC_F_Q0 + 4*c + m)[i];
      const double q0 = *q[0], q1 = *q[1], q2 = *q[2], q3 = *q[3];
      const double q3_sq = t2*q3*q3;                 /* hyperbolic coordinates */
      const double q_sq = q1*q1 + q2*q2 + q3_sq - q0*q0;
      if (q_sq <= 0.0) continue;
      const double q_mag = sqrt(q_sq);
      const double threshold = rq_max*n_abs*f_str;
      if (q_mag <= threshold) continue;
      const double scale = threshold/q_mag;
      *q[0] = q0*scale; *q[1] = q1*scale; *q[2] = q2*scale; *q[3] = q3*scale;
      for (int d = 0; d < 3; ++d) FLD(ev, ORC_F_EXT_Q0 + 3*c + d)[i] = *q[d + 1]/sigma;
    }
  }
}

/* set_diffusion (1441-1459): extensive q^i = q^i / sigma */
void orc_set_diffusion(orc_event* ev)
{
  for (int i = 0; i < ev->n; ++i) {
    const double sigma = FLD(ev, ORC_F_SIGMA)[i];
    for (int d = 0; d < 3; ++d)
      for (int c = 0; c < 3; ++c)
        FLD(ev, ORC_F_EXT_Q0 + 3*c + d)[i] = FLD(ev, ORC_F_Q0 + 4*c + d + 1)[i]/sigma;
  }
}

/* inverse of a 3x3 matrix by cofactors */
static void inv3(const double A[3][3], double R[3][3])
{
  const double c00 = A[1][1]*A[2][2] - A[1][2]*A[2][1];
  const double c01 = A[1][2]*A[2][0] - A[1][0]*A[2][2];
  const double c02 = A[1][0]*A[2][1] - A[1][1]*A[2][0];
  const double det = A[0][0]*c00 + A[0][1]*c01 + A[0][2]*c02;
  R[0][0] = c00/det; R[1][0] = c01/det; R[2][0] = c02/det;
  R[0][1] = (A[0][2]*A[2][1] - A[0][1]*A[2][2])/det;
  R[1][1] = (A[0][0]*A[2][2] - A[0][2]*A[2][0])/det;
  R[2][1] = (A[0][1]*A[2][0] - A[0][0]*A[2][1])/det;
  R[0][2] = (A[0][1]*A[1][2] - A[0][2]*A[1][1])/det;
  R[1][2] = (A[0][2]*A[1][0] - A[0][0]*A[1][2])/det;
  R[2][2] = (A[0][0]*A[1][1] - A[0][1]*A[1][0])/det;
}

/* calculate_MRF_diffusion (1416-1822) for particle i, D = 2.
 * In: the entropy MRFs M_S (D), F_S, R_S (3) of this evaluation.
 * Out: F_q0[c], R_q0[c][c'], M_q0[idir][c] (the q^0 MRFs stored as F_q0_a,
 * R_q0_a_b, M_q0_j_a) and F_q[idir][c], M_q[idir][c][jdir], R_q[idir][c][c']
 * (the extensive diffusion MRFs). */
void orc_mrf_diffusion(orc_event* ev, int i, const double M_S_in[3], double F_S, const double R_S[3],
                       double F_q0[3], double R_q0[3][3], double M_q0[3][3],
                       double F_q[3][3], double M_q[3][3][3], double R_q[3][3][3])
{
  const int D = 2;
  const double t = ev->t, t2 = t*t;
  const double cmd[4] = { 1.0, -1.0, -1.0, -1.0/t/t };
  const double sigma = FLD(ev, ORC_F_SIGMA)[i], gamma = FLD(ev, ORC_F_GAMMA)[i];
  const double s = FLD(ev, ORC_F_S)[i];
  const double rho[3] = { FLD(ev, ORC_F_RHOB)[i], FLD(ev, ORC_F_RHOS)[i], FLD(ev, ORC_F_RHOQ)[i] };
  double das[3], dar[3][3];
  for (int c = 0; c < 3; ++c) {
    das[c] = FLD(ev, ORC_F_DALPHA0 + 4*c)[i];
    for (int k = 0; k < 3; ++k) dar[c][k] = FLD(ev, ORC_F_DALPHA0 + 4*c + 1 + k)[i];
  }
  double M_S[3] = { 0, 0, 0 }, u[3] = { 0, 0, 0 }, fsu[3] = { 0, 0, 0 }, galpha[3][3] = {{0}};
  double divV = 0.0;
  for (int d = 0; d < D; ++d) {
    M_S[d] = M_S_in[d];
    u[d] = fsu[d] = FLD(ev, ORC_F_UX + d)[i];
    divV += FLD(ev, ORC_F_GRADV0 + 3*d + d)[i];
    for (int c = 0; c < 3; ++c) galpha[d][c] = FLD(ev, ORC_F_GRAD_ALPHA0 + 3*d + c)[i];
  }
  double tq[3][3], kq[3][3], q[3][4];
  for (int c = 0; c < 3; ++c) {
    for (int k = 0; k < 3; ++k) { tq[c][k] = FLD(ev, ORC_F_TAU_Q0 + 3*c + k)[i]; kq[c][k] = FLD(ev, ORC_F_KAPPA_Q0 + 3*c + k)[i]; }
    for (int m = 0; m < 4; ++m) q[c][m] = FLD(ev, ORC_F_Q0 + 4*c + m)[i];
  }
  /* covariant q: Vector<double,3>::make_covariant on the spatial part */
  double qc[3][4];
  for (int c = 0; c < 3; ++c) {
    qc[c][0] = q[c][0];
    qc[c][1] = -q[c][1]; qc[c][2] = -q[c][2]; qc[c][3] = q[c][3]*-t2;
  }
  const double uc[2] = { -u[0], -u[1] };
  const double gf = 0.0;                              /* delta_i_eta(D-1) = 0 for D = 2 */
  const double fsuc[3] = { -fsu[0], -fsu[1], fsu[2]*-t2 };
  double itq[3][3], itt[3][3], itk[3][3];
  inv3(tq, itq);
  for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) {
    itt[a][b] = 0.0; itk[a][b] = 0.0;
    for (int k = 0; k < 3; ++k) { itt[a][b] += itq[a][k]*tq[k][b]; itk[a][b] += itq[a][k]*kq[k][b]; }
  }
  const double g2 = gamma*gamma;
  const double expf = gamma*divV + gf + gamma/t;
  double F0D[3], F0a[3], A0a[3], R0a[3][3], M0a[3][2], M0D[3][2];
  for (int c = 0; c < 3; ++c) {
    F0D[c] = g2*gf*q[c][0] + 2.*g2*qc[c][3]*fsu[2]/t + t*gamma*qc[c][0]*fsu[2]*fsu[2] + t*fsu[2]*q[c][3];
    F0a[c] = (1. - g2)*sigma*das[c]*F_S;
    A0a[c] = -(1. - g2)*das[c]*s/gamma;
    for (int d = 0; d < D; ++d) {
      M0D[c][d] = g2*(qc[c][d + 1] - qc[c][0]*uc[d]/gamma);
      M0a[c][d] = (1. - g2)*sigma*das[c]*M_S[d];
      F0a[c] += -galpha[d][c]*u[d]/gamma;
    }
    for (int k = 0; k < 3; ++k) {
      R0a[c][k] = (1. - g2)*sigma*(das[c]*R_S[k] + dar[c][k]);
      A0a[c] += -(1. - g2)*dar[c][k]*rho[k]/gamma;
    }
  }
  for (int c = 0; c < 3; ++c) {
    F_q0[c] = 0.0;
    for (int d = 0; d < D; ++d) {
      double m = 0.0;
      for (int k = 0; k < 3; ++k)
        m += (-itt[c][k]*M0D[k][d] + itk[c][k]*M0a[k][d] - itk[c][k]*A0a[k]*uc[d]/gamma)/(gamma)
           + itt[c][k]*q[k][0]*uc[d]/gamma/gamma;
      M_q0[d][c] = m;
    }
    for (int k = 0; k < 3; ++k) {
      F_q0[c] += (-itt[c][k]*F0D[k] - itq[c][k]*q[k][0] + itk[c][k]*F0a[k] + itk[c][k]*A0a[k]*expf)/(gamma)
               - itt[c][k]*q[k][0]*expf/gamma;
      double r = 0.0;
      for (int kk = 0; kk < 3; ++kk) r += itk[c][kk]*R0a[kk][k]/(gamma);
      R_q0[c][k] = r;
    }
  }
  /* the spatial MRFs; F_D keeps only its first two terms: the reference ends
     that statement with a semicolon after them (src/eom_default.cpp:1744-1745) */
  double FD[3][3], Fa[3][3], Aa[3][3], MD[3][3][3], Ma[3][3][3], Ra[3][3][3];
  for (int d = 0; d < 3; ++d) for (int c = 0; c < 3; ++c) {
    FD[d][c] = fsu[d]*gamma*q[c][0]*gf + 2.*fsu[d]*qc[c][3]*fsu[2]*gamma/t;
    Fa[d][c] = -fsu[d]*gamma*sigma*das[c]*F_S + cmd[d + 1]*galpha[d][c];
    Aa[d][c] = fsu[d]*das[c]*s;
    for (int j = 0; j < 3; ++j) {
      MD[d][c][j] = fsu[d]*gamma*(qc[c][j + 1] - qc[c][0]*fsuc[j]/gamma);
      Ma[d][c][j] = -fsu[d]*gamma*sigma*das[c]*M_S[j];
    }
    for (int k = 0; k < 3; ++k) {
      Ra[d][c][k] = -fsu[d]*gamma*sigma*(das[c]*R_S[k] + dar[c][k]);
      Aa[d][c] += fsu[d]*dar[c][k]*rho[k];
    }
  }
  for (int d = 0; d < 3; ++d) for (int c = 0; c < 3; ++c) {
    F_q[d][c] = 0.0;
    for (int k = 0; k < 3; ++k) {
      F_q[d][c] += (-itt[c][k]*FD[d][k] - itq[c][k]*q[k][d + 1] + itk[c][k]*Fa[d][k]
                    + itk[c][k]*Aa[d][k]*expf)/(gamma*sigma);
      double r = 0.0;
      for (int kk = 0; kk < 3; ++kk) r += itk[c][kk]*Ra[d][kk][k]/(gamma*sigma);
      R_q[d][c][k] = r;
    }
    for (int j = 0; j < 3; ++j) {
      double m = 0.0;
      for (int k = 0; k < 3; ++k)
        m += (-itt[c][k]*MD[d][k][j] + itk[c][k]*Ma[d][k][j] - itk[c][k]*Aa[d][k]*fsuc[j]/gamma)/(gamma*sigma);
      M_q[d][c][j] = m;
    }
  }
}
