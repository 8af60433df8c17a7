/*
 * orc_eom.c -- CPU restatement of the per-particle physics of CCAKE's
 * time-step (test infrastructure only; see ccake_oracle.h):
 *   reset_pi_tensor          src/eom_default.cpp:56-172
 *   transport coefficients   src/transport_coefficients.cpp:9-912,
 *                            src/sph_workstation.cpp:1622-1699
 *   evaluate_time_derivatives src/eom_default.cpp:211-869 with
 *   calculate_MRF_shear      src/eom_default.cpp:882-1412
 * Milne index conventions follow include/milne.hpp: a 2-vector or 2x2
 * matrix is lowered/raised by a pure sign flip (milne.hpp:353-364,
 * 488-530); larger objects carry -tau^2 on the eta slot; 4x4 tensors use
 * the dedicated slot rules (milne.hpp:546-640).
 */
#include "orc_internal.h"

/* ---------------- Milne helpers (include/milne.hpp) ---------------- */
static void vec_cov(int D, double* v, double t2)
{
  if (D == 2) { v[0] = -v[0]; v[1] = -v[1]; return; }
  for (int i = 0; i < D-1; ++i) v[i] = -v[i];
  v[D-1] *= -t2;
}
static void vec_contra(int D, double* v, double t2)
{
  if (D == 2) { v[0] = -v[0]; v[1] = -v[1]; return; }
  for (int i = 0; i < D-1; ++i) v[i] = -v[i];
  v[D-1] *= -1.0/t2;
}
/* Vector<double,3>::make_covariant (generic rule) */
static void vec3_cov(double* v, double t2) { v[0] = -v[0]; v[1] = -v[1]; v[2] *= -t2; }
/* Matrix<double,4,4>::make_covariant(slot) */
static void mat4_cov(double M[4][4], int slot, double t2)
{
  if (slot == 0) {
    for (int i = 1; i < 3; ++i) for (int j = 0; j < 4; ++j) M[i][j] = -M[i][j];
    for (int j = 0; j < 4; ++j) M[3][j] *= -t2;
  } else {
    for (int i = 0; i < 4; ++i) for (int j = 1; j < 3; ++j) M[i][j] = -M[i][j];
    for (int i = 0; i < 4; ++i) M[i][3] *= -t2;
  }
}
/* Matrix<double,D,D>::make_contravariant(0) */
static void matD_contra0(int D, double M[3][3], double t2)
{
  if (D == 2) { for (int i = 0; i < 2; ++i) for (int j = 0; j < 2; ++j) M[i][j] = -M[i][j]; return; }
  for (int i = 0; i < D-1; ++i) for (int j = 0; j < D; ++j) M[i][j] = -M[i][j];
  for (int j = 0; j < D; ++j) M[D-1][j] *= -1.0/t2;
}
/* milne::inverse for 1x1, 2x2, 3x3 (milne.hpp:790-829) */
static void inverseD(int D, double A[3][3], double R[3][3])
{
  if (D == 1) { R[0][0] = 1.0/A[0][0]; return; }
  if (D == 2) {
    double det = A[0][0]*A[1][1] - A[0][1]*A[1][0];
    R[0][0] = A[1][1]/det; R[0][1] = -A[0][1]/det; R[1][0] = -A[1][0]/det; R[1][1] = A[0][0]/det;
    return;
  }
  double det = A[0][0]*A[1][1]*A[2][2] + A[0][1]*A[1][2]*A[2][0] + A[0][2]*A[1][0]*A[2][1]
             - A[0][2]*A[1][1]*A[2][0] - A[0][1]*A[1][0]*A[2][2] - A[0][0]*A[1][2]*A[2][1];
  R[0][0] = (A[1][1]*A[2][2] - A[1][2]*A[2][1])/det;
  R[0][1] = (A[0][2]*A[2][1] - A[0][1]*A[2][2])/det;
  R[0][2] = (A[0][1]*A[1][2] - A[0][2]*A[1][1])/det;
  R[1][0] = (A[1][2]*A[2][0] - A[1][0]*A[2][2])/det;
  R[1][1] = (A[0][0]*A[2][2] - A[0][2]*A[2][0])/det;
  R[1][2] = (A[0][2]*A[1][0] - A[0][0]*A[1][2])/det;
  R[2][0] = (A[1][0]*A[2][1] - A[1][1]*A[2][0])/det;
  R[2][1] = (A[0][1]*A[2][0] - A[0][0]*A[2][1])/det;
  R[2][2] = (A[0][0]*A[1][1] - A[0][1]*A[1][0])/det;
}
/* delta_i_eta<D> (milne.hpp:1095-1108); the 2D specialisation is zero. */
static double delta_eta(int D, int i) { return (D == 2) ? 0.0 : (i == D-1 ? 1.0 : 0.0); }

static void load_shv(orc_event* ev, int i, double S[4][4])
{
  for (int a = 0; a < 4; ++a) for (int b = 0; b < 4; ++b) S[a][b] = FLD(ev, ORC_F_SHV0 + 4*a + b)[i];
}

/* ------------------------------------------------------------------ */
/* EoM_default::reset_pi_tensor (src/eom_default.cpp:56-172) */
void orc_reset_pi_tensor(orc_event* ev)
{
  const int D = ev->p.dim;
  const double t = ev->t, t2 = t*t;
  #pragma omp parallel for schedule(static) num_threads(orc_nthreads())
  for (int i = 0; i < ev->n; ++i) {
    const double gamma = FLD(ev, ORC_F_GAMMA)[i];
    const double sigma = FLD(ev, ORC_F_SIGMA)[i];
    double fu[3] = { 0, 0, 0 };
    for (int d = 0; d < D; ++d) fu[d] = FLD(ev, ORC_F_UX + d)[i];
    double fc[3] = { fu[0], fu[1], fu[2] };
    vec3_cov(fc, t2);
    FLD(ev, ORC_F_BULK)[i] = FLD(ev, ORC_F_EXT_BULK)[i]*sigma;
    double S[4][4];
    memset(S, 0, sizeof(S));
    int k = 0;
    for (int a = 0; a < 2; ++a) for (int b = a; b < 3; ++b) S[a+1][b+1] = FLD(ev, ORC_F_EXT_SHV0 + k++)[i]*sigma;
    for (int a = 1; a < 4; ++a) for (int b = a+1; b < 4; ++b) S[b][a] = S[a][b];
    S[0][0] = ( fc[0]*fc[0]*S[1][1] + fc[1]*fc[1]*S[2][2] + 2.*fc[1]*fc[0]*S[2][1]
              + 2.*fc[2]*fc[0]*S[3][1] + 2.*fc[2]*fc[1]*S[3][2]
              - fc[2]*fc[2]*(S[1][1] + S[2][2])/t2 )/(gamma*gamma - fc[2]*fc[2]/t2);
    S[3][3] = (S[0][0] - S[1][1] - S[2][2])/t2;
    for (int a = 0; a < 3; ++a) {
      double p0 = 0.0;
      for (int b = 0; b < 3; ++b) p0 += -S[a+1][b+1]*fc[b]/gamma;
      S[0][a+1] = p0;
    }
    for (int a = 1; a < 4; ++a) S[a][0] = S[0][a];
    for (int a = 0; a < 4; ++a) for (int b = 0; b < 4; ++b) FLD(ev, ORC_F_SHV0 + 4*a + b)[i] = S[a][b];
    if (ev->p.using_diffusion) {
      /* q^i = extensive q^i sigma, q^0 = -q^i u_i / gamma (eom_default.cpp:139-163) */
      for (int c = 0; c < 3; ++c) {
        double qv[3];
        for (int d = 0; d < 3; ++d) qv[d] = FLD(ev, ORC_F_EXT_Q0 + 3*c + d)[i]*sigma;
        double q0 = 0.0;
        for (int d = 0; d < 3; ++d) q0 += -qv[d]*fc[d]/gamma;
        FLD(ev, ORC_F_Q0 + 4*c)[i] = q0;
        for (int d = 0; d < 3; ++d) FLD(ev, ORC_F_Q0 + 4*c + d + 1)[i] = qv[d];
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* transport coefficients (src/transport_coefficients.cpp:284-912).
 * th = 31 thermo doubles in thermo_info order. */
enum { TH_T = 0, TH_MUB, TH_MUS, TH_MUQ, TH_E, TH_S, TH_RB, TH_RS, TH_RQ, TH_P, TH_CS2, TH_W };

static double tc_eta(const orc_params* P, const double* th)
{
  const double w = th[TH_W], T = th[TH_T];
  if (P->eta_mode == 1) { if (T <= 0.0) return 0.0; return P->constant_eta_over_s*(w/T); }
  return 0.20*(w/T);
}
static double tc_tau_pi(const orc_params* P, const double* th)
{
  const double x = 5.0*tc_eta(P, th)/th[TH_W];
  if (P->shear_relax_mode == 2) return x;
  if (P->shear_relax_mode == 1) return fmax(x, 0.001);
  return fmax(x, 0.005);
}
static double tc_zeta(const orc_params* P, const double* th)
{
  const double w = th[TH_W], T = th[TH_T];
  switch (P->zeta_mode) {
    case 1: return P->constant_zeta_over_s*(w/T);
    case 2: {
      const double A = P->cs2_dependent_zeta_A, p = P->cs2_dependent_zeta_p;
      double factor = 1.0;
      if (P->modulate_zeta_with_tanh) {
        const double Ttr = 150.0/ORC_HBARC_MEVFM, Tsc = 10.0/ORC_HBARC_MEVFM;
        const double thx = tanh((T - Ttr)/Tsc);
        const double xp = pow(T/Ttr, p);
        factor = 0.5*(1.0 + xp) + 0.5*(1.0 - xp)*thx;
      }
      return A*factor*pow((1.0/3.0) - fmin(th[TH_CS2], 1.0), p)*w/T;
    }
    case 3: {
      const double d = T - P->zeta_gauss_T0;
      return P->zeta_gauss_A*exp(-d*d/(2.0*P->zeta_gauss_width*P->zeta_gauss_width))*w/T;
    }
    default: return 0.005*(w/T);
  }
}
static double tc_tau_Pi(const orc_params* P, const double* th)
{
  const double p = P->cs2_dependent_zeta_p;
  const double cc = th[TH_CS2] > 1 ? 1 : th[TH_CS2];
  if (pow(1.0/3.0 - cc, p) < 1e-10) return 1e10;
  return fmax(5.0*tc_zeta(P, th)/(pow(1.0/3.0 - cc, p)*th[TH_W]), 0.1);
}

static void transport_point(const orc_params* P, const double* th, double* out)
{
  const double tau_pi = tc_tau_pi(P, th);
  const double tau_Pi = tc_tau_Pi(P, th);
  const double cc = th[TH_CS2] > 1 ? 1 : th[TH_CS2];
  out[0] = tc_eta(P, th);
  out[1] = tau_pi;
  out[2] = tc_zeta(P, th);
  out[3] = tau_Pi;
  out[4] = P->delta_PiPi_mode == 1 ? 0.0 : (P->delta_PiPi_mode == 2 ? tau_Pi : 2.*tau_Pi/3.0);
  out[5] = P->delta_pipi_mode == 1 ? 0.0 : (P->delta_pipi_mode == 2 ? tau_pi : 4.*tau_pi/3.0);
  out[6] = P->lambda_piPi_mode == 1 ? 0.0 : 6./5.*tau_pi;
  out[7] = P->lambda_Pipi_mode == 1 ? 0.0 : 8.*(1./3. - cc)*tau_Pi/5.;
  out[8] = 0.0;                                  /* phi1 */
  out[9] = 0.0;                                  /* phi3 */
  out[10] = 0.0;                                 /* phi6 */
  out[11] = P->phi7_mode == 1 ? 0.0 : 9./(70.*th[TH_P]);
  out[12] = P->tau_pipi_mode == 1 ? 0.0 : 10.*tau_pi/7.0;
}

void orc_viscosities(orc_event* ev)
{
  #pragma omp parallel for schedule(static) num_threads(orc_nthreads())
  for (int i = 0; i < ev->n; ++i) {
    double th[31], out[13];
    for (int k = 0; k < 31; ++k) th[k] = FLD(ev, ORC_F_T + k)[i];
    transport_point(&ev->p, th, out);
    for (int k = 0; k < 13; ++k) FLD(ev, ORC_F_ETA_PI + k)[i] = out[k];
    if (ev->p.using_diffusion) {
      double kq[3][3], tq[3][3];
      orc_diffusion_transport(&ev->p, th, kq, tq);
      for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) {
        FLD(ev, ORC_F_KAPPA_Q0 + 3*a + b)[i] = kq[a][b];
        FLD(ev, ORC_F_TAU_Q0 + 3*a + b)[i] = tq[a][b];
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* EoM_default::evaluate_time_derivatives for one particle. */
static void eom_particle(orc_event* ev, int i)
{
  const orc_params* P = &ev->p;
  const int D = P->dim;
  const double t = ev->t, t2 = t*t;
  const double use_vort = P->using_vorticity ? 1.0 : 0.0;
  const int using_shear = P->using_shear;
  const double cmd[4] = { 1.0, -1.0, -1.0, -1.0/t/t };   /* get_contra_metric_diagonal<3> */
  double dl[3]; for (int a = 0; a < 3; ++a) dl[a] = (a < D) ? delta_eta(D, a) : 0.0;

  const double gamma = FLD(ev, ORC_F_GAMMA)[i];
  const double sigma = FLD(ev, ORC_F_SIGMA)[i];
  const double bulk = FLD(ev, ORC_F_BULK)[i];
  const double T = FLD(ev, ORC_F_T)[i], w = FLD(ev, ORC_F_W)[i], s = FLD(ev, ORC_F_S)[i];
  const double muB = FLD(ev, ORC_F_MUB)[i], muS = FLD(ev, ORC_F_MUS)[i], muQ = FLD(ev, ORC_F_MUQ)[i];
  const double rhoB = FLD(ev, ORC_F_RHOB)[i], rhoS = FLD(ev, ORC_F_RHOS)[i], rhoQ = FLD(ev, ORC_F_RHOQ)[i];
  const double dwds = FLD(ev, ORC_F_DWDS)[i], dwdB = FLD(ev, ORC_F_DWDB)[i];
  const double dwdS = FLD(ev, ORC_F_DWDS_S)[i], dwdQ = FLD(ev, ORC_F_DWDQ)[i];
  const double eta_pi = FLD(ev, ORC_F_ETA_PI)[i], tau_pi = FLD(ev, ORC_F_TAU_PI)[i];
  const double zeta = FLD(ev, ORC_F_ZETA_PI)[i], tau_Pi = FLD(ev, ORC_F_TAU_BULK)[i];
  const double delta_PiPi = FLD(ev, ORC_F_DELTA_BB)[i], delta_pipi = FLD(ev, ORC_F_DELTA_PIPI)[i];
  const double lambda_piPi = FLD(ev, ORC_F_LAMBDA_PIB)[i], lambda_Pipi = FLD(ev, ORC_F_LAMBDA_BPI)[i];
  const double phi1 = FLD(ev, ORC_F_PHI1)[i], phi3 = FLD(ev, ORC_F_PHI3)[i];
  const double phi6 = FLD(ev, ORC_F_PHI6)[i], phi7 = FLD(ev, ORC_F_PHI7)[i];
  const double tau_pipi = FLD(ev, ORC_F_TAU_PIPI)[i];

  double u[3] = {0,0,0}, v[3] = {0,0,0};
  for (int d = 0; d < D; ++d) { u[d] = FLD(ev, ORC_F_UX + d)[i]; v[d] = FLD(ev, ORC_F_VX + d)[i]; }
  double u_cov[3] = { u[0], u[1], u[2] };
  vec_cov(D, u_cov, t2);
  double fu[3] = { u[0], u[1], u[2] };
  if (D < 3) fu[2] = 0.0;
  double fc[3] = { fu[0], fu[1], fu[2] };
  vec3_cov(fc, t2);
  double gradV[3][3] = {{0}};
  double divV = 0.0;
  for (int a = 0; a < D; ++a) {
    divV += FLD(ev, ORC_F_GRADV0 + 3*a + a)[i];
    for (int b = 0; b < D; ++b) gradV[a][b] = FLD(ev, ORC_F_GRADV0 + 3*a + b)[i];
  }
  const double gf = dl[D-1]*u[D-1]*u[D-1]*t/gamma;   /* geometric_factor */
  double S[4][4], H[4][4], C[4][4];
  load_shv(ev, i, S);
  memcpy(H, S, sizeof(S)); mat4_cov(H, 1, t2);
  memcpy(C, H, sizeof(S)); mat4_cov(C, 0, t2);
  double shv_shvcov = 0.0;
  for (int a = 0; a < 4; ++a) for (int b = 0; b < 4; ++b) shv_shvcov += S[a][b]*C[a][b];

  /* ---------------- calculate_MRF_shear (882-1412) ---------------- */
  double M_shv_nabla_u[3] = {0,0,0}, F_shv_nabla_u = 0.0;
  double F_0i[3] = {0,0,0}, M_0i[3][3] = {{0}};
  double F_es[2][3] = {{0}}, M_es[2][3][3] = {{{0}}};
  double sigF[2][3] = {{0}}, sigM[2][3][3] = {{{0}}};   /* sigma^{ij} MRFs (compute_derivatives 784-808) */
  if (using_shear) {
    double grad_uj[3][3] = {{0}}, grad_u0[3] = {0,0,0};
    for (int a = 0; a < D; ++a) for (int b = 0; b < D; ++b) {
      grad_uj[a][b] = gamma*gradV[a][b];
      for (int k = 0; k < D; ++k) grad_uj[a][b] += -gamma*u[b]*u_cov[k]*gradV[a][k];
    }
    for (int a = 0; a < D; ++a) {
      double c = 0.0; for (int b = 0; b < D; ++b) c += grad_uj[a][b]*u_cov[b];
      grad_u0[a] = -1.*c/gamma;
    }
    double gu0c[3] = { grad_u0[0], grad_u0[1], grad_u0[2] };
    vec_contra(D, gu0c, t2);
    double cgu[3][3]; memcpy(cgu, grad_uj, sizeof(cgu));
    matD_contra0(D, cgu, t2);
    double cg3[3][3] = {{0}}, g3[3][3] = {{0}};
    for (int a = 0; a < D; ++a) for (int b = 0; b < D; ++b) { cg3[a][b] = cgu[a][b]; g3[a][b] = grad_uj[a][b]; }
    double guv[3] = {0,0,0};      /* contract(grad_uj, v, FirstIndex) */
    for (int b = 0; b < D; ++b) for (int a = 0; a < D; ++a) guv[b] += grad_uj[a][b]*v[a];
    double gu0v = 0.0; for (int a = 0; a < D; ++a) gu0v += grad_u0[a]*v[a];
    double v_gradu_shv[3] = {0,0,0}, gugs[3] = {0,0,0}, sgui[3] = {0,0,0}, vsgu[3] = {0,0,0};
    for (int a = 0; a < D; ++a) {
      for (int b = 0; b < D; ++b) v_gradu_shv[a] += guv[b]*H[a+1][b+1];
      v_gradu_shv[a] += gu0v*H[a+1][0];
      for (int b = 0; b < D; ++b) gugs[a] += cg3[a][b]*H[0][b+1];
      gugs[a] += cmd[a+1]*grad_u0[a]*H[0][0];
      for (int b = 0; b < D; ++b) sgui[a] += S[b+1][0]*grad_uj[b][a] + S[b+1][a+1]*grad_u0[b];
      for (int b = 0; b < D; ++b) vsgu[a] += v[b]*S[0][0]*grad_uj[b][a] + v[b]*S[0][a+1]*grad_u0[b];
    }
    double shv_grad_u = 0.0;
    for (int a = 0; a < D; ++a) {
      for (int b = 0; b < D; ++b) shv_grad_u += (H[a+1][b+1] - v[a]*H[0][b+1])*grad_uj[a][b];
      shv_grad_u += (H[a+1][0] - v[a]*H[0][0])*grad_u0[a];
    }
    double sl0l[3] = {0,0,0};
    for (int a = 0; a < D; ++a) for (int b = 0; b < 4; ++b) sl0l[a] += (S[0][b]*H[a+1][b] + S[a+1][b]*H[0][b])/2.;
    const double uD = u[D-1];
    const double expf = gf + gamma/t + gamma*divV;
    for (int a = 0; a < D; ++a) {
      double Fsig = -(guv[a] - gu0c[a])/2. + u[a]*gamma*(gf + gamma/t + gamma*divV)/3. - gamma*gf*u[a]/2.;
      double FD = gamma*(u[a]*H[0][0] + gamma*H[a+1][0])*gf + t*uD*S[3][a+1]*dl[a];
      double Fds = -(v_gradu_shv[a] - gugs[a] - 4.*gamma*u[a]*shv_grad_u/3. - sgui[a] + vsgu[a])/4.
                 + (H[a+1][0] - gamma*gamma*H[a+1][0] - gamma*u[a]*H[0][0] + S[0][a+1] + 4.*u[a]*gamma*H[0][0]/3.)*gf/4.
                 + (gamma*cmd[a+1]*H[0][3]*dl[a]/t + gamma*S[3][0]*dl[a]/t + 4.*gamma*u[a]*(gamma*H[3][3]/t)/3.)/4.
                 - S[0][a+1]*(gamma/t + gf + gamma*divV)/3.;
      double Fdo = -(v_gradu_shv[a] - gugs[a] + sgui[a] - vsgu[a])/4.
                 + (H[a+1][0] - gamma*gamma*H[a+1][0] - gamma*u[a]*H[0][0] - S[0][a+1])*gf/4.
                 + (gamma*cmd[a+1]*H[0][3]*dl[a]/t - gamma*S[3][0]*dl[a]/t)/4.;
      double Fdd = S[a+1][0]*expf;
      double Msig[3], MD[3], Mds[3], Mdo[3], Mdd[3];
      for (int b = 0; b < D; ++b) {
        Msig[b] = u[a]*u_cov[b]/(2.*3.);
        MD[b] = gamma*u[a]*(H[0][b+1] - H[0][0]*u_cov[b]/gamma) + gamma*gamma*(H[a+1][b+1] - H[a+1][0]*u_cov[b]/gamma);
        Mds[b] = (H[a+1][b+1] - H[a+1][0]*u_cov[b]/gamma - gamma*gamma*H[a+1][b+1] - gamma*u[a]*H[0][b+1]
                  + u_cov[b]*gamma*H[a+1][0] + u_cov[b]*u[a]*H[0][0]
                  + 4.*gamma*u[a]*(H[0][b+1] - H[0][0]*u_cov[b]/gamma)/3.
                  + (-S[0][a+1] + 4.*S[0][a+1]/3.)*u_cov[b]/gamma)/4.;
        Mdo[b] = (H[a+1][b+1] - H[a+1][0]*u_cov[b]/gamma - gamma*gamma*H[a+1][b+1] - gamma*u[a]*H[0][b+1]
                  + u_cov[b]*gamma*H[a+1][0] + u_cov[b]*u[a]*H[0][0] + S[0][a+1]*u_cov[b]/gamma)/4.;
        Mdd[b] = -S[a+1][0]*u_cov[b]/gamma;
      }
      Msig[a] += (1. - gamma*gamma)/2.;
      Mdo[a] += -S[0][0]/4.;
      Mds[a] += +S[0][0]/4.;
      Fsig += dl[a]*(-(-uD*t - uD/t + gamma*uD*uD*t + 2.*gamma*gamma*uD/t)/2.);
      FD += dl[a]*(uD*S[0][0]/t + gamma*S[3][0]/t);
      if (D != 2) {
        Fds += (t*uD*dl[a]*cmd[a+1]*H[0][0] + uD*H[a+1][3]*gamma/t/4. + H[a+1][3]*t*uD
                + S[0][0]*dl[a]*uD/t + t*uD*uD*(gamma*H[a+1][0] + u[a]*H[0][0])
                + 2.*gamma*uD*(gamma*H[a+1][3] + u[a]*S[0][3])/t)/4.
             + gamma*u[a]*(H[3][0]*uD*t + uD*H[0][3]/t)/3.;
        Fdo += (t*uD*dl[a]*cmd[a+1]*H[0][0] + uD*H[a+1][3]*gamma/t - H[3][a+1]*t*uD
                - S[0][0]*dl[a]*uD/t + t*uD*uD*(gamma*H[a+1][0] + u[a]*H[0][0])
                + 2.*gamma*uD*(gamma*H[a+1][3] + u[a]*S[0][3])/t)/4.;
      }
      F_0i[a] = (-S[a+1][0] - tau_pi*FD - delta_pipi*Fdd + 2.*tau_pi*Fdo*use_vort - tau_pipi*Fds
                 + (2.*eta_pi + lambda_piPi*bulk)*Fsig + phi6*bulk*S[a+1][0]
                 + phi7*(sl0l[a] + gamma*u[a]*shv_shvcov/3.))/(tau_pi*gamma);
      for (int b = 0; b < D; ++b)
        M_0i[a][b] = (-tau_pi*MD[b] - delta_pipi*Mdd[b] + 2.*tau_pi*Mdo[b]*use_vort - tau_pipi*Mds[b]
                      + (2.*eta_pi + lambda_piPi*bulk)*Msig[b])/(tau_pi*gamma);
    }
    /* shv_nabla_u MRFs */
    for (int a = 0; a < D; ++a) {
      M_shv_nabla_u[a] = H[0][a+1] - H[0][0]*u_cov[a]/gamma;
      for (int b = 0; b < D; ++b) F_shv_nabla_u += (H[a+1][b+1] - v[a]*H[0][b+1])*grad_uj[a][b];
      F_shv_nabla_u += (H[a+1][0] - v[a]*H[0][0])*grad_u0[a];
    }
    F_shv_nabla_u += +H[3][3]*gamma/t + H[0][0]*gf + (t*H[3][0] + H[0][3]/t)*u[D-1]*dl[D-1];

    /* extensive-shear MRFs */
    const double tilde_delta = delta_pipi - tau_pi;
    double gugj[2][3] = {{0}}, sgu[2][3] = {{0}}, vsg[2][3] = {{0}}, sls[2][3] = {{0}};
    for (int a = 0; a < 2; ++a) for (int b = a; b < 3; ++b) {
      for (int k = 0; k < D; ++k) {
        gugj[a][b] += cg3[a][k]*H[b+1][k+1] + cg3[b][k]*H[a+1][k+1];
        sgu[a][b] += S[k+1][a+1]*g3[k][b] + S[k+1][b+1]*g3[k][a];
        vsg[a][b] += v[k]*S[0][a+1]*g3[k][b] + v[k]*S[0][b+1]*g3[k][a];
      }
      for (int k = 0; k < 4; ++k) sls[a][b] += (S[a+1][k]*H[b+1][k] + S[b+1][k]*H[a+1][k])/2.;
    }
    double FDm[2][3] = {{0}}, Fsm[2][3] = {{0}}, Fdl[2][3] = {{0}}, Fdom[2][3] = {{0}}, Fdsm[2][3] = {{0}};
    double MDm[2][3][3] = {{{0}}}, Msm[2][3][3] = {{{0}}}, Mdl[2][3][3] = {{{0}}}, Mdom[2][3][3] = {{{0}}}, Mdsm[2][3][3] = {{{0}}};
    for (int a = 0; a < 2; ++a) {
      for (int b = a; b < 3; ++b) {
        FDm[a][b] = gamma*gf*(fu[a]*H[b+1][0] + fu[b]*H[a+1][0]);
        Fsm[a][b] = cg3[a][b]/2. + cg3[b][a]/2. + fu[a]*fu[b]*expf/3.;
        Fdl[a][b] = S[a+1][b+1]*expf;
        const double metric_terms =
            t*fu[2]*(cmd[a+1]*H[b+1][0]*dl[a] + cmd[b+1]*H[a+1][0]*dl[b])
          + gamma*(cmd[a+1]*H[b+1][3]*dl[a] + cmd[b+1]*H[a+1][3]*dl[b])/t
          + fu[2]*fu[2]*t*(fu[a]*H[b+1][0] + fu[b]*H[a+1][0])
          + 2.*gamma*fu[2]*(fu[a]*H[b+1][3] + fu[b]*H[a+1][3])/t;
        Fdom[a][b] = -(-gugj[a][b] + sgu[a][b] - vsg[a][b])/4.
                   + (-fu[a]*gamma*H[b+1][0] - fu[b]*gamma*H[a+1][0])*gf/4.
                   + (metric_terms
                      - gamma*(S[3][a+1]*dl[b] + S[3][b+1]*dl[a])/t
                      - fu[2]*(S[0][a+1]*dl[b] + S[0][b+1]*dl[a])/t)/4.;
        Fdsm[a][b] = -(-gugj[a][b] - sgu[a][b] + vsg[a][b] - 4.*(fu[a]*fu[b])*shv_grad_u/3.)/4.
                   + (-fu[a]*gamma*H[b+1][0] - fu[b]*gamma*H[a+1][0] + 4.*fu[a]*fu[b]*H[0][0]/3.)*gf/4.
                   + (metric_terms
                      + gamma*(S[3][a+1]*dl[b] + S[3][b+1]*dl[a])/t
                      + fu[2]*(S[0][a+1]*dl[b] + S[0][b+1]*dl[a])/t
                      + 4.*fu[a]*fu[b]*(H[3][0]*fu[2]*t + gamma*H[3][3]/t + fu[2]*H[0][3]/t)/3.)/4.
                   - S[a+1][b+1]*(gamma/t + gf + gamma*divV)/3.;
        for (int k = 0; k < D; ++k) {
          MDm[a][b][k] = gamma*(fu[a]*H[b+1][k+1] - fu[a]*H[b+1][0]*fc[k]/gamma
                                + fu[b]*H[a+1][k+1] - fu[b]*H[a+1][0]*fc[k]/gamma);
          Msm[a][b][k] = -fu[a]*fu[b]*fc[k]/(3.*gamma);
          Mdl[a][b][k] = -S[a+1][b+1]*fc[k]/gamma;
          Mdom[a][b][k] = (-fu[a]*H[b+1][k+1]*gamma - fu[b]*H[a+1][k+1]*gamma
                           + fc[k]*fu[a]*H[b+1][0] + fc[k]*fu[b]*H[a+1][0])/4.;
          Mdsm[a][b][k] = (-fu[a]*H[b+1][k+1]*gamma - fu[b]*H[a+1][k+1]*gamma
                           + fc[k]*fu[a]*H[b+1][0] + fc[k]*fu[b]*H[a+1][0]
                           + 4.*fu[a]*fu[b]*(H[0][k+1] - H[0][0]*fc[k]/gamma)/3.
                           + 4.*S[a+1][b+1]*fc[k]/3./gamma)/4.;
        }
        Mdsm[a][b][a] += S[0][b+1]/4.;
        Mdom[a][b][a] += -S[0][b+1]/4.;
        Msm[a][b][a] += -gamma*fu[b]/2.;
      }
      for (int b = a; b < D; ++b) {
        Msm[a][b][b] += -gamma*fu[a]/2.;
        Mdsm[a][b][b] += S[0][a+1]/4.;
        Mdom[a][b][b] += -S[0][a+1]/4.;
      }
      for (int k = 0; k < D; ++k) {
        Msm[a][a][k] += cmd[a+1]*fc[k]/(gamma*3.);
        Mdsm[a][a][k] += -cmd[a+1]*(H[0][k+1] - H[0][0]*fc[k]/gamma)/3.;
      }
      const double ueta = fu[2];
      FDm[a][2] += gamma*S[3][a+1]/t + ueta*S[0][a+1]/t;
      Fsm[a][a] += -cmd[a+1]*expf/3.;
      Fsm[a][2] += -fu[a]*ueta*ueta*t/2. - fu[a]*ueta*gamma/t;
      Fdsm[a][a] += -cmd[a+1]*shv_grad_u/3. - cmd[a+1]*H[0][0]*gf/3.
                    - cmd[a+1]*(H[3][0]*fu[2]*t + gamma*H[3][3]/t + fu[2]*H[0][3]/t)/3.;
    }
    for (int a = 0; a < 2; ++a) {
      for (int b = a; b < 3; ++b) {
        F_es[a][b] = (-S[a+1][b+1] - tilde_delta*Fdl[a][b] - tau_pi*FDm[a][b]
                      + 2.*tau_pi*Fdom[a][b]*use_vort - tau_pipi*Fdsm[a][b]
                      + (2.*eta_pi + lambda_piPi*bulk)*Fsm[a][b] + phi6*bulk*S[a+1][b+1]
                      + phi7*(sls[a][b] + fu[a]*fu[b]*shv_shvcov/3.))/(sigma*tau_pi*gamma);
        for (int k = 0; k < D; ++k)
          M_es[a][b][k] = (-tau_pi*MDm[a][b][k] - tilde_delta*Mdl[a][b][k] + 2.*tau_pi*Mdom[a][b][k]*use_vort
                           - tau_pipi*Mdsm[a][b][k] + (2.*eta_pi + lambda_piPi*bulk)*Msm[a][b][k])/(sigma*tau_pi*gamma);
      }
      F_es[a][a] += phi7*(-cmd[a+1]*shv_shvcov)/(3.*tau_pi*gamma*sigma);
    }
    memcpy(sigF, Fsm, sizeof(sigF));
    memcpy(sigM, Msm, sizeof(sigM));
  }

  /* ---------------- fill_auxiliary_variables (230-362) ---------------- */
  const double zeta_tilde = zeta + (delta_PiPi - tau_Pi)*bulk;
  double M_eb[3] = {0,0,0}, M_S[3] = {0,0,0};
  for (int a = 0; a < D; ++a) {
    M_eb[a] = (zeta_tilde*u_cov[a]/gamma)/(sigma*gamma*tau_Pi);
    M_S[a] = (bulk*u_cov[a]/gamma)/(sigma*gamma*T);
  }
  const double mu_vec[3] = { muB, muS, muQ };
  double R_S[3];
  for (int c = 0; c < 3; ++c) R_S[c] = -mu_vec[c]/T;
  const double R_eb[3] = { 0, 0, 0 };
  double F_eb = -(zeta_tilde*(gamma*divV + gamma/t + gf) + bulk - phi1*bulk*bulk)/(sigma*gamma*tau_Pi);
  double F_S = (-bulk*(gamma*divV + gamma/t + gf))/(sigma*gamma*T);   /* sources are disabled */
  if (using_shear) {
    F_S += F_shv_nabla_u/(sigma*gamma*T);
    F_eb += -(lambda_Pipi*F_shv_nabla_u - phi3*shv_shvcov)/(sigma*gamma*tau_Pi);
    for (int a = 0; a < D; ++a) {
      M_eb[a] += -lambda_Pipi*M_shv_nabla_u[a]/(sigma*gamma*tau_Pi);
      M_S[a] += M_shv_nabla_u[a]/(sigma*gamma*T);
    }
  }

  /* ---------------- calculate_MRF_diffusion (1416-1822) ---------------- */
  const int using_diffusion = P->using_diffusion;
  double F_q0[3] = {0,0,0}, R_q0[3][3] = {{0}}, M_q0[3][3] = {{0}};
  double F_q[3][3] = {{0}}, M_q[3][3][3] = {{{0}}}, R_q[3][3][3] = {{{0}}};
  if (using_diffusion)
    orc_mrf_diffusion(ev, i, M_S, F_S, R_S, F_q0, R_q0, M_q0, F_q, M_q, R_q);

  /* ---------------- compute_charge_derivative (369-453) ---------------- */
  double R_N[3][3] = {{0}}, F_N[3] = {0,0,0}, M_N[3][3] = {{0}};
  for (int c = 0; c < 3; ++c) R_N[c][c] += gamma*sigma;
  if (using_diffusion) {
    for (int d = 0; d < D; ++d) for (int c = 0; c < 3; ++c) M_N[d][c] += -M_q0[d][c];
    for (int c = 0; c < 3; ++c) {
      for (int k = 0; k < 3; ++k) R_N[c][k] += -R_q0[c][k];
      F_N[c] += -F_q0[c] + FLD(ev, ORC_F_GRAD_QA0 + c)[i] - FLD(ev, ORC_F_DIV_QA0 + c)[i]
              - FLD(ev, ORC_F_Q0 + 4*c)[i]/t;
    }
  }

  /* ---------------- compute_velocity_derivative (459-618) ---------------- */
  double gradP[3] = {0,0,0}, gradBulk[3] = {0,0,0}, divshear[3] = {0,0,0}, gradshear[3] = {0,0,0};
  for (int a = 0; a < D; ++a) {
    gradP[a] = FLD(ev, ORC_F_GRADP0 + a)[i];
    gradBulk[a] = FLD(ev, ORC_F_GRADBULK0 + a)[i];
    divshear[a] = FLD(ev, ORC_F_DIVSHEAR0 + a)[i];
    gradshear[a] = FLD(ev, ORC_F_GRADSHEAR0 + a)[i];
  }
  vec_contra(D, gradP, t2);
  vec_contra(D, gradBulk, t2);
  const double rho_vec[3] = { rhoB, rhoS, rhoQ }, dwdr[3] = { dwdB, dwdS, dwdQ };
  const double rdw = rho_vec[0]*dwdr[0] + rho_vec[1]*dwdr[1] + rho_vec[2]*dwdr[2];
  double M_w[3] = {0,0,0}, R_w[3];
  for (int a = 0; a < D; ++a) M_w[a] = (s*dwds/gamma + rdw/gamma)*u_cov[a]/gamma + sigma*dwds*M_S[a];
  for (int c = 0; c < 3; ++c) R_w[c] = sigma*dwdr[c] + sigma*dwds*R_S[c];
  const double F_w = -(s*dwds/gamma + rdw/gamma)*(gf + gamma/t + gamma*divV) + sigma*dwds*F_S;
  const double F_bulk = sigma*F_eb - bulk*(gamma*divV + gamma/t + gf)/gamma;
  double M_u[3][3] = {{0}}, F_u[3] = {0,0,0}, R_u[3][3] = {{0}};
  for (int a = 0; a < D; ++a) {
    for (int b = 0; b < D; ++b)
      M_u[a][b] = u[a]*gamma*(M_eb[b] + M_w[b]) - u[a]*u_cov[b]*(w + bulk)/gamma;
    M_u[a][a] += gamma*(w + bulk);
  }
  for (int a = 0; a < D; ++a)
    F_u[a] = (gradP[a] + gradBulk[a]) - gamma*(F_w + F_bulk)*u[a]
           - (w + bulk)*(gamma*divV + gamma/t + gf)*u[a]
           - 2.*(w + bulk)*gamma*u[D-1]*dl[a]/t;
  for (int a = 0; a < D; ++a) for (int c = 0; c < 3; ++c) R_u[a][c] = -u[a]*gamma*(R_w[c] + R_eb[c]*sigma);
  if (using_shear) {
    for (int a = 0; a < D; ++a) {
      F_u[a] += -S[a+1][0]/t - 2.*dl[a]*S[0][3]/t - divshear[a] + gradshear[a] - F_0i[a];
      for (int b = 0; b < D; ++b) M_u[a][b] += M_0i[a][b];
    }
  }
  double RI[3][3];
  inverseD(3, R_N, RI);
  double aux_MR[3][3] = {{0}}, aux_FR[3] = {0,0,0};
  for (int a = 0; a < D; ++a) {
    for (int b = 0; b < D; ++b)
      for (int c = 0; c < 3; ++c) for (int d = 0; d < 3; ++d) aux_MR[a][b] += R_u[a][c]*RI[c][d]*M_N[b][d];
    for (int c = 0; c < 3; ++c) for (int d = 0; d < 3; ++d) aux_FR[a] += R_u[a][c]*RI[c][d]*F_N[d];
  }
  double Mm[3][3] = {{0}}, MI[3][3];
  for (int a = 0; a < D; ++a) for (int b = 0; b < D; ++b) Mm[a][b] = M_u[a][b] - aux_MR[a][b];
  inverseD(D, Mm, MI);
  double du_dt[3] = {0,0,0};
  for (int a = 0; a < D; ++a) for (int b = 0; b < D; ++b) du_dt[a] += MI[a][b]*(F_u[b] + aux_FR[b]);

  /* ---------------- compute_derivatives (623-825) ---------------- */
  double MU_aux[3] = {0,0,0};
  for (int c = 0; c < 3; ++c) for (int a = 0; a < D; ++a) MU_aux[c] += M_N[a][c]*du_dt[a];
  double RNI[3][3];
  inverseD(3, R_N, RNI);
  double dN_dt[3] = {0,0,0};
  for (int c = 0; c < 3; ++c) for (int d = 0; d < 3; ++d) dN_dt[c] += RNI[c][d]*(F_N[d] + MU_aux[d]);
  double d_s = F_S, d_b = F_eb, snu = F_shv_nabla_u, ucdu = 0.0;
  for (int a = 0; a < D; ++a) {
    d_s += M_S[a]*du_dt[a];
    d_b += M_eb[a]*du_dt[a];
    snu += M_shv_nabla_u[a]*du_dt[a];
    ucdu += u_cov[a]*du_dt[a];
  }
  /* milne::contract(M_S, du_dt) is evaluated before adding F (order kept) */
  {
    double c1 = 0.0, c2 = 0.0, c3 = 0.0;
    for (int a = 0; a < D; ++a) { c1 += M_S[a]*du_dt[a]; c2 += M_eb[a]*du_dt[a]; c3 += M_shv_nabla_u[a]*du_dt[a]; }
    double rs = 0.0, rb = 0.0;
    for (int c = 0; c < 3; ++c) { rs += R_S[c]*dN_dt[c]; rb += R_eb[c]*dN_dt[c]; }
    d_s = c1 + F_S + rs;
    d_b = c2 + F_eb + rb;
    snu = c3 + F_shv_nabla_u;
  }
  const double theta = -1.*ucdu/gamma + gamma*divV + gamma/t + gf;
  for (int d = 0; d < 3; ++d) FLD(ev, ORC_F_DUDT0 + d)[i] = (d < D) ? du_dt[d] : 0.0;
  FLD(ev, ORC_F_DEXT_S)[i] = d_s;
  FLD(ev, ORC_F_DEXT_B)[i] = dN_dt[0];
  FLD(ev, ORC_F_DEXT_SS)[i] = dN_dt[1];
  FLD(ev, ORC_F_DEXT_Q)[i] = dN_dt[2];
  FLD(ev, ORC_F_DEXT_BULK)[i] = d_b;
  FLD(ev, ORC_F_THETA)[i] = theta;
  FLD(ev, ORC_F_SHV_NABLA_U)[i] = snu;
  if (using_diffusion) {
    for (int d = 0; d < 3; ++d) for (int c = 0; c < 3; ++c) {
      double v = F_q[d][c];
      for (int j = 0; j < D; ++j) v += M_q[d][c][j]*du_dt[j];
      for (int k = 0; k < 3; ++k) v += R_q[d][c][k]*dN_dt[k];
      FLD(ev, ORC_F_DEXT_Q0 + 3*c + d)[i] = v;
    }
  }
  if (using_shear) {
    int k = 0;
    for (int a = 0; a < 2; ++a) for (int b = a; b < 3; ++b) {
      double m = 0.0;
      for (int c = 0; c < D; ++c) m += M_es[a][b][c]*du_dt[c];
      FLD(ev, ORC_F_DEXT_SHV0 + k++)[i] = m + F_es[a][b];
    }
  }

  /* ---------------- compute_Ez_derivative (828-862) ---------------- */
  const double p = FLD(ev, ORC_F_P)[i], e = FLD(ev, ORC_F_E)[i];

  /* ---------------- compute_hydro_numbers (1831-1995), with shear or diffusion on ---------------- */
  if ((using_shear || using_diffusion) && D == 2) {
    /* the stored sigma tensor: (a+1, b+1) = M_sigma du/dt + F_sigma (784-808) */
    double stored[4][4] = {{0}};
    for (int a = 0; a < 2; ++a) for (int b = a; b < 3; ++b) {
      double m = 0.0;
      for (int k = 0; k < D; ++k) m += sigM[a][b][k]*du_dt[k];
      stored[a+1][b+1] = m + sigF[a][b];
    }
    /* read back into (a, b) (1893-1898) */
    double sg[4][4] = {{0}};
    for (int a = 0; a < 2; ++a) for (int b = a; b < 3; ++b) sg[a][b] = stored[a+1][b+1];
    for (int a = 1; a < 4; ++a) for (int b = a+1; b < 4; ++b) sg[b][a] = sg[a][b];
    double uc2[2] = { u[0], u[1] };
    vec_cov(D, uc2, t2);
    for (int a = 1; a < D+1; ++a) {
      double row[2];
      for (int b = 1; b < D+1; ++b) row[b-1] = sg[a][b];
      double c = 0.0;
      for (int k = 0; k < D; ++k) c += uc2[k]*row[k];
      sg[0][a] = -1.*c/gamma;
    }
    for (int a = 1; a < 4; ++a) sg[a][0] = sg[0][a];
    sg[0][0] = (fc[0]*fc[0]*sg[1][1] + fc[1]*fc[1]*sg[2][2] + 2.*fc[1]*fc[0]*sg[2][1]
                + 2.*fc[2]*fc[0]*sg[3][1] + 2.*fc[2]*fc[1]*sg[3][2]
                - fc[2]*fc[2]*(sg[1][1] + sg[2][2])/t2)/(gamma*gamma - fc[2]*fc[2]/t2);
    sg[3][3] = (sg[0][0] - sg[1][1] - sg[2][2])/t2;
    double sh[4][4], scov[4][4];
    memcpy(sh, sg, sizeof(sg)); mat4_cov(sh, 1, t2);
    memcpy(scov, sh, sizeof(sg)); mat4_cov(scov, 0, t2);
    double sigma_norm = 0.0, shv_norm = 0.0;      /* milne::contract: columns outer, rows inner */
    for (int c = 0; c < 4; ++c) for (int r = 0; r < 4; ++r) {
      sigma_norm += sg[r][c]*scov[r][c];
      shv_norm += S[r][c]*C[r][c];
    }
    FLD(ev, ORC_F_SHEAR_KNUDSEN)[i] = tau_pi*sqrt(fabs(sigma_norm));
    FLD(ev, ORC_F_BULK_KNUDSEN)[i] = tau_Pi*fabs(theta);
    FLD(ev, ORC_F_INV_REYNOLDS_SHEAR)[i] = sqrt(fabs(shv_norm))/p;
    FLD(ev, ORC_F_INV_REYNOLDS_BULK)[i] = fabs(bulk)/p;
    /* diffusion Knudsen tau_q^(aa) |grad alpha_a| and inverse Reynolds |q_a|/|n_a| */
    const double rho_n[3] = { rhoB, rhoS, rhoQ };
    for (int c = 0; c < 3; ++c) {
      double ga2 = 0.0;
      for (int d = 0; d < D; ++d) { const double g = FLD(ev, ORC_F_GRAD_ALPHA0 + 3*d + c)[i]; ga2 += g*g; }
      const double tqa = FLD(ev, ORC_F_TAU_Q0 + 4*c)[i];
      const double q0 = FLD(ev, ORC_F_Q0 + 4*c)[i], q1 = FLD(ev, ORC_F_Q0 + 4*c + 1)[i];
      const double q2 = FLD(ev, ORC_F_Q0 + 4*c + 2)[i], q3 = FLD(ev, ORC_F_Q0 + 4*c + 3)[i];
      const double q_sq = q1*q1 + q2*q2 + t2*q3*q3 - q0*q0;
      const double q_mag = q_sq > 0.0 ? sqrt(q_sq) : 0.0;
      const double n_abs = fabs(rho_n[c]);
      FLD(ev, ORC_F_KN_DIFF0 + c)[i] = tqa*sqrt(ga2);
      FLD(ev, ORC_F_RE_DIFF0 + c)[i] = (n_abs > 1e-15) ? q_mag/n_abs : 0.0;
    }
  }
  const double u3 = (D == 3) ? u[2] : 0.0;
  FLD(ev, ORC_F_DEZ)[i] = ((e + p + bulk)*u3*u3 + (p + bulk)/t2 + S[3][3])
                          * FLD(ev, ORC_F_MASS_S)[i]*t2/FLD(ev, ORC_F_SIGMA_LAB)[i];
}

void orc_eom(orc_event* ev)
{
  #pragma omp parallel for schedule(static) num_threads(orc_nthreads())
  for (int i = 0; i < ev->n; ++i) eom_particle(ev, i);
}
