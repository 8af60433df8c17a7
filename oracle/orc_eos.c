/*
 * orc_eos.c -- CPU restatement of CCAKE's equation-of-state layer
 * (test infrastructure only; see ccake_oracle.h).
 *
 *  - analytic members: conformal (include/eos_conformal.h:34-195),
 *    conformal_diagonal (include/eos_conformal_diagonal.h:39-204),
 *    tanh-modulated conformal (include/eos_tanh_conformal.h:22-304);
 *  - table member: 4D multilinear interpolation on a uniform grid
 *    (src/interpolatorND.cpp:250-392, src/eos_table.cpp:157-430);
 *  - chain set-up with matched fallback coefficients
 *    (src/eos_initialization.cpp:216-501);
 *  - inversion (s or e, rhoB, rhoS, rhoQ) -> (T, muB, muQ, muS) with the
 *    seed / perturbation / EoS-chain cascade of src/eos.cpp:629-975 and
 *    src/rootfinder.cpp:421-715.  GSL's `hybrids` solver is replaced by a
 *    damped Newton iteration on the same residual with the same absolute
 *    stopping rule (sum |F_i| < 1e-10) followed by one polishing step;
 *  - set_thermo and the enthalpy / alpha derivatives by Schur complements
 *    of the susceptibility matrix (include/eos.h:121-175, src/eos.cpp:276-492,
 *    src/eos_derivatives.cpp:16-555).
 */
#include "orc_internal.h"

static double sgn(double v) { return (double)((0.0 < v) - (v < 0.0)); }

/* ------------------------------------------------------------------ */
/* conformal: p = c T0^4 x^2, x = (T/T0)^2 + sum (mu/mu0)^2             */
static void conf_full(const orc_member* m, const double* pt, double* r)
{
  const double T = pt[0], muB = pt[1], muQ = pt[2], muS = pt[3];
  const double T0 = m->T0, B0 = m->muB0, S0 = m->muS0, Q0 = m->muQ0;
  const double x = (T/T0)*(T/T0) + (muB/B0)*(muB/B0) + (muS/S0)*(muS/S0) + (muQ/Q0)*(muQ/Q0);
  const double cp = m->c*T0*T0*T0*T0;
  const double p = cp*x*x;
  const double s = 4.0*cp*T*x/(T0*T0);
  const double B = 4.0*cp*muB*x/(B0*B0);
  const double S = 4.0*cp*muS*x/(S0*S0);
  const double Q = 4.0*cp*muQ*x/(Q0*Q0);
  r[X_P] = p; r[X_S] = s; r[X_B] = B; r[X_SS] = S; r[X_Q] = Q;
  r[X_E] = (s*T + muB*B + muQ*Q + muS*S) - p;
  r[X_CS2] = 1.0/3.0;
  r[X_BB] = cp*(8.0*muB*muB/(B0*B0*B0*B0) + 4.0*x/(B0*B0));
  r[X_QQ] = cp*(8.0*muQ*muQ/(Q0*Q0*Q0*Q0) + 4.0*x/(Q0*Q0));
  r[X_SSS] = cp*(8.0*muS*muS/(S0*S0*S0*S0) + 4.0*x/(S0*S0));
  r[X_BQ] = 8.0*cp*muB*muQ/(B0*B0*Q0*Q0);
  r[X_BS] = 8.0*cp*muB*muS/(B0*B0*S0*S0);
  r[X_QS] = 8.0*cp*muS*muQ/(S0*S0*Q0*Q0);
  r[X_TB] = 8.0*cp*T*muB/(T0*T0*B0*B0);
  r[X_TQ] = 8.0*cp*T*muQ/(T0*T0*Q0*Q0);
  r[X_TS] = 8.0*cp*T*muS/(T0*T0*S0*S0);
  r[X_TT] = cp*(8.0*T*T/(T0*T0*T0*T0) + 4.0*x/(T0*T0));
}

/* conformal_diagonal: p = c T0^4 sum (a/a0)^4                          */
static void diag_full(const orc_member* m, const double* pt, double* r)
{
  const double T = pt[0], muB = pt[1], muQ = pt[2], muS = pt[3];
  const double T0 = m->T0, B0 = m->muB0, S0 = m->muS0, Q0 = m->muQ0;
  const double cp = m->c*T0*T0*T0*T0;
  const double tT = T/T0, tB = muB/B0, tS = muS/S0, tQ = muQ/Q0;
  const double p = cp*(tT*tT*tT*tT + tB*tB*tB*tB + tS*tS*tS*tS + tQ*tQ*tQ*tQ);
  const double s = 4.0*cp*T*T*T/(T0*T0*T0*T0);
  const double B = 4.0*cp*muB*muB*muB/(B0*B0*B0*B0);
  const double S = 4.0*cp*muS*muS*muS/(S0*S0*S0*S0);
  const double Q = 4.0*cp*muQ*muQ*muQ/(Q0*Q0*Q0*Q0);
  r[X_P] = p; r[X_S] = s; r[X_B] = B; r[X_SS] = S; r[X_Q] = Q;
  r[X_E] = (s*T + muB*B + muQ*Q + muS*S) - p;
  r[X_CS2] = 1.0/3.0;
  r[X_BB] = 12.0*cp*muB*muB/(B0*B0*B0*B0);
  r[X_QQ] = 12.0*cp*muQ*muQ/(Q0*Q0*Q0*Q0);
  r[X_SSS] = 12.0*cp*muS*muS/(S0*S0*S0*S0);
  r[X_BQ] = ORC_TINY; r[X_BS] = ORC_TINY; r[X_QS] = ORC_TINY;
  r[X_TB] = ORC_TINY; r[X_TQ] = ORC_TINY; r[X_TS] = ORC_TINY;
  r[X_TT] = 12.0*cp*T*T/(T0*T0*T0*T0);
}

/* Adiabatic speed of sound from first and second derivatives of p(T, mu):
 * cs2 = (dp/de) along d(n_X/s) = 0.  dx solves H dx = (ds, dn) with
 * dn_X = (n_X/s) ds (the 1D direction fixed by the three constraints).
 * This is the quantity include/eos_tanh_conformal.h:243-301 writes out in
 * closed form; here it is computed from its definition. */
static double adiabatic_cs2(const double* pt, const double* r)
{
  /* H over (T, muB, muQ, muS) */
  double H[4][4] = {
    { r[X_TT], r[X_TB], r[X_TQ], r[X_TS] },
    { r[X_TB], r[X_BB], r[X_BQ], r[X_BS] },
    { r[X_TQ], r[X_BQ], r[X_QQ], r[X_QS] },
    { r[X_TS], r[X_BS], r[X_QS], r[X_SSS] } };
  const double s = r[X_S];
  double rhs[4] = { 1.0, r[X_B]/s, r[X_Q]/s, r[X_SS]/s };   /* ds = 1 */
  /* Gaussian elimination with partial pivoting */
  double A[4][5];
  for (int i = 0; i < 4; ++i) { for (int j = 0; j < 4; ++j) A[i][j] = H[i][j]; A[i][4] = rhs[i]; }
  for (int c = 0; c < 4; ++c) {
    int piv = c; for (int i = c+1; i < 4; ++i) if (fabs(A[i][c]) > fabs(A[piv][c])) piv = i;
    if (piv != c) for (int j = 0; j < 5; ++j) { double t = A[c][j]; A[c][j] = A[piv][j]; A[piv][j] = t; }
    if (A[c][c] == 0.0) return 1.0/3.0;
    for (int i = c+1; i < 4; ++i) { double f = A[i][c]/A[c][c]; for (int j = c; j < 5; ++j) A[i][j] -= f*A[c][j]; }
  }
  double dx[4];
  for (int i = 3; i >= 0; --i) { double v = A[i][4]; for (int j = i+1; j < 4; ++j) v -= A[i][j]*dx[j]; dx[i] = v/A[i][i]; }
  const double dp = r[X_S]*dx[0] + r[X_B]*dx[1] + r[X_Q]*dx[2] + r[X_SS]*dx[3];
  const double de = pt[0]*1.0 + pt[1]*rhs[1] + pt[2]*rhs[2] + pt[3]*rhs[3];
  return dp/de;
}

/* tanh-modulated conformal: p = t(T) p_conf, t = (1 + tanh((T-Tc)/Ts))/2 */
static void tanh_full(const orc_member* m, const double* pt, double* r)
{
  double rc[17];
  conf_full(m, pt, rc);
  const double T = pt[0], muB = pt[1], muQ = pt[2], muS = pt[3];
  const double arg = (T - m->Tc)/m->Ts;
  const double th = tanh(arg);
  const double sech = 1.0/cosh(arg);
  const double tf = 0.5*(1.0 + th);
  const double tp = 0.5*sech*sech/m->Ts;
  const double tpp = -2.0*th*tp/m->Ts;
  const double pc = rc[X_P], sc = rc[X_S], Bc = rc[X_B], Sc = rc[X_SS], Qc = rc[X_Q];
  r[X_P] = tf*pc;
  r[X_S] = tf*sc + tp*pc;
  r[X_B] = tf*Bc; r[X_SS] = tf*Sc; r[X_Q] = tf*Qc;
  r[X_E] = (r[X_S]*T + muB*r[X_B] + muQ*r[X_Q] + muS*r[X_SS]) - r[X_P];
  r[X_BB] = tf*rc[X_BB]; r[X_QQ] = tf*rc[X_QQ]; r[X_SSS] = tf*rc[X_SSS];
  r[X_BQ] = tf*rc[X_BQ]; r[X_BS] = tf*rc[X_BS]; r[X_QS] = tf*rc[X_QS];
  r[X_TB] = tf*rc[X_TB] + tp*Bc;
  r[X_TQ] = tf*rc[X_TQ] + tp*Qc;
  r[X_TS] = tf*rc[X_TS] + tp*Sc;
  r[X_TT] = tf*rc[X_TT] + 2.0*tp*sc + tpp*pc;
  r[X_CS2] = adiabatic_cs2(pt, r);
}

/* ------------------------------------------------------------------ */
/* table: src/interpolatorND.cpp:250-392 (index/fraction rule, corner
 * order T slowest, weight product in axis order). Out of range -> NaN.   */
static int table_locate(const orc_member* m, const double* pt, int* ind, double* frac)
{
  for (int a = 0; a < 4; ++a) {
    if (m->dims[a] <= 1 || fabs(m->gdel[a]) < 1e-30) { ind[a] = 0; frac[a] = 1.0; continue; }
    double u = (pt[a] - m->gmin[a])/m->gdel[a];
    double ip = trunc(u);
    frac[a] = 1.0 - (u - ip);
    ind[a] = (int)ip;
    if (ind[a] + 1 == m->dims[a]) { frac[a] = 1.0 - frac[a]; ind[a]--; }
    if (ind[a] < 0 || ind[a] >= m->dims[a]) return 0;
    if (ind[a] + 1 >= m->dims[a]) return 0;
  }
  return 1;
}

static void table_eval(const orc_member* m, const double* pt, const int* fl, int nf,
                       double* out, double* grad /* nf x 4, may be NULL */)
{
  int ind[4]; double fr[4];
  if (isnan(pt[0]) || isnan(pt[1]) || isnan(pt[2]) || isnan(pt[3]) || !table_locate(m, pt, ind, fr)) {
    for (int k = 0; k < nf; ++k) out[k] = NAN;
    if (grad) for (int k = 0; k < 4*nf; ++k) grad[k] = NAN;
    return;
  }
  for (int k = 0; k < nf; ++k) out[k] = 0.0;
  if (grad) for (int k = 0; k < 4*nf; ++k) grad[k] = 0.0;
  for (int h = 0; h < 16; ++h) {
    int bit[4] = { (h >> 3) & 1, (h >> 2) & 1, (h >> 1) & 1, h & 1 };
    double wa[4];
    double w = 1.0;
    for (int a = 0; a < 4; ++a) {
      wa[a] = fr[a] + bit[a] - 2.0*bit[a]*fr[a];
      w *= wa[a];
    }
    size_t node = (size_t)(ind[0] + bit[0]);
    for (int a = 1; a < 4; ++a) node = node*(size_t)m->dims[a] + (size_t)(ind[a] + bit[a]);
    const double* cell = m->tab + node*17;
    for (int k = 0; k < nf; ++k) out[k] += w*cell[fl[k]];
    if (grad) {
      for (int a = 0; a < 4; ++a) {
        if (m->dims[a] <= 1 || fabs(m->gdel[a]) < 1e-30) continue;
        double dw = (bit[a] ? 1.0 : -1.0)/m->gdel[a];
        for (int b = 0; b < 4; ++b) if (b != a) dw *= wa[b];
        for (int k = 0; k < nf; ++k) grad[4*k + a] += dw*cell[fl[k]];
      }
    }
  }
}

void orc_member_full(const orc_member* m, const double x[4], double r[17])
{
  switch (m->kind) {
    case ORC_M_CONF: conf_full(m, x, r); break;
    case ORC_M_DIAG: diag_full(m, x, r); break;
    case ORC_M_TANH: tanh_full(m, x, r); break;
    default: {
      static const int all[17] = {0,1,2,3,4,5,6,7,8,9,10,11,12,13,14,15,16};
      table_eval(m, x, all, 17, r, NULL);
    }
  }
}

/* Residual densities (s or e, B, S, Q) and their Jacobian with respect to
 * (T, muB, muQ, muS).  For analytic members the Jacobian is the exact
 * Hessian of p; for the table it is the derivative of the multilinear
 * interpolant.  Row order follows get_sBSQ / get_eBSQ (s|e, B, S, Q). */
static void member_densities(const orc_member* m, int mode, const double* x,
                             double* F, double* J)
{
  if (m->kind == ORC_M_TABLE) {
    const int fl[4] = { mode ? X_E : X_S, X_B, X_SS, X_Q };
    table_eval(m, x, fl, 4, F, J);
    return;
  }
  double r[17];
  orc_member_full(m, x, r);
  const double T = x[0], muB = x[1], muQ = x[2], muS = x[3];
  /* d(s,B,S,Q)/d(T,muB,muQ,muS) */
  double Js[4] = { r[X_TT], r[X_TB], r[X_TQ], r[X_TS] };
  double JB[4] = { r[X_TB], r[X_BB], r[X_BQ], r[X_BS] };
  double JS[4] = { r[X_TS], r[X_BS], r[X_QS], r[X_SSS] };
  double JQ[4] = { r[X_TQ], r[X_BQ], r[X_QQ], r[X_QS] };
  F[1] = r[X_B]; F[2] = r[X_SS]; F[3] = r[X_Q];
  if (mode == 0) {
    F[0] = r[X_S];
    for (int a = 0; a < 4; ++a) J[a] = Js[a];
  } else {
    F[0] = r[X_E];
    for (int a = 0; a < 4; ++a) J[a] = T*Js[a] + muB*JB[a] + muQ*JQ[a] + muS*JS[a];
  }
  for (int a = 0; a < 4; ++a) { J[4+a] = JB[a]; J[8+a] = JS[a]; J[12+a] = JQ[a]; }
}

static int in_box(const orc_member* m, const double* x)
{
  for (int a = 0; a < 4; ++a) if (!(x[a] >= m->mins[a] && x[a] <= m->maxs[a])) return 0;
  return 1;
}

static int solve4(const double* Jin, const double* b, double* x)
{
  double A[4][5];
  for (int i = 0; i < 4; ++i) { for (int j = 0; j < 4; ++j) A[i][j] = Jin[4*i+j]; A[i][4] = b[i]; }
  for (int c = 0; c < 4; ++c) {
    int piv = c; for (int i = c+1; i < 4; ++i) if (fabs(A[i][c]) > fabs(A[piv][c])) piv = i;
    if (A[piv][c] == 0.0 || !isfinite(A[piv][c])) return 0;
    if (piv != c) for (int j = 0; j < 5; ++j) { double t = A[c][j]; A[c][j] = A[piv][j]; A[piv][j] = t; }
    for (int i = c+1; i < 4; ++i) { double f = A[i][c]/A[c][c]; for (int j = c; j < 5; ++j) A[i][j] -= f*A[c][j]; }
  }
  for (int i = 3; i >= 0; --i) { double v = A[i][4]; for (int j = i+1; j < 4; ++j) v -= A[i][j]*x[j]; x[i] = v/A[i][i]; }
  for (int i = 0; i < 4; ++i) if (!isfinite(x[i])) return 0;
  return 1;
}

static double resid(const double* F, const double* tgt, double* R)
{
  double s = 0.0;
  for (int a = 0; a < 4; ++a) { R[a] = F[a] - tgt[a]; s += fabs(R[a]); }
  return s;
}

/* One rootfinder4D attempt (src/rootfinder.cpp:206-416): converge from x,
 * never leaving the member's box.  Returns 1 and the root in x. */
#define ORC_NEWTON_STEPS 60
static int newton(const orc_member* m, int mode, const double* tgt, double* x)
{
  double F[4], J[16], R[4], dx[4], xn[4], Fn[4], Jn[16], Rn[4];
  if (!in_box(m, x)) return 0;
  member_densities(m, mode, x, F, J);
  double res = resid(F, tgt, R);
  for (int it = 0; it < ORC_NEWTON_STEPS; ++it) {
    if (!isfinite(res)) return 0;
    const int converged = res < ORC_RF_TOL;
    double nr[4] = { -R[0], -R[1], -R[2], -R[3] };
    if (!solve4(J, nr, dx)) return converged;
    double lam = 1.0;
    int accepted = 0;
    for (int k = 0; k < 40; ++k) {
      for (int a = 0; a < 4; ++a) xn[a] = x[a] + lam*dx[a];
      if (in_box(m, xn)) {
        member_densities(m, mode, xn, Fn, Jn);
        double rn = resid(Fn, tgt, Rn);
        if (isfinite(rn) && (rn < res || (converged && rn <= res))) {
          for (int a = 0; a < 4; ++a) { x[a] = xn[a]; F[a] = Fn[a]; R[a] = Rn[a]; }
          for (int a = 0; a < 16; ++a) J[a] = Jn[a];
          res = rn; accepted = 1; break;
        }
      }
      if (converged) break;        /* polishing step: full step only */
      lam *= 0.5;
    }
    if (converged) return 1;
    if (!accepted) return 0;
  }
  return res < ORC_RF_TOL;
}

/* Rootfinder::find_root (src/rootfinder.cpp:421-715): seed adjustment and
 * the perturbed-seed attempts. */
static int find_root(const orc_member* m, int mode, const double* tgt, double* x)
{
  const double tol = ORC_RF_TOL;
  double seed[4] = { x[0], x[1], x[2], x[3] };
  const double sg[4] = { 1.0, sgn(tgt[1]), sgn(tgt[3]), sgn(tgt[2]) }; /* T, B, Q, S */
  for (int a = 0; a < 4; ++a)
    if (seed[a]*seed[a] < tol) seed[a] = fmin(0.5*m->maxs[a], 1.0)*sg[a];
  double t0 = seed[0], b0 = seed[1], q0 = seed[2], s0 = seed[3];
  double tries[10][4];
  int nt = 0;
#define TRY(a,b,c,d) do { tries[nt][0]=(a); tries[nt][1]=(b); tries[nt][2]=(c); tries[nt][3]=(d); ++nt; } while (0)
  TRY(t0, b0, q0, s0);
  TRY((t0 + 0.2*t0 > m->maxs[0]) ? m->maxs[0] - 1 : t0 + 0.2*t0, b0, q0, s0);
  TRY((t0 - 0.2*t0 < m->mins[0]) ? m->mins[0] + 1 : t0 - 0.2*t0, b0, q0, s0);
  TRY(t0, (b0 + 0.2*b0 > m->maxs[1]) ? m->maxs[1] - 1 : b0 + 0.2*b0, q0, s0);
  TRY(t0, (b0 - 0.2*b0 < m->mins[1]) ? m->mins[1] + 1 : b0 - 0.2*b0, q0, s0);
  TRY(t0, b0, q0, (s0 + 0.2*s0 > m->maxs[3]) ? m->maxs[3] - 1 : s0 + 0.2*s0);
  TRY(t0, b0, q0, (s0 - 0.2*s0 < m->mins[3]) ? m->mins[3] + 1 : s0 - 0.2*s0);
  TRY(t0, b0, (q0 + 0.2*q0 > m->maxs[2]) ? m->maxs[2] - 1 : q0 + 0.2*q0, s0);
  TRY(t0, b0, (q0 - 0.2*q0 < m->mins[2]) ? m->mins[2] + 1 : q0 - 0.2*q0, s0);
  TRY(t0, fmin(0.5*m->maxs[1], 1.0)*sg[1], fmin(0.5*m->maxs[2], 1.0)*sg[2], fmin(0.5*m->maxs[3], 1.0)*sg[3]);
#undef TRY
  for (int k = 0; k < nt; ++k) {
    double y[4] = { tries[k][0], tries[k][1], tries[k][2], tries[k][3] };
    if (newton(m, mode, tgt, y)) { for (int a = 0; a < 4; ++a) x[a] = y[a]; return 1; }
  }
  return 0;
}

/* EoS_conformal_diagonal::get_tbqs_seed_from_{s,e}BSQ
 * (include/eos_conformal_diagonal.h:39-80). tgt = (s|e, B, S, Q). */
static void diag_seed(const orc_member* d, int mode, const double* tgt, double* x)
{
  const double two23 = pow(2.0, 2.0/3.0), c13 = pow(d->c, 1.0/3.0);
  const double den = two23*c13;
  const double B = tgt[1], S = tgt[2], Q = tgt[3];
  const double B0 = d->muB0, S0 = d->muS0, Q0 = d->muQ0, T0 = d->T0;
  if (mode == 0) {
    x[0] = pow(T0*T0*T0*T0*tgt[0], 1.0/3.0)/den;
  } else {
    double disc = tgt[0] - 3.0*(pow(B0*fabs(B), 4.0/3.0) + pow(S0*fabs(S), 4.0/3.0)
                               + pow(Q0*fabs(Q), 4.0/3.0))/(4.0*two23*c13);
    x[0] = T0*pow(disc/(3.0*d->c), 0.25);
  }
  x[1] = sgn(B)*pow(B0*B0*B0*B0*fabs(B), 1.0/3.0)/den;
  x[2] = sgn(Q)*pow(Q0*Q0*Q0*Q0*fabs(Q), 1.0/3.0)/den;
  x[3] = sgn(S)*pow(S0*S0*S0*S0*fabs(S), 1.0/3.0)/den;
}

/* ------------------------------------------------------------------ */
/* Schur-complement derivatives (src/eos_derivatives.cpp:16-555).       */
/* a^T M^{-1} b with the protective clamps on M(1,1), M(2,2).            */
static double aTm1b(const double* a, const double* Min, const double* b)
{
  double M[9]; memcpy(M, Min, sizeof(M));
  const double EPS = 1e-3;
  if (fabs(M[4]) < EPS) M[4] = (M[4] >= 0.0 ? EPS : -EPS);
  if (fabs(M[8]) < EPS) M[8] = (M[8] >= 0.0 ? EPS : -EPS);
  /* y = M^{-1} b by LU with partial pivoting */
  double A[3][4];
  for (int i = 0; i < 3; ++i) { for (int j = 0; j < 3; ++j) A[i][j] = M[3*i+j]; A[i][3] = b[i]; }
  for (int c = 0; c < 3; ++c) {
    int piv = c; for (int i = c+1; i < 3; ++i) if (fabs(A[i][c]) > fabs(A[piv][c])) piv = i;
    if (piv != c) for (int j = 0; j < 4; ++j) { double t = A[c][j]; A[c][j] = A[piv][j]; A[piv][j] = t; }
    for (int i = c+1; i < 3; ++i) { double f = A[i][c]/A[c][c]; for (int j = c; j < 4; ++j) A[i][j] -= f*A[c][j]; }
  }
  double y[3];
  for (int i = 2; i >= 0; --i) { double v = A[i][3]; for (int j = i+1; j < 3; ++j) v -= A[i][j]*y[j]; y[i] = v/A[i][i]; }
  return a[0]*y[0] + a[1]*y[1] + a[2]*y[2];
}

/* Each derivative is  chi_lead - a^T M^{-1} b  with entries drawn from the
 * susceptibilities; the index triples below follow calc_term_1..4. */
typedef struct { int lead; int a[3]; int M[9]; int b[3]; } schur_t;
#define TT X_TT
#define TB X_TB
#define TQ X_TQ
#define TS X_TS
#define BB X_BB
#define QQ X_QQ
#define SS X_SSS
#define BQ X_BQ
#define BS X_BS
#define QS X_QS
static const schur_t K_T1   = { TT, {TB,TS,TQ}, {BB,BS,BQ, BS,SS,QS, BQ,QS,QQ}, {TB,TS,TQ} };
static const schur_t K_T2B  = { TB, {TT,TS,TQ}, {TB,BS,BQ, TS,SS,QS, TQ,QS,QQ}, {BB,BS,BQ} };
static const schur_t K_T2S  = { TS, {TT,TB,TQ}, {TB,BB,BQ, TS,BS,QS, TQ,BQ,QQ}, {BS,SS,QS} };
static const schur_t K_T2Q  = { TQ, {TT,TB,TS}, {TB,BB,BS, TS,BS,SS, TQ,BQ,QS}, {BQ,QS,QQ} };
static const schur_t K_T3B  = { TB, {BB,BS,BQ}, {TB,TS,TQ, BS,SS,QS, BQ,QS,QQ}, {TT,TS,TQ} };
static const schur_t K_T3S  = { TS, {BS,SS,QS}, {TB,TS,TQ, BB,BS,BQ, BQ,QS,QQ}, {TT,TB,TQ} };
static const schur_t K_T3Q  = { TQ, {BQ,QS,QQ}, {TB,TS,TQ, BB,BS,BQ, BS,SS,QS}, {TT,TB,TS} };
/* term4(j, i): outer i, inner j */
static const schur_t K_BB   = { BB, {TB,BS,BQ}, {TT,TS,TQ, TS,SS,QS, TQ,QS,QQ}, {TB,BS,BQ} }; /* db_dmub */
static const schur_t K_SB   = { BS, {TS,SS,QS}, {TT,TS,TQ, TB,BS,BQ, TQ,QS,QQ}, {TB,BB,BQ} }; /* ds_dmub */
static const schur_t K_QB   = { BQ, {TQ,QS,QQ}, {TT,TS,TQ, TB,BS,BQ, TS,SS,QS}, {TB,BB,BS} }; /* dq_dmub */
static const schur_t K_BS   = { BS, {TB,BB,BQ}, {TT,TB,TQ, TS,BS,QS, TQ,BQ,QQ}, {TS,SS,QS} }; /* db_dmus */
static const schur_t K_SS   = { SS, {TS,BS,QS}, {TT,TB,TQ, TB,BB,BQ, TQ,BQ,QQ}, {TS,BS,QS} }; /* ds_dmus */
static const schur_t K_QS   = { QS, {TQ,BQ,QQ}, {TT,TB,TQ, TB,BB,BQ, TS,BS,QS}, {TS,BS,SS} }; /* dq_dmus */
static const schur_t K_BQ   = { BQ, {TB,BB,BS}, {TT,TB,TS, TS,BS,SS, TQ,BQ,QS}, {TQ,QS,QQ} }; /* db_dmuq */
static const schur_t K_SQ   = { QS, {TS,BS,SS}, {TT,TB,TS, TB,BB,BS, TQ,BQ,QS}, {TQ,BQ,QQ} }; /* ds_dmuq */
static const schur_t K_QQ   = { QQ, {TQ,BQ,QS}, {TT,TB,TS, TB,BB,BS, TS,BS,SS}, {TQ,BQ,QS} }; /* dq_dmuq */
#undef TT
#undef TB
#undef TQ
#undef TS
#undef BB
#undef QQ
#undef SS
#undef BQ
#undef BS
#undef QS

static double schur(const schur_t* k, const double* r)
{
  double a[3], b[3], M[9];
  for (int i = 0; i < 3; ++i) { a[i] = r[k->a[i]]; b[i] = r[k->b[i]]; }
  for (int i = 0; i < 9; ++i) M[i] = r[k->M[i]];
  return r[k->lead] - aTm1b(a, M, b);
}

/* EquationOfState::set_thermo (include/eos.h:121-175) for the point x of
 * member m: fills the 31 thermodynamic_info doubles. */
static void set_thermo(const orc_eos* e, int member, const double* x, double* th)
{
  const orc_member* m = &e->m[member];
  double r[17];
  orc_member_full(m, x, r);
  const double T = x[0], muB = x[1], muQ = x[2], muS = x[3];
  const double s = r[X_S], B = r[X_B], S = r[X_SS], Q = r[X_Q];
  const int hB = fabs(B) > ORC_TINY, hS = fabs(S) > ORC_TINY, hQ = fabs(Q) > ORC_TINY;
  const double dsdt = schur(&K_T1, r);
  /* dwds (src/eos.cpp:276-292) */
  double ct = 0.0;
  if (hB) ct += B/schur(&K_T2B, r);
  if (hQ) ct += Q/schur(&K_T2Q, r);
  if (hS) ct += S/schur(&K_T2S, r);
  const double dwds = T + s/dsdt + ct;
  double dwdB = muB, dwdS = muS, dwdQ = muQ;
  if (hB) {
    double c = s/schur(&K_T3B, r) + B/schur(&K_BB, r);
    if (hQ) c += Q/schur(&K_BQ, r);
    if (hS) c += S/schur(&K_BS, r);
    dwdB += c;
  }
  if (hS) {
    double c = s/schur(&K_T3S, r) + S/schur(&K_SS, r);
    if (hB) c += B/schur(&K_SB, r);
    if (hQ) c += Q/schur(&K_SQ, r);
    dwdS += c;
  }
  if (hQ) {
    double c = s/schur(&K_T3Q, r) + Q/schur(&K_QQ, r);
    if (hB) c += B/schur(&K_QB, r);
    if (hS) c += S/schur(&K_QS, r);
    dwdQ += c;
  }
  /* dalpha_X d{s,B,S,Q} (src/eos.cpp:336-492) */
  double aBs = -muB/T/T/dsdt + (hB ? 1/T/schur(&K_T2B, r) : 0.0);
  double aSs = -muS/T/T/dsdt + (hS ? 1/T/schur(&K_T2S, r) : 0.0);
  double aQs = -muQ/T/T/dsdt + (hQ ? 1/T/schur(&K_T2Q, r) : 0.0);
  double aBB = hB ? 1/T/schur(&K_BB, r) - muB/T/T/schur(&K_T3B, r) : 0.0;
  double aSB = hB ? -muS/T/T/schur(&K_T3B, r) + (hS ? 1/T/schur(&K_BS, r) : 0.0) : 0.0;
  double aQB = hB ? -muQ/T/T/schur(&K_T3B, r) + (hQ ? 1/T/schur(&K_BQ, r) : 0.0) : 0.0;
  double aBS = hS ? -muB/T/T/schur(&K_T3S, r) + (hB ? 1/T/schur(&K_SB, r) : 0.0) : 0.0;
  double aSS = hS ? -muS/T/T/schur(&K_T3S, r) + 1/T/schur(&K_SS, r) : 0.0;
  double aQS = hS ? -muQ/T/T/schur(&K_T3S, r) + (hQ ? 1/T/schur(&K_SQ, r) : 0.0) : 0.0;
  double aBQ = hQ ? -muB/T/T/schur(&K_T3Q, r) + (hB ? 1/T/schur(&K_QB, r) : 0.0) : 0.0;
  double aSQ = hQ ? -muS/T/T/schur(&K_T3Q, r) + (hS ? 1/T/schur(&K_QS, r) : 0.0) : 0.0;
  double aQQ = hQ ? -muQ/T/T/schur(&K_T3Q, r) + 1/T/schur(&K_QQ, r) : 0.0;
  /* chiBB0 (src/eos.cpp:198-211) */
  double chiBB0 = r[X_BB];
  if (m->kind == ORC_M_TABLE && (muB != 0.0 || muQ != 0.0 || muS != 0.0)) {
    double z[4] = { T, 0.0, 0.0, 0.0 }, r0[17];
    orc_member_full(m, z, r0);
    chiBB0 = r0[X_BB];
  }
  th[0] = T; th[1] = muB; th[2] = muS; th[3] = muQ;
  th[4] = r[X_E]; th[5] = s; th[6] = B; th[7] = S; th[8] = Q;
  th[9] = r[X_P]; th[10] = r[X_CS2]; th[11] = r[X_E] + r[X_P];
  th[12] = dwds; th[13] = dwdB; th[14] = dwdS; th[15] = dwdQ;
  th[16] = aBs; th[17] = aBB; th[18] = aBS; th[19] = aBQ;
  th[20] = aSs; th[21] = aSB; th[22] = aSS; th[23] = aSQ;
  th[24] = aQs; th[25] = aQB; th[26] = aQS; th[27] = aQQ;
  th[28] = r[X_BB]; th[29] = chiBB0;
  th[30] = (m->kind == ORC_M_TABLE) ? 1.0 : 0.0;
}

/* ------------------------------------------------------------------ */
/* point-level inversion: EquationOfState::rootfinder_update_s
 * (src/eos.cpp:629-804) for mode 0 and rootfinder_s_out (813-975) for
 * mode 1.  target = (s|e, rhoB, rhoS, rhoQ); x = seed (T, muB, muQ, muS). */
int orc_eos_invert_point(orc_eos* e, int mode, const double target[4],
                         double x[4], int* member, double thermo31[31])
{
  const double hc = ORC_HBARC_MEVFM;
  const orc_member* diag = &e->m[e->n - 1];
  int found = 0, used = -1;
  double res[4];
  int start = (mode == 0) ? *member : 0;
  if (start < 0 || start >= e->n) start = 0;
  for (int k = start; k < e->n; ++k) {
    const orc_member* m = &e->m[k];
    double seeds[4][4];
    int ns = 0;
    if (mode == 0) {
      for (int a = 0; a < 4; ++a) seeds[ns][a] = x[a];
      ++ns;
      seeds[ns][0] = x[0]; seeds[ns][1] = 0.0; seeds[ns][2] = 0.0; seeds[ns][3] = 0.0; ++ns;
    }
    seeds[ns][0] = 580.0/hc; seeds[ns][1] = sgn(target[1]); seeds[ns][2] = sgn(target[3]); seeds[ns][3] = sgn(target[2]); ++ns;
    diag_seed(diag, mode, target, seeds[ns]); ++ns;
    found = 0;
    for (int q = 0; q < ns && !found; ++q) {
      double y[4] = { seeds[q][0], seeds[q][1], seeds[q][2], seeds[q][3] };
      if (find_root(m, mode, target, y)) { found = 1; for (int a = 0; a < 4; ++a) res[a] = y[a]; }
    }
    if (found) {
      used = k;
      double r[17];
      orc_member_full(m, res, r);
      /* cs2 checks (src/eos.cpp:775-786); a rejected solution becomes the
       * current location that seeds the next member (tbqs(result, eos)). */
      if (r[X_CS2] < 0.0 || r[X_CS2] > 1.0) {
        for (int a = 0; a < 4; ++a) x[a] = res[a];
        continue;
      }
      break;
    }
  }
  if (!found) return 0;
  for (int a = 0; a < 4; ++a) x[a] = res[a];
  *member = used;
  set_thermo(e, used, res, thermo31);
  return 1;
}

/* ------------------------------------------------------------------ */
static void init_analytic(orc_member* m, int kind, double c, double T0, double muB0,
                          double muQ0, double muS0, double Tc, double Ts)
{
  memset(m, 0, sizeof(*m));
  m->kind = kind; m->c = c; m->T0 = T0; m->muB0 = muB0; m->muQ0 = muQ0; m->muS0 = muS0;
  m->Tc = Tc; m->Ts = Ts;
  m->mins[0] = 0.0; m->maxs[0] = ORC_TBQS_INF;
  for (int a = 1; a < 4; ++a) { m->mins[a] = -ORC_TBQS_INF; m->maxs[a] = ORC_TBQS_INF; }
}

orc_eos* orc_eos_create(const orc_eos_desc* d)
{
  orc_eos* e = (orc_eos*)calloc(1, sizeof(orc_eos));
  e->kind = d->kind;
  if (d->kind == ORC_EOS_PREINVERTED) {      /* EoS_Interpolator, flat (s, nB) format */
    e->n = 0;
    e->pre = orc_preinv_create(d);
    if (!e->pre) { free(e); return NULL; }
    return e;
  }
  const double Nc = 3.0, Nf = 2.5;
  const double c0 = ORC_PI*ORC_PI*(2.0*(Nc*Nc - 1.0) + (7.0/2.0)*Nc*Nf)/90.0;
  if (d->kind == ORC_EOS_CONFORMAL) {
    e->n = 2;
    init_analytic(&e->m[0], ORC_M_CONF, c0, 1.0, 1.0, 1.0, 1.0, 0, 0);
    init_analytic(&e->m[1], ORC_M_DIAG, c0, 1.0, 1.0, 1.0, 1.0, 0, 0);
    return e;
  }
  /* table chain */
  e->n = 4;
  orc_member* t = &e->m[0];
  memset(t, 0, sizeof(*t));
  t->kind = ORC_M_TABLE;
  size_t nn = 1;
  for (int a = 0; a < 4; ++a) {
    t->dims[a] = d->dims[a]; nn *= (size_t)d->dims[a];
    t->gmin[a] = d->mins[a]; t->gmax[a] = d->maxs[a];
    t->gdel[a] = (d->dims[a] <= 1) ? 0.0 : (d->maxs[a] - d->mins[a])/(d->dims[a] - 1.0);
    t->mins[a] = d->mins[a]; t->maxs[a] = d->maxs[a];
  }
  t->tab = (double*)malloc(nn*17*sizeof(double));
  memcpy(t->tab, d->table, nn*17*sizeof(double));
  /* matched fallbacks at Tmax = 1.1 T_FO (src/eos_initialization.cpp:270-488) */
  const double Tmax = 1.1*d->T_freeze_out;
  double r[17], pt[4] = { Tmax, 0, 0, 0 };
  orc_member_full(t, pt, r);
  const double pT = r[X_P];
  const double c = pT/(Tmax*Tmax*Tmax*Tmax);
  double mu0[4] = { 1.0, 1.0, 1.0, 1.0 }, mud[4] = { 1.0, 1.0, 1.0, 1.0 };
  for (int a = 1; a < 4; ++a) {
    double mmax = t->maxs[a];
    if (fabs(mmax) < 1e-30) continue;
    double q[4] = { Tmax, 0, 0, 0 }; q[a] = mmax;
    orc_member_full(t, q, r);
    mu0[a] = pow(c, 0.25)*mmax/sqrt(sqrt(r[X_P]) - sqrt(pT));
    mud[a] = pow(c/(r[X_P] - pT), 0.25)*mmax;
  }
  const double Tc = 220.0/ORC_HBARC_MEVFM, Ts = 120.0/ORC_HBARC_MEVFM;
  /* axes: 1 = muB, 2 = muQ, 3 = muS */
  init_analytic(&e->m[1], ORC_M_TANH, c, 1.0, mu0[1], mu0[2], mu0[3], Tc, Ts);
  init_analytic(&e->m[2], ORC_M_CONF, c, 1.0, mu0[1], mu0[2], mu0[3], 0, 0);
  init_analytic(&e->m[3], ORC_M_DIAG, c, 1.0, mud[1], mud[2], mud[3], 0, 0);
  return e;
}

void orc_eos_destroy(orc_eos* e)
{
  if (!e) return;
  for (int k = 0; k < e->n; ++k) free(e->m[k].tab);
  orc_preinv_destroy(e->pre);
  free(e);
}

int orc_eos_members(orc_eos* e) { return e->n; }

void orc_eos_full_thermo(orc_eos* e, int member, const double tbqs[4], double out[17])
{
  orc_member_full(&e->m[member], tbqs, out);
}

/* EquationOfState::efreeze: e(T_FO, 0, 0, 0) of the default member. */
double orc_eos_efreeze(orc_eos* e, double T_fo)
{
  if (e->pre) return orc_preinv_efreeze(e->pre, T_fo);   /* EoS_Interpolator::efreeze_at_T */
  double r[17], pt[4] = { T_fo, 0, 0, 0 };
  orc_member_full(&e->m[0], pt, r);
  return r[X_E];
}
