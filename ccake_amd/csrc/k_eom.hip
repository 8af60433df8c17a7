// Equations of motion (Milne, Israel-Stewart, MRF form) and the RK update,
// one thread per particle, D = 2.
// Replaces EoM_default::evaluate_time_derivatives (src/eom_default.cpp:211-869)
// with calculate_MRF_shear (882-1412) -- seven Kokkos kernels with a global
// round trip between each -- and Evolver::step_rk's update (src/evolver.cpp:363-484).
// Index conventions (include/milne.hpp): 2-vectors and 2x2 matrices are
// lowered / raised by a sign flip; pi^{mu nu} (4x4) lowers slots 1,2 by a
// sign and slot 3 by -tau^2; the 2+1D geometric factor and delta_{i eta}
// vanish.
#include "ccake_internal.h"
#include "grad_sum.h"

namespace ccake {

struct EomIn {
  double t;
  int shear, vort;
};

// Column accessor for particle i.
struct Cols {
  double* F; int npad; int i;
  __device__ __forceinline__ double& operator()(int c) const { return F[(size_t)c * npad + i]; }
};

// EoM_default::compute_hydro_numbers (src/eom_default.cpp:1831-1995) for
// D = 2 without diffusion: shear and bulk Knudsen numbers and inverse
// Reynolds numbers.  s11, s12 are sigma^{ij} at (i, j) = (1, 1), (1, 2) of
// the MRF (M_sigma du/dt + F_sigma).  The reference stores sigma^{ij} at
// (i+1, j+1) of its 4x4 and reads it back into (i, j); the time components
// it then derives from u overwrite (0, 0..2), so only (1, 1) and (1, 2)
// survive the read, and (2, 2) stays zero.  Both contractions lower with
// Matrix<double,4,4>::make_covariant (include/milne.hpp:550-590): slots 1, 2
// by -1, slot 3 by -t^2, and contract() runs over columns, then rows.
__device__ __forceinline__ void hydro_numbers(const double S[4][4], double s11, double s12, double ux, double uy,
                                              double gamma, double t2, double tau_pi, double tau_Pi,
                                              double theta, double bulk, double pressure, double* hn)
{
  double sg[4][4];
  #pragma unroll
  for (int a = 0; a < 4; ++a)
    #pragma unroll
    for (int b = 0; b < 4; ++b) sg[a][b] = 0.0;
  sg[1][1] = s11; sg[1][2] = s12; sg[2][1] = s12;
  const double uc[2] = { -ux, -uy };
  sg[0][1] = -1. * ((0.0 + uc[0] * sg[1][1]) + uc[1] * sg[1][2]) / gamma;
  sg[0][2] = -1. * ((0.0 + uc[0] * sg[2][1]) + uc[1] * sg[2][2]) / gamma;
  sg[1][0] = sg[0][1]; sg[2][0] = sg[0][2];
  const double fc0 = -ux, fc1 = -uy, fc2 = 0.0 * -t2;
  sg[0][0] = (fc0 * fc0 * sg[1][1] + fc1 * fc1 * sg[2][2] + 2. * fc1 * fc0 * sg[2][1] + 2. * fc2 * fc0 * sg[3][1]
              + 2. * fc2 * fc1 * sg[3][2] - fc2 * fc2 * (sg[1][1] + sg[2][2]) / t2)
           / (gamma * gamma - fc2 * fc2 / t2);
  sg[3][3] = (sg[0][0] - sg[1][1] - sg[2][2]) / t2;
  const double g[4] = { 1.0, -1.0, -1.0, -t2 };
  double sn = 0.0, pn = 0.0;
  #pragma unroll
  for (int c = 0; c < 4; ++c)
    #pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double sc = (c == 0 ? sg[r][c] : sg[r][c] * g[c]);
      const double pc = (c == 0 ? S[r][c] : S[r][c] * g[c]);
      sn += sg[r][c] * (r == 0 ? sc : sc * g[r]);
      pn += S[r][c] * (r == 0 ? pc : pc * g[r]);
    }
  hn[0] = tau_pi * sqrt(fabs(sn));
  hn[1] = tau_Pi * fabs(theta);
  hn[2] = sqrt(fabs(pn)) / pressure;
  hn[3] = fabs(bulk) / pressure;
}

// ---------------- calculate_MRF_diffusion (src/eom_default.cpp:1416-1822), D = 2 ----------------
// Shared set-up: tau_q^{-1}, tau_q^{-1} tau_q, tau_q^{-1} kappa_q, q^mu and
// its covariant form, the EoS derivatives d alpha_c / d(s, n_c') and
// grad alpha.  Charges run B, S, Q.
struct DiffCtx {
  double itq[3][3], itt[3][3], itk[3][3];
  double q[3][4], qc[3][4];
  double das[3], dar[3][3];
  double galpha[2][3];
};

// inverse of a 3x3 matrix by cofactors
__device__ __forceinline__ void inv3_cof(const double A[3][3], double R[3][3])
{
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  R[0][0] = c00 / det; R[1][0] = c01 / det; R[2][0] = c02 / det;
  R[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
  R[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
  R[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
  R[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
  R[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
  R[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
}

// inverse of a 3x3 matrix through its explicit determinant (used for R_N)
__device__ __forceinline__ void inv3_det(const double A[3][3], double R[3][3])
{
  const double det = A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1]
                   - A[0][2] * A[1][1] * A[2][0] - A[0][1] * A[1][0] * A[2][2] - A[0][0] * A[1][2] * A[2][1];
  R[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
  R[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
  R[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
  R[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
  R[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
  R[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
  R[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
  R[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
  R[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
}

__device__ __forceinline__ void diff_ctx(const Cols& col, double t2, DiffCtx& x)
{
  double tq[3][3], kq[3][3];
  #pragma unroll
  for (int c = 0; c < 3; ++c) {
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
      tq[c][k] = col(CCAKE_F_TAU_Q0 + 3 * c + k);
      kq[c][k] = col(CCAKE_F_KAPPA_Q0 + 3 * c + k);
      x.dar[c][k] = col(CCAKE_F_DALPHA0 + 4 * c + 1 + k);
    }
    x.das[c] = col(CCAKE_F_DALPHA0 + 4 * c);
    #pragma unroll
    for (int m = 0; m < 4; ++m) x.q[c][m] = col(CCAKE_F_Q0 + 4 * c + m);
    x.qc[c][0] = x.q[c][0]; x.qc[c][1] = -x.q[c][1]; x.qc[c][2] = -x.q[c][2]; x.qc[c][3] = x.q[c][3] * -t2;
    x.galpha[0][c] = col(CCAKE_F_GRAD_ALPHA0 + c);
    x.galpha[1][c] = col(CCAKE_F_GRAD_ALPHA0 + 3 + c);
  }
  inv3_cof(tq, x.itq);
  #pragma unroll
  for (int a = 0; a < 3; ++a)
    #pragma unroll
    for (int b = 0; b < 3; ++b) {
      double tt = 0.0, tk = 0.0;
      #pragma unroll
      for (int k = 0; k < 3; ++k) { tt += x.itq[a][k] * tq[k][b]; tk += x.itq[a][k] * kq[k][b]; }
      x.itt[a][b] = tt; x.itk[a][b] = tk;
    }
}

// The q^0 MRFs: F_q0[c], R_q0[c][c'], M_q0[j][c].
__device__ __forceinline__ void diff_q0(const DiffCtx& x, double t, double gamma, double sigma, double s,
                                        const double rho[3], const double u[2], double divV, const double M_S[2],
                                        double F_S, const double R_S[3], double F_q0[3], double R_q0[3][3],
                                        double M_q0[2][3])
{
  const double g2 = gamma * gamma, gf = 0.0, ueta = 0.0;
  const double uc[2] = { -u[0], -u[1] };
  const double expf = gamma * divV + gf + gamma / t;
  double F0D[3], F0a[3], A0a[3], R0a[3][3], M0a[3][2], M0D[3][2];
  #pragma unroll
  for (int c = 0; c < 3; ++c) {
    F0D[c] = g2 * gf * x.q[c][0] + 2. * g2 * x.qc[c][3] * ueta / t + t * gamma * x.qc[c][0] * ueta * ueta
           + t * ueta * x.q[c][3];
    F0a[c] = (1. - g2) * sigma * x.das[c] * F_S;
    A0a[c] = -(1. - g2) * x.das[c] * s / gamma;
    #pragma unroll
    for (int d = 0; d < 2; ++d) {
      M0D[c][d] = g2 * (x.qc[c][d + 1] - x.qc[c][0] * uc[d] / gamma);
      M0a[c][d] = (1. - g2) * sigma * x.das[c] * M_S[d];
      F0a[c] += -x.galpha[d][c] * u[d] / gamma;
    }
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
      R0a[c][k] = (1. - g2) * sigma * (x.das[c] * R_S[k] + x.dar[c][k]);
      A0a[c] += -(1. - g2) * x.dar[c][k] * rho[k] / gamma;
    }
  }
  #pragma unroll
  for (int c = 0; c < 3; ++c) {
    #pragma unroll
    for (int d = 0; d < 2; ++d) {
      double m = 0.0;
      #pragma unroll
      for (int k = 0; k < 3; ++k)
        m += (-x.itt[c][k] * M0D[k][d] + x.itk[c][k] * M0a[k][d] - x.itk[c][k] * A0a[k] * uc[d] / gamma) / (gamma)
           + x.itt[c][k] * x.q[k][0] * uc[d] / gamma / gamma;
      M_q0[d][c] = m;
    }
    double f = 0.0;
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
      f += (-x.itt[c][k] * F0D[k] - x.itq[c][k] * x.q[k][0] + x.itk[c][k] * F0a[k] + x.itk[c][k] * A0a[k] * expf)
           / (gamma)
         - x.itt[c][k] * x.q[k][0] * expf / gamma;
      double r = 0.0;
      #pragma unroll
      for (int kk = 0; kk < 3; ++kk) r += x.itk[c][kk] * R0a[kk][k] / (gamma);
      R_q0[c][k] = r;
    }
    F_q0[c] = f;
  }
}

// d(extensive q^i_c)/dt = F_q + M_q du/dt + R_q dN/dt, out[3 c + i].  F_D
// keeps only its first two terms: the reference ends that statement with a
// semicolon after them (src/eom_default.cpp:1744-1745).
__device__ __forceinline__ void diff_dext_q(const DiffCtx& x, double t, double t2, double gamma, double sigma,
                                            double s, const double rho[3], const double u[2], double divV,
                                            const double M_S[2], double F_S, const double R_S[3], double du0,
                                            double du1, const double dN[3], double* out)
{
  const double cmd[3] = { -1.0, -1.0, -1.0 / t / t };
  const double gf = 0.0;
  const double fsu[3] = { u[0], u[1], 0.0 };
  const double fsuc[2] = { -fsu[0], -fsu[1] };
  const double expf = gamma * divV + gf + gamma / t;
  const double du[2] = { du0, du1 };
  #pragma unroll
  for (int d = 0; d < 3; ++d) {
    double FD[3], Fa[3], Aa[3], MD[3][2], Ma[3][2], Ra[3][3];
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double ga = (d < 2) ? x.galpha[d][c] : 0.0;
      FD[c] = fsu[d] * gamma * x.q[c][0] * gf + 2. * fsu[d] * x.qc[c][3] * fsu[2] * gamma / t;
      Fa[c] = -fsu[d] * gamma * sigma * x.das[c] * F_S + cmd[d] * ga;
      Aa[c] = fsu[d] * x.das[c] * s;
      #pragma unroll
      for (int j = 0; j < 2; ++j) {
        MD[c][j] = fsu[d] * gamma * (x.qc[c][j + 1] - x.qc[c][0] * fsuc[j] / gamma);
        Ma[c][j] = -fsu[d] * gamma * sigma * x.das[c] * M_S[j];
      }
      #pragma unroll
      for (int k = 0; k < 3; ++k) {
        Ra[c][k] = -fsu[d] * gamma * sigma * (x.das[c] * R_S[k] + x.dar[c][k]);
        Aa[c] += fsu[d] * x.dar[c][k] * rho[k];
      }
    }
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      #pragma unroll
      for (int k = 0; k < 3; ++k)
        v += (-x.itt[c][k] * FD[k] - x.itq[c][k] * x.q[k][d + 1] + x.itk[c][k] * Fa[k]
              + x.itk[c][k] * Aa[k] * expf) / (gamma * sigma);
      #pragma unroll
      for (int j = 0; j < 2; ++j) {
        double m = 0.0;
        #pragma unroll
        for (int k = 0; k < 3; ++k)
          m += (-x.itt[c][k] * MD[k][j] + x.itk[c][k] * Ma[k][j] - x.itk[c][k] * Aa[k] * fsuc[j] / gamma)
               / (gamma * sigma);
        v += m * du[j];
      }
      #pragma unroll
      for (int k = 0; k < 3; ++k) {
        double r = 0.0;
        #pragma unroll
        for (int kk = 0; kk < 3; ++kk) r += x.itk[c][kk] * Ra[kk][k] / (gamma * sigma);
        v += r * dN[k];
      }
      out[3 * c + d] = v;
    }
  }
}

template <bool SHEAR, bool DIFF>
__device__ __forceinline__ void eom_d2(const Cols& col, double t, double use_vort, double* du_out,
                                       double* d_s_out, double* dN, double* d_b_out, double* d_shv_out,
                                       double* theta_out, double* snu_out, double* dez_out, double* hn_out,
                                       double* dq_out)
{
  const double t2 = t * t;
  const double cmd3 = -1.0 / t / t;                      // contra metric (1,-1,-1,-1/t^2)
  const double cmd[4] = { 1.0, -1.0, -1.0, cmd3 };
  const double gamma = col(CCAKE_F_GAMMA), sigma = col(CCAKE_F_SIGMA), bulk = col(CCAKE_F_BULK);
  const double T = col(CCAKE_F_T), w = col(CCAKE_F_W), s = col(CCAKE_F_S);
  const double muB = col(CCAKE_F_MUB), muS = col(CCAKE_F_MUS), muQ = col(CCAKE_F_MUQ);
  const double rB = col(CCAKE_F_RHOB), rS = col(CCAKE_F_RHOS), rQ = col(CCAKE_F_RHOQ);
  const double dwds = col(CCAKE_F_DWDS), dwdB = col(CCAKE_F_DWDB), dwdS = col(CCAKE_F_DWDS_S), dwdQ = col(CCAKE_F_DWDQ);
  const double eta_pi = col(CCAKE_F_ETA_PI), tau_pi = col(CCAKE_F_TAU_PI);
  const double zeta = col(CCAKE_F_ZETA_PI), tau_Pi = col(CCAKE_F_TAU_BULK);
  const double delta_PiPi = col(CCAKE_F_DELTA_BB), delta_pipi = col(CCAKE_F_DELTA_PIPI);
  const double lambda_piPi = col(CCAKE_F_LAMBDA_PIB), lambda_Pipi = col(CCAKE_F_LAMBDA_BPI);
  const double phi1 = col(CCAKE_F_PHI1), phi3 = col(CCAKE_F_PHI3), phi6 = col(CCAKE_F_PHI6), phi7 = col(CCAKE_F_PHI7);
  const double tau_pipi = col(CCAKE_F_TAU_PIPI);
  const double u[2] = { col(CCAKE_F_UX), col(CCAKE_F_UY) };
  const double v[2] = { col(CCAKE_F_VX), col(CCAKE_F_VY) };
  const double uc[2] = { -u[0], -u[1] };
  const double fu[3] = { u[0], u[1], 0.0 };
  const double fc[3] = { -u[0], -u[1], -t2 * 0.0 };
  double gV[2][2] = { { col(CCAKE_F_GRADV0 + 0), col(CCAKE_F_GRADV0 + 1) },
                      { col(CCAKE_F_GRADV0 + 3), col(CCAKE_F_GRADV0 + 4) } };
  const double divV = gV[0][0] + gV[1][1];
  double S[4][4], H[4][4];
  #pragma unroll
  for (int a = 0; a < 4; ++a)
    #pragma unroll
    for (int b = 0; b < 4; ++b) S[a][b] = col(CCAKE_F_SHV0 + 4 * a + b);
  // H = pi^mu_nu (slot 1 lowered); C = pi_{mu nu}
  #pragma unroll
  for (int a = 0; a < 4; ++a) { H[a][0] = S[a][0]; H[a][1] = -S[a][1]; H[a][2] = -S[a][2]; H[a][3] = S[a][3] * -t2; }
  double shv_shvcov = 0.0;
  #pragma unroll
  for (int a = 0; a < 4; ++a)
    #pragma unroll
    for (int b = 0; b < 4; ++b) {
      const double c = (a == 0) ? H[a][b] : ((a == 3) ? H[a][b] * -t2 : -H[a][b]);
      shv_shvcov += S[a][b] * c;
    }
  const double expf = gamma / t + gamma * divV;   // gf + gamma/t + gamma*divV, gf = 0

  double M_snu[2] = { 0.0, 0.0 }, F_snu = 0.0;
  double F_0i[2] = { 0.0, 0.0 }, M_0i[2][2] = { { 0.0, 0.0 }, { 0.0, 0.0 } };
  double F_es[2][3] = { { 0, 0, 0 }, { 0, 0, 0 } };
  double M_es[2][3][2] = { { { 0, 0 }, { 0, 0 }, { 0, 0 } }, { { 0, 0 }, { 0, 0 }, { 0, 0 } } };
  double sgF[2] = { 0.0, 0.0 }, sgM[2][2] = { { 0.0, 0.0 }, { 0.0, 0.0 } };   // sigma^{11}, sigma^{12} MRFs
  if (SHEAR) {
    // ---------------- calculate_MRF_shear ----------------
    double gu[2][2];
    #pragma unroll
    for (int a = 0; a < 2; ++a)
      #pragma unroll
      for (int b = 0; b < 2; ++b) {
        gu[a][b] = gamma * gV[a][b];
        gu[a][b] += -gamma * u[b] * uc[0] * gV[a][0];
        gu[a][b] += -gamma * u[b] * uc[1] * gV[a][1];
      }
    double gu0[2];
    #pragma unroll
    for (int a = 0; a < 2; ++a) gu0[a] = -1. * (gu[a][0] * uc[0] + gu[a][1] * uc[1]) / gamma;
    const double gu0c[2] = { -gu0[0], -gu0[1] };
    double cg3[3][3] = { { -gu[0][0], -gu[0][1], 0.0 }, { -gu[1][0], -gu[1][1], 0.0 }, { 0.0, 0.0, 0.0 } };
    double g3[3][3] = { { gu[0][0], gu[0][1], 0.0 }, { gu[1][0], gu[1][1], 0.0 }, { 0.0, 0.0, 0.0 } };
    double guv[2];
    #pragma unroll
    for (int b = 0; b < 2; ++b) guv[b] = gu[0][b] * v[0] + gu[1][b] * v[1];
    const double gu0v = gu0[0] * v[0] + gu0[1] * v[1];
    double vgs[2], gugs[2], sgui[2], vsgu[2];
    #pragma unroll
    for (int a = 0; a < 2; ++a) {
      vgs[a] = guv[0] * H[a + 1][1] + guv[1] * H[a + 1][2] + gu0v * H[a + 1][0];
      gugs[a] = cg3[a][0] * H[0][1] + cg3[a][1] * H[0][2] + cmd[a + 1] * gu0[a] * H[0][0];
      sgui[a] = 0.0; vsgu[a] = 0.0;
      #pragma unroll
      for (int b = 0; b < 2; ++b) {
        sgui[a] += S[b + 1][0] * gu[b][a] + S[b + 1][a + 1] * gu0[b];
        vsgu[a] += v[b] * S[0][0] * gu[b][a] + v[b] * S[0][a + 1] * gu0[b];
      }
    }
    double sgu = 0.0;
    #pragma unroll
    for (int a = 0; a < 2; ++a) {
      #pragma unroll
      for (int b = 0; b < 2; ++b) sgu += (H[a + 1][b + 1] - v[a] * H[0][b + 1]) * gu[a][b];
      sgu += (H[a + 1][0] - v[a] * H[0][0]) * gu0[a];
    }
    double sl0l[2];
    #pragma unroll
    for (int a = 0; a < 2; ++a) {
      sl0l[a] = 0.0;
      #pragma unroll
      for (int b = 0; b < 4; ++b) sl0l[a] += (S[0][b] * H[a + 1][b] + S[a + 1][b] * H[0][b]) / 2.;
    }
    #pragma unroll
    for (int a = 0; a < 2; ++a) {
      const double Fsig = -(guv[a] - gu0c[a]) / 2. + u[a] * gamma * expf / 3.;
      const double FD = 0.0;
      const double Fds = -(vgs[a] - gugs[a] - 4. * gamma * u[a] * sgu / 3. - sgui[a] + vsgu[a]) / 4.
                       + (4. * gamma * u[a] * (gamma * H[3][3] / t) / 3.) / 4.
                       - S[0][a + 1] * expf / 3.;
      const double Fdo = -(vgs[a] - gugs[a] + sgui[a] - vsgu[a]) / 4.;
      const double Fdd = S[a + 1][0] * expf;
      double Msig[2], MD[2], Mds[2], Mdo[2], Mdd[2];
      #pragma unroll
      for (int b = 0; b < 2; ++b) {
        Msig[b] = u[a] * uc[b] / (2. * 3.);
        MD[b] = gamma * u[a] * (H[0][b + 1] - H[0][0] * uc[b] / gamma)
              + gamma * gamma * (H[a + 1][b + 1] - H[a + 1][0] * uc[b] / gamma);
        Mds[b] = (H[a + 1][b + 1] - H[a + 1][0] * uc[b] / gamma - gamma * gamma * H[a + 1][b + 1]
                  - gamma * u[a] * H[0][b + 1] + uc[b] * gamma * H[a + 1][0] + uc[b] * u[a] * H[0][0]
                  + 4. * gamma * u[a] * (H[0][b + 1] - H[0][0] * uc[b] / gamma) / 3.
                  + (-S[0][a + 1] + 4. * S[0][a + 1] / 3.) * uc[b] / gamma) / 4.;
        Mdo[b] = (H[a + 1][b + 1] - H[a + 1][0] * uc[b] / gamma - gamma * gamma * H[a + 1][b + 1]
                  - gamma * u[a] * H[0][b + 1] + uc[b] * gamma * H[a + 1][0] + uc[b] * u[a] * H[0][0]
                  + S[0][a + 1] * uc[b] / gamma) / 4.;
        Mdd[b] = -S[a + 1][0] * uc[b] / gamma;
      }
      Msig[a] += (1. - gamma * gamma) / 2.;
      Mdo[a] += -S[0][0] / 4.;
      Mds[a] += +S[0][0] / 4.;
      F_0i[a] = (-S[a + 1][0] - tau_pi * FD - delta_pipi * Fdd + 2. * tau_pi * Fdo * use_vort - tau_pipi * Fds
                 + (2. * eta_pi + lambda_piPi * bulk) * Fsig + phi6 * bulk * S[a + 1][0]
                 + phi7 * (sl0l[a] + gamma * u[a] * shv_shvcov / 3.)) / (tau_pi * gamma);
      #pragma unroll
      for (int b = 0; b < 2; ++b)
        M_0i[a][b] = (-tau_pi * MD[b] - delta_pipi * Mdd[b] + 2. * tau_pi * Mdo[b] * use_vort - tau_pipi * Mds[b]
                      + (2. * eta_pi + lambda_piPi * bulk) * Msig[b]) / (tau_pi * gamma);
    }
    #pragma unroll
    for (int a = 0; a < 2; ++a) {
      M_snu[a] = H[0][a + 1] - H[0][0] * uc[a] / gamma;
      #pragma unroll
      for (int b = 0; b < 2; ++b) F_snu += (H[a + 1][b + 1] - v[a] * H[0][b + 1]) * gu[a][b];
      F_snu += (H[a + 1][0] - v[a] * H[0][0]) * gu0[a];
    }
    F_snu += +H[3][3] * gamma / t;

    const double tilde_delta = delta_pipi - tau_pi;
    double Fdl[2][3], FDm[2][3], Fsm[2][3], Fdom[2][3], Fdsm[2][3], sls[2][3];
    double MDm[2][3][2], Msm[2][3][2], Mdl[2][3][2], Mdom[2][3][2], Mdsm[2][3][2];
    #pragma unroll
    for (int a = 0; a < 2; ++a) {
      #pragma unroll
      for (int b = 0; b < 3; ++b) {
        if (b < a) continue;
        double gugj = 0.0, sgj = 0.0, vsg = 0.0, sl_ = 0.0;
        #pragma unroll
        for (int k = 0; k < 2; ++k) {
          gugj += cg3[a][k] * H[b + 1][k + 1] + cg3[b][k] * H[a + 1][k + 1];
          sgj += S[k + 1][a + 1] * g3[k][b] + S[k + 1][b + 1] * g3[k][a];
          vsg += v[k] * S[0][a + 1] * g3[k][b] + v[k] * S[0][b + 1] * g3[k][a];
        }
        #pragma unroll
        for (int k = 0; k < 4; ++k) sl_ += (S[a + 1][k] * H[b + 1][k] + S[b + 1][k] * H[a + 1][k]) / 2.;
        sls[a][b] = sl_;
        FDm[a][b] = 0.0;
        Fsm[a][b] = cg3[a][b] / 2. + cg3[b][a] / 2. + fu[a] * fu[b] * expf / 3.;
        Fdl[a][b] = S[a + 1][b + 1] * expf;
        Fdom[a][b] = -(-gugj + sgj - vsg) / 4.;
        Fdsm[a][b] = -(-gugj - sgj + vsg - 4. * (fu[a] * fu[b]) * sgu / 3.) / 4.
                   + (4. * fu[a] * fu[b] * (gamma * H[3][3] / t) / 3.) / 4.
                   - S[a + 1][b + 1] * expf / 3.;
        #pragma unroll
        for (int k = 0; k < 2; ++k) {
          MDm[a][b][k] = gamma * (fu[a] * H[b + 1][k + 1] - fu[a] * H[b + 1][0] * fc[k] / gamma
                                  + fu[b] * H[a + 1][k + 1] - fu[b] * H[a + 1][0] * fc[k] / gamma);
          Msm[a][b][k] = -fu[a] * fu[b] * fc[k] / (3. * gamma);
          Mdl[a][b][k] = -S[a + 1][b + 1] * fc[k] / gamma;
          Mdom[a][b][k] = (-fu[a] * H[b + 1][k + 1] * gamma - fu[b] * H[a + 1][k + 1] * gamma
                           + fc[k] * fu[a] * H[b + 1][0] + fc[k] * fu[b] * H[a + 1][0]) / 4.;
          Mdsm[a][b][k] = (-fu[a] * H[b + 1][k + 1] * gamma - fu[b] * H[a + 1][k + 1] * gamma
                           + fc[k] * fu[a] * H[b + 1][0] + fc[k] * fu[b] * H[a + 1][0]
                           + 4. * fu[a] * fu[b] * (H[0][k + 1] - H[0][0] * fc[k] / gamma) / 3.
                           + 4. * S[a + 1][b + 1] * fc[k] / 3. / gamma) / 4.;
        }
        Mdsm[a][b][a] += S[0][b + 1] / 4.;
        Mdom[a][b][a] += -S[0][b + 1] / 4.;
        Msm[a][b][a] += -gamma * fu[b] / 2.;
      }
      #pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (b < a) continue;
        Msm[a][b][b] += -gamma * fu[a] / 2.;
        Mdsm[a][b][b] += S[0][a + 1] / 4.;
        Mdom[a][b][b] += -S[0][a + 1] / 4.;
      }
      #pragma unroll
      for (int k = 0; k < 2; ++k) {
        Msm[a][a][k] += cmd[a + 1] * fc[k] / (gamma * 3.);
        Mdsm[a][a][k] += -cmd[a + 1] * (H[0][k + 1] - H[0][0] * fc[k] / gamma) / 3.;
      }
      FDm[a][2] += gamma * S[3][a + 1] / t + 0.0 * S[0][a + 1] / t;
      Fsm[a][a] += -cmd[a + 1] * expf / 3.;
      Fsm[a][2] += -fu[a] * 0.0 * 0.0 * t / 2. - fu[a] * 0.0 * gamma / t;
      Fdsm[a][a] += -cmd[a + 1] * sgu / 3. - cmd[a + 1] * (gamma * H[3][3] / t) / 3.;
    }
    sgF[0] = Fsm[1][1]; sgF[1] = Fsm[1][2];
    sgM[0][0] = Msm[1][1][0]; sgM[0][1] = Msm[1][1][1];
    sgM[1][0] = Msm[1][2][0]; sgM[1][1] = Msm[1][2][1];
    #pragma unroll
    for (int a = 0; a < 2; ++a) {
      #pragma unroll
      for (int b = 0; b < 3; ++b) {
        if (b < a) continue;
        F_es[a][b] = (-S[a + 1][b + 1] - tilde_delta * Fdl[a][b] - tau_pi * FDm[a][b]
                      + 2. * tau_pi * Fdom[a][b] * use_vort - tau_pipi * Fdsm[a][b]
                      + (2. * eta_pi + lambda_piPi * bulk) * Fsm[a][b] + phi6 * bulk * S[a + 1][b + 1]
                      + phi7 * (sls[a][b] + fu[a] * fu[b] * shv_shvcov / 3.)) / (sigma * tau_pi * gamma);
        #pragma unroll
        for (int k = 0; k < 2; ++k)
          M_es[a][b][k] = (-tau_pi * MDm[a][b][k] - tilde_delta * Mdl[a][b][k] + 2. * tau_pi * Mdom[a][b][k] * use_vort
                           - tau_pipi * Mdsm[a][b][k] + (2. * eta_pi + lambda_piPi * bulk) * Msm[a][b][k])
                          / (sigma * tau_pi * gamma);
      }
      F_es[a][a] += phi7 * (-cmd[a + 1] * shv_shvcov) / (3. * tau_pi * gamma * sigma);
    }
  }

  // ---------------- fill_auxiliary_variables ----------------
  const double zeta_tilde = zeta + (delta_PiPi - tau_Pi) * bulk;
  double M_eb[2], M_S[2];
  #pragma unroll
  for (int a = 0; a < 2; ++a) {
    M_eb[a] = (zeta_tilde * uc[a] / gamma) / (sigma * gamma * tau_Pi);
    M_S[a] = (bulk * uc[a] / gamma) / (sigma * gamma * T);
  }
  const double R_S[3] = { -muB / T, -muS / T, -muQ / T };
  double F_eb = -(zeta_tilde * (gamma * divV + gamma / t) + bulk - phi1 * bulk * bulk) / (sigma * gamma * tau_Pi);
  double F_S = (-bulk * (gamma * divV + gamma / t)) / (sigma * gamma * T);
  if (SHEAR) {
    F_S += F_snu / (sigma * gamma * T);
    F_eb += -(lambda_Pipi * F_snu - phi3 * shv_shvcov) / (sigma * gamma * tau_Pi);
    #pragma unroll
    for (int a = 0; a < 2; ++a) {
      M_eb[a] += -lambda_Pipi * M_snu[a] / (sigma * gamma * tau_Pi);
      M_S[a] += M_snu[a] / (sigma * gamma * T);
    }
  }

  // ---------------- calculate_MRF_diffusion, compute_charge_derivative ----------------
  // Without diffusion R_N = gamma sigma I, F_N = 0 and M_N = 0 (no sources),
  // so the Schur terms R_u R_N^{-1} M_N and R_u R_N^{-1} F_N vanish.
  const double rho[3] = { rB, rS, rQ };
  DiffCtx dx;
  double RNI[3][3], F_N[3], M_N[2][3];
  if constexpr (DIFF) {
    diff_ctx(col, t2, dx);
    double F_q0[3], R_q0[3][3];
    diff_q0(dx, t, gamma, sigma, s, rho, u, divV, M_S, F_S, R_S, F_q0, R_q0, M_N);
    double R_N[3][3];
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
      #pragma unroll
      for (int k = 0; k < 3; ++k) R_N[c][k] = (c == k ? gamma * sigma : 0.0) + -R_q0[c][k];
      F_N[c] = 0.0 + (-F_q0[c] + col(CCAKE_F_GRAD_QA0 + c) - col(CCAKE_F_DIV_QA0 + c) - dx.q[c][0] / t);
    }
    #pragma unroll
    for (int d = 0; d < 2; ++d)
      #pragma unroll
      for (int c = 0; c < 3; ++c) M_N[d][c] = 0.0 + -M_N[d][c];
    inv3_det(R_N, RNI);
  }

  // ---------------- compute_velocity_derivative ----------------
  const double gP[2] = { -col(CCAKE_F_GRADP0 + 0), -col(CCAKE_F_GRADP0 + 1) };
  const double gB[2] = { -col(CCAKE_F_GRADBULK0 + 0), -col(CCAKE_F_GRADBULK0 + 1) };
  const double rdw = rB * dwdB + rS * dwdS + rQ * dwdQ;
  double M_w[2];
  #pragma unroll
  for (int a = 0; a < 2; ++a) M_w[a] = (s * dwds / gamma + rdw / gamma) * uc[a] / gamma + sigma * dwds * M_S[a];
  const double F_w = -(s * dwds / gamma + rdw / gamma) * (gamma / t + gamma * divV) + sigma * dwds * F_S;
  const double F_bulk = sigma * F_eb - bulk * (gamma * divV + gamma / t) / gamma;
  double Mu[2][2], Fu[2];
  #pragma unroll
  for (int a = 0; a < 2; ++a) {
    #pragma unroll
    for (int b = 0; b < 2; ++b) Mu[a][b] = u[a] * gamma * (M_eb[b] + M_w[b]) - u[a] * uc[b] * (w + bulk) / gamma;
    Mu[a][a] += gamma * (w + bulk);
    Fu[a] = (gP[a] + gB[a]) - gamma * (F_w + F_bulk) * u[a] - (w + bulk) * (gamma * divV + gamma / t) * u[a];
  }
  if (SHEAR) {
    const double ds_[2] = { col(CCAKE_F_DIVSHEAR0 + 0), col(CCAKE_F_DIVSHEAR0 + 1) };
    const double gs_[2] = { col(CCAKE_F_GRADSHEAR0 + 0), col(CCAKE_F_GRADSHEAR0 + 1) };
    #pragma unroll
    for (int a = 0; a < 2; ++a) {
      Fu[a] += -S[a + 1][0] / t - ds_[a] + gs_[a] - F_0i[a];
      Mu[a][0] += M_0i[a][0];
      Mu[a][1] += M_0i[a][1];
    }
  }
  if constexpr (DIFF) {
    // Schur complement of the charge block: M_u - R_u R_N^{-1} M_N, F_u + R_u R_N^{-1} F_N
    const double dwdr[3] = { dwdB, dwdS, dwdQ };
    double R_u[2][3];
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double R_w = sigma * dwdr[c] + sigma * dwds * R_S[c];
      #pragma unroll
      for (int a = 0; a < 2; ++a) R_u[a][c] = -u[a] * gamma * (R_w + 0.0 * sigma);
    }
    #pragma unroll
    for (int a = 0; a < 2; ++a) {
      double afr = 0.0;
      #pragma unroll
      for (int b = 0; b < 2; ++b) {
        double amr = 0.0;
        #pragma unroll
        for (int c = 0; c < 3; ++c)
          #pragma unroll
          for (int d = 0; d < 3; ++d) amr += R_u[a][c] * RNI[c][d] * M_N[b][d];
        Mu[a][b] = Mu[a][b] - amr;
      }
      #pragma unroll
      for (int c = 0; c < 3; ++c)
        #pragma unroll
        for (int d = 0; d < 3; ++d) afr += R_u[a][c] * RNI[c][d] * F_N[d];
      Fu[a] = Fu[a] + afr;
    }
  }
  const double det = Mu[0][0] * Mu[1][1] - Mu[0][1] * Mu[1][0];
  const double du0 = (Mu[1][1] / det) * Fu[0] + (-Mu[0][1] / det) * Fu[1];
  const double du1 = (-Mu[1][0] / det) * Fu[0] + (Mu[0][0] / det) * Fu[1];
  du_out[0] = du0; du_out[1] = du1;

  // ---------------- compute_derivatives ----------------
  dN[0] = 0.0; dN[1] = 0.0; dN[2] = 0.0;
  if constexpr (DIFF) {
    double rhs[3];
    #pragma unroll
    for (int c = 0; c < 3; ++c) rhs[c] = F_N[c] + ((0.0 + M_N[0][c] * du0) + M_N[1][c] * du1);
    #pragma unroll
    for (int c = 0; c < 3; ++c) dN[c] = ((0.0 + RNI[c][0] * rhs[0]) + RNI[c][1] * rhs[1]) + RNI[c][2] * rhs[2];
  }
  *d_s_out = (M_S[0] * du0 + M_S[1] * du1) + F_S + ((0.0 + R_S[0] * dN[0]) + R_S[1] * dN[1]) + R_S[2] * dN[2];
  *d_b_out = (M_eb[0] * du0 + M_eb[1] * du1) + F_eb;
  *snu_out = (M_snu[0] * du0 + M_snu[1] * du1) + F_snu;
  *theta_out = -1. * (uc[0] * du0 + uc[1] * du1) / gamma + gamma * divV + gamma / t;
  if (SHEAR) {
    int q = 0;
    #pragma unroll
    for (int a = 0; a < 2; ++a)
      #pragma unroll
      for (int b = 0; b < 3; ++b) {
        if (b < a) continue;
        d_shv_out[q++] = (M_es[a][b][0] * du0 + M_es[a][b][1] * du1) + F_es[a][b];
      }
  }
  // ---------------- compute_Ez_derivative ----------------
  const double p = col(CCAKE_F_P), e = col(CCAKE_F_E);
  if (SHEAR || DIFF) {
    const double s11 = ((0.0 + sgM[0][0] * du0) + sgM[0][1] * du1) + sgF[0];
    const double s12 = ((0.0 + sgM[1][0] * du0) + sgM[1][1] * du1) + sgF[1];
    hydro_numbers(S, s11, s12, u[0], u[1], gamma, t2, tau_pi, tau_Pi, *theta_out, bulk, p, hn_out);
  }
  if constexpr (DIFF) {
    diff_dext_q(dx, t, t2, gamma, sigma, s, rho, u, divV, M_S, F_S, R_S, du0, du1, dN, dq_out);
    // diffusion Knudsen numbers tau_q^(cc) |grad alpha_c| and inverse
    // Reynolds numbers |q_c| / |n_c| (compute_hydro_numbers, 1831-1995)
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double g0 = dx.galpha[0][c], g1 = dx.galpha[1][c];
      const double ga2 = (0.0 + g0 * g0) + g1 * g1;
      const double q_sq = dx.q[c][1] * dx.q[c][1] + dx.q[c][2] * dx.q[c][2] + t2 * dx.q[c][3] * dx.q[c][3]
                        - dx.q[c][0] * dx.q[c][0];
      const double q_mag = q_sq > 0.0 ? sqrt(q_sq) : 0.0;
      const double n_abs = fabs(rho[c]);
      hn_out[4 + c] = col(CCAKE_F_TAU_Q0 + 4 * c) * sqrt(ga2);
      hn_out[7 + c] = (n_abs > 1e-15) ? q_mag / n_abs : 0.0;
    }
  }
  *dez_out = ((e + p + bulk) * 0.0 * 0.0 + (p + bulk) / t2 + S[3][3]) * col(CCAKE_F_MASS_S) * t2 / col(CCAKE_F_SIGMA_LAB);
  (void)e;
}

// Evolver::set_current_timestep_quantities (src/evolver.cpp:124-168) for
// one particle: the evolved variables into the RK cache columns.
__device__ __forceinline__ void cache_state(const Cols& col, int diff)
{
  for (int q = 0; q < 5; ++q) col(P_C_SHV0 + q) = col(CCAKE_F_EXT_SHV0 + q);
  col(P_C_UX) = col(CCAKE_F_UX); col(P_C_UY) = col(CCAKE_F_UY);
  col(P_C_X) = col(CCAKE_F_X); col(P_C_Y) = col(CCAKE_F_Y);
  col(P_C_S) = col(CCAKE_F_EXT_S); col(P_C_BULK) = col(CCAKE_F_EXT_BULK);
  col(P_C_B) = col(CCAKE_F_EXT_B); col(P_C_SS) = col(CCAKE_F_EXT_SS); col(P_C_Q) = col(CCAKE_F_EXT_Q);
  col(P_C_EZ) = col(CCAKE_F_EZ);
  if (diff) for (int q = 0; q < 9; ++q) col(P_C_Q0 + q) = col(CCAKE_F_EXT_Q0 + q);
}

// mode 0: store derivatives; mode 1: store derivatives and apply the RK
// update y = y0 + dt f (src/evolver.cpp:363-484); returns the negative
// entropy flag through err.  With `cache` (the first stage of a step) the
// RK cache is taken from the current state first, as k_cache would at the
// start of the step: nothing in the stage before this kernel changes the
// evolved variables (the diffusion regulator does, so events with diffusion
// keep the separate k_cache).
// The gradient pass's sum folded in front (gs.perm set): thread k is sorted
// position k and works on particle perm[k]; otherwise thread i is particle i.
// With `next_cand` (cell-order walk only) the kernel also begins the next
// stage when that stage keeps the cell sort: gamma and v from the updated u
// and the smoothing record at the updated positions (k_stage_gather's work),
// and the stage's pair count cleared.
struct GradSum {
  const int32_t* perm = nullptr;
  const double* cand = nullptr;
  const double* part = nullptr;
  int nslices = 0;
  double C = 0.0;
  double* next_cand = nullptr;
  unsigned long long* pair_count = nullptr;
};

template <bool SHEAR, bool RK, bool DIFF>
__global__ __launch_bounds__(256) void k_eom(double* __restrict__ F, int npad, int n,
                                             const int32_t* __restrict__ freeze, double t, double use_vort,
                                             double dt, int32_t* __restrict__ err, int cache, GradSum gs,
                                             Live live)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int i = k;
  if (!gs.perm && live.retired(i)) return;      // an ended event of a batch keeps its state
  if (gs.perm) {
    i = gs.perm[k];
    if (gs.nslices == 1) grad_sum_into<SHEAR, 1>(gs.cand, k, i, gs.C, gs.part, 1, npad, F);
    else grad_sum_into<SHEAR, 0>(gs.cand, k, i, gs.C, gs.part, gs.nslices, npad, F);
  }
  const Cols col{ F, npad, i };
  double du[2], ds, dN[3], db, dshv[5] = { 0, 0, 0, 0, 0 }, theta, snu, dez, hn[10], dq[9];
  eom_d2<SHEAR, DIFF>(col, t, use_vort, du, &ds, dN, &db, dshv, &theta, &snu, &dez, hn, dq);
  col(CCAKE_F_DUDT0 + 0) = du[0]; col(CCAKE_F_DUDT0 + 1) = du[1]; col(CCAKE_F_DUDT0 + 2) = 0.0;
  col(CCAKE_F_DEXT_S) = ds;
  col(CCAKE_F_DEXT_B) = dN[0]; col(CCAKE_F_DEXT_SS) = dN[1]; col(CCAKE_F_DEXT_Q) = dN[2];
  col(CCAKE_F_DEXT_BULK) = db;
  if (SHEAR) for (int q = 0; q < 5; ++q) col(CCAKE_F_DEXT_SHV0 + q) = dshv[q];
  col(CCAKE_F_THETA) = theta;
  col(CCAKE_F_SHV_NABLA_U) = snu;
  col(CCAKE_F_DEZ) = dez;
  if (SHEAR || DIFF) {
    col(CCAKE_F_SHEAR_KNUDSEN) = hn[0]; col(CCAKE_F_BULK_KNUDSEN) = hn[1];
    col(CCAKE_F_INV_REYNOLDS_SHEAR) = hn[2]; col(CCAKE_F_INV_REYNOLDS_BULK) = hn[3];
  }
  if (DIFF) {
    for (int q = 0; q < 9; ++q) col(CCAKE_F_DEXT_Q0 + q) = dq[q];
    for (int c = 0; c < 3; ++c) { col(CCAKE_F_KN_DIFF0 + c) = hn[4 + c]; col(CCAKE_F_RE_DIFF0 + c) = hn[7 + c]; }
  }
  if (!RK) return;
  if (cache) cache_state(col, DIFF);
  double es = col(P_C_S) + dt * ds;
  if (es < 0.0 && freeze[i] > 3) es = 1.e-3;
  else if (es < 0.0) atomicOr(err, DE_NEG_ENTROPY);
  col(CCAKE_F_EXT_S) = es;
  col(CCAKE_F_EXT_BULK) = col(P_C_BULK) + dt * db;
  col(CCAKE_F_EXT_B) = col(P_C_B) + dt * dN[0];
  col(CCAKE_F_EXT_SS) = col(P_C_SS) + dt * dN[1];
  col(CCAKE_F_EXT_Q) = col(P_C_Q) + dt * dN[2];
  col(CCAKE_F_EZ) = col(P_C_EZ) + dt * dez;
  col(CCAKE_F_X) = col(P_C_X) + dt * col(CCAKE_F_VX);
  col(CCAKE_F_Y) = col(P_C_Y) + dt * col(CCAKE_F_VY);
  col(CCAKE_F_UX) = col(P_C_UX) + dt * du[0];
  col(CCAKE_F_UY) = col(P_C_UY) + dt * du[1];
  for (int q = 0; q < 5; ++q) col(CCAKE_F_EXT_SHV0 + q) = col(P_C_SHV0 + q) + dt * col(CCAKE_F_DEXT_SHV0 + q);
  if (DIFF) for (int q = 0; q < 9; ++q) col(CCAKE_F_EXT_Q0 + q) = col(P_C_Q0 + q) + dt * dq[q];
  col(CCAKE_F_ETA) = 0.0;
  col(CCAKE_F_UETA) = 0.0;
  if (gs.next_cand) {
    if (k == 0) gs.pair_count[0] = 0ull;
    const double ux = col(CCAKE_F_UX), uy = col(CCAKE_F_UY);
    const double c = (-ux) * ux + (-uy) * uy;
    const double g = sqrt(1.0 - c);
    col(CCAKE_F_GAMMA) = g;
    col(CCAKE_F_VX) = ux / g;
    col(CCAKE_F_VY) = uy / g;
    smooth_record(F, npad, i, gs.next_cand + (size_t)k * kCandS);
  }
}

// Stand-alone RK update from stored derivatives.
__global__ __launch_bounds__(256) void k_rk(double* __restrict__ F, int npad, int n,
                                            const int32_t* __restrict__ freeze, double dt, int32_t* __restrict__ err,
                                            int diff, Live live)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || live.retired(i)) return;
  const Cols col{ F, npad, i };
  double es = col(P_C_S) + dt * col(CCAKE_F_DEXT_S);
  if (es < 0.0 && freeze[i] > 3) es = 1.e-3;
  else if (es < 0.0) atomicOr(err, DE_NEG_ENTROPY);
  col(CCAKE_F_EXT_S) = es;
  col(CCAKE_F_EXT_BULK) = col(P_C_BULK) + dt * col(CCAKE_F_DEXT_BULK);
  col(CCAKE_F_EXT_B) = col(P_C_B) + dt * col(CCAKE_F_DEXT_B);
  col(CCAKE_F_EXT_SS) = col(P_C_SS) + dt * col(CCAKE_F_DEXT_SS);
  col(CCAKE_F_EXT_Q) = col(P_C_Q) + dt * col(CCAKE_F_DEXT_Q);
  col(CCAKE_F_EZ) = col(P_C_EZ) + dt * col(CCAKE_F_DEZ);
  col(CCAKE_F_X) = col(P_C_X) + dt * col(CCAKE_F_VX);
  col(CCAKE_F_Y) = col(P_C_Y) + dt * col(CCAKE_F_VY);
  col(CCAKE_F_UX) = col(P_C_UX) + dt * col(CCAKE_F_DUDT0 + 0);
  col(CCAKE_F_UY) = col(P_C_UY) + dt * col(CCAKE_F_DUDT0 + 1);
  for (int q = 0; q < 5; ++q) col(CCAKE_F_EXT_SHV0 + q) = col(P_C_SHV0 + q) + dt * col(CCAKE_F_DEXT_SHV0 + q);
  if (diff) for (int q = 0; q < 9; ++q) col(CCAKE_F_EXT_Q0 + q) = col(P_C_Q0 + q) + dt * col(CCAKE_F_DEXT_Q0 + q);
  col(CCAKE_F_ETA) = 0.0;
  col(CCAKE_F_UETA) = 0.0;
}

// Evolver::set_current_timestep_quantities (src/evolver.cpp:124-168)
__global__ __launch_bounds__(256) void k_cache(double* __restrict__ F, int npad, int n, int diff, Live live)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || live.retired(i)) return;
  cache_state(Cols{ F, npad, i }, diff);
}

// ---- the reference's fourth-order step (Evolver::update_k and update_rk4,
// src/evolver.cpp:171-341): the accumulator k has 24 columns in ev->k4,
// K4_* below.  update_k(n) adds w_n times the stage's derivatives
// (w = 1/6, 2/6, 2/6, 1/6 for n = 1..4; position from v); update_rk4 sets
// y = y0 + dt k from the RK cache, with the negative-entropy rule, and
// writes the extensive bulk pressure into the bulk field as well, as the
// reference does (318).
enum { K4_UX = 0, K4_UY, K4_X, K4_Y, K4_S, K4_B, K4_SS, K4_Q, K4_BULK, K4_EZ, K4_SHV0, K4_Q0 = K4_SHV0 + 5,
       K4_NUM = K4_Q0 + 9 };

__global__ __launch_bounds__(256) void k_rk4_update_k(double* __restrict__ F, int npad, int n, double* __restrict__ K,
                                                      double w, Live live)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || live.retired(i)) return;
  const Cols col{ F, npad, i };
  auto k = [&](int c) -> double& { return K[(size_t)c * npad + i]; };
  for (int q = 0; q < 5; ++q) k(K4_SHV0 + q) += w * col(CCAKE_F_DEXT_SHV0 + q);
  for (int q = 0; q < 9; ++q) k(K4_Q0 + q) += w * col(CCAKE_F_DEXT_Q0 + q);
  k(K4_UX) += w * col(CCAKE_F_DUDT0 + 0); k(K4_UY) += w * col(CCAKE_F_DUDT0 + 1);
  k(K4_X) += w * col(CCAKE_F_VX); k(K4_Y) += w * col(CCAKE_F_VY);
  k(K4_B) += w * col(CCAKE_F_DEXT_B); k(K4_SS) += w * col(CCAKE_F_DEXT_SS); k(K4_Q) += w * col(CCAKE_F_DEXT_Q);
  k(K4_S) += w * col(CCAKE_F_DEXT_S);
  k(K4_BULK) += w * col(CCAKE_F_DEXT_BULK);
  k(K4_EZ) += w * col(CCAKE_F_DEZ);
}

__global__ __launch_bounds__(256) void k_rk4_update(double* __restrict__ F, int npad, int n,
                                                    const double* __restrict__ K, const int32_t* __restrict__ freeze,
                                                    double dt, int32_t* __restrict__ err, Live live)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || live.retired(i)) return;
  const Cols col{ F, npad, i };
  auto k = [&](int c) { return K[(size_t)c * npad + i]; };
  double es = col(P_C_S) + dt * k(K4_S);
  if (es < 0.0 && freeze[i] > 3) es = 1.e-3;
  else if (es < 0.0) atomicOr(err, DE_NEG_ENTROPY);
  const double bulk = col(P_C_BULK) + dt * k(K4_BULK);
  col(CCAKE_F_EXT_S) = es;
  col(CCAKE_F_EXT_BULK) = bulk;
  col(CCAKE_F_EXT_B) = col(P_C_B) + dt * k(K4_B);
  col(CCAKE_F_EXT_SS) = col(P_C_SS) + dt * k(K4_SS);
  col(CCAKE_F_EXT_Q) = col(P_C_Q) + dt * k(K4_Q);
  col(CCAKE_F_BULK) = bulk;
  col(CCAKE_F_EZ) = col(P_C_EZ) + dt * k(K4_EZ);
  col(CCAKE_F_X) = col(P_C_X) + dt * k(K4_X);
  col(CCAKE_F_Y) = col(P_C_Y) + dt * k(K4_Y);
  col(CCAKE_F_UX) = col(P_C_UX) + dt * k(K4_UX);
  col(CCAKE_F_UY) = col(P_C_UY) + dt * k(K4_UY);
  for (int q = 0; q < 5; ++q) col(CCAKE_F_EXT_SHV0 + q) = col(P_C_SHV0 + q) + dt * k(K4_SHV0 + q);
  for (int q = 0; q < 9; ++q) col(CCAKE_F_EXT_Q0 + q) = col(P_C_Q0 + q) + dt * k(K4_Q0 + q);
  col(CCAKE_F_ETA) = 0.0;
  col(CCAKE_F_UETA) = 0.0;
}

int launch_rk4_update_k(ccake_event* ev, double w)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k_rk4_update_k, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n, ev->k4, w,
                     live_of(ev));
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch_rk4_update(ccake_event* ev, double dt)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k_rk4_update, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n, ev->k4,
                     ev->I + (size_t)CCAKE_I_FREEZE * ev->npad, dt, ev->err, live_of(ev));
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int rk4_columns() { return K4_NUM; }

int launch_eom(ccake_event* ev, int fuse_rk, double dt, double t0, bool fold_grad_sum, bool begin_next)
{
  (void)t0;
  if (ev->p.dim == 3) return launch3_eom(ev, fuse_rk, dt);
  GradSum gs;
  if (fold_grad_sum) {
    gs.perm = ev->perm;
    gs.cand = ev->cand_g;
    gs.part = ev->part;
    gs.nslices = ev->grad_sum_slices;
    gs.C = grad_sum_scale(ev->p.hT);
    if (begin_next && fuse_rk) {
      gs.next_cand = ev->cand_s;
      gs.pair_count = (unsigned long long*)ev->pair_count;
    }
  }
  // sorted positions of the live events with the folded gradient sums,
  // particle ids (ended events skipped by the kernel) without
  const int n = fold_grad_sum ? ev->n_live : ev->n;
  const Live live = live_of(ev);
  const dim3 g((n + 255) / 256), b(256);
  hipStream_t st = ev->stream;
  const int32_t* fz = ev->I + (size_t)CCAKE_I_FREEZE * ev->npad;
  const double uv = ev->p.using_vorticity ? 1.0 : 0.0;
  // fuse_rk: 0 derivatives only, 1 with the RK update, 2 with the RK cache
  // taken first and the update (the first stage of a step)
  const int cache = fuse_rk == 2 ? 1 : 0;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, g, b, 0, st, ev->F, ev->npad, n, fz, ev->t, uv, dt, ev->err, cache, gs, live);
  };
  const int key = (ev->p.using_shear ? 4 : 0) | (fuse_rk ? 2 : 0) | (ev->p.using_diffusion ? 1 : 0);
  switch (key) {
    case 0: go(k_eom<false, false, false>); break;
    case 1: go(k_eom<false, false, true>); break;
    case 2: go(k_eom<false, true, false>); break;
    case 3: go(k_eom<false, true, true>); break;
    case 4: go(k_eom<true, false, false>); break;
    case 5: go(k_eom<true, false, true>); break;
    case 6: go(k_eom<true, true, false>); break;
    default: go(k_eom<true, true, true>); break;
  }
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch_rk(ccake_event* ev, double dt, double t0)
{
  (void)t0;
  if (ev->p.dim == 3) return launch3_rk(ev, dt);
  const int n = ev->n;
  hipLaunchKernelGGL(k_rk, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n,
                     ev->I + (size_t)CCAKE_I_FREEZE * ev->npad, dt, ev->err, ev->p.using_diffusion, live_of(ev));
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch_cache(ccake_event* ev)
{
  if (ev->p.dim == 3) return launch3_cache(ev);
  const int n = ev->n;
  hipLaunchKernelGGL(k_cache, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n,
                     ev->p.using_diffusion, live_of(ev));
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

}  // namespace ccake
