// Equation of state on the device (and host): forward fields of the chain
// members, Newton inversion, and the thermodynamic derivatives.
//   conformal            include/eos_conformal.h:34-195
//   conformal_diagonal   include/eos_conformal_diagonal.h:39-204
//   tanh-conformal       include/eos_tanh_conformal.h:22-304
//   table (multilinear)  src/interpolatorND.cpp:250-392
//   inversion cascade    src/eos.cpp:629-975, src/rootfinder.cpp:421-715
//   set_thermo           include/eos.h:121-175, src/eos.cpp:276-492,
//                        src/eos_derivatives.cpp:16-555
// The GSL `hybrids` solver of the reference is replaced by a damped Newton
// iteration on the same residual with the same absolute stopping rule
// (sum |F_i| < 1e-10), followed by one polishing step.
#pragma once
#include "ccake_internal.h"
#include <cmath>

namespace ccake {

#define CK_HD __host__ __device__ __forceinline__

constexpr double kTiny = 1e-25;        // include/eos_header.h:16
constexpr double kRfTol = 1e-10;       // include/rootfinder.h:37
constexpr int kNewtonSteps = 60;

enum { M_TABLE = 0, M_TANH = 1, M_CONF = 2, M_DIAG = 3 };
// Member-kind masks (bit k: kind k may occur).  The solver templates take the
// mask of the chain they run on, so a chain without a table compiles without
// the table's gathers and registers; the arithmetic is the same either way.
constexpr int kAllMembers = 15;
constexpr int kAnalyticMembers = (1 << M_TANH) | (1 << M_CONF) | (1 << M_DIAG);
template <int KM, int K> constexpr bool has_kind() { return (KM >> K) & 1; }
// 17-field order (src/eos_table.cpp:30-34)
enum { X_P = 0, X_S, X_B, X_SS, X_Q, X_E, X_CS2, X_BB, X_QQ, X_SSS, X_BQ, X_BS, X_QS, X_TB, X_TQ, X_TS, X_TT };

CK_HD double sgnd(double v) { return (double)((0.0 < v) - (v < 0.0)); }

CK_HD void conf_fields(const EosMember& m, const double* x, double* r)
{
  const double T = x[0], mB = x[1], mQ = x[2], mS = x[3];
  const double T0 = m.T0, B0 = m.muB0, S0 = m.muS0, Q0 = m.muQ0;
  const double X = (T / T0) * (T / T0) + (mB / B0) * (mB / B0) + (mS / S0) * (mS / S0) + (mQ / Q0) * (mQ / Q0);
  const double cp = m.c * T0 * T0 * T0 * T0;
  const double p = cp * X * X;
  const double s = 4.0 * cp * T * X / (T0 * T0);
  const double B = 4.0 * cp * mB * X / (B0 * B0);
  const double S = 4.0 * cp * mS * X / (S0 * S0);
  const double Q = 4.0 * cp * mQ * X / (Q0 * Q0);
  r[X_P] = p; r[X_S] = s; r[X_B] = B; r[X_SS] = S; r[X_Q] = Q;
  r[X_E] = (s * T + mB * B + mQ * Q + mS * S) - p;
  r[X_CS2] = 1.0 / 3.0;
  r[X_BB] = cp * (8.0 * mB * mB / (B0 * B0 * B0 * B0) + 4.0 * X / (B0 * B0));
  r[X_QQ] = cp * (8.0 * mQ * mQ / (Q0 * Q0 * Q0 * Q0) + 4.0 * X / (Q0 * Q0));
  r[X_SSS] = cp * (8.0 * mS * mS / (S0 * S0 * S0 * S0) + 4.0 * X / (S0 * S0));
  r[X_BQ] = 8.0 * cp * mB * mQ / (B0 * B0 * Q0 * Q0);
  r[X_BS] = 8.0 * cp * mB * mS / (B0 * B0 * S0 * S0);
  r[X_QS] = 8.0 * cp * mS * mQ / (S0 * S0 * Q0 * Q0);
  r[X_TB] = 8.0 * cp * T * mB / (T0 * T0 * B0 * B0);
  r[X_TQ] = 8.0 * cp * T * mQ / (T0 * T0 * Q0 * Q0);
  r[X_TS] = 8.0 * cp * T * mS / (T0 * T0 * S0 * S0);
  r[X_TT] = cp * (8.0 * T * T / (T0 * T0 * T0 * T0) + 4.0 * X / (T0 * T0));
}

CK_HD void diag_fields(const EosMember& m, const double* x, double* r)
{
  const double T = x[0], mB = x[1], mQ = x[2], mS = x[3];
  const double T0 = m.T0, B0 = m.muB0, S0 = m.muS0, Q0 = m.muQ0;
  const double cp = m.c * T0 * T0 * T0 * T0;
  const double a = T / T0, b = mB / B0, s_ = mS / S0, q = mQ / Q0;
  r[X_P] = cp * (a * a * a * a + b * b * b * b + s_ * s_ * s_ * s_ + q * q * q * q);
  r[X_S] = 4.0 * cp * T * T * T / (T0 * T0 * T0 * T0);
  r[X_B] = 4.0 * cp * mB * mB * mB / (B0 * B0 * B0 * B0);
  r[X_SS] = 4.0 * cp * mS * mS * mS / (S0 * S0 * S0 * S0);
  r[X_Q] = 4.0 * cp * mQ * mQ * mQ / (Q0 * Q0 * Q0 * Q0);
  r[X_E] = (r[X_S] * T + mB * r[X_B] + mQ * r[X_Q] + mS * r[X_SS]) - r[X_P];
  r[X_CS2] = 1.0 / 3.0;
  r[X_BB] = 12.0 * cp * mB * mB / (B0 * B0 * B0 * B0);
  r[X_QQ] = 12.0 * cp * mQ * mQ / (Q0 * Q0 * Q0 * Q0);
  r[X_SSS] = 12.0 * cp * mS * mS / (S0 * S0 * S0 * S0);
  r[X_BQ] = r[X_BS] = r[X_QS] = r[X_TB] = r[X_TQ] = r[X_TS] = kTiny;
  r[X_TT] = 12.0 * cp * T * T / (T0 * T0 * T0 * T0);
}

// Swap rows c and pv (pv >= c) of a small matrix with every index known at
// compile time after unrolling: the pivot row is selected, not addressed,
// so the matrix stays in registers (an array indexed by a run-time row
// would live in scratch memory).  Same values as the indexed swap.
template <int R, int C>
CK_HD void swap_rows_sel(double (&A)[R][C], int c, int pv)
{
  #pragma unroll
  for (int j = 0; j < C; ++j) {
    double top = A[c][j];
    #pragma unroll
    for (int i = 0; i < R; ++i)
      if (i > c) top = (pv == i) ? A[i][j] : top;
    #pragma unroll
    for (int i = 0; i < R; ++i)
      if (i > c) A[i][j] = (pv == i) ? A[c][j] : A[i][j];
    A[c][j] = top;
  }
}

// 4x4 linear solve with partial pivoting; false if singular.
CK_HD bool solve_4x4(const double* J, const double* b, double* x)
{
  double A[4][5];
  #pragma unroll
  for (int i = 0; i < 4; ++i) {
    #pragma unroll
    for (int j = 0; j < 4; ++j) A[i][j] = J[4 * i + j];
    A[i][4] = b[i];
  }
  #pragma unroll
  for (int c = 0; c < 4; ++c) {
    int pv = c;
    double best = fabs(A[c][c]);
    #pragma unroll
    for (int i = c + 1; i < 4; ++i) if (fabs(A[i][c]) > best) { pv = i; best = fabs(A[i][c]); }
    swap_rows_sel(A, c, pv);
    if (A[c][c] == 0.0 || !__builtin_isfinite(A[c][c])) return false;
    #pragma unroll
    for (int i = c + 1; i < 4; ++i) {
      const double f = A[i][c] / A[c][c];
      for (int j = c; j < 5; ++j) A[i][j] -= f * A[c][j];
    }
  }
  for (int i = 3; i >= 0; --i) {
    double v = A[i][4];
    for (int j = i + 1; j < 4; ++j) v -= A[i][j] * x[j];
    x[i] = v / A[i][i];
  }
  for (int i = 0; i < 4; ++i) if (!__builtin_isfinite(x[i])) return false;
  return true;
}

// (dp/de) along fixed n_X/s, from the Hessian of p over (T, muB, muQ, muS).
CK_HD double cs2_adiabatic(const double* x, const double* r)
{
  const double H[16] = { r[X_TT], r[X_TB], r[X_TQ], r[X_TS],
                         r[X_TB], r[X_BB], r[X_BQ], r[X_BS],
                         r[X_TQ], r[X_BQ], r[X_QQ], r[X_QS],
                         r[X_TS], r[X_BS], r[X_QS], r[X_SSS] };
  const double s = r[X_S];
  const double rhs[4] = { 1.0, r[X_B] / s, r[X_Q] / s, r[X_SS] / s };
  double d[4];
  if (!solve_4x4(H, rhs, d)) return 1.0 / 3.0;
  const double dp = r[X_S] * d[0] + r[X_B] * d[1] + r[X_Q] * d[2] + r[X_SS] * d[3];
  const double de = x[0] * 1.0 + x[1] * rhs[1] + x[2] * rhs[2] + x[3] * rhs[3];
  return dp / de;
}

CK_HD void tanh_fields(const EosMember& m, const double* x, double* r)
{
  double c[17];
  conf_fields(m, x, c);
  const double T = x[0], mB = x[1], mQ = x[2], mS = x[3];
  const double arg = (T - m.Tc) / m.Ts;
  const double th = tanh(arg);
  const double sh = 1.0 / cosh(arg);
  const double tf = 0.5 * (1.0 + th);
  const double tp = 0.5 * sh * sh / m.Ts;
  const double tpp = -2.0 * th * tp / m.Ts;
  r[X_P] = tf * c[X_P];
  r[X_S] = tf * c[X_S] + tp * c[X_P];
  r[X_B] = tf * c[X_B]; r[X_SS] = tf * c[X_SS]; r[X_Q] = tf * c[X_Q];
  r[X_E] = (r[X_S] * T + mB * r[X_B] + mQ * r[X_Q] + mS * r[X_SS]) - r[X_P];
  r[X_BB] = tf * c[X_BB]; r[X_QQ] = tf * c[X_QQ]; r[X_SSS] = tf * c[X_SSS];
  r[X_BQ] = tf * c[X_BQ]; r[X_BS] = tf * c[X_BS]; r[X_QS] = tf * c[X_QS];
  r[X_TB] = tf * c[X_TB] + tp * c[X_B];
  r[X_TQ] = tf * c[X_TQ] + tp * c[X_Q];
  r[X_TS] = tf * c[X_TS] + tp * c[X_SS];
  r[X_TT] = tf * c[X_TT] + 2.0 * tp * c[X_S] + tpp * c[X_P];
  r[X_CS2] = cs2_adiabatic(x, r);
}

// Multilinear cell location (src/interpolatorND.cpp:265-285).
CK_HD bool table_cell(const EosMember& m, const double* x, int* ind, double* fr)
{
  for (int a = 0; a < 4; ++a) {
    if (m.dims[a] <= 1 || fabs(m.gdel[a]) < 1e-30) { ind[a] = 0; fr[a] = 1.0; continue; }
    const double u = (x[a] - m.gmin[a]) / m.gdel[a];
    const double ip = trunc(u);
    fr[a] = 1.0 - (u - ip);
    ind[a] = (int)ip;
    if (ind[a] + 1 == m.dims[a]) { fr[a] = 1.0 - fr[a]; ind[a]--; }
    if (ind[a] < 0 || ind[a] + 1 >= m.dims[a]) return false;
  }
  return true;
}

// One table field at x: the same weights and the same accumulation order as
// table_fields, so the value is the one table_fields gives; NaN outside.
CK_HD double table_field_one(const EosMember& m, const double* tab, const double* x, int field)
{
  int ind[4]; double fr[4];
  if (__builtin_isnan(x[0]) || __builtin_isnan(x[1]) || __builtin_isnan(x[2]) || __builtin_isnan(x[3]) || !table_cell(m, x, ind, fr))
    return NAN;
  double r = 0.0;
  for (int h = 0; h < 16; ++h) {
    const int b0 = (h >> 3) & 1, b1 = (h >> 2) & 1, b2 = (h >> 1) & 1, b3 = h & 1;
    double w = 1.0;
    w *= fr[0] + b0 - 2.0 * b0 * fr[0];
    w *= fr[1] + b1 - 2.0 * b1 * fr[1];
    w *= fr[2] + b2 - 2.0 * b2 * fr[2];
    w *= fr[3] + b3 - 2.0 * b3 * fr[3];
    const size_t node = (((size_t)(ind[0] + b0) * m.dims[1] + (ind[1] + b1)) * m.dims[2] + (ind[2] + b2)) * m.dims[3]
                        + (ind[3] + b3);
    r += w * tab[node * 17 + field];
  }
  return r;
}

// Table fields at x: all 17 (tab, stride 17) into r; NaN outside.
CK_HD void table_fields(const EosMember& m, const double* tab, const double* x, double* r)
{
  int ind[4]; double fr[4];
  if (__builtin_isnan(x[0]) || __builtin_isnan(x[1]) || __builtin_isnan(x[2]) || __builtin_isnan(x[3]) || !table_cell(m, x, ind, fr)) {
    for (int k = 0; k < 17; ++k) r[k] = NAN;
    return;
  }
  for (int k = 0; k < 17; ++k) r[k] = 0.0;
  for (int h = 0; h < 16; ++h) {
    const int b0 = (h >> 3) & 1, b1 = (h >> 2) & 1, b2 = (h >> 1) & 1, b3 = h & 1;
    double w = 1.0;
    w *= fr[0] + b0 - 2.0 * b0 * fr[0];
    w *= fr[1] + b1 - 2.0 * b1 * fr[1];
    w *= fr[2] + b2 - 2.0 * b2 * fr[2];
    w *= fr[3] + b3 - 2.0 * b3 * fr[3];
    const size_t node = (((size_t)(ind[0] + b0) * m.dims[1] + (ind[1] + b1)) * m.dims[2] + (ind[2] + b2)) * m.dims[3]
                        + (ind[3] + b3);
    const double* c = tab + node * 17;
    for (int k = 0; k < 17; ++k) r[k] += w * c[k];
  }
}

template <int KM = kAllMembers>
CK_HD void member_fields(const EosMember& m, const double* tab, const double* x, double* r)
{
  if (has_kind<KM, M_CONF>() && m.kind == M_CONF) conf_fields(m, x, r);
  else if (has_kind<KM, M_DIAG>() && m.kind == M_DIAG) diag_fields(m, x, r);
  else if (has_kind<KM, M_TANH>() && m.kind == M_TANH) tanh_fields(m, x, r);
  else if (has_kind<KM, M_TABLE>()) table_fields(m, tab, x, r);
  else for (int k = 0; k < 17; ++k) r[k] = NAN;          // a kind outside the mask: never dispatched
}

// Residual densities (s|e, B, S, Q) at x and their Jacobian in (T, muB, muQ, muS).
// Table: packed tab5 (s, B, S, Q, e per node) and the derivative of the
// multilinear interpolant; analytic members: Hessian of p.
template <int KM = kAllMembers>
CK_HD void member_dens(const EosMember& m, const double* tab5, int mode, const double* x, double* F, double* J)
{
  if (has_kind<KM, M_TABLE>() && m.kind == M_TABLE) {
    int ind[4]; double fr[4];
    if (__builtin_isnan(x[0]) || __builtin_isnan(x[1]) || __builtin_isnan(x[2]) || __builtin_isnan(x[3]) || !table_cell(m, x, ind, fr)) {
      for (int k = 0; k < 4; ++k) F[k] = NAN;
      for (int k = 0; k < 16; ++k) J[k] = NAN;
      return;
    }
    for (int k = 0; k < 4; ++k) F[k] = 0.0;
    for (int k = 0; k < 16; ++k) J[k] = 0.0;
    double ga[4];
    for (int a = 0; a < 4; ++a) ga[a] = (m.dims[a] <= 1 || fabs(m.gdel[a]) < 1e-30) ? 0.0 : 1.0 / m.gdel[a];
    for (int h = 0; h < 16; ++h) {
      const int bt[4] = { (h >> 3) & 1, (h >> 2) & 1, (h >> 1) & 1, h & 1 };
      double wa[4];
      double w = 1.0;
      for (int a = 0; a < 4; ++a) { wa[a] = fr[a] + bt[a] - 2.0 * bt[a] * fr[a]; w *= wa[a]; }
      const size_t node = (((size_t)(ind[0] + bt[0]) * m.dims[1] + (ind[1] + bt[1])) * m.dims[2] + (ind[2] + bt[2]))
                          * m.dims[3] + (ind[3] + bt[3]);
      const double* c = tab5 + node * 5;
      const double v[4] = { mode ? c[4] : c[0], c[1], c[2], c[3] };
      for (int k = 0; k < 4; ++k) F[k] += w * v[k];
      for (int a = 0; a < 4; ++a) {
        if (ga[a] == 0.0) continue;
        double dw = bt[a] ? ga[a] : -ga[a];
        for (int b = 0; b < 4; ++b) if (b != a) dw *= wa[b];
        for (int k = 0; k < 4; ++k) J[4 * k + a] += dw * v[k];
      }
    }
    return;
  }
  double r[17];
  member_fields<KM & kAnalyticMembers>(m, nullptr, x, r);
  const double T = x[0], mB = x[1], mQ = x[2], mS = x[3];
  const double Js[4] = { r[X_TT], r[X_TB], r[X_TQ], r[X_TS] };
  const double JB[4] = { r[X_TB], r[X_BB], r[X_BQ], r[X_BS] };
  const double JS[4] = { r[X_TS], r[X_BS], r[X_QS], r[X_SSS] };
  const double JQ[4] = { r[X_TQ], r[X_BQ], r[X_QQ], r[X_QS] };
  F[1] = r[X_B]; F[2] = r[X_SS]; F[3] = r[X_Q];
  if (mode == 0) {
    F[0] = r[X_S];
    for (int a = 0; a < 4; ++a) J[a] = Js[a];
  } else {
    F[0] = r[X_E];
    for (int a = 0; a < 4; ++a) J[a] = T * Js[a] + mB * JB[a] + mQ * JQ[a] + mS * JS[a];
  }
  for (int a = 0; a < 4; ++a) { J[4 + a] = JB[a]; J[8 + a] = JS[a]; J[12 + a] = JQ[a]; }
}

CK_HD bool in_range(const EosMember& m, const double* x)
{
  for (int a = 0; a < 4; ++a) if (!(x[a] >= m.mins[a] && x[a] <= m.maxs[a])) return false;
  return true;
}

// A target the member cannot reach: the multilinear interpolant lies within
// its node values, so a density more than the stopping tolerance outside
// their range leaves |F| >= kRfTol everywhere in the table, and every
// attempt on the member fails.  Skipping the member's attempts then gives
// invert_point's result.
CK_HD bool unreachable(const EosMember& m, int mode, const double* tgt)
{
  if (!(m.dmin[0] <= m.dmax[0])) return false;
  const double v[4] = { tgt[0], tgt[1], tgt[2], tgt[3] };
  const int idx[4] = { mode ? 4 : 0, 1, 2, 3 };    // packed (s, B, S, Q, e)
  for (int a = 0; a < 4; ++a)
    if (v[a] < m.dmin[idx[a]] - kRfTol || v[a] > m.dmax[idx[a]] + kRfTol) return true;
  return false;
}

CK_HD double residual(const double* F, const double* tgt, double* R)
{
  double s = 0.0;
  for (int a = 0; a < 4; ++a) { R[a] = F[a] - tgt[a]; s += fabs(R[a]); }
  return s;
}

// One rootfinder attempt from x (src/rootfinder.cpp:206-416): damped Newton
// that never leaves the member's range; one polishing step after the
// absolute residual test passes.  With budget < kNewtonSteps the attempt
// stops undecided (-1) after `budget` iterations; otherwise it returns 1
// (converged, x the root) or 0.  A decided outcome within the budget is
// the outcome of the full attempt: the iterations are the same.
template <int KM = kAllMembers>
CK_HD int newton_attempt_b(const EosMember& m, const double* tab5, int mode, const double* tgt, double* x,
                           int budget)
{
  double F[4], J[16], R[4], dx[4], xn[4], Fn[4], Jn[16], Rn[4];
  if (!in_range(m, x)) return 0;
  member_dens<KM>(m, tab5, mode, x, F, J);
  double res = residual(F, tgt, R);
  for (int it = 0; it < budget; ++it) {
    if (!__builtin_isfinite(res)) return 0;
    const bool conv = res < kRfTol;
    const double nr[4] = { -R[0], -R[1], -R[2], -R[3] };
    if (!solve_4x4(J, nr, dx)) return conv ? 1 : 0;
    double lam = 1.0;
    bool acc = false;
    for (int k = 0; k < 40; ++k) {
      for (int a = 0; a < 4; ++a) xn[a] = x[a] + lam * dx[a];
      if (in_range(m, xn)) {
        member_dens<KM>(m, tab5, mode, xn, Fn, Jn);
        const double rn = residual(Fn, tgt, Rn);
        if (__builtin_isfinite(rn) && (rn < res || (conv && rn <= res))) {
          for (int a = 0; a < 4; ++a) { x[a] = xn[a]; R[a] = Rn[a]; }
          for (int a = 0; a < 16; ++a) J[a] = Jn[a];
          res = rn;
          acc = true;
          break;
        }
      }
      if (conv) break;
      lam *= 0.5;
    }
    if (conv) return 1;
    if (!acc) return 0;
  }
  if (budget < kNewtonSteps) return -1;
  return res < kRfTol ? 1 : 0;
}

template <int KM = kAllMembers>
CK_HD bool newton_attempt(const EosMember& m, const double* tab5, int mode, const double* tgt, double* x)
{
  return newton_attempt_b<KM>(m, tab5, mode, tgt, x, kNewtonSteps) == 1;
}

// Attempt `k` (0..9) of Rootfinder::find_root from the seed x
// (src/rootfinder.cpp:421-715): the seed adjusted away from zero, then
// attempt 0 as it is and attempts 1-8 with one coordinate moved by +-20 %
// (clamped inside the member's range), attempt 9 with the chemical
// potentials reset.  tgt = (s|e, B, S, Q); x = (T, muB, muQ, muS).
CK_HD void root_seed(const EosMember& m, const double* tgt, const double* x, int k, double* y)
{
  double sd[4] = { x[0], x[1], x[2], x[3] };
  const double sg[4] = { 1.0, sgnd(tgt[1]), sgnd(tgt[3]), sgnd(tgt[2]) };
  for (int a = 0; a < 4; ++a)
    if (sd[a] * sd[a] < kRfTol) sd[a] = fmin(0.5 * m.maxs[a], 1.0) * sg[a];
  const double t0 = sd[0], b0 = sd[1], q0 = sd[2], s0 = sd[3];
  y[0] = t0; y[1] = b0; y[2] = q0; y[3] = s0;
  switch (k) {
    case 1: y[0] = (t0 + 0.2 * t0 > m.maxs[0]) ? m.maxs[0] - 1 : t0 + 0.2 * t0; break;
    case 2: y[0] = (t0 - 0.2 * t0 < m.mins[0]) ? m.mins[0] + 1 : t0 - 0.2 * t0; break;
    case 3: y[1] = (b0 + 0.2 * b0 > m.maxs[1]) ? m.maxs[1] - 1 : b0 + 0.2 * b0; break;
    case 4: y[1] = (b0 - 0.2 * b0 < m.mins[1]) ? m.mins[1] + 1 : b0 - 0.2 * b0; break;
    case 5: y[3] = (s0 + 0.2 * s0 > m.maxs[3]) ? m.maxs[3] - 1 : s0 + 0.2 * s0; break;
    case 6: y[3] = (s0 - 0.2 * s0 < m.mins[3]) ? m.mins[3] + 1 : s0 - 0.2 * s0; break;
    case 7: y[2] = (q0 + 0.2 * q0 > m.maxs[2]) ? m.maxs[2] - 1 : q0 + 0.2 * q0; break;
    case 8: y[2] = (q0 - 0.2 * q0 < m.mins[2]) ? m.mins[2] + 1 : q0 - 0.2 * q0; break;
    case 9:
      y[1] = fmin(0.5 * m.maxs[1], 1.0) * sg[1];
      y[2] = fmin(0.5 * m.maxs[2], 1.0) * sg[2];
      y[3] = fmin(0.5 * m.maxs[3], 1.0) * sg[3];
      break;
    default: break;
  }
}

constexpr int kRootAttempts = 10;

// Rootfinder::find_root: the attempts in order, the first that converges.
template <int KM = kAllMembers>
CK_HD bool find_root(const EosMember& m, const double* tab5, int mode, const double* tgt, double* x)
{
  for (int k = 0; k < kRootAttempts; ++k) {
    double y[4];
    root_seed(m, tgt, x, k, y);
    if (newton_attempt<KM>(m, tab5, mode, tgt, y)) { for (int a = 0; a < 4; ++a) x[a] = y[a]; return true; }
  }
  return false;
}

// EoS_conformal_diagonal::get_tbqs_seed_from_{s,e}BSQ (include/eos_conformal_diagonal.h:39-80)
CK_HD void diag_seed(const EosMember& d, int mode, const double* tgt, double* x)
{
  const double two23 = pow(2.0, 2.0 / 3.0), c13 = pow(d.c, 1.0 / 3.0);
  const double den = two23 * c13;
  const double B = tgt[1], S = tgt[2], Q = tgt[3];
  const double B0 = d.muB0, S0 = d.muS0, Q0 = d.muQ0, T0 = d.T0;
  if (mode == 0) {
    x[0] = pow(T0 * T0 * T0 * T0 * tgt[0], 1.0 / 3.0) / den;
  } else {
    const double disc = tgt[0] - 3.0 * (pow(B0 * fabs(B), 4.0 / 3.0) + pow(S0 * fabs(S), 4.0 / 3.0)
                                        + pow(Q0 * fabs(Q), 4.0 / 3.0)) / (4.0 * two23 * c13);
    x[0] = T0 * pow(disc / (3.0 * d.c), 0.25);
  }
  x[1] = sgnd(B) * pow(B0 * B0 * B0 * B0 * fabs(B), 1.0 / 3.0) / den;
  x[2] = sgnd(Q) * pow(Q0 * Q0 * Q0 * Q0 * fabs(Q), 1.0 / 3.0) / den;
  x[3] = sgnd(S) * pow(S0 * S0 * S0 * S0 * fabs(S), 1.0 / 3.0) / den;
}

// ---- Schur complements of the susceptibility matrix -----------------------
// chi_lead - a^T M^{-1} b; protective clamps on M(1,1), M(2,2) at |.| < 1e-3.
CK_HD double schur3(double lead, double a0, double a1, double a2,
                    double m00, double m01, double m02, double m10, double m11, double m12,
                    double m20, double m21, double m22, double b0, double b1, double b2)
{
  const double EPS = 1e-3;
  if (fabs(m11) < EPS) m11 = (m11 >= 0.0 ? EPS : -EPS);
  if (fabs(m22) < EPS) m22 = (m22 >= 0.0 ? EPS : -EPS);
  double A[3][4] = { { m00, m01, m02, b0 }, { m10, m11, m12, b1 }, { m20, m21, m22, b2 } };
  #pragma unroll
  for (int c = 0; c < 3; ++c) {
    int pv = c;
    double best = fabs(A[c][c]);
    #pragma unroll
    for (int i = c + 1; i < 3; ++i) if (fabs(A[i][c]) > best) { pv = i; best = fabs(A[i][c]); }
    swap_rows_sel(A, c, pv);
    #pragma unroll
    for (int i = c + 1; i < 3; ++i) {
      const double f = A[i][c] / A[c][c];
      for (int j = c; j < 4; ++j) A[i][j] -= f * A[c][j];
    }
  }
  const double y2 = A[2][3] / A[2][2];
  const double y1 = (A[1][3] - A[1][2] * y2) / A[1][1];
  const double y0 = (A[0][3] - A[0][1] * y1 - A[0][2] * y2) / A[0][0];
  return lead - (a0 * y0 + a1 * y1 + a2 * y2);
}

// The 31 thermodynamic_info doubles at x for member `mi` of the chain.
template <int KM = kAllMembers>
CK_HD void set_thermo(const EosChain& ch, const double* tab, int mi, const double* x, double* th)
{
  const EosMember& m = ch.m[mi];
  double r[17];
  member_fields<KM>(m, tab, x, r);
  const double T = x[0], mB = x[1], mQ = x[2], mS = x[3];
  const double s = r[X_S], B = r[X_B], S = r[X_SS], Q = r[X_Q];
  const double tt = r[X_TT], tb = r[X_TB], tq = r[X_TQ], ts = r[X_TS];
  const double bb = r[X_BB], qq = r[X_QQ], ss = r[X_SSS], bq = r[X_BQ], bs = r[X_BS], qs = r[X_QS];
  const bool hB = fabs(B) > kTiny, hS = fabs(S) > kTiny, hQ = fabs(Q) > kTiny;
  // calc_term_1..4 (src/eos_derivatives.cpp)
  const double dsdt = schur3(tt, tb, ts, tq, bb, bs, bq, bs, ss, qs, bq, qs, qq, tb, ts, tq);
  auto t2B = [&] { return schur3(tb, tt, ts, tq, tb, bs, bq, ts, ss, qs, tq, qs, qq, bb, bs, bq); };
  auto t2S = [&] { return schur3(ts, tt, tb, tq, tb, bb, bq, ts, bs, qs, tq, bq, qq, bs, ss, qs); };
  auto t2Q = [&] { return schur3(tq, tt, tb, ts, tb, bb, bs, ts, bs, ss, tq, bq, qs, bq, qs, qq); };
  auto t3B = [&] { return schur3(tb, bb, bs, bq, tb, ts, tq, bs, ss, qs, bq, qs, qq, tt, ts, tq); };
  auto t3S = [&] { return schur3(ts, bs, ss, qs, tb, ts, tq, bb, bs, bq, bq, qs, qq, tt, tb, tq); };
  auto t3Q = [&] { return schur3(tq, bq, qs, qq, tb, ts, tq, bb, bs, bq, bs, ss, qs, tt, tb, ts); };
  auto kBB = [&] { return schur3(bb, tb, bs, bq, tt, ts, tq, ts, ss, qs, tq, qs, qq, tb, bs, bq); };
  auto kSB = [&] { return schur3(bs, ts, ss, qs, tt, ts, tq, tb, bs, bq, tq, qs, qq, tb, bb, bq); };
  auto kQB = [&] { return schur3(bq, tq, qs, qq, tt, ts, tq, tb, bs, bq, ts, ss, qs, tb, bb, bs); };
  auto kBS = [&] { return schur3(bs, tb, bb, bq, tt, tb, tq, ts, bs, qs, tq, bq, qq, ts, ss, qs); };
  auto kSS = [&] { return schur3(ss, ts, bs, qs, tt, tb, tq, tb, bb, bq, tq, bq, qq, ts, bs, qs); };
  auto kQS = [&] { return schur3(qs, tq, bq, qq, tt, tb, tq, tb, bb, bq, ts, bs, qs, ts, bs, ss); };
  auto kBQ = [&] { return schur3(bq, tb, bb, bs, tt, tb, ts, ts, bs, ss, tq, bq, qs, tq, qs, qq); };
  auto kSQ = [&] { return schur3(qs, ts, bs, ss, tt, tb, ts, tb, bb, bs, tq, bq, qs, tq, bq, qq); };
  auto kQQ = [&] { return schur3(qq, tq, bq, qs, tt, tb, ts, tb, bb, bs, ts, bs, ss, tq, bq, qs); };
  // each complement once, under the condition of its uses (pure functions:
  // the same values as evaluating them at every use)
  const double vt2B = hB ? t2B() : 0.0, vt3B = hB ? t3B() : 0.0, vkBB = hB ? kBB() : 0.0;
  const double vt2S = hS ? t2S() : 0.0, vt3S = hS ? t3S() : 0.0, vkSS = hS ? kSS() : 0.0;
  const double vt2Q = hQ ? t2Q() : 0.0, vt3Q = hQ ? t3Q() : 0.0, vkQQ = hQ ? kQQ() : 0.0;
  const double vkBQ = (hB && hQ) ? kBQ() : 0.0, vkBS = (hB && hS) ? kBS() : 0.0;
  const double vkSB = (hS && hB) ? kSB() : 0.0, vkSQ = (hS && hQ) ? kSQ() : 0.0;
  const double vkQB = (hQ && hB) ? kQB() : 0.0, vkQS = (hQ && hS) ? kQS() : 0.0;
  double ct = 0.0;
  if (hB) ct += B / vt2B;
  if (hQ) ct += Q / vt2Q;
  if (hS) ct += S / vt2S;
  const double dwds = T + s / dsdt + ct;
  double dwdB = mB, dwdS = mS, dwdQ = mQ;
  if (hB) { double c = s / vt3B + B / vkBB; if (hQ) c += Q / vkBQ; if (hS) c += S / vkBS; dwdB += c; }
  if (hS) { double c = s / vt3S + S / vkSS; if (hB) c += B / vkSB; if (hQ) c += Q / vkSQ; dwdS += c; }
  if (hQ) { double c = s / vt3Q + Q / vkQQ; if (hB) c += B / vkQB; if (hS) c += S / vkQS; dwdQ += c; }
  const double aBs = -mB / T / T / dsdt + (hB ? 1 / T / vt2B : 0.0);
  const double aSs = -mS / T / T / dsdt + (hS ? 1 / T / vt2S : 0.0);
  const double aQs = -mQ / T / T / dsdt + (hQ ? 1 / T / vt2Q : 0.0);
  const double aBB = hB ? 1 / T / vkBB - mB / T / T / vt3B : 0.0;
  const double aSB = hB ? -mS / T / T / vt3B + (hS ? 1 / T / vkBS : 0.0) : 0.0;
  const double aQB = hB ? -mQ / T / T / vt3B + (hQ ? 1 / T / vkBQ : 0.0) : 0.0;
  const double aBS = hS ? -mB / T / T / vt3S + (hB ? 1 / T / vkSB : 0.0) : 0.0;
  const double aSS = hS ? -mS / T / T / vt3S + 1 / T / vkSS : 0.0;
  const double aQS = hS ? -mQ / T / T / vt3S + (hQ ? 1 / T / vkSQ : 0.0) : 0.0;
  const double aBQ = hQ ? -mB / T / T / vt3Q + (hB ? 1 / T / vkQB : 0.0) : 0.0;
  const double aSQ = hQ ? -mS / T / T / vt3Q + (hS ? 1 / T / vkQS : 0.0) : 0.0;
  const double aQQ = hQ ? -mQ / T / T / vt3Q + 1 / T / vkQQ : 0.0;
  double chiBB0 = r[X_BB];
  if (has_kind<KM, M_TABLE>() && m.kind == M_TABLE && (mB != 0.0 || mQ != 0.0 || mS != 0.0)) {
    const double z[4] = { T, 0.0, 0.0, 0.0 };   // chi_BB at mu = 0: one field of the 17
    chiBB0 = table_field_one(m, tab, z, X_BB);
  }
  th[0] = T; th[1] = mB; th[2] = mS; th[3] = mQ;
  th[4] = r[X_E]; th[5] = s; th[6] = B; th[7] = S; th[8] = Q;
  th[9] = r[X_P]; th[10] = r[X_CS2]; th[11] = r[X_E] + r[X_P];
  th[12] = dwds; th[13] = dwdB; th[14] = dwdS; th[15] = dwdQ;
  th[16] = aBs; th[17] = aBB; th[18] = aBS; th[19] = aBQ;
  th[20] = aSs; th[21] = aSB; th[22] = aSS; th[23] = aSQ;
  th[24] = aQs; th[25] = aQB; th[26] = aQS; th[27] = aQQ;
  th[28] = r[X_BB]; th[29] = chiBB0;
  th[30] = (m.kind == M_TABLE) ? 1.0 : 0.0;
}

// Seed of kind `kind` for a chain member (src/eos.cpp:629-975): 0 the
// current location, 1 (T, 0, 0, 0), 2 (580 MeV, sgn rho), 3 the
// conformal-diagonal seed of the chain's last member.
CK_HD void seed_of_kind(const EosChain& ch, int mode, int kind, const double* tgt, const double* x, double* y)
{
  if (kind == 0) { y[0] = x[0]; y[1] = x[1]; y[2] = x[2]; y[3] = x[3]; }
  else if (kind == 1) { y[0] = x[0]; y[1] = 0.0; y[2] = 0.0; y[3] = 0.0; }
  else if (kind == 2) { y[0] = 580.0 / kHbarcMeVfm; y[1] = sgnd(tgt[1]); y[2] = sgnd(tgt[3]); y[3] = sgnd(tgt[2]); }
  else diag_seed(ch.m[ch.n - 1], mode, tgt, y);
}
// Seed kinds per member: mode 0 (s) tries kinds 0-3, mode 1 (e) kinds 2-3.
CK_HD int seed_kinds(int mode) { return mode == 0 ? 4 : 2; }
CK_HD int first_kind(int mode) { return mode == 0 ? 0 : 2; }

// Point inversion over the chain (src/eos.cpp:629-975,
// src/sph_workstation.cpp:1713-1851), "invert_point" below.  mode 0:
// (s, B, S, Q) starting at the particle's current member, seeds: current
// location, (T, 0, 0, 0), (580 MeV, sgn), conformal-diagonal seed.  mode 1:
// (e, B, S, Q) from the first member, seeds: (580 MeV, sgn),
// conformal-diagonal seed.  Within a member the seed kinds are tried in
// order, each with find_root's attempts in order, and the first attempt
// that converges is taken; a solution with cs2 outside [0, 1] moves on to
// the next member, seeded from it.  The device runs it in two parts:
// invert_first (the first attempt alone) and invert_wave (all of it, one
// wave per particle).

// The first attempt of invert_point alone, with a Newton budget: the
// starting member, the first seed kind, attempt 0.  Returns 1 when that
// attempt converges within `budget` iterations to a point with cs2 in
// [0, 1]: invert_point's result then is that point (x, *member set).
// Returns -1 otherwise (the full inversion decides).
template <int KM = kAllMembers>
CK_HD int invert_first(const EosChain& ch, const double* tab, const double* tab5, int mode, const double* tgt,
                       double* x, int* member, int budget)
{
  int start = (mode == 0) ? *member : 0;
  if (start < 0 || start >= ch.n) start = 0;
  const EosMember& m = ch.m[start];
  if (has_kind<KM, M_TABLE>() && m.kind == M_TABLE && unreachable(m, mode, tgt)) return -1;
  double y0[4], y[4];
  seed_of_kind(ch, mode, first_kind(mode), tgt, x, y0);
  root_seed(m, tgt, y0, 0, y);
  if (newton_attempt_b<KM>(m, tab5, mode, tgt, y, budget) != 1) return -1;
  double r[17];
  member_fields<KM>(m, tab, y, r);
  if (r[X_CS2] < 0.0 || r[X_CS2] > 1.0) return -1;
  for (int a = 0; a < 4; ++a) x[a] = y[a];
  *member = start;
  return 1;
}

// The whole of invert_point for one particle by one wave (every lane
// active, the particle's values the same on every lane): the attempts of a
// member -- seed kind q, find_root attempt k, at most 4 x 10 -- run on lanes
// q * 10 + k at once, and the first that converges in that order is taken,
// which is the attempt invert_point's loops stop at.  Members are tried in
// order as there.
constexpr int kHardLanes = 64;
template <int KM = kAllMembers>
__device__ bool invert_wave(const EosChain& ch, const double* tab, const double* tab5, int mode, const double* tgt,
                            double* x, int* member)
{
  const int lane = threadIdx.x & (kHardLanes - 1);
  int start = (mode == 0) ? *member : 0;
  if (start < 0 || start >= ch.n) start = 0;
  const int na = seed_kinds(mode) * kRootAttempts;
  bool found = false;
  int used = -1;
  double res[4] = { 0, 0, 0, 0 };
  for (int k = start; k < ch.n; ++k) {
    const EosMember& m = ch.m[k];
    if (has_kind<KM, M_TABLE>() && m.kind == M_TABLE && unreachable(m, mode, tgt)) { found = false; continue; }
    double y[4] = { 0, 0, 0, 0 };
    bool ok = false;
    if (lane < na) {
      double y0[4];
      seed_of_kind(ch, mode, first_kind(mode) + lane / kRootAttempts, tgt, x, y0);
      root_seed(m, tgt, y0, lane % kRootAttempts, y);
      ok = newton_attempt<KM>(m, tab5, mode, tgt, y);
    }
    const unsigned long long mask = __ballot(ok);
    found = mask != 0ull;
    if (found) {
      const int first = __builtin_ctzll(mask);
      for (int a = 0; a < 4; ++a) res[a] = __shfl(y[a], first);
      used = k;
      double r[17];
      member_fields<KM>(m, tab, res, r);
      if (r[X_CS2] < 0.0 || r[X_CS2] > 1.0) { for (int a = 0; a < 4; ++a) x[a] = res[a]; continue; }
      break;
    }
  }
  if (!found) return false;
  for (int a = 0; a < 4; ++a) x[a] = res[a];
  *member = used;
  return true;
}

#undef CK_HD
}  // namespace ccake
