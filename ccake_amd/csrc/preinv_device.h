// Pre-inverted EoS, flat (s, nB) format: the device side of CCAKE's
// EoS_Interpolator (include/eos_interpolator.h, src/eos_interpolator.cpp)
// used when online_inverter_enabled is false.  Eleven variables on a grid
// uniform in (log10 s, log10 nB); each particle reads four nodes per
// variable, so the lookup replaces the Newton solve with 44 gathers from a
// table that stays in L2.
#pragma once
#include <hip/hip_runtime.h>

namespace ccake {

// variables in the reference's eos_variables order (include/eos_interpolator.h:29-35)
enum PreinvVar { PI_T = 0, PI_MUB, PI_MUQ, PI_MUS, PI_E, PI_P, PI_CS2, PI_DWDS, PI_DWDB, PI_DWDQ, PI_DWDS_S, PI_NUM };

struct Preinv {
  int Ns = 0, NB = 0;
  double ls0 = 0, dls = 0, lb0 = 0, dlb = 0;   // log10 start and stride of the s and nB axes
  double s_first = 0, s_last = 0, nB_first = 0;  // fm^-3
  const double* v = nullptr;                   // PI_NUM x Ns x NB, s slowest
  const double* s_axis = nullptr;              // Ns
};

__host__ __device__ inline double pv(const Preinv& p, int k, int is, int iB)
{
  return p.v[((size_t)k * p.Ns + is) * p.NB + iB];
}

__host__ __device__ inline int clamp_cell(int i, int n) { return i < 0 ? 0 : (i > n - 2 ? n - 2 : i); }

// fill_thermodynamics (src/eos_interpolator.cpp:531-627) through the bilinear
// branch of interpolate4D (685-711): the S and Q axes of the flat slot are
// degenerate.  o: T, muB, muS, muQ, e, p, cs2, dwds, dwdB, dwdS, dwdQ, w.
// The cell index is clamped to N-2 (the reference clamps to N-1 and then
// reads node N at the upper edge): the same value on the last node,
// extrapolation of the last cell beyond it.
__device__ inline void preinv_fill(const Preinv& p, double s_in, double rhoB_in, double* o)
{
  const double s = fmax(.001, s_in);
  const double rhoB = fmax(.001, fabs(rhoB_in));
  const double pos0 = log10(fmax(s, 1e-30));
  const double pos1 = log10(fmax(rhoB, 1e-30));
  const int i0 = clamp_cell(fabs(p.dls) < 1e-30 ? 0 : (int)floor((pos0 - p.ls0) / p.dls), p.Ns);
  const int i1 = clamp_cell(fabs(p.dlb) < 1e-30 ? 0 : (int)floor((pos1 - p.lb0) / p.dlb), p.NB);
  const double x0 = (pos0 - p.ls0 - i0 * p.dls) / p.dls;
  const double x1 = (pos1 - p.lb0 - i1 * p.dlb) / p.dlb;
  // phi_c for corner c = (bit 0: s, bit 1: nB), in the reference's order
  const double ph[4] = { (1.0 * (1.0 - x0)) * (1.0 - x1), (1.0 * x0) * (1.0 - x1),
                         (1.0 * (1.0 - x0)) * x1, (1.0 * x0) * x1 };
  double f[PI_NUM];
  #pragma unroll
  for (int k = 0; k < PI_NUM; ++k) {
    const double* row0 = p.v + ((size_t)k * p.Ns + i0) * p.NB + i1;
    const double* row1 = row0 + p.NB;
    double acc = 0.0;
    acc += ph[0] * row0[0];
    acc += ph[1] * row1[0];
    acc += ph[2] * row0[1];
    acc += ph[3] * row1[1];
    f[k] = acc;
  }
  o[0] = f[PI_T]; o[1] = f[PI_MUB]; o[2] = f[PI_MUS]; o[3] = f[PI_MUQ];
  o[4] = f[PI_E]; o[5] = f[PI_P]; o[6] = f[PI_CS2];
  o[7] = f[PI_DWDS]; o[8] = f[PI_DWDB]; o[9] = f[PI_DWDS_S]; o[10] = f[PI_DWDQ];
  o[11] = o[4] + o[5];
}

// query_thermo_host (src/eos_interpolator.cpp:751-811): bilinear in
// (log10 s, log10 |rhoB|) with both clamped into the table.
// o: T, muB, muS, muQ, e, p, cs2, dwds, dwdB, w.
__host__ __device__ inline void preinv_query(const Preinv& p, double s, double rhoB, double* o)
{
  const double abs_rhoB = fmax(fabs(rhoB), p.nB_first);
  const double s_clamped = fmax(s, p.s_first);
  const double L_s_max = p.ls0 + p.dls * (p.Ns - 1);
  const double L_nB_max = p.lb0 + p.dlb * (p.NB - 1);
  const double L_s = fmin(fmax(log10(s_clamped), p.ls0), L_s_max);
  const double L_nB = fmin(fmax(log10(abs_rhoB), p.lb0), L_nB_max);
  const double xs = (L_s - p.ls0) / p.dls;
  const double xb = (L_nB - p.lb0) / p.dlb;
  const int is = clamp_cell((int)floor(xs), p.Ns);
  const int iB = clamp_cell((int)floor(xb), p.NB);
  const double a = xs - is, b = xb - iB;
  auto bil = [&](int k) {
    return (1.0 - a) * (1.0 - b) * pv(p, k, is, iB) + a * (1.0 - b) * pv(p, k, is + 1, iB)
         + (1.0 - a) * b * pv(p, k, is, iB + 1) + a * b * pv(p, k, is + 1, iB + 1);
  };
  o[0] = bil(PI_T); o[1] = bil(PI_MUB); o[2] = bil(PI_MUS); o[3] = bil(PI_MUQ);
  o[4] = bil(PI_E); o[5] = bil(PI_P); o[6] = bil(PI_CS2);
  o[7] = bil(PI_DWDS); o[8] = bil(PI_DWDB);
  o[9] = o[4] + o[5];
}

// invert_e_to_s_host (src/eos_interpolator.cpp:815-861): bisection of e along
// the s axis at fixed |rhoB|, then log10-linear in s inside the bracketing
// cell; the ends clamp to the first and last s node.
__host__ __device__ inline double preinv_e_to_s(const Preinv& p, double e_input, double rhoB)
{
  const double L_nB_max = p.lb0 + p.dlb * (p.NB - 1);
  const double abs_rhoB = fmax(fabs(rhoB), p.nB_first);
  const double L_nB = fmin(log10(abs_rhoB), L_nB_max);
  const double xi_B = (L_nB - p.lb0) / p.dlb;
  const int iB = clamp_cell((int)floor(xi_B), p.NB);
  const double b = xi_B - iB;
  auto e_at = [&](int is) { return (1.0 - b) * pv(p, PI_E, is, iB) + b * pv(p, PI_E, is, iB + 1); };
  if (e_input <= e_at(0)) return p.s_first;
  if (e_input >= e_at(p.Ns - 1)) return p.s_last;
  int lo = 0, hi = p.Ns - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (e_at(mid) <= e_input) lo = mid; else hi = mid;
  }
  const double e_lo = e_at(lo), e_hi = e_at(hi);
  const double alpha = (e_hi > e_lo) ? (e_input - e_lo) / (e_hi - e_lo) : 0.0;
  const double L_lo = log10(p.s_axis[lo]), L_hi = log10(p.s_axis[hi]);
  return pow(10.0, L_lo + alpha * (L_hi - L_lo));
}

}  // namespace ccake
