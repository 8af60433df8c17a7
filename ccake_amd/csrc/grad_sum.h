// The gradient pass's sum for one particle: the slices' partial sums of
// sorted position k in slice order, scaled by C = kgrad / h^4, then the
// owner terms (v_a, p_a/sl_a^2, e_a, Pi_a/sl_a^2, pi_a/sl_a^2, sl_a from the
// particle's candidate record), written to particle id.  S > 0: S slices,
// all loads in flight at once; S = 0: `nslices` slices in blocks of 8 (a
// clamped slice index past the last adds an exact 0.0), the same sums.
// Shared by k_grad_sum and the equations-of-motion kernel, which folds it in.
#pragma once
#include "pair_common.h"

namespace ccake {

constexpr int kCandG = 16;   // doubles per gradient candidate record

template <bool SHEAR>
constexpr int grad_nacc() { return SHEAR ? 17 : 12; }

template <bool SHEAR, int S>
__device__ __forceinline__ void grad_sum_into(const double* __restrict__ cand, int k, int id, double C,
                                              const double* __restrict__ part, int nslices, int npad,
                                              double* __restrict__ F)
{
  constexpr int NACC = grad_nacc<SHEAR>();
  double acc[NACC];
  #pragma unroll
  for (int q = 0; q < NACC; ++q) acc[q] = 0.0;
  if constexpr (S > 0) {
    #pragma unroll
    for (int sli = 0; sli < S; ++sli)
      #pragma unroll
      for (int q = 0; q < NACC; ++q) acc[q] += part[((size_t)sli * NACC + q) * npad + k];
  } else {
    for (int s0 = 0; s0 < nslices; s0 += 8) {
      double v[8][NACC];
      #pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int sli = s0 + u < nslices ? s0 + u : nslices - 1;
        #pragma unroll
        for (int q = 0; q < NACC; ++q) v[u][q] = part[((size_t)sli * NACC + q) * npad + k];
      }
      #pragma unroll
      for (int u = 0; u < 8; ++u)
        #pragma unroll
        for (int q = 0; q < NACC; ++q) acc[q] += (s0 + u < nslices) ? v[u][q] : 0.0;
    }
  }
  #pragma unroll
  for (int q = 0; q < NACC; ++q) acc[q] *= C;
  const double* me = cand + (size_t)k * kCandG;
  const double vax = me[6], vay = me[7];

  const double a1x = acc[0], a1y = acc[1], apx = acc[2], apy = acc[3], aex = acc[4], aey = acc[5];
  const double abx = acc[6], aby = acc[7], vxx = acc[8], vxy = acc[9], vyx = acc[10], vyy = acc[11];
  const double sla = me[13], isla = 1.0 / sla;
  auto put = [&](int c, double v) { F[(size_t)c * npad + id] = v; };
  put(CCAKE_F_GRADP0 + 0, sla * (apx + me[3] * a1x));
  put(CCAKE_F_GRADP0 + 1, sla * (apy + me[3] * a1y));
  put(CCAKE_F_GRADP0 + 2, 0.0);
  put(CCAKE_F_GRADE0 + 0, (aex - me[4] * a1x) * isla);
  put(CCAKE_F_GRADE0 + 1, (aey - me[4] * a1y) * isla);
  put(CCAKE_F_GRADE0 + 2, 0.0);
  put(CCAKE_F_GRADBULK0 + 0, sla * (abx + me[5] * a1x));
  put(CCAKE_F_GRADBULK0 + 1, sla * (aby + me[5] * a1y));
  put(CCAKE_F_GRADBULK0 + 2, 0.0);
  put(CCAKE_F_GRADV0 + 0, (vxx - vax * a1x) * isla);
  put(CCAKE_F_GRADV0 + 1, (vxy - vay * a1x) * isla);
  put(CCAKE_F_GRADV0 + 2, 0.0);
  put(CCAKE_F_GRADV0 + 3, (vyx - vax * a1y) * isla);
  put(CCAKE_F_GRADV0 + 4, (vyy - vay * a1y) * isla);
  put(CCAKE_F_GRADV0 + 5, 0.0);
  put(CCAKE_F_GRADV0 + 6, 0.0); put(CCAKE_F_GRADV0 + 7, 0.0); put(CCAKE_F_GRADV0 + 8, 0.0);
  if (SHEAR) {
    const double acv = acc[SHEAR ? 12 : 0], gsx = acc[SHEAR ? 13 : 0], gsy = acc[SHEAR ? 14 : 0];
    const double dsx = acc[SHEAR ? 15 : 0], dsy = acc[SHEAR ? 16 : 0];
    put(CCAKE_F_GRADSHEAR0 + 0, sla * (gsx + me[8] * acv));
    put(CCAKE_F_GRADSHEAR0 + 1, sla * (gsy + me[9] * acv));
    put(CCAKE_F_DIVSHEAR0 + 0, sla * (dsx + me[10] * a1x + me[11] * a1y));
    put(CCAKE_F_DIVSHEAR0 + 1, sla * (dsy + me[11] * a1x + me[12] * a1y));
  } else {
    put(CCAKE_F_GRADSHEAR0 + 0, 0.0); put(CCAKE_F_GRADSHEAR0 + 1, 0.0);
    put(CCAKE_F_DIVSHEAR0 + 0, 0.0); put(CCAKE_F_DIVSHEAR0 + 1, 0.0);
  }
  put(CCAKE_F_GRADSHEAR0 + 2, 0.0);
  put(CCAKE_F_DIVSHEAR0 + 2, 0.0);
}

}  // namespace ccake
