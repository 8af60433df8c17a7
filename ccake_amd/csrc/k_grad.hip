// Gradient pair kernel on gfx950 (owner computes, no atomics).
// Replaces SPHWorkstation::smooth_all_particle_gradients
// (src/sph_workstation.cpp:796-986), D = 2, diffusion terms off.
//
// With K_ab = m_b sl_a grad W_ab the reference sums, per neighbour b,
//   gradP    += (p_b/sl_b^2 + p_a/sl_a^2) K_ab
//   gradE    += (m_b/sl_a)(e_b - e_a) grad W_ab
//   gradBulk += (Pi_b/sl_b^2 + Pi_a/sl_a^2) K_ab
//   gradV_ij += (m_b/sl_a)(v_b^j - v_a^j) d_i W_ab
//   gradshear+= (K_ab . v_a)(pi^{0i}_b/sl_b^2 + pi^{0i}_a/sl_a^2)
//   divshear_j += sum_i K_ab,i (pi^{ji}_b/sl_b^2 + pi^{ji}_a/sl_a^2).
// Every owner-side factor is constant over b, so the kernel accumulates
// only neighbour-side sums of w = m_b grad W_ab (17 accumulators) and
// applies the owner terms once at the end.
#include "pair_common.h"

namespace ccake {

constexpr int kCandG = 16;

struct GradW {
  double hinv, ng, ng2;   // kgrad/h^3, kgrad2/h^3 (src/kernel.cpp:95-118)
  // scalar g with grad W = g * rel for a neighbour (coincident particles sit
  // at q = 0, on the inner branch, which does not divide by r)
  __device__ __forceinline__ double operator()(double r2) const
  {
    double r, ir;
    sqrt_rsqrt(r2, r, ir);
    const double q = r * hinv;
    const double t = 2.0 - q;
    const double far = ng * ir * t * t;
    const double near = ng2 * (1.5 * q - 2.0) * hinv;
    // only neighbours (r <= 2h) use the value: at q >= 2, t^2 = 0 to rounding
    return (q >= 1.0) ? far : near;
  }
};

// One 64-particle tile per workgroup of kSplit waves; each wave walks a
// slice of the tile's candidates and wave 0 adds the slices in wave order.
template <bool SHEAR>
__global__ __launch_bounds__(kSplit * kWave) void k_grad(const double* __restrict__ cand,
                                                         const int32_t* __restrict__ skey,
                                                         const int32_t* __restrict__ perm,
                                                         const int32_t* __restrict__ cell_off,
                                                         const Grid* __restrict__ gridp, int n,
                                                         double rc2, double rcs, GradW G,
                                                         double* __restrict__ F, int npad)
{
  constexpr int NACC = SHEAR ? 17 : 12;
  __shared__ double red[kSplit - 1][NACC][kWave];
  const Grid g = *gridp;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = uni((int)(threadIdx.x / kWave));   // wave-uniform
  const int k = xcd_tile(blockIdx.x, gridDim.x) * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand + (size_t)kk * kCandG;
  const double xa = me[0], ya = me[1];
  const double vax = me[6], vay = me[7];
  const int myrow = act ? skey[kk] / g.nc[0] : INT_MAX;

  double acc[NACC];
  for (int q = 0; q < NACC; ++q) acc[q] = 0.0;

  for_segments(act, myrow, xa, ya, [&](const Seg& seg, bool in) {
    for_spans_slice(g, cell_off, seg, rcs, wave, [&](int j0, int j1) {
      // Candidate records are wave-uniform (scalar loads into SGPRs); two
      // records are requested together so the second arrives while the
      // first is being used.
      constexpr int NR = SHEAR ? 13 : 8;
      auto one = [&](const double* c, int j) {
        const double dx = xa - c[0], dy = ya - c[1];
        const double r2 = dist2_nocontract(dx, dy);
        const bool nb = in && (r2 <= rc2) && (j != k);
        if (kSkipEmpty && __ballot(nb) == 0ull) return;
        const double gv = G(r2);
        const double gg = nb ? gv : 0.0;
        const double wx = c[2] * (gg * dx), wy = c[2] * (gg * dy);
        acc[0] += wx; acc[1] += wy;                    // sum m_b grad W
        acc[2] += c[3] * wx; acc[3] += c[3] * wy;      // p_b / sl_b^2
        acc[4] += c[4] * wx; acc[5] += c[4] * wy;      // e_b
        acc[6] += c[5] * wx; acc[7] += c[5] * wy;      // Pi_b / sl_b^2
        acc[8] += c[6] * wx; acc[9] += c[7] * wx;      // v_b d_x W
        acc[10] += c[6] * wy; acc[11] += c[7] * wy;    // v_b d_y W
        if (SHEAR) {
          const double cv = wx * vax + wy * vay;
          acc[12] += cv;
          acc[13] += cv * c[SHEAR ? 8 : 0]; acc[14] += cv * c[SHEAR ? 9 : 0];
          acc[15] += wx * c[SHEAR ? 10 : 0] + wy * c[SHEAR ? 11 : 0];
          acc[16] += wx * c[SHEAR ? 11 : 0] + wy * c[SHEAR ? 12 : 0];
        }
      };
      int j = j0;
      for (; j + 1 < j1; j += 2) {
        double ca[NR], cb[NR];
        for (int q = 0; q < NR; ++q) ca[q] = cand[(size_t)j * kCandG + q];
        for (int q = 0; q < NR; ++q) cb[q] = cand[(size_t)(j + 1) * kCandG + q];
        one(ca, j);
        one(cb, j + 1);
      }
      if (j < j1) {
        double ca[NR];
        for (int q = 0; q < NR; ++q) ca[q] = cand[(size_t)j * kCandG + q];
        one(ca, j);
      }
    });
  });

  if (wave > 0)
    for (int q = 0; q < NACC; ++q) red[wave - 1][q][lane] = acc[q];
  __syncthreads();
  if (wave != 0 || !act) return;
  for (int w = 0; w < kSplit - 1; ++w)
    for (int q = 0; q < NACC; ++q) acc[q] += red[w][q][lane];

  const double a1x = acc[0], a1y = acc[1], apx = acc[2], apy = acc[3], aex = acc[4], aey = acc[5];
  const double abx = acc[6], aby = acc[7], vxx = acc[8], vxy = acc[9], vyx = acc[10], vyy = acc[11];
  const int id = perm[kk];
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

int launch_grad(ccake_event* ev)
{
  const int n = ev->n;
  const double h = ev->p.hT;
  const double rc = 2.0 * h;
  GradW G;
  G.hinv = 1.0 / h;
  G.ng = -15.0 / (14.0 * kPi) / pow(h, 3);
  G.ng2 = 15.0 / (7.0 * kPi) / pow(h, 3);
  const dim3 grid((n + kTile - 1) / kTile), blk(kSplit * kWave);
  const double rcs = rc * (1.0 + 1e-12) + 1e-12;
  if (ev->p.using_shear)
    hipLaunchKernelGGL(k_grad<true>, grid, blk, 0, ev->dev->stream, ev->cand_g, ev->skey, ev->perm, ev->cell_off,
                       ev->grid, n, rc * rc, rcs, G, ev->F, ev->npad);
  else
    hipLaunchKernelGGL(k_grad<false>, grid, blk, 0, ev->dev->stream, ev->cand_g, ev->skey, ev->perm, ev->cell_off,
                       ev->grid, n, rc * rc, rcs, G, ev->F, ev->npad);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

}  // namespace ccake
