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
  // scalar g with grad W = g * rel
  __device__ __forceinline__ double operator()(double r2) const
  {
    const double r = sqrt(r2);
    const double q = r * hinv;
    const double t = 2.0 - q;
    const double far = (ng / r) * t * t;
    const double near = ng2 * (1.5 * q - 2.0) * hinv;
    return q >= 2.0 ? 0.0 : (q >= 1.0 ? far : near);
  }
};

template <bool SHEAR>
__global__ __launch_bounds__(64) void k_grad(const double* __restrict__ cand,
                                             const int32_t* __restrict__ skey,
                                             const int32_t* __restrict__ perm,
                                             const int32_t* __restrict__ cell_off,
                                             const Grid* __restrict__ gridp, int n,
                                             double rc2, double rcs, GradW G,
                                             double* __restrict__ F, int npad)
{
  const Grid g = *gridp;
  const int lane = threadIdx.x;
  const int k = blockIdx.x * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand + (size_t)kk * kCandG;
  const double xa = me[0], ya = me[1];
  const double vax = me[6], vay = me[7];
  const int myrow = act ? skey[kk] / g.nc[0] : INT_MAX;

  double a1x = 0, a1y = 0, apx = 0, apy = 0, aex = 0, aey = 0, abx = 0, aby = 0;
  double vxx = 0, vxy = 0, vyx = 0, vyy = 0;
  double gsx = 0, gsy = 0, acv = 0, dsx = 0, dsy = 0;

  for_segments(act, myrow, xa, ya, [&](const Seg& seg, bool in) {
    for_spans(g, cell_off, seg, rcs, [&](int j) {
      const double* c = cand + (size_t)j * kCandG;
      const double dx = xa - c[0], dy = ya - c[1];
      const double r2 = dist2_nocontract(dx, dy);
      const bool nb = in && (r2 <= rc2) && (j != k);
      const double gg = nb ? G(r2) : 0.0;
      const double wx = c[2] * (gg * dx), wy = c[2] * (gg * dy);
      a1x += wx; a1y += wy;
      apx += c[3] * wx; apy += c[3] * wy;
      aex += c[4] * wx; aey += c[4] * wy;
      abx += c[5] * wx; aby += c[5] * wy;
      vxx += c[6] * wx; vxy += c[7] * wx;
      vyx += c[6] * wy; vyy += c[7] * wy;
      if (SHEAR) {
        const double cv = wx * vax + wy * vay;
        acv += cv;
        gsx += cv * c[8]; gsy += cv * c[9];
        dsx += wx * c[10] + wy * c[11];
        dsy += wx * c[11] + wy * c[12];
      }
    });
  });

  if (!act) return;
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
  const dim3 grid((n + kTile - 1) / kTile), blk(kTile);
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
