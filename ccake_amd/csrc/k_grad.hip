// Gradient pair kernels on gfx950 (owner computes, no atomics).
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
#include "grad_sum.h"

namespace ccake {

// The walk of one (tile, slice) unit: slice `slice` of `nslices` of the
// tile's candidates, one wave; fin(acc) receives the neighbour-side sums.
// BULK false: the event's bulk pressure is zero on every particle (zeta = 0
// and no initial bulk, checked at initialize), so the grad Pi sums are the
// exact zeros the full walk would give and are not accumulated.
template <bool SHEAR, bool BULK, class Fin>
__device__ __forceinline__ void grad_walk(const double* __restrict__ cand, const int32_t* __restrict__ skey,
                                          const int32_t* __restrict__ cell_off, const Grid* __restrict__ gridp,
                                          int n, double rc2, double rcs, double hinv2, int tile, int slice,
                                          int nslices, Fin&& fin)
{
  constexpr int NACC = grad_nacc<SHEAR>();
  const Grid g0 = *gridp;
  const int lane = threadIdx.x & (kWave - 1);
  const int k = tile * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand + (size_t)kk * kCandG;
  const double xa = me[0], ya = me[1];
  const double vax = me[6], vay = me[7];
  const int myrow = act ? skey[kk] / g0.stride : INT_MAX;

  double acc[NACC];
  for (int q = 0; q < NACC; ++q) acc[q] = 0.0;

  for_segments(act, myrow, xa, ya, [&](const Seg& seg, bool in) {
    const Grid& g = gridp[event_index(gridp, g0, seg.row)];
    // lanes of other cell rows sit this segment's walk out
    for_spans_slice(g, cell_off, seg, rcs, slice, nslices, in, [&](int j0, int j1) {
      // Candidate records are wave-uniform (scalar loads into SGPRs); the
      // shear record is too large to hold two steps of two in SGPRs.
      constexpr int NR = SHEAR ? 13 : 8;
      walk_records<NR, kCandG, SHEAR ? 1 : 2>(cand, xa, ya, j0, j1,
                                              [&](const double* c, double dx, double dy, int j, bool valid) {
        const double r2 = dist2_nocontract(dx, dy);
        if (kSkipEmpty && wave_none(valid && r2 <= rc2)) return;
        double q, iq;
        q_and_inverse(r2, hinv2, q, iq);
        const double gm = valid ? c[2] * spline_g(q, iq) : 0.0;
        const double wx = gm * dx, wy = gm * dy;
        acc[0] += wx; acc[1] += wy;                    // sum m_b grad W / C
        acc[2] += c[3] * wx; acc[3] += c[3] * wy;      // p_b / sl_b^2
        acc[4] += c[4] * wx; acc[5] += c[4] * wy;      // e_b
        if (BULK) { acc[6] += c[5] * wx; acc[7] += c[5] * wy; }   // Pi_b / sl_b^2
        acc[8] += c[6] * wx; acc[9] += c[7] * wx;      // v_b d_x W
        acc[10] += c[6] * wy; acc[11] += c[7] * wy;    // v_b d_y W
        if (SHEAR) {
          const double cv = wx * vax + wy * vay;
          acc[SHEAR ? 12 : 0] += cv;
          acc[SHEAR ? 13 : 0] += cv * c[SHEAR ? 8 : 0]; acc[SHEAR ? 14 : 0] += cv * c[SHEAR ? 9 : 0];
          acc[SHEAR ? 15 : 0] += wx * c[SHEAR ? 10 : 0] + wy * c[SHEAR ? 11 : 0];
          acc[SHEAR ? 16 : 0] += wx * c[SHEAR ? 11 : 0] + wy * c[SHEAR ? 12 : 0];
        }
      });
    });
  });
  fin(acc);
}

// Pair part: one single-wave workgroup per (tile, slice) unit walks its
// slice of the tile's candidates and stores the slice's neighbour-side sums;
// k_grad_sum adds the slices in slice order and applies the owner terms.
template <bool SHEAR, bool BULK>
__global__ __launch_bounds__(kWave) void k_grad_slice(const double* __restrict__ cand,
                                                      const int32_t* __restrict__ skey,
                                                      const int32_t* __restrict__ cell_off,
                                                      const Grid* __restrict__ gridp, int n,
                                                      double rc2, double rcs, double hinv2, int nslices,
                                                      double* __restrict__ part, int npad)
{
  constexpr int NACC = grad_nacc<SHEAR>();
  const int unit = xcd_tile(blockIdx.x, gridDim.x);
  const int tile = unit / nslices;
  const int slice = unit - tile * nslices;
  grad_walk<SHEAR, BULK>(cand, skey, cell_off, gridp, n, rc2, rcs, hinv2, tile, slice, nslices,
                         [&](const double* acc) {
    double* p = part + (size_t)slice * NACC * npad + tile * kTile + threadIdx.x;
    for (int q = 0; q < NACC; ++q) p[(size_t)q * npad] = acc[q];
  });
}

// The same walk with the W slices of a tile in one workgroup of W waves
// (wave w walks slice w): the slices' sums are added in slice order in LDS
// and stored as the sums of a single slice, so the sum pass reads NACC
// values per particle instead of W * NACC and no partial sum of a slice
// goes through HBM.
template <bool SHEAR, bool BULK, int W>
__global__ __launch_bounds__(kWave * W) void k_grad_tile(const double* __restrict__ cand,
                                                         const int32_t* __restrict__ skey,
                                                         const int32_t* __restrict__ cell_off,
                                                         const Grid* __restrict__ gridp, int n,
                                                         double rc2, double rcs, double hinv2,
                                                         double* __restrict__ part, int npad)
{
  constexpr int NACC = grad_nacc<SHEAR>();
  __shared__ double red[W][NACC][kWave];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int slice = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);   // wave-uniform
  const int lane = threadIdx.x & (kWave - 1);
  grad_walk<SHEAR, BULK>(cand, skey, cell_off, gridp, n, rc2, rcs, hinv2, tile, slice, W, [&](const double* acc) {
    #pragma unroll
    for (int q = 0; q < NACC; ++q) red[slice][q][lane] = acc[q];
  });
  __syncthreads();
  for (int t = threadIdx.x; t < NACC * kWave; t += kWave * W) {
    const int q = t / kWave, l = t - q * kWave;
    double s = 0.0;
    #pragma unroll
    for (int w = 0; w < W; ++w) s += red[w][q][l];
    part[(size_t)q * npad + tile * kTile + l] = s;
  }
}

// Sum pass: the slices' sums in slice order, scaled by C = kgrad / h^4,
// then the owner terms, written by particle id (grad_sum_into).  S is the
// slice count, so every load of a particle is in flight at once.
template <bool SHEAR, int S>
__global__ __launch_bounds__(256) void k_grad_sum(const double* __restrict__ cand,
                                                  const int32_t* __restrict__ perm, int n, double C,
                                                  const double* __restrict__ part, int npad,
                                                  double* __restrict__ F)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  grad_sum_into<SHEAR, S>(cand, k, perm[k], C, part, S, npad, F);
}

int launch_grad_pairs(ccake_event* ev)
{
  if (ev->p.pair_precision == 1) return launch_grad_pairs32(ev);
  const int n = ev->n_live;          // sorted positions of the live events' particles
  const double h = ev->p.hT;
  const double rc = 2.0 * h;
  const double hinv2 = 1.0 / (h * h);
  const int ntiles = (n + kTile - 1) / kTile;
  const double rcs = search_radius(ev);
  // tiles of 4 waves: 1.318 ms per launch on the 200k event against 1.343
  // with 8 (the smoothing walk prefers 8)
  const int W = lds_slices(ev) ? 4 : 0;
  if (W) {
    // the tile's slices reduced in LDS: the sum pass sees one slice
    ev->grad_sum_slices = 1;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(ntiles), dim3(kWave * W), 0, ev->stream, ev->cand_g, ev->skey, ev->cell_off,
                         ev->grid, n, rc * rc, rcs, hinv2, ev->part, ev->npad);
    };
    const bool sh = ev->p.using_shear, bu = !ev->no_bulk;
    if (sh) { if (bu) go(k_grad_tile<true, true, 4>); else go(k_grad_tile<true, false, 4>); }
    else { if (bu) go(k_grad_tile<false, true, 4>); else go(k_grad_tile<false, false, 4>); }
    return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
  }
  ev->grad_sum_slices = ev->nslices;
  const dim3 grid(ntiles * ev->nslices), blk(kWave);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, blk, 0, ev->stream, ev->cand_g, ev->skey, ev->cell_off, ev->grid, n, rc * rc,
                       rcs, hinv2, ev->nslices, ev->part, ev->npad);
  };
  if (ev->p.using_shear) {
    if (ev->no_bulk) go(k_grad_slice<true, false>); else go(k_grad_slice<true, true>);
  } else {
    if (ev->no_bulk) go(k_grad_slice<false, false>); else go(k_grad_slice<false, true>);
  }
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

template <bool SHEAR, int S = 1>
static int grad_sum_dispatch(ccake_event* ev, double C)
{
  if constexpr (S > kMaxSlices) {
    set_error("slice count out of range");
    return CCAKE_ERR_INVALID;
  } else {
    if (ev->grad_sum_slices != S) return grad_sum_dispatch<SHEAR, S + 1>(ev, C);
    const int n = ev->n_live;
    hipLaunchKernelGGL((k_grad_sum<SHEAR, S>), dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->cand_g,
                       ev->perm, n, C, ev->part, ev->npad, ev->F);
    return 0;
  }
}

double grad_sum_scale(double h) { return -0.5 * (15.0 / (7.0 * kPi)) / pow(h, 4); }   // kgrad / h^4

int launch_grad_sum(ccake_event* ev)
{
  const double C = grad_sum_scale(ev->p.hT);
  const int rc = ev->p.using_shear ? grad_sum_dispatch<true>(ev, C) : grad_sum_dispatch<false>(ev, C);
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch_grad(ccake_event* ev)
{
  if (ev->p.dim == 3) return launch3_grad(ev);
  int rc = launch_grad_pairs(ev);
  return rc ? rc : launch_grad_sum(ev);
}

}  // namespace ccake
