// Smoothing pair kernel on gfx950 (owner computes, no atomics).
// Replaces SPHWorkstation::smooth_all_particle_fields
// (src/sph_workstation.cpp:494-578) and calculate_intial_sigma (582-640);
// the same walk counts neighbours (btrack, src/system_state.cpp:712-715)
// and exports the pair set for parity checks.
#include "pair_common.h"
#include <vector>
#include <algorithm>

namespace ccake {

constexpr int kCandS = 8;   // doubles per smoothing candidate record

// Candidate records in sorted order: x, y, m_s, m_s*ext_s, m_B*ext_B,
// m_S*ext_S, m_Q*ext_Q, 0.
__global__ __launch_bounds__(256) void k_gather_smooth(const double* __restrict__ F, int npad, int n,
                                                       const int32_t* __restrict__ perm,
                                                       double* __restrict__ cand)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm[k];
  auto f = [&](int c) { return F[(size_t)c * npad + i]; };
  double* o = cand + (size_t)k * kCandS;
  const double ms = f(CCAKE_F_MASS_S);
  o[0] = f(CCAKE_F_X);
  o[1] = f(CCAKE_F_Y);
  o[2] = ms;
  o[3] = ms * f(CCAKE_F_EXT_S);
  o[4] = f(CCAKE_F_MASS_B) * f(CCAKE_F_EXT_B);
  o[5] = f(CCAKE_F_MASS_SS) * f(CCAKE_F_EXT_SS);
  o[6] = f(CCAKE_F_MASS_Q) * f(CCAKE_F_EXT_Q);
  o[7] = 0.0;
}

struct KernelW {
  double hinv, norm, norm2;
  // Cubic spline W(r, h) (src/kernel.cpp:34-48) from r^2.
  __device__ __forceinline__ double of_q(double q) const
  {
    const double qq = q * q;
    const double a = norm * (1.0 - 1.5 * qq + 0.75 * q * qq);
    const double t = 2.0 - q;
    const double b = norm2 * t * t * t;
    const double v = (q >= 1.0) ? b : a;
    return (q >= 2.0) ? 0.0 : v;
  }
  // For neighbours only (r <= 2h): the q >= 2 case is then (2 - q)^3 = 0 up
  // to rounding, so the outer branch needs no cut.
  __device__ __forceinline__ double of_q_inside(double q) const
  {
    const double qq = q * q;
    const double a = norm * (1.0 - 1.5 * qq + 0.75 * q * qq);
    const double t = 2.0 - q;
    const double b = norm2 * t * t * t;
    return (q >= 1.0) ? b : a;
  }
  __device__ __forceinline__ double operator()(double r2) const { return of_q(sqrt(r2) * hinv); }
  __device__ __forceinline__ double fast(double r2) const
  {
    double r, ir;
    sqrt_rsqrt(r2, r, ir);
    return of_q_inside(r * hinv);
  }
};

// Neighbour walk without kernel sums, one wave per tile.  MODE 2: neighbour
// counts (btrack) and the pair total; MODE 3: write the pair rows (parity
// export) at the row offsets row_off.
template <int MODE>
__global__ __launch_bounds__(64) void k_walk(const double* __restrict__ cand,
                                             const int32_t* __restrict__ skey,
                                             const int32_t* __restrict__ perm,
                                             const int32_t* __restrict__ cell_off,
                                             const Grid* __restrict__ gridp, int n,
                                             double rc2, double rcs,
                                             int32_t* __restrict__ btrack,
                                             unsigned long long* __restrict__ pair_total,
                                             const int64_t* __restrict__ row_off,
                                             int32_t* __restrict__ out_i, int32_t* __restrict__ out_j)
{
  static_assert(MODE == 2 || MODE == 3, "walk modes");
  const Grid g = *gridp;
  const int lane = threadIdx.x;
  const int k = blockIdx.x * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand + (size_t)kk * kCandS;
  const double xi = me[0], yi = me[1];
  const int myrow = act ? skey[kk] / g.nc[0] : INT_MAX;
  const int id = act ? perm[kk] : 0;
  int cnt = 0;
  int64_t wpos = 0;
  if (MODE == 3 && act) wpos = row_off[id];

  for_segments(act, myrow, xi, yi, [&](const Seg& seg, bool in) {
    for_spans(g, cell_off, seg, rcs, [&](int j) {
      const double* c = cand + (size_t)j * kCandS;
      const double dx = xi - c[0], dy = yi - c[1];
      const double r2 = dist2_nocontract(dx, dy);
      const bool nb = in && (r2 <= rc2) && (j != k);
      if (MODE == 2) cnt += nb;
      if (MODE == 3 && nb) { out_i[wpos] = id; out_j[wpos] = perm[j]; ++wpos; }
    });
  });

  if (!act || MODE == 3) return;
  unsigned long long tot = (unsigned long long)cnt;
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
  if (lane == 0 && pair_total) atomicAdd(pair_total, tot);
  const int b = btrack[id];
  btrack[id] = (b == -1) ? -1 : cnt;
}

// Hot smoothing pass (MODE 0: fields + LRF + btrack + pair total; MODE 1:
// initial sigma): one 64-particle tile per workgroup of kSplit waves, each
// wave walking a slice of the tile's candidates.  Wave 0 starts from the
// self terms (src/sph_workstation.cpp:514-523) and adds the other waves'
// partial sums in wave order.
template <int MODE>
__global__ __launch_bounds__(kSplit * kWave) void k_smooth_tile(const double* __restrict__ cand,
                                                                const int32_t* __restrict__ skey,
                                                                const int32_t* __restrict__ perm,
                                                                const int32_t* __restrict__ cell_off,
                                                                const Grid* __restrict__ gridp, int n,
                                                                double rc2, double rcs, KernelW W, double kern0,
                                                                double t, double* __restrict__ F, int npad,
                                                                int32_t* __restrict__ btrack,
                                                                unsigned long long* __restrict__ pair_total)
{
  constexpr int NACC = 6;
  __shared__ double red[kSplit - 1][NACC][kWave];
  const Grid g = *gridp;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = uni((int)(threadIdx.x / kWave));   // wave-uniform
  const int k = xcd_tile(blockIdx.x, gridDim.x) * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand + (size_t)kk * kCandS;
  const double xi = me[0], yi = me[1];
  const int myrow = act ? skey[kk] / g.nc[0] : INT_MAX;

  const double self = (wave == 0) ? kern0 : 0.0;
  double sl = me[2] * self;
  double ss = me[3] * self, sB = me[4] * self, sS = me[5] * self, sQ = me[6] * self;
  int cnt = 0;

  for_segments(act, myrow, xi, yi, [&](const Seg& seg, bool in) {
    for_spans_slice(g, cell_off, seg, rcs, wave, [&](int j0, int j1) {
      // wave-uniform candidate records (scalar loads into SGPRs), two
      // requested together
      constexpr int NR = (MODE == 0) ? 7 : 3;
      auto one = [&](const double* c, int j) {
        const double dx = xi - c[0], dy = yi - c[1];
        const double r2 = dist2_nocontract(dx, dy);
        const bool nb = in && (r2 <= rc2) && (j != k);
        if (kSkipEmpty && __ballot(nb) == 0ull) return;
        const double wv = W.fast(r2);
        const double w = nb ? wv : 0.0;
        cnt += nb;
        sl += c[2] * w;
        if (MODE == 0) {
          ss += c[MODE == 0 ? 3 : 0] * w; sB += c[MODE == 0 ? 4 : 0] * w;
          sS += c[MODE == 0 ? 5 : 0] * w; sQ += c[MODE == 0 ? 6 : 0] * w;
        }
      };
      int j = j0;
      for (; j + 1 < j1; j += 2) {
        double ca[NR], cb[NR];
        for (int q = 0; q < NR; ++q) ca[q] = cand[(size_t)j * kCandS + q];
        for (int q = 0; q < NR; ++q) cb[q] = cand[(size_t)(j + 1) * kCandS + q];
        one(ca, j);
        one(cb, j + 1);
      }
      if (j < j1) {
        double ca[NR];
        for (int q = 0; q < NR; ++q) ca[q] = cand[(size_t)j * kCandS + q];
        one(ca, j);
      }
    });
  });

  if (wave > 0) {
    double* r = &red[wave - 1][0][0];
    r[0 * kWave + lane] = sl; r[1 * kWave + lane] = ss; r[2 * kWave + lane] = sB;
    r[3 * kWave + lane] = sS; r[4 * kWave + lane] = sQ; r[5 * kWave + lane] = (double)cnt;
  }
  __syncthreads();
  if (wave != 0) return;
  for (int w = 0; w < kSplit - 1; ++w) {
    sl += red[w][0][lane]; ss += red[w][1][lane]; sB += red[w][2][lane];
    sS += red[w][3][lane]; sQ += red[w][4][lane]; cnt += (int)red[w][5][lane];
  }
  if (MODE == 0) {
    // pair total for the roofline: one atomic per tile
    unsigned long long tot = act ? (unsigned long long)cnt : 0ull;
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
    if (lane == 0 && pair_total) { atomicAdd(pair_total, tot); atomicAdd(pair_total + 1, tot); }
  }
  if (!act) return;
  const int id = perm[kk];
  if (MODE == 0) {
    const int b = btrack[id];
    btrack[id] = (b == -1) ? -1 : cnt;
  }
  const double gam = F[(size_t)CCAKE_F_GAMMA * npad + id];
  F[(size_t)CCAKE_F_SIGMA_LAB * npad + id] = sl;
  // get_LRF = lab/gamma/t (src/eom_default.cpp:198-203)
  F[(size_t)CCAKE_F_SIGMA * npad + id] = sl / gam / t;
  if (MODE == 0) {
    F[(size_t)CCAKE_F_SM_S * npad + id] = ss;
    F[(size_t)CCAKE_F_SM_B * npad + id] = sB;
    F[(size_t)CCAKE_F_SM_SS * npad + id] = sS;
    F[(size_t)CCAKE_F_SM_Q * npad + id] = sQ;
    F[(size_t)CCAKE_F_S * npad + id] = ss / gam / t;
    F[(size_t)CCAKE_F_RHOB * npad + id] = sB / gam / t;
    F[(size_t)CCAKE_F_RHOS * npad + id] = sS / gam / t;
    F[(size_t)CCAKE_F_RHOQ * npad + id] = sQ / gam / t;
  }
}

static KernelW make_w(double hT)
{
  KernelW w;
  w.hinv = 1.0 / hT;
  w.norm = 10.0 / (7.0 * kPi) / pow(hT, 2);     // src/kernel.cpp:16-17
  w.norm2 = 5.0 / (14.0 * kPi) / pow(hT, 2);
  return w;
}

int launch_gather_smooth(ccake_event* ev)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k_gather_smooth, dim3((n + 255) / 256), dim3(256), 0, ev->dev->stream, ev->F, ev->npad,
                     n, ev->perm, ev->cand_s);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch_smooth(ccake_event* ev, int mode, bool count_pairs)
{
  const int n = ev->n;
  const double rc = 2.0 * ev->p.hT;
  const double rc2 = rc * rc;
  const double rcs = rc * (1.0 + 1e-12) + 1e-12;   // search radius with a rounding margin
  const KernelW W = make_w(ev->p.hT);
  const double kern0 = W.norm;                      // W(0) = norm
  const dim3 grid((n + kTile - 1) / kTile), blk(kTile);
  hipStream_t st = ev->dev->stream;
  int32_t* bt = ev->I + (size_t)CCAKE_I_BTRACK * ev->npad;
  unsigned long long* pc = count_pairs ? (unsigned long long*)ev->pair_count : nullptr;
  const dim3 tblk(kSplit * kWave);
  if (mode == 0)
    hipLaunchKernelGGL(k_smooth_tile<0>, grid, tblk, 0, st, ev->cand_s, ev->skey, ev->perm, ev->cell_off, ev->grid,
                       n, rc2, rcs, W, kern0, ev->t, ev->F, ev->npad, bt, pc);
  else if (mode == 1)
    hipLaunchKernelGGL(k_smooth_tile<1>, grid, tblk, 0, st, ev->cand_s, ev->skey, ev->perm, ev->cell_off, ev->grid,
                       n, rc2, rcs, W, kern0, ev->t, ev->F, ev->npad, bt, pc);
  else
    hipLaunchKernelGGL(k_walk<2>, grid, blk, 0, st, ev->cand_s, ev->skey, ev->perm, ev->cell_off, ev->grid, n,
                       rc2, rcs, bt, pc, nullptr, nullptr, nullptr);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

// Parity export: counts are in btrack (sticky -1 rows are recounted here).
int launch_export_pairs(ccake_event* ev, int32_t* pi, int32_t* pj, size_t cap, size_t* n_out)
{
  const int n = ev->n;
  hipStream_t st = ev->dev->stream;
  // fresh counts through btrack, which is saved and restored around the count
  std::vector<int32_t> saved_bt(n), cnt(n);
  int32_t* bt = ev->I + (size_t)CCAKE_I_BTRACK * ev->npad;
  CCAKE_HIP(hipMemcpyAsync(saved_bt.data(), bt, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipMemsetAsync(bt, 0, n * sizeof(int32_t), st));
  const int st_count = launch_smooth(ev, 2, false);   // leaves the stage's pair total alone
  if (st_count) return st_count;
  CCAKE_HIP(hipMemcpyAsync(cnt.data(), bt, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipStreamSynchronize(st));
  CCAKE_HIP(hipMemcpyAsync(bt, saved_bt.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  std::vector<int64_t> off(n + 1, 0);
  for (int i = 0; i < n; ++i) off[i + 1] = off[i] + cnt[i];
  const size_t total = (size_t)off[n];
  *n_out = total;
  if (total > cap) { CCAKE_HIP(hipStreamSynchronize(st)); return CCAKE_ERR_CAPACITY; }
  int64_t* d_off = nullptr; int32_t* d_i = nullptr; int32_t* d_j = nullptr;
  CCAKE_HIP(hipMalloc(&d_off, (n + 1) * sizeof(int64_t)));
  CCAKE_HIP(hipMalloc(&d_i, (total + 1) * sizeof(int32_t)));
  CCAKE_HIP(hipMalloc(&d_j, (total + 1) * sizeof(int32_t)));
  CCAKE_HIP(hipMemcpyAsync(d_off, off.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  const double rc = 2.0 * ev->p.hT;
  hipLaunchKernelGGL(k_walk<3>, dim3((n + kTile - 1) / kTile), dim3(kTile), 0, st, ev->cand_s, ev->skey,
                     ev->perm, ev->cell_off, ev->grid, n, rc * rc, rc * (1.0 + 1e-12) + 1e-12, bt, nullptr, d_off,
                     d_i, d_j);
  CCAKE_HIP(hipGetLastError());
  CCAKE_HIP(hipMemcpyAsync(pi, d_i, total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipMemcpyAsync(pj, d_j, total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipStreamSynchronize(st));
  CCAKE_HIP(hipFree(d_off));
  CCAKE_HIP(hipFree(d_i));
  CCAKE_HIP(hipFree(d_j));
  // canonical order: rows by i, j ascending inside a row
  for (int i = 0; i < n; ++i) std::sort(pj + off[i], pj + off[i + 1]);
  return 0;
}

}  // namespace ccake
