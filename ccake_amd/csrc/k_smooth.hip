// Smoothing pair kernel on gfx950 (owner computes, no atomics).
// Replaces SPHWorkstation::smooth_all_particle_fields
// (src/sph_workstation.cpp:494-578) and calculate_intial_sigma (582-640);
// the same walk counts neighbours (btrack, src/system_state.cpp:712-715)
// and exports the pair set for parity checks.
#include "pair_common.h"
#include <vector>
#include <algorithm>

namespace ccake {

// Candidate records in sorted order (smooth_record).
__global__ __launch_bounds__(256) void k_gather_smooth(const double* __restrict__ F, int npad, int n,
                                                       const int32_t* __restrict__ perm,
                                                       double* __restrict__ cand, int dim)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (dim == 3) smooth_record3(F, npad, perm[k], cand + (size_t)k * kCandS);
  else smooth_record(F, npad, perm[k], cand + (size_t)k * kCandS);
}

// A stage that keeps the cell sort (stage begin plus record gather in one
// kernel): gamma and v of particle perm[k] (EoM_default::update_velocity,
// src/eom_default.cpp:17-38, as k_stage_begin does), its smoothing record at
// sorted position k, and the stage's pair count cleared.  The bounding box is
// not needed: the kept grid stays.
__global__ __launch_bounds__(256) void k_stage_gather(double* __restrict__ F, int npad, int n,
                                                      const int32_t* __restrict__ perm,
                                                      double* __restrict__ cand,
                                                      unsigned long long* __restrict__ pair_count)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) pair_count[0] = 0ull;
  if (k >= n) return;
  const int i = perm[k];
  const double ux = F[(size_t)CCAKE_F_UX * npad + i], uy = F[(size_t)CCAKE_F_UY * npad + i];
  const double c = (-ux) * ux + (-uy) * uy;
  const double g = sqrt(1.0 - c);
  F[(size_t)CCAKE_F_GAMMA * npad + i] = g;
  F[(size_t)CCAKE_F_VX * npad + i] = ux / g;
  F[(size_t)CCAKE_F_VY * npad + i] = uy / g;
  smooth_record(F, npad, i, cand + (size_t)k * kCandS);
}

// Which particles carry a charge (bit 1: a smoothing record's charge entry,
// mass_X ext_X, is not zero) or a bulk pressure (bit 2: ext_Pi is not zero).
__global__ __launch_bounds__(256) void k_any_charge(const double* __restrict__ F, int npad, int n,
                                                    int32_t* __restrict__ flag)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool ch = false, bu = false;
  if (i < n) {
    auto f = [&](int c) { return F[(size_t)c * npad + i]; };
    ch = f(CCAKE_F_MASS_B) * f(CCAKE_F_EXT_B) != 0.0 || f(CCAKE_F_MASS_SS) * f(CCAKE_F_EXT_SS) != 0.0
      || f(CCAKE_F_MASS_Q) * f(CCAKE_F_EXT_Q) != 0.0;
    bu = f(CCAKE_F_EXT_BULK) != 0.0;
  }
  const int bits = (__ballot(ch) != 0ull ? 1 : 0) | (__ballot(bu) != 0ull ? 2 : 0);
  if (bits && (threadIdx.x & (kWave - 1)) == 0) atomicOr(flag, bits);
}

int launch_charge_check(ccake_event* ev, bool* any_charge, bool* any_bulk)
{
  const int n = ev->n;
  int32_t* d = ev->err + 1;
  CCAKE_HIP(hipMemsetAsync(d, 0, sizeof(int32_t), ev->stream));
  hipLaunchKernelGGL(k_any_charge, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n, d);
  CCAKE_HIP(hipGetLastError());
  int32_t h = 0;
  CCAKE_HIP(hipMemcpyAsync(&h, d, sizeof(int32_t), hipMemcpyDeviceToHost, ev->stream));
  CCAKE_HIP(hipStreamSynchronize(ev->stream));
  *any_charge = (h & 1) != 0;
  *any_bulk = (h & 2) != 0;
  return 0;
}

int launch_stage_gather(ccake_event* ev)
{
  if (ev->p.dim == 3) return launch3_stage_gather(ev);
  const int n = ev->n_live;
  hipLaunchKernelGGL(k_stage_gather, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n,
                     ev->perm, ev->cand_s, (unsigned long long*)ev->pair_count);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

struct KernelW {
  double norm, norm2;   // src/kernel.cpp:16-17 (D = 2): W(0) = norm, norm / norm2 = 4
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
  const Grid g0 = *gridp;
  const int lane = threadIdx.x;
  const int k = blockIdx.x * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand + (size_t)kk * kCandS;
  const double xi = me[0], yi = me[1];
  const int myrow = act ? skey[kk] / g0.stride : INT_MAX;
  const int id = act ? perm[kk] : 0;
  int cnt = 0;
  int64_t wpos = 0;
  if (MODE == 3 && act) wpos = row_off[id];

  for_segments(act, myrow, xi, yi, [&](const Seg& seg, bool in) {
    const Grid& g = gridp[event_index(gridp, g0, seg.row)];
    for_spans(g, cell_off, seg, rcs, [&](int j) {
      const double* c = cand + (size_t)j * kCandS;
      const double dx = xi - c[0], dy = yi - c[1];
      const double r2 = dist2_nocontract(dx, dy);
      const bool nb = in && (r2 <= rc2) && (j != k);
      if (MODE == 2) cnt += nb;
      if (MODE == 3 && nb) { out_i[wpos] = id; out_j[wpos] = perm[j]; ++wpos; }
    });
  });

  if (MODE == 3) return;
  // the wave's pair total, every lane taking part (lanes past n add 0)
  unsigned long long tot = act ? (unsigned long long)cnt : 0ull;
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
  if (lane == 0 && pair_total) atomicAdd(pair_total, tot);
  if (!act) return;
  const int b = btrack[id];
  btrack[id] = (b == -1) ? -1 : cnt;
}

// Hot smoothing pass, pair part (MODE 0: fields + LRF + btrack + pair
// total; MODE 1: initial sigma).  One single-wave workgroup per (tile,
// slice) unit walks its slice of the tile's candidates and stores the
// slice's partial sums of m_b w(q) (the self pair included, W = norm2 w);
// k_smooth_sum adds them in slice order.
// MODE 0: sigma_lab, the entropy and the three charges; 3: the same sums for
// an event whose charges are zero on every particle (the charge sums are
// written as the exact zeros the full walk would give, and the records are
// read four doubles at a time); 1: sigma_lab only (initial sigma).
template <int MODE>
constexpr int smooth_nacc() { return (MODE == 0 || MODE == 3) ? 6 : 2; }

// The walk of one (tile, slice) unit (one wave); fin(acc) receives the
// slice's sums (sigma_lab, the pair count, then s, B, S, Q).
template <int MODE, class Fin>
__device__ __forceinline__ void smooth_walk(const double* __restrict__ cand, const int32_t* __restrict__ skey,
                                            const int32_t* __restrict__ cell_off, const Grid* __restrict__ gridp,
                                            int n, double rc2, double rcs, double hinv2, int tile, int slice,
                                            int nslices, Fin&& fin)
{
  const Grid g0 = *gridp;
  const int lane = threadIdx.x & (kWave - 1);
  const int k = tile * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand + (size_t)kk * kCandS;
  const double xi = me[0], yi = me[1];
  const int myrow = act ? skey[kk] / g0.stride : INT_MAX;

  double sl = 0.0, ss = 0.0, sB = 0.0, sS = 0.0, sQ = 0.0;
  int cnt = 0;   // r2 <= rc2, the self pair included

  for_segments(act, myrow, xi, yi, [&](const Seg& seg, bool in) {
    const Grid& g = gridp[event_index(gridp, g0, seg.row)];
    // lanes of other cell rows sit this segment's walk out
    for_spans_slice(g, cell_off, seg, rcs, slice, nslices, in, [&](int j0, int j1) {
      // wave-uniform candidate records (scalar loads into SGPRs)
      constexpr int NR = (MODE == 0) ? 7 : (MODE == 3 ? 4 : 3);
      walk_records<NR, kCandS, 2>(cand, xi, yi, j0, j1,
                                  [&](const double* c, double dx, double dy, int j, bool valid) {
        const double r2 = dist2_nocontract(dx, dy);
        const bool nb = valid && r2 <= rc2;
        cnt += nb ? 1 : 0;
        if (kSkipEmpty && wave_none(nb)) return;
        double q, iq;
        q_and_inverse(r2, hinv2, q, iq);
        const double w = valid ? spline_w(q) : 0.0;
        sl += c[2] * w;
        if (MODE == 0 || MODE == 3) ss += c[MODE == 1 ? 0 : 3] * w;
        if (MODE == 0) {
          sB += c[MODE == 0 ? 4 : 0] * w;
          sS += c[MODE == 0 ? 5 : 0] * w; sQ += c[MODE == 0 ? 6 : 0] * w;
        }
      });
    });
  });
  const double acc[6] = { sl, (double)cnt, ss, sB, sS, sQ };
  fin(acc);
}

template <int MODE>
__global__ __launch_bounds__(kWave) void k_smooth_slice(const double* __restrict__ cand,
                                                        const int32_t* __restrict__ skey,
                                                        const int32_t* __restrict__ cell_off,
                                                        const Grid* __restrict__ gridp, int n,
                                                        double rc2, double rcs, double hinv2, int nslices,
                                                        double* __restrict__ part, int npad)
{
  constexpr int NACC = smooth_nacc<MODE>();
  const int unit = xcd_tile(blockIdx.x, gridDim.x);
  const int tile = unit / nslices;
  const int slice = unit - tile * nslices;
  smooth_walk<MODE>(cand, skey, cell_off, gridp, n, rc2, rcs, hinv2, tile, slice, nslices, [&](const double* acc) {
    double* p = part + (size_t)slice * NACC * npad + tile * kTile + threadIdx.x;
    #pragma unroll
    for (int q = 0; q < NACC; ++q) p[(size_t)q * npad] = acc[q];
  });
}

// The W slices of a tile in one workgroup of W waves, their sums added in
// slice order in LDS and stored as one slice's (as k_grad_tile).
template <int MODE, int W>
__global__ __launch_bounds__(kWave * W) void k_smooth_tile(const double* __restrict__ cand,
                                                           const int32_t* __restrict__ skey,
                                                           const int32_t* __restrict__ cell_off,
                                                           const Grid* __restrict__ gridp, int n,
                                                           double rc2, double rcs, double hinv2,
                                                           double* __restrict__ part, int npad)
{
  constexpr int NACC = smooth_nacc<MODE>();
  __shared__ double red[W][NACC][kWave];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int slice = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);   // wave-uniform
  const int lane = threadIdx.x & (kWave - 1);
  smooth_walk<MODE>(cand, skey, cell_off, gridp, n, rc2, rcs, hinv2, tile, slice, W, [&](const double* acc) {
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

// Sum pass: the slices' partial sums in slice order (the self terms of
// src/sph_workstation.cpp:514-523 are among them), scaled by norm2, then the
// lab and LRF fields by particle id.  S is the slice count, so every load
// of a particle is in flight at once.
constexpr int kSumThreads = 256;   // small workgroups spread the pass over every CU
template <int MODE, int S>
__global__ __launch_bounds__(kSumThreads) void k_smooth_sum(const int32_t* __restrict__ perm, int n, double norm2,
                                                            double t, const double* __restrict__ part, int npad,
                                                            double* __restrict__ F,
                                                            int32_t* __restrict__ btrack,
                                                            unsigned long long* __restrict__ pair_total)
{
  constexpr int NACC = smooth_nacc<MODE>();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = k < n;
  double acc[NACC];
  #pragma unroll
  for (int f = 0; f < NACC; ++f) acc[f] = 0.0;
  if (act) {
    #pragma unroll
    for (int sli = 0; sli < S; ++sli)
      #pragma unroll
      for (int f = 0; f < NACC; ++f) acc[f] += part[((size_t)sli * NACC + f) * npad + k];
  }
  const int cnt = act ? (int)acc[1] - 1 : 0;    // neighbours, self excluded
  if (MODE == 0 && pair_total) {
    // pair total for the roofline: one pair of atomics per workgroup (a
    // burst of same-address atomics from every wave costs more than the
    // whole sum pass)
    __shared__ unsigned long long wsum[kSumThreads / kWave];
    unsigned long long tot = (unsigned long long)cnt;
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
    if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long b = 0;
      for (int w = 0; w < kSumThreads / kWave; ++w) b += wsum[w];
      if (b) { atomicAdd(pair_total, b); atomicAdd(pair_total + 1, b); }
    }
  }
  if (!act) return;
  const int id = perm[k];
  if (MODE == 0) {
    const int b = btrack[id];
    btrack[id] = (b == -1) ? -1 : cnt;
  }
  const double sl = acc[0] * norm2;
  const double gam = F[(size_t)CCAKE_F_GAMMA * npad + id];
  F[(size_t)CCAKE_F_SIGMA_LAB * npad + id] = sl;
  // get_LRF = lab/gamma/t (src/eom_default.cpp:198-203)
  F[(size_t)CCAKE_F_SIGMA * npad + id] = sl / gam / t;
  if (MODE == 0) {
    const double ss = acc[MODE == 0 ? 2 : 0] * norm2, sB = acc[MODE == 0 ? 3 : 0] * norm2;
    const double sS = acc[MODE == 0 ? 4 : 0] * norm2, sQ = acc[MODE == 0 ? 5 : 0] * norm2;
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
  w.norm = 10.0 / (7.0 * kPi) / pow(hT, 2);     // src/kernel.cpp:16-17
  w.norm2 = 5.0 / (14.0 * kPi) / pow(hT, 2);
  return w;
}

int launch_gather_smooth(ccake_event* ev)
{
  const int n = ev->n_live;
  hipLaunchKernelGGL(k_gather_smooth, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad,
                     n, ev->perm, ev->cand_s, ev->p.dim);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

// Pair part of a smoothing pass (modes 0, 1) or the neighbour count walk
// (mode 2, which is complete by itself).
int launch_smooth_pairs(ccake_event* ev, int mode, bool count_pairs)
{
  if (ev->p.dim == 3) return launch3_smooth_pairs(ev, mode, count_pairs);
  const int n = ev->n_live;          // sorted positions of the live events' particles
  const double rc = 2.0 * ev->p.hT;
  const double rc2 = rc * rc;
  const double rcs = search_radius(ev);
  const int ntiles = (n + kTile - 1) / kTile;
  hipStream_t st = ev->stream;
  if (mode == 2) {
    int32_t* bt = ev->I + (size_t)CCAKE_I_BTRACK * ev->npad;
    unsigned long long* pc = count_pairs ? (unsigned long long*)ev->pair_count : nullptr;
    hipLaunchKernelGGL(k_walk<2>, dim3(ntiles), dim3(kTile), 0, st, ev->cand_s, ev->skey, ev->perm, ev->cell_off,
                       ev->grid, n, rc2, rcs, bt, pc, nullptr, nullptr, nullptr);
    return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
  }
  if (mode == 0 && ev->p.pair_precision == 1) return launch_smooth_pairs32(ev);
  const double hinv2 = 1.0 / (ev->p.hT * ev->p.hT);
  const int W = lds_slices(ev);
  if (W) {
    // the tile's slices reduced in LDS: the sum pass sees one slice
    ev->smooth_sum_slices = 1;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(ntiles), dim3(kWave * W), 0, st, ev->cand_s, ev->skey, ev->cell_off, ev->grid, n,
                         rc2, rcs, hinv2, ev->part, ev->npad);
    };
    const int m = (mode == 0 && ev->no_charges) ? 3 : mode;
    if (W == 8) { if (m == 3) go(k_smooth_tile<3, 8>); else if (m == 0) go(k_smooth_tile<0, 8>); else go(k_smooth_tile<1, 8>); }
    else { if (m == 3) go(k_smooth_tile<3, 4>); else if (m == 0) go(k_smooth_tile<0, 4>); else go(k_smooth_tile<1, 4>); }
    return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
  }
  ev->smooth_sum_slices = ev->nslices;
  const dim3 grid(ntiles * ev->nslices), blk(kWave);
  if (mode == 0 && ev->no_charges)
    hipLaunchKernelGGL(k_smooth_slice<3>, grid, blk, 0, st, ev->cand_s, ev->skey, ev->cell_off, ev->grid, n, rc2,
                       rcs, hinv2, ev->nslices, ev->part, ev->npad);
  else if (mode == 0)
    hipLaunchKernelGGL(k_smooth_slice<0>, grid, blk, 0, st, ev->cand_s, ev->skey, ev->cell_off, ev->grid, n, rc2,
                       rcs, hinv2, ev->nslices, ev->part, ev->npad);
  else
    hipLaunchKernelGGL(k_smooth_slice<1>, grid, blk, 0, st, ev->cand_s, ev->skey, ev->cell_off, ev->grid, n, rc2,
                       rcs, hinv2, ev->nslices, ev->part, ev->npad);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

template <int MODE, int S>
static void smooth_sum_s(ccake_event* ev, double norm2, int32_t* bt, unsigned long long* pc)
{
  const int n = ev->n_live;
  const dim3 grid((n + kSumThreads - 1) / kSumThreads), blk(kSumThreads);
  hipLaunchKernelGGL((k_smooth_sum<MODE, S>), grid, blk, 0, ev->stream, ev->perm, n, norm2, ev->t, ev->part,
                     ev->npad, ev->F, bt, pc);
}

template <int MODE, int S = 1>
static int smooth_sum_dispatch(ccake_event* ev, double norm2, int32_t* bt, unsigned long long* pc)
{
  if constexpr (S > kMaxSlices) {
    set_error("slice count out of range");
    return CCAKE_ERR_INVALID;
  } else {
    if (ev->smooth_sum_slices == S) { smooth_sum_s<MODE, S>(ev, norm2, bt, pc); return 0; }
    return smooth_sum_dispatch<MODE, S + 1>(ev, norm2, bt, pc);
  }
}

// Sum part of a smoothing pass (modes 0, 1).
double smooth_norm2(double hT) { return make_w(hT).norm2; }

int launch_smooth_sum(ccake_event* ev, int mode, bool count_pairs)
{
  const double norm2 = smooth_norm2_event(ev);
  int32_t* bt = ev->I + (size_t)CCAKE_I_BTRACK * ev->npad;
  unsigned long long* pc = count_pairs ? (unsigned long long*)ev->pair_count : nullptr;
  const int rc = (mode == 0) ? smooth_sum_dispatch<0>(ev, norm2, bt, pc) : smooth_sum_dispatch<1>(ev, norm2, bt, pc);
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch_smooth(ccake_event* ev, int mode, bool count_pairs)
{
  int rc = launch_smooth_pairs(ev, mode, count_pairs);
  if (rc || mode == 2) return rc;
  return launch_smooth_sum(ev, mode, count_pairs);
}

// Parity export: counts are in btrack (sticky -1 rows are recounted here).
int launch_export_pairs(ccake_event* ev, int32_t* pi, int32_t* pj, size_t cap, size_t* n_out)
{
  if (ev->p.dim == 3) return launch3_export_pairs(ev, pi, pj, cap, n_out);
  const int n = ev->n;
  hipStream_t st = ev->stream;
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
  // the restore reads saved_bt: it completes before any return below
  CCAKE_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> off(n + 1, 0);
  for (int i = 0; i < n; ++i) off[i + 1] = off[i] + cnt[i];
  const size_t total = (size_t)off[n];
  *n_out = total;
  if (total > cap) return CCAKE_ERR_CAPACITY;
  struct DevBuf {                     // freed on every return
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
  } b_off, b_i, b_j;
  CCAKE_HIP(hipMalloc(&b_off.p, (n + 1) * sizeof(int64_t)));
  CCAKE_HIP(hipMalloc(&b_i.p, (total + 1) * sizeof(int32_t)));
  CCAKE_HIP(hipMalloc(&b_j.p, (total + 1) * sizeof(int32_t)));
  int64_t* d_off = static_cast<int64_t*>(b_off.p);
  int32_t* d_i = static_cast<int32_t*>(b_i.p);
  int32_t* d_j = static_cast<int32_t*>(b_j.p);
  CCAKE_HIP(hipMemcpyAsync(d_off, off.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  const double rc = 2.0 * ev->p.hT;
  // the live events' rows (an ended event of a batch exports none)
  hipLaunchKernelGGL(k_walk<3>, dim3((ev->n_live + kTile - 1) / kTile), dim3(kTile), 0, st, ev->cand_s, ev->skey,
                     ev->perm, ev->cell_off, ev->grid, ev->n_live, rc * rc, search_radius(ev), bt, nullptr, d_off,
                     d_i, d_j);
  CCAKE_HIP(hipGetLastError());
  CCAKE_HIP(hipMemcpyAsync(pi, d_i, total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipMemcpyAsync(pj, d_j, total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipStreamSynchronize(st));
  // canonical order: rows by i, j ascending inside a row
  for (int i = 0; i < n; ++i) std::sort(pj + off[i], pj + off[i + 1]);
  return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free pair export.  The device side of the export (the rows'
// neighbour counts, their offsets and the pair rows) lives in `scratch`, a
// file-scope device array of fixed size, instead of buffers allocated per
// call, so that callers that cannot allocate device memory can take the
// pair set.  Each device holds its own copy at its own address, which
// ccake_dev_create obtains once per device (scratch_address) and keeps in
// the device handle; host code uses that address, never the array's name.
// Calls on one device share the array: two calls that overlap in time on the
// same device race.
constexpr size_t kScratchBytes = size_t(128) << 20;                 // 128 MiB per device
constexpr int kScratchMaxParticles = 1 << 20;
constexpr size_t kScratchCountOff = 0;                                // int32 counts [n]
constexpr size_t kScratchRowOffOff = size_t(kScratchMaxParticles) * 4;   // int64 offsets [n + 1]
constexpr size_t kScratchRowsOff = kScratchRowOffOff + (size_t(kScratchMaxParticles) + 1) * 8 + 248;  // 256-aligned
constexpr size_t kScratchMaxPairs = (kScratchBytes - kScratchRowsOff) / 8;   // int32 (i, j) per pair
static_assert(kScratchRowsOff % 256 == 0, "scratch row area alignment");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

int scratch_address(void** out)
{
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  *out = p;
  return 0;
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t* __restrict__ a, int n, int32_t v)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Row offsets off[0..n] (int64) of the counts cnt[0..n) by one workgroup:
// thread t owns the run [t*per, min((t+1)*per, n)), per = ceil(n / 1024).
constexpr int kRowScanThreads = 1024;
__global__ __launch_bounds__(kRowScanThreads) void k_row_offsets(const int32_t* __restrict__ cnt, int n,
                                                                 int64_t* __restrict__ off)
{
  __shared__ int64_t wtot[kRowScanThreads / kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int per = (n + kRowScanThreads - 1) / kRowScanThreads;
  const int b = threadIdx.x * per;
  const int e = min(b + per, n);
  int64_t s = 0;
  for (int i = b; i < e; ++i) s += cnt[i];
  int64_t x = s;                               // inclusive scan of the run sums within the wave
  for (int o = 1; o < kWave; o <<= 1) {
    const int64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) wtot[w] = x;
  __syncthreads();
  int64_t run = x - s;
  for (int q = 0; q < w; ++q) run += wtot[q];
  for (int i = b; i < e; ++i) { off[i] = run; run += cnt[i]; }
  if (threadIdx.x == kRowScanThreads - 1) off[n] = run;   // the last thread's run ends at n
}

// The pair export of launch_export_pairs with every device buffer in
// `scratch` (no allocation, host or device): counts, offsets and rows.
int launch_export_pairs_noalloc(ccake_event* ev, int32_t* pi, int32_t* pj, size_t cap, size_t* n_out)
{
  const int n = ev->n;
  if (n > kScratchMaxParticles) {
    set_error("ccake_export_pairs_noalloc: at most 1,048,576 particles (the scratch array's size)");
    return CCAKE_ERR_CAPACITY;
  }
  unsigned char* base = static_cast<unsigned char*>(ev->dev->scratch);
  if (!base) { set_error("ccake_export_pairs_noalloc: no scratch address for this device"); return CCAKE_ERR_INVALID; }
  int32_t* cnt = reinterpret_cast<int32_t*>(base + kScratchCountOff);
  int64_t* off = reinterpret_cast<int64_t*>(base + kScratchRowOffOff);
  int32_t* ri = reinterpret_cast<int32_t*>(base + kScratchRowsOff);
  hipStream_t st = ev->stream;
  const double rc = 2.0 * ev->p.hT;
  const int nl = ev->n_live;            // the walks cover the live events' sorted positions
  const int ntiles = (nl + kTile - 1) / kTile;
  // counts (self excluded) into scratch, cleared first: the last call's
  // contents are still there
  hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, st, cnt, n, 0);
  hipLaunchKernelGGL(k_walk<2>, dim3(ntiles), dim3(kTile), 0, st, ev->cand_s, ev->skey, ev->perm, ev->cell_off,
                     ev->grid, nl, rc * rc, search_radius(ev), cnt, nullptr, nullptr, nullptr, nullptr);
  hipLaunchKernelGGL(k_row_offsets, dim3(1), dim3(kRowScanThreads), 0, st, cnt, n, off);
  CCAKE_HIP(hipGetLastError());
  int64_t total = 0;
  CCAKE_HIP(hipMemcpyAsync(&total, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipStreamSynchronize(st));
  *n_out = (size_t)total;
  if ((size_t)total > kScratchMaxPairs) {
    set_error("ccake_export_pairs_noalloc: more pairs than the scratch array holds (15,204,320)");
    return CCAKE_ERR_CAPACITY;
  }
  if ((size_t)total > cap) return CCAKE_ERR_CAPACITY;
  int32_t* rj = ri + total;
  hipLaunchKernelGGL(k_walk<3>, dim3(ntiles), dim3(kTile), 0, st, ev->cand_s, ev->skey, ev->perm, ev->cell_off,
                     ev->grid, nl, rc * rc, search_radius(ev), cnt, nullptr, off, ri, rj);
  CCAKE_HIP(hipGetLastError());
  if (total) {
    CCAKE_HIP(hipMemcpyAsync(pi, ri, (size_t)total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CCAKE_HIP(hipMemcpyAsync(pj, rj, (size_t)total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  }
  CCAKE_HIP(hipStreamSynchronize(st));
  // canonical order: rows by i (already), j ascending inside a row
  for (size_t a = 0; a < (size_t)total;) {
    size_t b = a + 1;
    while (b < (size_t)total && pi[b] == pi[a]) ++b;
    std::sort(pj + a, pj + b);
    a = b;
  }
  return 0;
}

}  // namespace ccake
