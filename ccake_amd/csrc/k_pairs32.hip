// Mixed-precision pair kernels (pair_precision 1): the smoothing and
// gradient pair sums of a time step with FP32 pair terms and compensated
// FP32 accumulation.  Replaces, for that precision, the same reference loops
// as k_smooth_slice and k_grad_slice: SPHWorkstation::smooth_all_particle_fields
// (src/sph_workstation.cpp:525-553) and smooth_all_particle_gradients
// (src/sph_workstation.cpp:796-986).
//
// What stays FP64: positions, the offsets dx, dy and the pair predicate
// fl(fl(dx^2) + fl(dy^2)) <= rc^2 (so the pair set and btrack are the FP64
// path's, bit for bit), the slices' sum pass and owner terms, and everything
// outside the pair loops (EoS inversion, EoM, RK accumulation, shear
// update).  What is FP32: q = r/h and 1/q (v_rsq_f32), the spline, the
// neighbour's record entries and the products that form each pair's terms.
//
// Accumulation.  Each of a lane's sums is kept as a block sum b and a
// compensated pair (hi, lo): the pair terms are added to b, and every
// FOLD candidates (16 by default; and at the end of every span) b is folded into
// (hi, lo) with the error-free TwoSum (Knuth):
//   s = hi + b;  z = s - hi;  err = (hi - (s - z)) + (b - z);  hi = s;  lo += err.
// A block holds at most FOLD terms, so its rounding is that of a short
// FP32 sum; the running total carries its rounding error in lo.  The slice's
// partial sum written for the sum pass is (double)hi + (double)lo.
#include <cstdlib>

#include "pair_common.h"
#include "grad_sum.h"

namespace ccake {

// Default candidates per FP32 block between folds.  On the 200k lumpy event
// to freeze-out, 16, 32 and 64 gave 5.62, 5.49 and 5.43 ms per step and
// freeze-out errors against FP64 (largest per particle, same snapshot) of
// 4.0e-7, 7.9e-7 and 1.9e-6 in u^mu, with the charge densities' whole-field
// errors 6.5e-7, 8.6e-7 and 1.2e-6.  In a batch of 512 events of 24k to
// 216k particles, 32 left one event's u^mu 1.8e-6 off its FP64 run alone;
// 16 halves the errors for 2 % of the step and keeps T, u^mu and the
// charges within 1e-6.
constexpr int kFoldEvery = 16;

struct Comp {
  float b = 0.0f, hi = 0.0f, lo = 0.0f;
  __device__ __forceinline__ void add(float x) { b += x; }
  __device__ __forceinline__ void fold()
  {
    #pragma clang fp contract(off)
    const float s = hi + b;
    const float z = s - hi;
    const float e = (hi - (s - z)) + (b - z);
    hi = s;
    lo += e;
    b = 0.0f;
  }
  __device__ __forceinline__ double value() const { return (double)hi + (double)lo + (double)b; }
};

__device__ __forceinline__ double get_f64(const uint32_t* c)
{
  return __hiloint2double((int)c[1], (int)c[0]);
}
__device__ __forceinline__ float get_f32(const uint32_t* c) { return __uint_as_float(c[0]); }

// FP32 records in sorted order from the FP64 ones (positions copied, the
// rest rounded); KIND 0 smoothing, 1 gradients.
template <int KIND>
__global__ __launch_bounds__(256) void k_records32(const double* __restrict__ cand, int n,
                                                   uint32_t* __restrict__ out)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (KIND == 0) smooth_record32(cand + (size_t)k * kCandS, out + (size_t)k * kCandS32);
  else grad_record32(cand + (size_t)k * kCandG, out + (size_t)k * kCandG32);
}

// walk_records over FP32 records: NR dwords of each record (wave-uniform
// scalar loads), U records per step with the next U requested while the
// current ones are evaluated (the same ordering as walk_records), and
// fold() after every FOLD candidates and at the end of the span.
template <int NR, int STRIDE, int U, int FOLD, class Body, class Fold>
__device__ __forceinline__ void walk_records32(const uint32_t* __restrict__ cand, double xi, double yi, int j0,
                                               int j1, Body&& body, Fold&& fold)
{
  static_assert(U == 1 || U == 2 || U == 4, "one, two or four records per step");
  static_assert(FOLD % U == 0, "a fold falls between steps");
  if (j0 >= j1) return;
  uint32_t cur[U][NR];
  #pragma unroll
  for (int u = 0; u < U; ++u) {
    const int ju = j0 + u < j1 ? j0 + u : j1 - 1;
    #pragma unroll
    for (int q = 0; q < NR; ++q) cur[u][q] = cand[(size_t)ju * STRIDE + q];
  }
  int since = 0;
  for (int j = j0; j < j1; j += U) {
    double dx[U], dy[U];
    #pragma unroll
    for (int u = 0; u < U; ++u) { dx[u] = xi - get_f64(cur[u]); dy[u] = yi - get_f64(cur[u] + 2); }
    __builtin_amdgcn_sched_barrier(0);
    uint32_t nxt[U][NR];
    #pragma unroll
    for (int u = 0; u < U; ++u) {
      const int ju = j + U + u < j1 ? j + U + u : j1 - 1;
      #pragma unroll
      for (int q = 0; q < NR; ++q) nxt[u][q] = cand[(size_t)ju * STRIDE + q];
    }
    __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
    for (int u = 0; u < U; ++u) body(cur[u], dx[u], dy[u], j + u, j + u < j1);
    since += U;
    if (since == FOLD) { fold(); since = 0; }
    #pragma unroll
    for (int u = 0; u < U; ++u)
      #pragma unroll
      for (int q = 0; q < NR; ++q) cur[u][q] = nxt[u][q];
  }
  fold();
}

// q = r/h and 1/q in FP32 from the FP64 r^2: v_rsq_f32 of q^2 (1e-30 added
// so that coincident particles get q ~ 1e-15 and a zero gradient).
__device__ __forceinline__ void q_and_inverse32(double r2, float hinv2, float& q, float& iq)
{
  const float x = fmaf((float)r2, hinv2, 1e-30f);
  iq = __builtin_amdgcn_rsqf(x);
  q = x * iq;
}
__device__ __forceinline__ float spline_w32(float q)
{
  const float t = fmaxf(2.0f - q, 0.0f), u = fmaxf(1.0f - q, 0.0f);
  return fmaf(-4.0f, u * u * u, t * t * t);
}
__device__ __forceinline__ float spline_g32(float q, float iq)
{
  const float t = fmaxf(2.0f - q, 0.0f), u = fmaxf(1.0f - q, 0.0f);
  return fmaf(-4.0f, u * u, t * t) * iq;
}

// Smoothing pairs, MODE 0 (charges) or 3 (no charges: the charge sums are
// the exact zeros the full walk gives).  Same partial-sum layout as
// k_smooth_slice: sigma_lab, neighbour count (FP64 predicate), s, B, S, Q.
// W = 0: one single-wave workgroup per (tile, slice) unit, each slice's sums
// stored; W > 0: the W slices of a tile in one workgroup, added in slice
// order in LDS and stored as one slice's (as k_smooth_tile).
template <int MODE, int U, int FOLD, int W>
__global__ __launch_bounds__(kWave * (W > 0 ? W : 1)) void k_smooth_slice32(const double* __restrict__ cand64,
                                                          const uint32_t* __restrict__ cand,
                                                          const int32_t* __restrict__ skey,
                                                          const int32_t* __restrict__ cell_off,
                                                          const Grid* __restrict__ gridp, int n, double rc2,
                                                          double rcs, float hinv2, int nslices,
                                                          double* __restrict__ part, int npad)
{
  static_assert(MODE == 0 || MODE == 3, "time-step smoothing modes");
  constexpr int NACC = 6;
  const Grid g0 = *gridp;
  const int lane = threadIdx.x & (kWave - 1);
  int tile, slice;
  if constexpr (W == 0) {
    const int unit = xcd_tile(blockIdx.x, gridDim.x);
    tile = unit / nslices;
    slice = unit - tile * nslices;
  } else {
    tile = xcd_tile(blockIdx.x, gridDim.x);
    slice = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);   // wave-uniform
    nslices = W;
  }
  const int k = tile * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double xi = cand64[(size_t)kk * kCandS], yi = cand64[(size_t)kk * kCandS + 1];
  const int myrow = act ? skey[kk] / g0.stride : INT_MAX;

  Comp sl, ss, sB, sS, sQ;
  int cnt = 0;

  for_segments(act, myrow, xi, yi, [&](const Seg& seg, bool in) {
    const Grid& g = gridp[event_index(gridp, g0, seg.row)];
    for_spans_slice(g, cell_off, seg, rcs, slice, nslices, in, [&](int j0, int j1) {
      constexpr int NR = (MODE == 0) ? 9 : 6;
      walk_records32<NR, kCandS32, U, FOLD>(cand, xi, yi, j0, j1,
                                      [&](const uint32_t* c, double dx, double dy, int j, bool valid) {
        const double r2 = dist2_nocontract(dx, dy);
        const bool nb = valid && r2 <= rc2;
        cnt += nb ? 1 : 0;
        if (kSkipEmpty && wave_none(nb)) return;
        float q, iq;
        q_and_inverse32(r2, hinv2, q, iq);
        const float w = valid ? spline_w32(q) : 0.0f;
        sl.add(get_f32(c + 4) * w);
        ss.add(get_f32(c + 5) * w);
        if (MODE == 0) {
          sB.add(get_f32(c + (MODE == 0 ? 6 : 4)) * w);
          sS.add(get_f32(c + (MODE == 0 ? 7 : 4)) * w);
          sQ.add(get_f32(c + (MODE == 0 ? 8 : 4)) * w);
        }
      }, [&]() {
        sl.fold(); ss.fold();
        if (MODE == 0) { sB.fold(); sS.fold(); sQ.fold(); }
      });
    });
  });

  const double v[NACC] = { sl.value(), (double)cnt, ss.value(), MODE == 0 ? sB.value() : 0.0,
                           MODE == 0 ? sS.value() : 0.0, MODE == 0 ? sQ.value() : 0.0 };
  if constexpr (W == 0) {
    double* p = part + (size_t)slice * NACC * npad + k;
    #pragma unroll
    for (int q = 0; q < NACC; ++q) p[(size_t)q * npad] = v[q];
  } else {
    __shared__ double red[W > 0 ? W : 1][NACC][kWave];
    #pragma unroll
    for (int q = 0; q < NACC; ++q) red[slice][q][lane] = v[q];
    __syncthreads();
    for (int t = threadIdx.x; t < NACC * kWave; t += kWave * W) {
      const int q = t / kWave, l = t - q * kWave;
      double s = 0.0;
      #pragma unroll
      for (int w = 0; w < W; ++w) s += red[w][q][l];
      part[(size_t)q * npad + tile * kTile + l] = s;
    }
  }
}

// Gradient pairs: the 17 (12 without shear) neighbour-side sums of
// w = m_b grad W_ab / C of k_grad_slice, same partial-sum layout; W as in
// k_smooth_slice32.
// Register budget of the FP32 gradient walk: 4 waves per SIMD (128 VGPRs,
// 11 values spilled) measured 4.36 ms per mixed step on the 200k event
// against 4.56 at the compiler's choice (147 VGPRs, 3 waves).
#ifndef CCAKE_GRAD32_WAVES_PER_SIMD
#define CCAKE_GRAD32_WAVES_PER_SIMD 4
#endif
template <bool SHEAR, bool BULK, int FOLD, int W>
__global__ __launch_bounds__(kWave * (W > 0 ? W : 1)) __attribute__((amdgpu_waves_per_eu(CCAKE_GRAD32_WAVES_PER_SIMD)))
void k_grad_slice32(const double* __restrict__ cand64,
                                                        const uint32_t* __restrict__ cand,
                                                        const int32_t* __restrict__ skey,
                                                        const int32_t* __restrict__ cell_off,
                                                        const Grid* __restrict__ gridp, int n, double rc2,
                                                        double rcs, float hinv2, int nslices,
                                                        double* __restrict__ part, int npad)
{
  constexpr int NACC = grad_nacc<SHEAR>();
  const Grid g0 = *gridp;
  const int lane = threadIdx.x & (kWave - 1);
  int tile, slice;
  if constexpr (W == 0) {
    const int unit = xcd_tile(blockIdx.x, gridDim.x);
    tile = unit / nslices;
    slice = unit - tile * nslices;
  } else {
    tile = xcd_tile(blockIdx.x, gridDim.x);
    slice = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);   // wave-uniform
    nslices = W;
  }
  const int k = tile * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand64 + (size_t)kk * kCandG;
  const double xa = me[0], ya = me[1];
  const float vax = (float)me[6], vay = (float)me[7];
  const int myrow = act ? skey[kk] / g0.stride : INT_MAX;

  Comp acc[NACC];

  for_segments(act, myrow, xa, ya, [&](const Seg& seg, bool in) {
    const Grid& g = gridp[event_index(gridp, g0, seg.row)];
    for_spans_slice(g, cell_off, seg, rcs, slice, nslices, in, [&](int j0, int j1) {
      constexpr int NR = SHEAR ? 15 : 10;
      walk_records32<NR, kCandG32, 2, FOLD>(cand, xa, ya, j0, j1,
                                      [&](const uint32_t* c, double dx, double dy, int j, bool valid) {
        const double r2 = dist2_nocontract(dx, dy);
        if (kSkipEmpty && wave_none(valid && r2 <= rc2)) return;
        float q, iq;
        q_and_inverse32(r2, hinv2, q, iq);
        const float gm = valid ? get_f32(c + 4) * spline_g32(q, iq) : 0.0f;
        const float wx = gm * (float)dx, wy = gm * (float)dy;
        const float pb = get_f32(c + 5), eb = get_f32(c + 6);
        const float vbx = get_f32(c + 8), vby = get_f32(c + 9);
        acc[0].add(wx); acc[1].add(wy);
        acc[2].add(pb * wx); acc[3].add(pb * wy);
        acc[4].add(eb * wx); acc[5].add(eb * wy);
        if (BULK) { const float bb = get_f32(c + 7); acc[6].add(bb * wx); acc[7].add(bb * wy); }
        acc[8].add(vbx * wx); acc[9].add(vby * wx);
        acc[10].add(vbx * wy); acc[11].add(vby * wy);
        if (SHEAR) {
          const float cv = wx * vax + wy * vay;
          acc[SHEAR ? 12 : 0].add(cv);
          acc[SHEAR ? 13 : 0].add(cv * get_f32(c + 10));
          acc[SHEAR ? 14 : 0].add(cv * get_f32(c + 11));
          acc[SHEAR ? 15 : 0].add(wx * get_f32(c + 12) + wy * get_f32(c + 13));
          acc[SHEAR ? 16 : 0].add(wx * get_f32(c + 13) + wy * get_f32(c + 14));
        }
      }, [&]() {
        #pragma unroll
        for (int q = 0; q < NACC; ++q)
          if (BULK || (q != 6 && q != 7)) acc[q].fold();
      });
    });
  });

  if constexpr (W == 0) {
    double* p = part + (size_t)slice * NACC * npad + k;
    #pragma unroll
    for (int q = 0; q < NACC; ++q) p[(size_t)q * npad] = acc[q].value();
  } else {
    __shared__ double red[W > 0 ? W : 1][NACC][kWave];
    #pragma unroll
    for (int q = 0; q < NACC; ++q) red[slice][q][lane] = acc[q].value();
    __syncthreads();
    for (int t = threadIdx.x; t < NACC * kWave; t += kWave * W) {
      const int q = t / kWave, l = t - q * kWave;
      double s = 0.0;
      #pragma unroll
      for (int w = 0; w < W; ++w) s += red[w][q][l];
      part[(size_t)q * npad + tile * kTile + l] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Tuning overrides: CCAKE_MP_FOLD (16, 32 or 64 candidates per block) and
// CCAKE_MP_SMOOTH_U (2 or 4 smoothing records per step).
static int mp_fold()
{
  static const int v = [] {
    const char* e = std::getenv("CCAKE_MP_FOLD");
    const int f = e ? std::atoi(e) : kFoldEvery;
    return (f == 32 || f == 64) ? f : 16;
  }();
  return v;
}
static int mp_smooth_u()
{
  static const int v = [] {
    const char* e = std::getenv("CCAKE_MP_SMOOTH_U");
    return (e && std::atoi(e) == 4) ? 4 : 2;
  }();
  return v;
}

// The tile-workgroup form (slices added in LDS) runs with the default fold
// and records per step; the tuning overrides use one workgroup per slice.
template <int MODE, int U>
static void smooth32_fold(ccake_event* ev, int ntiles, double rc2, double rcs, float hinv2)
{
  const int f = mp_fold();
  const int W = (f == kFoldEvery && U == 2) ? lds_slices(ev) : 0;
  ev->smooth_sum_slices = W ? 1 : ev->nslices;
  auto go = [&](auto kern, int wg, int waves) {
    hipLaunchKernelGGL(kern, dim3(wg), dim3(kWave * waves), 0, ev->stream, ev->cand_s, ev->cand_s32, ev->skey,
                       ev->cell_off, ev->grid, ev->n_live, rc2, rcs, hinv2, ev->nslices, ev->part, ev->npad);
  };
  if (W == 8) go(k_smooth_slice32<MODE, U, kFoldEvery, 8>, ntiles, 8);
  else if (W == 4) go(k_smooth_slice32<MODE, U, kFoldEvery, 4>, ntiles, 4);
  else if (f == 64) go(k_smooth_slice32<MODE, U, 64, 0>, ntiles * ev->nslices, 1);
  else if (f == 32) go(k_smooth_slice32<MODE, U, 32, 0>, ntiles * ev->nslices, 1);
  else go(k_smooth_slice32<MODE, U, 16, 0>, ntiles * ev->nslices, 1);
}

int launch_smooth_pairs32(ccake_event* ev)
{
  const int n = ev->n_live;          // sorted positions of the live events' particles
  const double rc = 2.0 * ev->p.hT;
  const double rcs = search_radius(ev);
  const int ntiles = (n + kTile - 1) / kTile;
  hipStream_t st = ev->stream;
  hipLaunchKernelGGL(k_records32<0>, dim3((n + 255) / 256), dim3(256), 0, st, ev->cand_s, n, ev->cand_s32);
  const float hinv2 = (float)(1.0 / (ev->p.hT * ev->p.hT));
  const bool four = mp_smooth_u() == 4;
  if (ev->no_charges) {
    if (four) smooth32_fold<3, 4>(ev, ntiles, rc * rc, rcs, hinv2); else smooth32_fold<3, 2>(ev, ntiles, rc * rc, rcs, hinv2);
  } else {
    if (four) smooth32_fold<0, 4>(ev, ntiles, rc * rc, rcs, hinv2); else smooth32_fold<0, 2>(ev, ntiles, rc * rc, rcs, hinv2);
  }
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

// Waves per tile workgroup of the FP32 gradient walk (at most 4 unless
// CCAKE_MP_GRAD_WAVES says otherwise; 4 and 8 measured 4.36 and 4.40 ms per
// step, and without the register budget 8-wave workgroups left one per CU).
static int mp_grad_waves(const ccake_event* ev)
{
  static const int v = [] {
    const char* e = std::getenv("CCAKE_MP_GRAD_WAVES");
    return e ? std::atoi(e) : 4;
  }();
  const int w = lds_slices(ev);
  return v <= 0 ? 0 : (w < v ? w : (v >= 8 ? 8 : 4));
}

template <bool SHEAR, bool BULK>
static void grad32_fold(ccake_event* ev, int ntiles, double rc2, double rcs, float hinv2)
{
  const int f = mp_fold();
  const int W = f == kFoldEvery ? mp_grad_waves(ev) : 0;
  ev->grad_sum_slices = W ? 1 : ev->nslices;
  auto go = [&](auto kern, int wg, int waves) {
    hipLaunchKernelGGL(kern, dim3(wg), dim3(kWave * waves), 0, ev->stream, ev->cand_g, ev->cand_g32, ev->skey,
                       ev->cell_off, ev->grid, ev->n_live, rc2, rcs, hinv2, ev->nslices, ev->part, ev->npad);
  };
  if (W == 8) go(k_grad_slice32<SHEAR, BULK, kFoldEvery, 8>, ntiles, 8);
  else if (W == 4) go(k_grad_slice32<SHEAR, BULK, kFoldEvery, 4>, ntiles, 4);
  else if (f == 64) go(k_grad_slice32<SHEAR, BULK, 64, 0>, ntiles * ev->nslices, 1);
  else if (f == 32) go(k_grad_slice32<SHEAR, BULK, 32, 0>, ntiles * ev->nslices, 1);
  else go(k_grad_slice32<SHEAR, BULK, 16, 0>, ntiles * ev->nslices, 1);
}

int launch_grad_pairs32(ccake_event* ev)
{
  const int n = ev->n_live;
  const double rc = 2.0 * ev->p.hT;
  const double rcs = search_radius(ev);
  const int ntiles = (n + kTile - 1) / kTile;
  hipStream_t st = ev->stream;
  hipLaunchKernelGGL(k_records32<1>, dim3((n + 255) / 256), dim3(256), 0, st, ev->cand_g, n, ev->cand_g32);
  const float hinv2 = (float)(1.0 / (ev->p.hT * ev->p.hT));
  if (ev->p.using_shear) {
    if (ev->no_bulk) grad32_fold<true, false>(ev, ntiles, rc * rc, rcs, hinv2);
    else grad32_fold<true, true>(ev, ntiles, rc * rc, rcs, hinv2);
  } else {
    if (ev->no_bulk) grad32_fold<false, false>(ev, ntiles, rc * rc, rcs, hinv2);
    else grad32_fold<false, true>(ev, ntiles, rc * rc, rcs, hinv2);
  }
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

}  // namespace ccake
