// Shared device helpers of the pair kernels (smoothing, gradients, pair
// export).  A workgroup is one wave; its 64 lanes own 64 consecutive
// particles of the cell-sorted order (a "tile").  The tile's bounding box
// selects, row by row, the contiguous run of cells within the interaction
// radius; every lane then walks that run in the same order, so the
// candidate index is wave-uniform and candidate records are fetched with
// scalar loads into SGPRs instead of being staged through LDS.
#pragma once
#include "ccake_internal.h"
#include <climits>

namespace ccake {

__device__ __forceinline__ int wave_min_i(int v)
{
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_min_d(double v)
{
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_max_d(double v)
{
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni(double v)
{
  union { double d; int i[2]; } u;
  u.d = v;
  u.i[0] = __builtin_amdgcn_readfirstlane(u.i[0]);
  u.i[1] = __builtin_amdgcn_readfirstlane(u.i[1]);
  return u.d;
}
__device__ __forceinline__ int cell_of(double x, double lo, double inv, int nc)
{
  int c = (int)floor((x - lo) * inv);
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

// Pair predicate of the build contract: fl(fl(dx*dx) + fl(dy*dy)) <= rc2,
// evaluated without FMA contraction so that it rounds exactly like the
// CPU oracle.
__device__ __forceinline__ double dist2_nocontract(double dx, double dy)
{
  #pragma clang fp contract(off)
  const double a = dx * dx;
  const double b = dy * dy;
  return a + b;
}

// Segment of a tile: lanes whose particle sits in cell row `row`.
struct Seg {
  int row;
  double xmn, xmx, ymn, ymx;
};

// The event that owns key row `row` (wave-uniform): 0 for a single event,
// else a binary search over the batch's first rows.  Kernels then read that
// event's grid through a reference (scalar loads): a by-value copy selected
// between two sources cost the gradient walk 12 VGPRs and 64 bytes of
// scratch per lane (one wave less per SIMD).
__device__ __forceinline__ int event_index(const Grid* __restrict__ gp, const Grid& g0, int row)
{
  if (g0.nev <= 1) return 0;
  int lo = 0, hi = g0.nev - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (gp[mid].row0 <= row) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Walk the candidate spans of one tile segment.  `body(j)` is called with a
// wave-uniform sorted index j for every candidate in range of the segment's
// bounding box (rows ascending, cells ascending within a row).
template <class Body>
__device__ __forceinline__ void for_spans(const Grid& g, const int32_t* __restrict__ cell_off,
                                          const Seg& s, double rcs, Body&& body)
{
  const int ry0 = uni(cell_of(s.ymn - rcs, g.lo[1], g.inv_cell, g.nc[1]));
  const int ry1 = uni(cell_of(s.ymx + rcs, g.lo[1], g.inv_cell, g.nc[1]));
  for (int ry = ry0; ry <= ry1; ++ry) {
    const double ylo = g.lo[1] + ry * g.cell, yhi = ylo + g.cell;
    const double dy = fmax(0.0, fmax(ylo - s.ymx, s.ymn - yhi));
    if (dy > rcs) continue;
    const double hw = sqrt(fmax(rcs * rcs - dy * dy, 0.0));
    const int cx0 = cell_of(s.xmn - hw, g.lo[0], g.inv_cellx, g.nc[0]);
    const int cx1 = cell_of(s.xmx + hw, g.lo[0], g.inv_cellx, g.nc[0]);
    const int kr = (g.row0 + ry) * g.stride;
    const int beg = uni(cell_off[kr + cx0]);
    const int end = uni(cell_off[kr + cx1 + 1]);
    for (int j = beg; j < end; ++j) body(j);
  }
}

// Workgroups are dispatched round-robin over the chip's 8 XCDs, each with
// its own L2.  Remap the workgroup index so that every XCD walks one
// contiguous range of tiles (neighbouring tiles share most candidates).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int bid, int nblk)
{
  const int x = bid % kXcds, r = bid / kXcds;
  const int per = nblk / kXcds, rem = nblk % kXcds;
  return x * per + (x < rem ? x : rem) + r;
}

// Hot pair kernels split a tile's flattened candidate list into `nslices`
// contiguous slices; slice `slice` is walked by one wave (one workgroup), so
// the work units are small enough to spread evenly over the chip's SIMDs
// even when there are few tiles.  Every wave of a tile computes the same
// segments and spans (wave-uniform values).

__device__ __forceinline__ int wave_sum_i(int v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ int wave_excl_scan_i(int v)
{
  const int lane = threadIdx.x & (kWave - 1);
  int x = v;
  for (int o = 1; o < kWave; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x - v;
}

// Candidate run of cell row ry0 + r for a segment: [b, e) in sorted order
// (empty when r >= nrow or the row is out of reach).  Evaluated per lane.
__device__ __forceinline__ void row_run(const Grid& g, const int32_t* __restrict__ cell_off, const Seg& s,
                                        double rcs, int ry0, int nrow, int r, int& b, int& e)
{
  b = e = 0;
  if (r >= nrow) return;
  const int ry = ry0 + r;
  const double ylo = g.lo[1] + ry * g.cell, yhi = ylo + g.cell;
  const double dy = fmax(0.0, fmax(ylo - s.ymx, s.ymn - yhi));
  if (dy > rcs) return;
  const double hw = sqrt(fmax(rcs * rcs - dy * dy, 0.0));
  const int cx0 = cell_of(s.xmn - hw, g.lo[0], g.inv_cellx, g.nc[0]);
  const int cx1 = cell_of(s.xmx + hw, g.lo[0], g.inv_cellx, g.nc[0]);
  const int kr = (g.row0 + ry) * g.stride;
  b = cell_off[kr + cx0];
  e = cell_off[kr + cx1 + 1];
}

// Walk slice `slice` of `nslices` of a segment's flattened candidate list
// (rows ascending, cells ascending within a row).  The rows are resolved
// lane-parallel, lane r holding cell row ry0 + r (64 rows at a time), so
// the cell offsets of every row arrive with one vector load; the walk reads
// a row's run back with readlane.  Called with every lane of the wave
// active; body(j0, j1) runs on the lanes with `in` set, with wave-uniform
// sorted indices j0 < j1.
// Skip a candidate (wave-uniform branch) when it is out of range for every
// active lane: such candidates lie in the cell-rounding margin of the
// tile's region, and every lane would add exact zeros.  On by default
// (measured 10 % faster at 40k particles); the records are requested before
// the branch, so it does not hold up the loads.
#ifndef CCAKE_PAIR_SKIP
#define CCAKE_PAIR_SKIP 1
#endif
constexpr bool kSkipEmpty = CCAKE_PAIR_SKIP != 0;

// The skip test: no lane of the wave has `p` set.  It runs inside the walk's
// body, on the lanes of one segment.  The host emulation build (kernels
// compiled for the CPU to check memory accesses) runs every candidate
// instead, since its lanes do not execute the body in lockstep; the records
// are read before the test either way, so the accesses are the same.
__device__ __forceinline__ bool wave_none(bool p)
{
#ifdef CCAKE_HOST_EMU
  (void)p;
  return false;
#else
  return __ballot(p) == 0ull;
#endif
}

template <class Body>
__device__ __forceinline__ void for_spans_slice(const Grid& g, const int32_t* __restrict__ cell_off,
                                                const Seg& s, double rcs, int slice, int nslices, bool in,
                                                Body&& body)
{
  const int lane = threadIdx.x & (kWave - 1);
  const int ry0 = uni(cell_of(s.ymn - rcs, g.lo[1], g.inv_cell, g.nc[1]));
  const int ry1 = uni(cell_of(s.ymx + rcs, g.lo[1], g.inv_cell, g.nc[1]));
  const int nrow = ry1 - ry0 + 1;
  int b, e;
  row_run(g, cell_off, s, rcs, ry0, nrow, lane, b, e);
  int tot = uni(wave_sum_i(e - b));
  for (int r0 = kWave; r0 < nrow; r0 += kWave) {     // more than 64 rows: rare
    int b2, e2;
    row_run(g, cell_off, s, rcs, ry0, nrow, r0 + lane, b2, e2);
    tot += uni(wave_sum_i(e2 - b2));
  }
  const int lo = (int)(((long long)tot * slice) / nslices);
  const int hi = (int)(((long long)tot * (slice + 1)) / nslices);
  int base = 0;
  for (int r0 = 0; r0 < nrow && base < hi; r0 += kWave) {
    if (r0 > 0) row_run(g, cell_off, s, rcs, ry0, nrow, r0 + lane, b, e);
    const int len = e - b;
    const int pre = wave_excl_scan_i(len);
    const int ctot = uni(wave_sum_i(len));
    const int nr = nrow - r0 < kWave ? nrow - r0 : kWave;
    for (int r = 0; r < nr; ++r) {
      const int rbase = base + __builtin_amdgcn_readlane(pre, r);
      if (rbase >= hi) break;
      const int rb = __builtin_amdgcn_readlane(b, r);
      const int rlen = __builtin_amdgcn_readlane(len, r);
      const int t0 = lo > rbase ? lo - rbase : 0;
      const int t1 = hi - rbase < rlen ? hi - rbase : rlen;
      if (t0 < t1 && in) body(rb + t0, rb + t1);
    }
    base += ctot;
  }
}

// Number of slices per tile for `ntiles` tiles on `simds` SIMDs: the
// smallest that gives near-whole rounds of single-wave work units and at
// least about four waves per SIMD (to hide load latency), at most 16.
inline int choose_slices(int ntiles, int simds)
{
  int best = 1;
  double best_score = -1.0;
  for (int s = 1; s <= 16; ++s) {
    const double w = (double)ntiles * s / (simds > 0 ? simds : 1);
    const double rounds = w <= 1.0 ? 1.0 : __builtin_ceil(w);
    const double score = (w / rounds) * (w < 4.0 ? w / 4.0 : 1.0);
    if (score > best_score + 1e-9) { best_score = score; best = s; }
  }
  return best;
}

// Per-slice partial sums: part[(slice * nacc + f) * npad + k] for sorted
// position k (every tile lane, k < ntiles * 64 <= npad, has a slot).
constexpr int kMaxAcc = 17;
constexpr int kMaxSlices = 16;

// Walk candidates j0 .. j1-1 (wave-uniform) U at a time with the next U
// records requested while the current ones are evaluated.  Scalar loads may
// return out of order, so a wait always covers every load in flight; the
// first use of the current records (their offsets dx, dy) is therefore
// pinned before the next request, and the wait for the next records falls
// after the current arithmetic.  Indices past j1 - 1 are clamped (those
// loads are in bounds and their results unused).  body(c, dx, dy, j, valid)
// evaluates one candidate; `valid` is false for the padding slot of a short
// last step.
template <int NR, int STRIDE, int U, class Body>
__device__ __forceinline__ void walk_records(const double* __restrict__ cand, double xi, double yi, int j0, int j1,
                                             Body&& body)
{
  static_assert(U == 1 || U == 2, "one or two records per step");
  if (j0 >= j1) return;
  double cur[U][NR];
  #pragma unroll
  for (int u = 0; u < U; ++u) {
    const int ju = j0 + u < j1 ? j0 + u : j1 - 1;
    #pragma unroll
    for (int q = 0; q < NR; ++q) cur[u][q] = cand[(size_t)ju * STRIDE + q];
  }
  for (int j = j0; j < j1; j += U) {
    double dx[U], dy[U];
    #pragma unroll
    for (int u = 0; u < U; ++u) { dx[u] = xi - cur[u][0]; dy[u] = yi - cur[u][1]; }
    __builtin_amdgcn_sched_barrier(0);
    double nxt[U][NR];
    #pragma unroll
    for (int u = 0; u < U; ++u) {
      const int ju = j + U + u < j1 ? j + U + u : j1 - 1;
      #pragma unroll
      for (int q = 0; q < NR; ++q) nxt[u][q] = cand[(size_t)ju * STRIDE + q];
    }
    __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
    for (int u = 0; u < U; ++u) body(cur[u], dx[u], dy[u], j + u, j + u < j1);
    #pragma unroll
    for (int u = 0; u < U; ++u)
      #pragma unroll
      for (int q = 0; q < NR; ++q) cur[u][q] = nxt[u][q];
  }
}

// q = r/h and 1/q from r^2: the hardware reciprocal square root
// (v_rsq_f64) of q^2 refined by one Newton step y <- y (3 - x y^2) / 2,
// relative error about 1e-14, which the pair sums tolerate.  1e-300 is added
// to q^2 so that coincident particles get q = 1e-150 (0 to rounding)
// instead of 0 * inf.  The pair predicate never uses it.
__device__ __forceinline__ void q_and_inverse(double r2, double hinv2, double& q, double& iq)
{
  const double x = fma(r2, hinv2, 1e-300);
  const double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  const double e = fma(-(hx * y), y, 1.5);
  iq = y * e;
  q = x * iq;
}

// Cubic spline in branch-free form (src/kernel.cpp:34-48 and 73-118, D = 2),
// with t = max(2 - q, 0) and u = max(1 - q, 0):
//   W      = norm2 * (t^3 - 4 u^3)       [4 - 6q^2 + 3q^3 for q < 1, since
//                                         norm / norm2 = 4; (2 - q)^3 for
//                                         1 <= q < 2; 0 beyond]
//   grad W = C * rel * (t^2 - 4 u^2) / q [4 - 3q for q < 1; (2 - q)^2 / q
//                                         for 1 <= q < 2; 0 beyond], with
//   C = kgrad / h^4 = -kgrad2 / (2 h^4), kgrad = -15/(14 pi), kgrad2 = 15/(7 pi).
// Non-neighbours weigh exactly 0 and the self pair (q = 0 to rounding)
// weighs W(0) and has zero gradient, so the pair loops need no selects.
__device__ __forceinline__ double spline_w(double q)
{
  const double t = fmax(2.0 - q, 0.0), u = fmax(1.0 - q, 0.0);
  return fma(-4.0, u * u * u, t * t * t);
}
__device__ __forceinline__ double spline_g(double q, double iq)
{
  const double t = fmax(2.0 - q, 0.0), u = fmax(1.0 - q, 0.0);
  return fma(-4.0, u * u, t * t) * iq;
}

// Iterate over the row segments of the tile owned by this wave.  `fn(seg,
// in_seg)` runs once per segment with in_seg true on the lanes of the segment.
template <class Fn>
__device__ __forceinline__ void for_segments(bool act, int myrow, double xi, double yi, Fn&& fn)
{
  bool done = !act;
  for (;;) {
    const int r = uni(wave_min_i(done ? INT_MAX : myrow));
    if (r == INT_MAX) break;
    const bool in = !done && myrow == r;
    Seg s;
    s.row = r;
    s.xmn = uni(wave_min_d(in ? xi : INFINITY));
    s.xmx = uni(wave_max_d(in ? xi : -INFINITY));
    s.ymn = uni(wave_min_d(in ? yi : INFINITY));
    s.ymx = uni(wave_max_d(in ? yi : -INFINITY));
    fn(s, in);
    done = done || in;
  }
}

}  // namespace ccake
