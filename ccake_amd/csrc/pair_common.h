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
    const int cx0 = cell_of(s.xmn - hw, g.lo[0], g.inv_cell, g.nc[0]);
    const int cx1 = cell_of(s.xmx + hw, g.lo[0], g.inv_cell, g.nc[0]);
    const int beg = uni(cell_off[ry * g.nc[0] + cx0]);
    const int end = uni(cell_off[ry * g.nc[0] + cx1 + 1]);
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

// Hot pair kernels give one tile to kSplit waves: every wave of the
// workgroup computes the same segments and spans (wave-uniform values) and
// walks a contiguous slice of the flattened candidate list; the slices'
// partial sums are added in wave order, so results do not depend on
// scheduling.
#ifndef CCAKE_PAIR_SPLIT
#define CCAKE_PAIR_SPLIT 8
#endif
constexpr int kSplit = CCAKE_PAIR_SPLIT;   // 8 measured fastest at 40k-250k particles
// Optionally skip a candidate (wave-uniform branch) when it is out of range
// for every lane of the tile; only adds of exact zeros are skipped, the sums
// are unchanged.  Off by default: the branch keeps the compiler from issuing
// the candidate loads ahead of their use, which costs more than it saves.
#ifndef CCAKE_PAIR_SKIP
#define CCAKE_PAIR_SKIP 0
#endif
constexpr bool kSkipEmpty = CCAKE_PAIR_SKIP != 0;

template <class Body>
__device__ __forceinline__ void for_spans_slice(const Grid& g, const int32_t* __restrict__ cell_off,
                                                const Seg& s, double rcs, int wave, Body&& body)
{
  const int ry0 = uni(cell_of(s.ymn - rcs, g.lo[1], g.inv_cell, g.nc[1]));
  const int ry1 = uni(cell_of(s.ymx + rcs, g.lo[1], g.inv_cell, g.nc[1]));
  auto span = [&](int ry, int& beg, int& end) {
    const double ylo = g.lo[1] + ry * g.cell, yhi = ylo + g.cell;
    const double dy = fmax(0.0, fmax(ylo - s.ymx, s.ymn - yhi));
    if (dy > rcs) { beg = end = 0; return; }
    const double hw = sqrt(fmax(rcs * rcs - dy * dy, 0.0));
    const int cx0 = cell_of(s.xmn - hw, g.lo[0], g.inv_cell, g.nc[0]);
    const int cx1 = cell_of(s.xmx + hw, g.lo[0], g.inv_cell, g.nc[0]);
    beg = uni(cell_off[ry * g.nc[0] + cx0]);
    end = uni(cell_off[ry * g.nc[0] + cx1 + 1]);
  };
  int tot = 0;
  for (int ry = ry0; ry <= ry1; ++ry) { int b, e; span(ry, b, e); tot += e - b; }
  const int lo = (int)(((long long)tot * wave) / kSplit);
  const int hi = (int)(((long long)tot * (wave + 1)) / kSplit);
  int base = 0;
  for (int ry = ry0; ry <= ry1 && base < hi; ++ry) {
    int b, e; span(ry, b, e);
    const int len = e - b;
    const int t0 = lo > base ? lo - base : 0;
    const int t1 = hi - base < len ? hi - base : len;
    if (t0 < t1) body(b + t0, b + t1);
    base += len;
  }
}

// sqrt(x) and 1/sqrt(x) for x >= 0: the hardware reciprocal square root
// (v_rsq_f64) refined by one Newton step y <- y (3 - x y^2) / 2, relative
// error about 1e-14, which the pair sums tolerate.  x is floored at 1e-300
// so that coincident particles get r = 1e-150 (q = 0 to rounding) instead of
// 0 * inf.  The pair predicate never uses it.
__device__ __forceinline__ void sqrt_rsqrt(double x, double& r, double& ir)
{
  const double xc = fmax(x, 1e-300);
  const double y = __builtin_amdgcn_rsq(xc);
  const double h = 0.5 * xc;
  const double e = fma(-(h * y), y, 1.5);
  ir = y * e;
  r = xc * ir;
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
