// Linked-cell neighbour build on gfx950: cell keys, a stable LSD radix sort
// (256 buckets per pass, wave-ballot ranking, device-wide scans), and cell
// offsets.  Replaces SystemState::reset_neighbour_list
// (src/system_state.cpp:650-719), which rebuilds a Cabana VerletList every
// RK stage.  The pair list itself is never materialised on the hot path:
// pair kernels walk the sorted cells directly.
#include "ccake_internal.h"
#include <climits>
#include <vector>
#include <algorithm>

namespace ccake {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 4;                       // items per thread per block
constexpr int kSortBlockItems = kSortThreads * kSortItems;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;

// ---------------------------------------------------------------------------
// Phase 02 fused with the bounding-box partials.
// EoM_default::update_velocity (src/eom_default.cpp:17-38), D = 2:
// gamma = sqrt(1 - u_cov.u) with u_cov = -u, v = u/gamma.
__global__ __launch_bounds__(256) void k_stage_begin(double* __restrict__ F, int npad, int n,
                                                     double* __restrict__ bbox_part, int do_gamma)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double xmn = INFINITY, xmx = -INFINITY, ymn = INFINITY, ymx = -INFINITY;
  if (i < n && do_gamma) {
    const double ux = F[(size_t)CCAKE_F_UX * npad + i], uy = F[(size_t)CCAKE_F_UY * npad + i];
    const double c = (-ux) * ux + (-uy) * uy;
    const double g = sqrt(1.0 - c);
    F[(size_t)CCAKE_F_GAMMA * npad + i] = g;
    F[(size_t)CCAKE_F_VX * npad + i] = ux / g;
    F[(size_t)CCAKE_F_VY * npad + i] = uy / g;
  }
  if (i < n) {
    const double x = F[(size_t)CCAKE_F_X * npad + i], y = F[(size_t)CCAKE_F_Y * npad + i];
    xmn = xmx = x; ymn = ymx = y;
  }
  for (int o = 32; o > 0; o >>= 1) {
    xmn = fmin(xmn, __shfl_xor(xmn, o)); xmx = fmax(xmx, __shfl_xor(xmx, o));
    ymn = fmin(ymn, __shfl_xor(ymn, o)); ymx = fmax(ymx, __shfl_xor(ymx, o));
  }
  __shared__ double s[4][4];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { s[w][0] = xmn; s[w][1] = xmx; s[w][2] = ymn; s[w][3] = ymx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      xmn = fmin(xmn, s[k][0]); xmx = fmax(xmx, s[k][1]); ymn = fmin(ymn, s[k][2]); ymx = fmax(ymx, s[k][3]);
    }
    double* b = bbox_part + 4 * blockIdx.x;
    b[0] = xmn; b[1] = xmx; b[2] = ymn; b[3] = ymx;
  }
}

// Reduce the partials and lay out the cell grid.  cell = rc / cells_per_rc,
// doubled until the grid fits the key capacity.
// x cells per row cell: the requested count, cut so that the key space
// (cells already laid out, times the count) stays within `cap` cells.
__device__ __forceinline__ int x_subdivision(int xsub, long long cap, long long cells)
{
  long long xs = xsub > 1 ? xsub : 1;
  const long long room = cells > 0 ? cap / cells : 1;
  if (xs > room) xs = room;
  return xs < 1 ? 1 : (int)xs;
}

// Cells along x of width `wx` for x cells of cell / xs, at most xs times
// the count for cells of `cell` (which the key space was sized for); a
// position past the last cell is clamped into it by the sort and the walks
// alike.
__device__ __forceinline__ int x_cells(double wx, double cell, int xs, int nx_coarse, double& cellx)
{
  cellx = cell / (double)xs;
  const double f = floor(wx / cellx) + 1.0;
  const double lim = (double)xs * (double)nx_coarse;
  return (int)(f < lim ? f : lim);
}

__global__ __launch_bounds__(256) void k_grid_setup(const double* __restrict__ bbox_part, int nparts,
                                                    double rc, int cells_per_rc, int xsub, int cap,
                                                    Grid* __restrict__ grid, int32_t* __restrict__ err,
                                                    unsigned long long* __restrict__ pair_count)
{
  if (threadIdx.x == 0) *pair_count = 0ull;
  double xmn = INFINITY, xmx = -INFINITY, ymn = INFINITY, ymx = -INFINITY;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
    const double* b = bbox_part + 4 * k;
    xmn = fmin(xmn, b[0]); xmx = fmax(xmx, b[1]); ymn = fmin(ymn, b[2]); ymx = fmax(ymx, b[3]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    xmn = fmin(xmn, __shfl_xor(xmn, o)); xmx = fmax(xmx, __shfl_xor(xmx, o));
    ymn = fmin(ymn, __shfl_xor(ymn, o)); ymx = fmax(ymx, __shfl_xor(ymx, o));
  }
  __shared__ double s[4][4];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { s[w][0] = xmn; s[w][1] = xmx; s[w][2] = ymn; s[w][3] = ymx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      xmn = fmin(xmn, s[k][0]); xmx = fmax(xmx, s[k][1]); ymn = fmin(ymn, s[k][2]); ymx = fmax(ymx, s[k][3]);
    }
    double cell = rc / (double)(cells_per_rc > 0 ? cells_per_rc : 1);
    double cellx = cell;
    int nx = 1, ny = 1;
    const double wx = xmx - xmn, wy = ymx - ymn;
    if (!(wx >= 0.0 && wx < 1e300) || !(wy >= 0.0 && wy < 1e300)) {
      // non-finite positions: one cell, flagged
      atomicOr(err, DE_GRID_OVERFLOW);
      xmn = 0.0; ymn = 0.0;
    } else {
      for (;;) {
        const double fx = floor(wx / cell) + 1.0, fy = floor(wy / cell) + 1.0;
        if (fx * fy <= (double)cap) { nx = (int)fx; ny = (int)fy; break; }
        cell *= 2.0;
      }
      nx = x_cells(wx, cell, x_subdivision(xsub, cap, (long long)nx * ny), nx, cellx);
    }
    Grid g;
    g.lo[0] = xmn; g.lo[1] = ymn; g.lo[2] = 0.0; g.cell = cell; g.inv_cell = 1.0 / cell;
    g.cellx = cellx; g.inv_cellx = 1.0 / cellx;
    g.nc[0] = nx; g.nc[1] = ny; g.nc[2] = 1; g.ncells = nx * ny;
    g.stride = nx; g.row0 = 0; g.nev = 1; g.dim = 2;
    *grid = g;
  }
}

// 3+1D: the bounding box in (x, y, eta) from the six-value partials of
// k3_stage_begin, and a grid of cubic cells of edge rc / cells_per_rc over
// it, doubled until the key space fits `cap` cells.
__global__ __launch_bounds__(256) void k3_grid_setup(const double* __restrict__ bbox_part, int nparts,
                                                     double rc, int cells_per_rc, int xsub, int cap,
                                                     Grid* __restrict__ grid, int32_t* __restrict__ err,
                                                     unsigned long long* __restrict__ pair_count)
{
  if (threadIdx.x == 0) *pair_count = 0ull;
  double mn[3] = { INFINITY, INFINITY, INFINITY }, mx[3] = { -INFINITY, -INFINITY, -INFINITY };
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
    const double* b = bbox_part + 6 * k;
    for (int d = 0; d < 3; ++d) { mn[d] = fmin(mn[d], b[2 * d]); mx[d] = fmax(mx[d], b[2 * d + 1]); }
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int d = 0; d < 3; ++d) { mn[d] = fmin(mn[d], __shfl_xor(mn[d], o)); mx[d] = fmax(mx[d], __shfl_xor(mx[d], o)); }
  __shared__ double s[4][6];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) for (int d = 0; d < 3; ++d) { s[w][2 * d] = mn[d]; s[w][2 * d + 1] = mx[d]; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 1; k < 4; ++k)
    for (int d = 0; d < 3; ++d) { mn[d] = fmin(mn[d], s[k][2 * d]); mx[d] = fmax(mx[d], s[k][2 * d + 1]); }
  double cell = rc / (double)(cells_per_rc > 0 ? cells_per_rc : 1);
  int nc[3] = { 1, 1, 1 };
  bool finite = true;
  for (int d = 0; d < 3; ++d) {
    const double wd = mx[d] - mn[d];
    if (!(wd >= 0.0 && wd < 1e300)) finite = false;
  }
  if (!finite) {
    atomicOr(err, DE_GRID_OVERFLOW);
    mn[0] = mn[1] = mn[2] = 0.0;
  } else {
    for (;;) {
      double f[3], tot = 1.0;
      for (int d = 0; d < 3; ++d) { f[d] = floor((mx[d] - mn[d]) / cell) + 1.0; tot *= f[d]; }
      if (tot <= (double)cap) { for (int d = 0; d < 3; ++d) nc[d] = (int)f[d]; break; }
      cell *= 2.0;
    }
  }
  double cellx = cell;
  if (finite)
    nc[0] = x_cells(mx[0] - mn[0], cell, x_subdivision(xsub, cap, (long long)nc[0] * nc[1] * nc[2]), nc[0], cellx);
  Grid g;
  for (int d = 0; d < 3; ++d) { g.lo[d] = mn[d]; g.nc[d] = nc[d]; }
  g.cell = cell; g.inv_cell = 1.0 / cell;
  g.cellx = cellx; g.inv_cellx = 1.0 / cellx;
  g.ncells = nc[0] * nc[1] * nc[2];
  g.stride = nc[0]; g.row0 = 0; g.nev = 1; g.dim = 3;
  *grid = g;
}

// A batch: the bounding box of each event (workgroup e reduces the ids
// [off[e], off[e+1])).
__global__ __launch_bounds__(256) void k_bbox_events(const double* __restrict__ F, int npad,
                                                     const int32_t* __restrict__ off, double* __restrict__ bbox)
{
  const int e = blockIdx.x;
  double xmn = INFINITY, xmx = -INFINITY, ymn = INFINITY, ymx = -INFINITY;
  for (int i = off[e] + threadIdx.x; i < off[e + 1]; i += blockDim.x) {
    const double x = F[(size_t)CCAKE_F_X * npad + i], y = F[(size_t)CCAKE_F_Y * npad + i];
    xmn = fmin(xmn, x); xmx = fmax(xmx, x); ymn = fmin(ymn, y); ymx = fmax(ymx, y);
  }
  for (int o = 32; o > 0; o >>= 1) {
    xmn = fmin(xmn, __shfl_xor(xmn, o)); xmx = fmax(xmx, __shfl_xor(xmx, o));
    ymn = fmin(ymn, __shfl_xor(ymn, o)); ymx = fmax(ymx, __shfl_xor(ymx, o));
  }
  __shared__ double s[4][4];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { s[w][0] = xmn; s[w][1] = xmx; s[w][2] = ymn; s[w][3] = ymx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      xmn = fmin(xmn, s[k][0]); xmx = fmax(xmx, s[k][1]); ymn = fmin(ymn, s[k][2]); ymx = fmax(ymx, s[k][3]);
    }
    double* b = bbox + 4 * e;
    b[0] = xmn; b[1] = xmx; b[2] = ymn; b[3] = ymx;
  }
}

// A batch: each event's grid as k_grid_setup lays out one event's, with the
// cell doubled until both axes have at most kBatchAxisCells cells, then the
// key rows event after event (row0) and the common row stride (the widest
// event).  One workgroup; thread e lays out event e, thread 0 the rows.
__global__ __launch_bounds__(256) void k_grid_events(const double* __restrict__ bbox, int nev, double rc,
                                                     int cells_per_rc, int xsub, int cap, Grid* __restrict__ grid,
                                                     int32_t* __restrict__ err,
                                                     unsigned long long* __restrict__ pair_count)
{
  if (threadIdx.x == 0) *pair_count = 0ull;
  for (int e = threadIdx.x; e < nev; e += blockDim.x) {
    double xmn = bbox[4 * e], xmx = bbox[4 * e + 1], ymn = bbox[4 * e + 2], ymx = bbox[4 * e + 3];
    double cell = rc / (double)(cells_per_rc > 0 ? cells_per_rc : 1);
    int nx = 1, ny = 1;
    const double wx = xmx - xmn, wy = ymx - ymn;
    if (!(wx >= 0.0 && wx < 1e300) || !(wy >= 0.0 && wy < 1e300)) {
      atomicOr(err, DE_GRID_OVERFLOW);
      xmn = 0.0; ymn = 0.0;
    } else {
      for (;;) {
        const double fx = floor(wx / cell) + 1.0, fy = floor(wy / cell) + 1.0;
        if (fx <= (double)kBatchAxisCells && fy <= (double)kBatchAxisCells) { nx = (int)fx; ny = (int)fy; break; }
        cell *= 2.0;
      }
    }
    Grid g;
    g.lo[0] = xmn; g.lo[1] = ymn; g.lo[2] = 0.0; g.cell = cell; g.inv_cell = 1.0 / cell;
    g.cellx = cell; g.inv_cellx = 1.0 / cell;
    g.nc[0] = nx; g.nc[1] = ny; g.nc[2] = 1; g.ncells = 0; g.stride = 0; g.row0 = 0; g.nev = nev; g.dim = 2;
    grid[e] = g;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int rows = 0, stride = 1;
    for (int e = 0; e < nev; ++e) { rows += grid[e].nc[1]; stride = max(stride, grid[e].nc[0]); }
    // x cells: one subdivision for the whole batch, cut so that the key
    // space (rows times the widest event's x cells) stays within cap
    const int xs = x_subdivision(xsub, cap, (long long)rows * stride);
    rows = 0;
    stride = 1;
    for (int e = 0; e < nev; ++e) {
      Grid& g = grid[e];
      const double wx = bbox[4 * e + 1] - bbox[4 * e];
      if (wx >= 0.0 && wx < 1e300) g.nc[0] = x_cells(wx, g.cell, xs, g.nc[0], g.cellx);
      g.inv_cellx = 1.0 / g.cellx;
      g.row0 = rows;
      rows += g.nc[1];
      stride = max(stride, g.nc[0]);
    }
    for (int e = 0; e < nev; ++e) { grid[e].stride = stride; grid[e].ncells = rows * stride; }
  }
}

__device__ __forceinline__ int cell_coord(double x, double lo, double inv, int nc)
{
  int c = (int)floor((x - lo) * inv);
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.  Block b owns items
// [b*kSortBlockItems, (b+1)*kSortBlockItems) in input order; hist is laid
// out digit-major (hist[d*nblk + b]) so that one exclusive scan yields every
// (digit, block) output base.
// KEYS: the first pass also computes the cell key of every particle from
// its position and the grid, and stores it (id order) for the scatter.  A
// particle of a retired event of a batch (live[eid] == 0) gets the key
// ncells, past every cell, so it sorts after the live events' particles.
template <bool KEYS>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const int32_t* __restrict__ keys_in, int n,
                                                             int shift, int nblk, int32_t* __restrict__ hist,
                                                             const double* __restrict__ F, int npad,
                                                             const Grid* __restrict__ grid,
                                                             int32_t* __restrict__ keys_out,
                                                             const int32_t* __restrict__ eid,
                                                             const int32_t* __restrict__ live)
{
  __shared__ int32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * kSortBlockItems;
  for (int r = 0; r < kSortItems; ++r) {
    const int idx = base + r * kSortThreads + threadIdx.x;
    if (idx >= n) continue;
    int key;
    if (KEYS) {
      const Grid g = grid[eid ? eid[idx] : 0];
      const int cx = cell_coord(F[(size_t)CCAKE_F_X * npad + idx], g.lo[0], g.inv_cellx, g.nc[0]);
      const int cy = cell_coord(F[(size_t)CCAKE_F_Y * npad + idx], g.lo[1], g.inv_cell, g.nc[1]);
      // 3+1D: key rows run over (eta, y) cell pairs
      const int cz = g.dim == 3 ? cell_coord(F[(size_t)CCAKE_F_ETA * npad + idx], g.lo[2], g.inv_cell, g.nc[2]) : 0;
      key = (g.row0 + cz * g.nc[1] + cy) * g.stride + cx;
      if (live && live[eid[idx]] == 0) key = g.ncells;
      keys_out[idx] = key;
    } else {
      key = keys_in[idx];
    }
    atomicAdd(&h[(key >> shift) & 255], 1);
  }
  __syncthreads();
  hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Peers of this lane inside its wave: lanes holding the same 8-bit digit.
__device__ __forceinline__ uint64_t digit_peers(int digit, bool active)
{
  uint64_t m = __ballot(active);
  #pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (digit >> b) & 1;
    const uint64_t bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const int32_t* __restrict__ keys_in,
                                                                const int32_t* __restrict__ vals_in,
                                                                int n, int shift, int nblk,
                                                                const int32_t* __restrict__ hist_scan,
                                                                int32_t* __restrict__ keys_out,
                                                                int32_t* __restrict__ vals_out)
{
  __shared__ int32_t base[256];
  __shared__ int32_t running[256];
  __shared__ int32_t wcnt[kSortThreads / kWave][256];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  base[threadIdx.x] = hist_scan[threadIdx.x * nblk + blockIdx.x];
  running[threadIdx.x] = 0;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int blk0 = blockIdx.x * kSortBlockItems;
  for (int r = 0; r < kSortItems; ++r) {
    for (int k = 0; k < kSortThreads / kWave; ++k) wcnt[k][threadIdx.x] = 0;
    __syncthreads();
    const int idx = blk0 + r * kSortThreads + threadIdx.x;
    const bool active = idx < n;
    int key = 0, val = 0, digit = 0;
    if (active) { key = keys_in[idx]; val = vals_in ? vals_in[idx] : idx; digit = (key >> shift) & 255; }
    const uint64_t peers = digit_peers(digit, active);
    const int rank = __popcll(peers & lt);
    if (active && rank == 0) wcnt[w][digit] = __popcll(peers);
    __syncthreads();
    if (active) {
      int off = running[digit] + rank;
      for (int k = 0; k < w; ++k) off += wcnt[k][digit];
      const int dst = base[digit] + off;
      keys_out[dst] = key;
      vals_out[dst] = val;
    }
    __syncthreads();
    int add = 0;
    for (int k = 0; k < kSortThreads / kWave; ++k) add += wcnt[k][threadIdx.x];
    running[threadIdx.x] += add;
    __syncthreads();
  }
}

// ---- device-wide exclusive scan of int32 (block sums, scan of sums, down-sweep)
__device__ int32_t block_exclusive_scan(int32_t v, int32_t* total)
{
  __shared__ int32_t ws[kScanThreads / kWave];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[w] = x;
  __syncthreads();
  int32_t pre = 0, tot = 0;
  for (int k = 0; k < kScanThreads / kWave; ++k) { if (k < w) pre += ws[k]; tot += ws[k]; }
  __syncthreads();
  *total = tot;
  return pre + x - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int32_t* __restrict__ in, int n,
                                                              int32_t* __restrict__ part)
{
  const int base = blockIdx.x * kScanThreads * kScanItems;
  int32_t s = 0;
  for (int r = 0; r < kScanItems; ++r) {
    const int idx = base + r * kScanThreads + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  int32_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_partials(int32_t* __restrict__ part, int np)
{
  int32_t carry = 0;
  for (int b0 = 0; b0 < np; b0 += kScanThreads) {
    const int idx = b0 + threadIdx.x;
    const int32_t v = idx < np ? part[idx] : 0;
    int32_t tot;
    const int32_t ex = block_exclusive_scan(v, &tot);
    if (idx < np) part[idx] = carry + ex;
    carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(const int32_t* __restrict__ in, int n,
                                                            const int32_t* __restrict__ part,
                                                            int32_t* __restrict__ out)
{
  const int base = blockIdx.x * kScanThreads * kScanItems;
  // each thread owns kScanItems consecutive items
  int32_t v[kScanItems];
  int32_t s = 0;
  const int t0 = base + threadIdx.x * kScanItems;
  for (int r = 0; r < kScanItems; ++r) { const int idx = t0 + r; v[r] = idx < n ? in[idx] : 0; s += v[r]; }
  int32_t tot;
  int32_t ex = block_exclusive_scan(s, &tot) + part[blockIdx.x];
  for (int r = 0; r < kScanItems; ++r) { const int idx = t0 + r; if (idx < n) out[idx] = ex; ex += v[r]; }
}

// In-place exclusive scan of a[0, n) by one workgroup: thread t owns the
// contiguous run [t*per, min((t+1)*per, n)) with per = ceil(n / 1024), so
// every load of a thread is issued at once; one block scan of the run sums
// gives each run's offset.
constexpr int kScanOneThreads = 1024;
constexpr int kScanOneMax = 65536;   // histograms of up to 256 radix blocks
__global__ __launch_bounds__(kScanOneThreads) void k_scan_one(int32_t* __restrict__ a, int n)
{
  __shared__ int32_t wtot[kScanOneThreads / kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int per = (n + kScanOneThreads - 1) / kScanOneThreads;
  const int b = threadIdx.x * per;
  const int e = min(b + per, n);                    // e <= b (empty run) past the end
  int32_t s = 0;
  #pragma unroll 16
  for (int i = b; i < e; ++i) s += a[i];
  int32_t x = s;                                    // inclusive scan of run sums within the wave
  for (int o = 1; o < kWave; o <<= 1) {
    const int32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) wtot[w] = x;
  __syncthreads();
  int32_t run = x - s;
  for (int q = 0; q < w; ++q) run += wtot[q];
  #pragma unroll 16
  for (int i = b; i < e; ++i) { const int32_t v = a[i]; a[i] = run; run += v; }
}

static int scan_exclusive(ccake_event* ev, const int32_t* in, int32_t* out, int n)
{
  const int per = kScanThreads * kScanItems;
  const int nb = (n + per - 1) / per;
  hipStream_t st = ev->stream;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanThreads), 0, st, in, n, ev->scan_tmp);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, st, ev->scan_tmp, nb);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanThreads), 0, st, in, n, ev->scan_tmp, out);
  return 0;
}

// cell_off[c] = first sorted position whose key >= c, for c in [0, ncells],
// and the smoothing candidate records in sorted order.
__global__ __launch_bounds__(256) void k_cells_records(const int32_t* __restrict__ skey,
                                                       const int32_t* __restrict__ perm, int n,
                                                       const Grid* __restrict__ grid,
                                                       int32_t* __restrict__ cell_off,
                                                       const double* __restrict__ F, int npad,
                                                       double* __restrict__ cand)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = grid->ncells;
  if (n == 0) {
    if (k <= nc) cell_off[k] = 0;
    return;
  }
  if (k >= n) return;
  const int cur = skey[k];
  const int prev = (k == 0) ? -1 : skey[k - 1];
  for (int c = prev + 1; c <= cur; ++c) cell_off[c] = k;
  if (k == n - 1) for (int c = cur + 1; c <= nc; ++c) cell_off[c] = n;
  if (grid->dim == 3) smooth_record3(F, npad, perm[k], cand + (size_t)k * kCandS);
  else smooth_record(F, npad, perm[k], cand + (size_t)k * kCandS);
}

// ---------------------------------------------------------------------------
static int stage_begin(ccake_event* ev, int do_gamma)
{
  const int n = ev->n;
  const int nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_stage_begin, dim3(nb), dim3(256), 0, ev->stream, ev->F, ev->npad, n, ev->bbox, do_gamma);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch_stage_begin(ccake_event* ev) { return ev->p.dim == 3 ? launch3_stage_begin(ev, 1) : stage_begin(ev, 1); }
int launch_bbox(ccake_event* ev) { return ev->p.dim == 3 ? launch3_stage_begin(ev, 0) : stage_begin(ev, 0); }
int launch_gamma_velocity(ccake_event* ev) { return launch_stage_begin(ev); }

static void scan_hist(ccake_event* ev, int nscan)
{
  if (nscan <= kScanOneMax)
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(kScanOneThreads), 0, ev->stream, ev->hist, nscan);
  else
    scan_exclusive(ev, ev->hist, ev->hist, nscan);
}

// Grid, keys, stable LSD radix sort (ev->sort_passes 8-bit passes: two
// while the key space fits 16 bits, three for a batch's larger one), cell
// offsets and the smoothing candidate records.
int launch_neighbor_sort(ccake_event* ev)
{
  const int n = ev->n;
  hipStream_t st = ev->stream;
  const int nb = (n + 255) / 256;
  const double rc = 2.0 * ev->p.hT;
  if (ev->p.dim == 3) {
    hipLaunchKernelGGL(k3_grid_setup, dim3(1), dim3(256), 0, st, ev->bbox, nb, rc, ev->p.cells_per_rc,
                       ev->p.x_subcells, ev->cap_cells, ev->grid, ev->err, (unsigned long long*)ev->pair_count);
  } else if (ev->nev > 1) {
    hipLaunchKernelGGL(k_bbox_events, dim3(ev->nev), dim3(256), 0, st, ev->F, ev->npad, ev->ev_off, ev->bbox);
    hipLaunchKernelGGL(k_grid_events, dim3(1), dim3(256), 0, st, ev->bbox, ev->nev, rc, ev->p.cells_per_rc,
                       ev->p.x_subcells, ev->cap_cells, ev->grid, ev->err, (unsigned long long*)ev->pair_count);
  } else {
    hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(256), 0, st, ev->bbox, nb, rc, ev->p.cells_per_rc,
                       ev->p.x_subcells, ev->cap_cells, ev->grid, ev->err, (unsigned long long*)ev->pair_count);
  }
  const int nblk = ev->radix_blocks;
  const int nscan = 256 * nblk;
  int32_t* kin = ev->key;
  int32_t* vin = nullptr;
  int32_t* kbuf[2] = { ev->key_a, ev->key_b };
  int32_t* vbuf[2] = { ev->val_a, ev->val_b };
  for (int pass = 0; pass < ev->sort_passes; ++pass) {
    const int shift = 8 * pass;
    if (pass == 0)
      hipLaunchKernelGGL(k_radix_hist<true>, dim3(nblk), dim3(kSortThreads), 0, st, nullptr, n, 0, nblk, ev->hist,
                         ev->F, ev->npad, ev->grid, ev->key, ev->eid, live_of(ev).flag);
    else
      hipLaunchKernelGGL(k_radix_hist<false>, dim3(nblk), dim3(kSortThreads), 0, st, kin, n, shift, nblk, ev->hist,
                         nullptr, 0, nullptr, nullptr, nullptr, nullptr);
    scan_hist(ev, nscan);
    int32_t* kout = kbuf[pass & 1];
    int32_t* vout = vbuf[pass & 1];
    hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(kSortThreads), 0, st, kin, vin, n, shift, nblk,
                       ev->hist, kout, vout);
    kin = kout;
    vin = vout;
  }
  ev->skey = kin;
  ev->perm = vin;
  const int nbc = (n + 255) / 256 + 1;
  hipLaunchKernelGGL(k_cells_records, dim3(nbc), dim3(256), 0, st, ev->skey, ev->perm, n, ev->grid, ev->cell_off,
                     ev->F, ev->npad, ev->cand_s);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

}  // namespace ccake
