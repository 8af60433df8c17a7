// The 3+1D Milne path (tau, x, y, eta_s; D = 3) on gfx950: the stages whose
// geometry or data layout differ from 2+1D.  The EoS inversion, the
// transport coefficients, the smoothing sum pass, the freeze-out status
// machine and the reductions are the 2+1D kernels (they are per particle and
// do not depend on D).
//
//   gamma and v          EoM_default::update_velocity (src/eom_default.cpp:17-38):
//                        gamma = sqrt(1 - u_cov . u), u_cov = (-u^x, -u^y, -tau^2 u^eta)
//   neighbour build      SystemState::reset_neighbour_list (src/system_state.cpp:650-719):
//                        isotropic radius 2 hT in (x, y, eta) (also with the factorized
//                        kernel, 689); cubic cells, key rows over (eta, y) cell pairs
//   smoothing            smooth_all_particle_fields / calculate_intial_sigma
//                        (src/sph_workstation.cpp:494-640) with SPHkernel<3>
//                        (src/kernel.cpp:10-118), or the factorized kernel
//                        W_2D(r_perp, hT) W_1D(|d eta|, hEta) (129-170, chosen at 502)
//   gradients            smooth_all_particle_gradients (796-986), D = 3, with the
//                        factorized gradKernel branch (843)
//   equations of motion  EoM_default::evaluate_time_derivatives (src/eom_default.cpp:211-869)
//                        and calculate_MRF_shear (882-1412) with the tau^2 metric factors
//                        and the delta_{i eta} terms (include/milne.hpp:85-107, 415-530,
//                        1093-1108), diffusion off
//   RK update            Evolver::step_rk (src/evolver.cpp:363-484), three positions and
//                        three velocity components
//
// The pair kernels follow the 2+1D design: owner computes, a tile of 64
// cell-sorted particles per wave, candidate records read wave-uniformly
// into SGPRs, a tile's flattened candidate list cut into slices (one
// single-wave workgroup each) whose partial sums a per-particle pass adds in
// slice order.  A tile segment is the set of lanes in one (y, eta) cell row;
// the cell rows within reach of its bounding box are walked as x runs.
#include "pair_common.h"

namespace ccake {

// ---------------------------------------------------------------------------
// Phase 02 with the (x, y, eta) bounding-box partials (six per workgroup).
__device__ __forceinline__ void gamma_v3(double* __restrict__ F, int npad, int i, double t2)
{
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
  const double ux = col(CCAKE_F_UX), uy = col(CCAKE_F_UY), ue = col(CCAKE_F_UETA);
  // milne::contract(u_cov, u), u_cov = (-u^x, -u^y, -tau^2 u^eta)
  double c = 0.0;
  c += (-ux) * ux;
  c += (-uy) * uy;
  c += (ue * -t2) * ue;
  const double g = sqrt(1.0 - c);
  col(CCAKE_F_GAMMA) = g;
  col(CCAKE_F_VX) = ux / g;
  col(CCAKE_F_VY) = uy / g;
  col(CCAKE_F_VETA) = ue / g;
}

__global__ __launch_bounds__(256) void k3_stage_begin(double* __restrict__ F, int npad, int n,
                                                      double* __restrict__ bbox_part, int do_gamma, double t2)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double mn[3] = { INFINITY, INFINITY, INFINITY }, mx[3] = { -INFINITY, -INFINITY, -INFINITY };
  if (i < n) {
    if (do_gamma) gamma_v3(F, npad, i, t2);
    const double r[3] = { F[(size_t)CCAKE_F_X * npad + i], F[(size_t)CCAKE_F_Y * npad + i],
                          F[(size_t)CCAKE_F_ETA * npad + i] };
    for (int d = 0; d < 3; ++d) mn[d] = mx[d] = r[d];
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int d = 0; d < 3; ++d) { mn[d] = fmin(mn[d], __shfl_xor(mn[d], o)); mx[d] = fmax(mx[d], __shfl_xor(mx[d], o)); }
  __shared__ double s[4][6];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) for (int d = 0; d < 3; ++d) { s[w][2 * d] = mn[d]; s[w][2 * d + 1] = mx[d]; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k)
      for (int d = 0; d < 3; ++d) { mn[d] = fmin(mn[d], s[k][2 * d]); mx[d] = fmax(mx[d], s[k][2 * d + 1]); }
    double* b = bbox_part + 6 * blockIdx.x;
    for (int d = 0; d < 3; ++d) { b[2 * d] = mn[d]; b[2 * d + 1] = mx[d]; }
  }
}

// A stage that keeps the cell sort: gamma, v and the smoothing record of
// particle perm[k] at sorted position k, and the stage's pair count cleared.
__global__ __launch_bounds__(256) void k3_stage_gather(double* __restrict__ F, int npad, int n,
                                                       const int32_t* __restrict__ perm, double* __restrict__ cand,
                                                       unsigned long long* __restrict__ pair_count, double t2)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) pair_count[0] = 0ull;
  if (k >= n) return;
  const int i = perm[k];
  gamma_v3(F, npad, i, t2);
  smooth_record3(F, npad, i, cand + (size_t)k * kCandS);
}

// ---------------------------------------------------------------------------
// The pair predicate of the build contract in 3D:
// fl(fl(fl(dx*dx) + fl(dy*dy)) + fl(dz*dz)) <= rc2, no FMA contraction.
__device__ __forceinline__ double dist2_3(double dx, double dy, double dz)
{
  #pragma clang fp contract(off)
  const double a = dx * dx;
  const double b = dy * dy;
  const double c = dz * dz;
  return (a + b) + c;
}
__device__ __forceinline__ double dist2_perp(double dx, double dy)
{
  #pragma clang fp contract(off)
  const double a = dx * dx;
  const double b = dy * dy;
  return a + b;
}

struct Seg3 {
  int row;
  double mn[3], mx[3];
};

// The (y, eta) cell-row segments of the tile owned by this wave.
template <class Fn>
__device__ __forceinline__ void for_segments3(bool act, int myrow, const double* r, Fn&& fn)
{
  bool done = !act;
  for (;;) {
    const int row = uni(wave_min_i(done ? INT_MAX : myrow));
    if (row == INT_MAX) break;
    const bool in = !done && myrow == row;
    Seg3 s;
    s.row = row;
    #pragma unroll
    for (int d = 0; d < 3; ++d) {
      s.mn[d] = uni(wave_min_d(in ? r[d] : INFINITY));
      s.mx[d] = uni(wave_max_d(in ? r[d] : -INFINITY));
    }
    fn(s, in);
    done = done || in;
  }
}

// Candidate run of flattened cell row r (y fastest) of a segment: [b, e)
// in sorted order, empty when r >= nrow or the row is out of reach.
__device__ __forceinline__ void row_run3(const Grid& g, const int32_t* __restrict__ cell_off, const Seg3& s,
                                         double rcs, int ry0, int nry, int rz0, int nrow, int r, int& b, int& e)
{
  b = e = 0;
  if (r >= nrow) return;
  const int ry = ry0 + r % nry, rz = rz0 + r / nry;
  const double ylo = g.lo[1] + ry * g.cell, yhi = ylo + g.cell;
  const double zlo = g.lo[2] + rz * g.cell, zhi = zlo + g.cell;
  const double dy = fmax(0.0, fmax(ylo - s.mx[1], s.mn[1] - yhi));
  const double dz = fmax(0.0, fmax(zlo - s.mx[2], s.mn[2] - zhi));
  const double d2 = dy * dy + dz * dz;
  if (d2 > rcs * rcs) return;
  const double hw = sqrt(fmax(rcs * rcs - d2, 0.0));
  const int cx0 = cell_of(s.mn[0] - hw, g.lo[0], g.inv_cellx, g.nc[0]);
  const int cx1 = cell_of(s.mx[0] + hw, g.lo[0], g.inv_cellx, g.nc[0]);
  const int kr = (rz * g.nc[1] + ry) * g.stride;
  b = cell_off[kr + cx0];
  e = cell_off[kr + cx1 + 1];
}

__device__ __forceinline__ void row_range3(const Grid& g, const Seg3& s, double rcs, int& ry0, int& nry, int& rz0,
                                           int& nrow)
{
  ry0 = uni(cell_of(s.mn[1] - rcs, g.lo[1], g.inv_cell, g.nc[1]));
  const int ry1 = uni(cell_of(s.mx[1] + rcs, g.lo[1], g.inv_cell, g.nc[1]));
  rz0 = uni(cell_of(s.mn[2] - rcs, g.lo[2], g.inv_cell, g.nc[2]));
  const int rz1 = uni(cell_of(s.mx[2] + rcs, g.lo[2], g.inv_cell, g.nc[2]));
  nry = ry1 - ry0 + 1;
  nrow = nry * (rz1 - rz0 + 1);
}

// Slice `slice` of `nslices` of a segment's flattened candidate list, rows
// resolved 64 at a time lane-parallel (as for_spans_slice in 2+1D).  The
// body runs on every lane (wave-uniform control flow, so the compiler keeps
// one copy of the accumulators): it masks the lanes outside the segment
// (`in` false) itself.
template <class Body>
__device__ __forceinline__ void for_spans_slice3(const Grid& g, const int32_t* __restrict__ cell_off,
                                                 const Seg3& s, double rcs, int slice, int nslices, bool in,
                                                 Body&& body)
{
  const int lane = threadIdx.x & (kWave - 1);
  int ry0, nry, rz0, nrow;
  row_range3(g, s, rcs, ry0, nry, rz0, nrow);
  int tot = 0;
  for (int r0 = 0; r0 < nrow; r0 += kWave) {
    int b2, e2;
    row_run3(g, cell_off, s, rcs, ry0, nry, rz0, nrow, r0 + lane, b2, e2);
    tot += uni(wave_sum_i(e2 - b2));
  }
  const int lo = (int)(((long long)tot * slice) / nslices);
  const int hi = (int)(((long long)tot * (slice + 1)) / nslices);
  int base = 0;
  for (int r0 = 0; r0 < nrow && base < hi; r0 += kWave) {
    int b, e;
    row_run3(g, cell_off, s, rcs, ry0, nry, rz0, nrow, r0 + lane, b, e);
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
      if (t0 < t1) body(rb + t0, rb + t1);
    }
    base += ctot;
  }
}

// Every candidate of a segment, one at a time (the count and export walks).
template <class Body>
__device__ __forceinline__ void for_spans3(const Grid& g, const int32_t* __restrict__ cell_off, const Seg3& s,
                                           double rcs, Body&& body)
{
  int ry0, nry, rz0, nrow;
  row_range3(g, s, rcs, ry0, nry, rz0, nrow);
  for (int r = 0; r < nrow; ++r) {
    int b, e;
    row_run3(g, cell_off, s, rcs, ry0, nry, rz0, nrow, r, b, e);
    b = uni(b); e = uni(e);
    for (int j = b; j < e; ++j) body(j);
  }
}

// Candidates j0 .. j1-1 one at a time, the next record requested while the
// current one is evaluated (scalar loads; the wait falls after the current
// arithmetic).  body(c, dx, dy, dz, j, valid).
template <int NR, int STRIDE, class Body>
__device__ __forceinline__ void walk_records3(const double* __restrict__ cand, const double* r, int j0, int j1,
                                              Body&& body)
{
  if (j0 >= j1) return;
  double cur[NR];
  #pragma unroll
  for (int q = 0; q < NR; ++q) cur[q] = cand[(size_t)j0 * STRIDE + q];
  for (int j = j0; j < j1; ++j) {
    const double dx = r[0] - cur[0], dy = r[1] - cur[1], dz = r[2] - cur[2];
    __builtin_amdgcn_sched_barrier(0);
    double nxt[NR];
    const int jn = j + 1 < j1 ? j + 1 : j1 - 1;
    #pragma unroll
    for (int q = 0; q < NR; ++q) nxt[q] = cand[(size_t)jn * STRIDE + q];
    __builtin_amdgcn_sched_barrier(0);
    body(cur, dx, dy, dz, j, true);
    #pragma unroll
    for (int q = 0; q < NR; ++q) cur[q] = nxt[q];
  }
}

// Candidates j0 .. j1-1 one at a time for records too large to double
// buffer in SGPRs: the head of a record (position and mass, 4 doubles) is
// requested one candidate ahead, the rest of it (entries 4 .. NR-1) when
// its candidate comes up, together with the next head.  pair(c, dx, dy,
// dz, w) computes the kernel factors w[3] from the head and returns false
// when no lane has the pair in range (the rest is then not needed); sums(c,
// w) adds the terms that need the rest.  The wait for the rest falls after
// the kernel arithmetic, and 2*4 + NR - 4 doubles are in flight at most
// (46 SGPRs for the 19-double shear record, against 76 for two records).
template <int NR, int STRIDE, class Pair, class Sums>
__device__ __forceinline__ void walk_records3_split(const double* __restrict__ cand, const double* r, int j0,
                                                    int j1, Pair&& pair, Sums&& sums)
{
  if (j0 >= j1) return;
  constexpr int NH = 4;
  double head[NH];
  #pragma unroll
  for (int q = 0; q < NH; ++q) head[q] = cand[(size_t)j0 * STRIDE + q];
  for (int j = j0; j < j1; ++j) {
    double c[NR];
    #pragma unroll
    for (int q = 0; q < NH; ++q) c[q] = head[q];
    const double dx = r[0] - c[0], dy = r[1] - c[1], dz = r[2] - c[2];
    __builtin_amdgcn_sched_barrier(0);
    const int jn = j + 1 < j1 ? j + 1 : j1 - 1;
    #pragma unroll
    for (int q = NH; q < NR; ++q) c[q] = cand[(size_t)j * STRIDE + q];
    #pragma unroll
    for (int q = 0; q < NH; ++q) head[q] = cand[(size_t)jn * STRIDE + q];
    __builtin_amdgcn_sched_barrier(0);
    double w[3];
    if (pair(c, dx, dy, dz, w)) sums(c, w);
  }
}

// Kernel constants of a 3+1D event.  Isotropic: W = norm2 w(q),
// grad W = C rel g(q) with q = r / hT (SPHkernel<3>: norm2 = 1/(4 pi hT^3),
// C = kgrad / hT^5 = -3/(4 pi hT^5)).  Factorized: W = norm2 w(q_perp) w(q_eta)
// (norm2 = 5/(14 pi hT^2) * 1/(6 hEta)), and with C = kgrad_2D / hT^4 *
// 1/(6 hEta) the gradient is C (rel_x g_perp w_eta, rel_y g_perp w_eta,
// kfac w_perp rel_eta g_eta), kfac = (5/(14 pi hT^2)) (-1/(2 hEta^3)) / C.
struct Kern3 {
  double norm2, C, kfac, hinv2, heinv2;
};
static Kern3 kern3(const ccake_params& p)
{
  Kern3 k;
  const double h = p.hT, he = p.hEta;
  k.hinv2 = 1.0 / (h * h);
  k.heinv2 = he > 0.0 ? 1.0 / (he * he) : 0.0;
  if (p.factorized_kernel) {
    const double n2 = 5.0 / (14.0 * kPi) / (h * h), n1 = (1.0 / 6.0) / he;
    const double C2 = -15.0 / (14.0 * kPi) / pow(h, 4), C1 = -0.5 / pow(he, 3);
    k.norm2 = n2 * n1;
    k.C = C2 * n1;
    k.kfac = n2 * C1 / k.C;
  } else {
    k.norm2 = 1.0 / (4.0 * kPi) / pow(h, 3);
    k.C = -3.0 / (4.0 * kPi) / pow(h, 5);
    k.kfac = 1.0;
  }
  return k;
}

// The weight w of one pair (W = norm2 w) for the isotropic or factorized
// kernel; `nb` masks pairs outside the list (the factorized kernel can be
// non-zero there when hEta > hT).
template <bool FACT>
__device__ __forceinline__ double pair_w3(double dx, double dy, double dz, double r2, bool nb, double hinv2,
                                          double heinv2)
{
  if (FACT) {
    double qp, iqp, qe, iqe;
    q_and_inverse(dist2_perp(dx, dy), hinv2, qp, iqp);
    q_and_inverse(dz * dz, heinv2, qe, iqe);
    return nb ? spline_w(qp) * spline_w(qe) : 0.0;
  } else {
    double q, iq;
    q_and_inverse(r2, hinv2, q, iq);
    return nb ? spline_w(q) : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Smoothing, pair part: MODE 0 sigma_lab, s, B, S, Q and the neighbour count
// (6 partial sums); MODE 1 sigma_lab and the count (initial sigma).  Same
// partial-sum layout as the 2+1D kernel, so k_smooth_sum adds them.
template <int MODE, bool FACT>
__global__ __launch_bounds__(kWave) void k3_smooth_slice(const double* __restrict__ cand,
                                                         const int32_t* __restrict__ skey,
                                                         const int32_t* __restrict__ cell_off,
                                                         const Grid* __restrict__ gridp, int n, double rc2,
                                                         double rcs, double hinv2, double heinv2, int nslices,
                                                         double* __restrict__ part, int npad)
{
  constexpr int NACC = MODE == 0 ? 6 : 2;
  const Grid& g = *gridp;
  const int lane = threadIdx.x;
  const int unit = xcd_tile(blockIdx.x, gridDim.x);
  const int tile = unit / nslices;
  const int slice = unit - tile * nslices;
  const int k = tile * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand + (size_t)kk * kCandS;
  const double r[3] = { me[0], me[1], me[2] };
  const int myrow = act ? skey[kk] / g.stride : INT_MAX;

  double sl = 0.0, ss = 0.0, sB = 0.0, sS = 0.0, sQ = 0.0;
  int cnt = 0;
  for_segments3(act, myrow, r, [&](const Seg3& seg, bool in) {
    for_spans_slice3(g, cell_off, seg, rcs, slice, nslices, in, [&](int j0, int j1) {
      constexpr int NR = MODE == 0 ? 8 : 4;
      walk_records3<NR, kCandS>(cand, r, j0, j1, [&](const double* c, double dx, double dy, double dz, int j,
                                                     bool valid) {
        const double r2 = dist2_3(dx, dy, dz);
        const bool nb = in && valid && r2 <= rc2;
        cnt += nb ? 1 : 0;
        if (kSkipEmpty && wave_none(nb)) return;
        const double w = pair_w3<FACT>(dx, dy, dz, r2, nb, hinv2, heinv2);
        sl += c[3] * w;
        if (MODE == 0) {
          ss += c[MODE == 0 ? 4 : 0] * w; sB += c[MODE == 0 ? 5 : 0] * w;
          sS += c[MODE == 0 ? 6 : 0] * w; sQ += c[MODE == 0 ? 7 : 0] * w;
        }
      });
    });
  });
  double* p = part + (size_t)slice * NACC * npad + k;
  p[0] = sl;
  p[(size_t)npad] = (double)cnt;
  if (MODE == 0) {
    p[(size_t)2 * npad] = ss; p[(size_t)3 * npad] = sB;
    p[(size_t)4 * npad] = sS; p[(size_t)5 * npad] = sQ;
  }
}

// Neighbour walk without sums: MODE 2 counts (btrack) and the pair total;
// MODE 3 writes the pair rows (parity export) at the row offsets.
template <int MODE>
__global__ __launch_bounds__(64) void k3_walk(const double* __restrict__ cand, const int32_t* __restrict__ skey,
                                              const int32_t* __restrict__ perm, const int32_t* __restrict__ cell_off,
                                              const Grid* __restrict__ gridp, int n, double rc2, double rcs,
                                              int32_t* __restrict__ btrack, unsigned long long* __restrict__ pair_total,
                                              const int64_t* __restrict__ row_off, int32_t* __restrict__ out_i,
                                              int32_t* __restrict__ out_j)
{
  const Grid& g = *gridp;
  const int lane = threadIdx.x;
  const int k = blockIdx.x * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand + (size_t)kk * kCandS;
  const double r[3] = { me[0], me[1], me[2] };
  const int myrow = act ? skey[kk] / g.stride : INT_MAX;
  const int id = act ? perm[kk] : 0;
  int cnt = 0;
  int64_t wpos = 0;
  if (MODE == 3 && act) wpos = row_off[id];
  for_segments3(act, myrow, r, [&](const Seg3& seg, bool in) {
    for_spans3(g, cell_off, seg, rcs, [&](int j) {
      const double* c = cand + (size_t)j * kCandS;
      const double r2 = dist2_3(r[0] - c[0], r[1] - c[1], r[2] - c[2]);
      const bool nb = in && (r2 <= rc2) && (j != k);
      if (MODE == 2) cnt += nb;
      if (MODE == 3 && nb) { out_i[wpos] = id; out_j[wpos] = perm[j]; ++wpos; }
    });
  });
  if (MODE == 3) return;
  unsigned long long tot = act ? (unsigned long long)cnt : 0ull;
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
  if (lane == 0 && pair_total) atomicAdd(pair_total, tot);
  if (!act) return;
  const int b = btrack[id];
  btrack[id] = (b == -1) ? -1 : cnt;
}

// ---------------------------------------------------------------------------
// Gradients, pair part.  With w = m_b grad W_ab / C the kernel accumulates
// neighbour-side sums (as in 2+1D): sum w (3), (p_b/sl_b^2) w (3), e_b w (3),
// (Pi_b/sl_b^2) w (3), v_b^j w_i (9) and, with shear, sum w.v_a (1),
// (w.v_a) pi^{0j}_b/sl_b^2 (3), sum_i w_i pi^{ji}_b/sl_b^2 (3).
template <bool SHEAR>
constexpr int grad3_nacc() { return SHEAR ? 28 : 21; }

// Register budget of the gradient walk: 4 waves per SIMD (128 VGPRs, a few
// values spilled to scratch) measured 1.115 ms per launch on the 126k event
// against 1.131 at 3 and 1.402 at the compiler's choice (183 VGPRs, 2 waves).
#ifndef CCAKE_GRAD3_WAVES_PER_SIMD
#define CCAKE_GRAD3_WAVES_PER_SIMD 4
#endif
template <bool SHEAR, bool FACT>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(CCAKE_GRAD3_WAVES_PER_SIMD)))
void k3_grad_slice(const double* __restrict__ cand,
                                                       const int32_t* __restrict__ skey,
                                                       const int32_t* __restrict__ cell_off,
                                                       const Grid* __restrict__ gridp, int n, double rc2, double rcs,
                                                       double hinv2, double heinv2, double kfac, int nslices,
                                                       double* __restrict__ part, int npad)
{
  constexpr int NACC = grad3_nacc<SHEAR>();
  const Grid& g = *gridp;
  const int lane = threadIdx.x;
  const int unit = xcd_tile(blockIdx.x, gridDim.x);
  const int tile = unit / nslices;
  const int slice = unit - tile * nslices;
  const int k = tile * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand + (size_t)kk * kCandG3;
  const double r[3] = { me[0], me[1], me[2] };
  const double va[3] = { me[7], me[8], me[9] };
  const int myrow = act ? skey[kk] / g.stride : INT_MAX;

  double acc[NACC];
  #pragma unroll
  for (int q = 0; q < NACC; ++q) acc[q] = 0.0;
  for_segments3(act, myrow, r, [&](const Seg3& seg, bool in) {
    for_spans_slice3(g, cell_off, seg, rcs, slice, nslices, in, [&](int j0, int j1) {
      constexpr int NR = SHEAR ? 19 : 10;
      auto pair = [&](const double* c, double dx, double dy, double dz, double* w) -> bool {
        const double r2 = dist2_3(dx, dy, dz);
        const bool nb = in && r2 <= rc2;
        if (kSkipEmpty && wave_none(nb)) return false;
        if (FACT) {
          double qp, iqp, qe, iqe;
          q_and_inverse(dist2_perp(dx, dy), hinv2, qp, iqp);
          q_and_inverse(dz * dz, heinv2, qe, iqe);
          const double m = nb ? c[3] : 0.0;
          const double gp = m * spline_g(qp, iqp) * spline_w(qe);
          w[0] = gp * dx; w[1] = gp * dy;
          w[2] = m * kfac * spline_w(qp) * spline_g(qe, iqe) * dz;
        } else {
          double q, iq;
          q_and_inverse(r2, hinv2, q, iq);
          const double gm = nb ? c[3] * spline_g(q, iq) : 0.0;
          w[0] = gm * dx; w[1] = gm * dy; w[2] = gm * dz;
        }
        return true;
      };
      auto sums = [&](const double* c, const double* w) {
        #pragma unroll
        for (int d = 0; d < 3; ++d) {
          acc[d] += w[d];
          acc[3 + d] += c[4] * w[d];
          acc[6 + d] += c[5] * w[d];
          acc[9 + d] += c[6] * w[d];
          #pragma unroll
          for (int e = 0; e < 3; ++e) acc[12 + 3 * d + e] += c[7 + e] * w[d];
        }
        if (SHEAR) {
          const double cv = w[0] * va[0] + w[1] * va[1] + w[2] * va[2];
          acc[SHEAR ? 21 : 0] += cv;
          #pragma unroll
          for (int e = 0; e < 3; ++e) acc[SHEAR ? 22 + e : 0] += cv * c[SHEAR ? 10 + e : 0];
          // pi^{ji}, i summed with w_i: rows of (xx, xy, x eta; xy, yy, y eta; x eta, y eta, eta eta)
          acc[SHEAR ? 25 : 0] += w[0] * c[SHEAR ? 13 : 0] + w[1] * c[SHEAR ? 14 : 0] + w[2] * c[SHEAR ? 15 : 0];
          acc[SHEAR ? 26 : 0] += w[0] * c[SHEAR ? 14 : 0] + w[1] * c[SHEAR ? 16 : 0] + w[2] * c[SHEAR ? 17 : 0];
          acc[SHEAR ? 27 : 0] += w[0] * c[SHEAR ? 15 : 0] + w[1] * c[SHEAR ? 17 : 0] + w[2] * c[SHEAR ? 18 : 0];
        }
      };
      walk_records3_split<NR, kCandG3>(cand, r, j0, j1, pair, sums);
    });
  });
  double* p = part + (size_t)slice * NACC * npad + k;
  #pragma unroll
  for (int q = 0; q < NACC; ++q) p[(size_t)q * npad] = acc[q];
}

// Sum pass: the slices in slice order, scaled by C, then the owner terms
// (from the particle's own record), written by particle id.
template <bool SHEAR>
__global__ __launch_bounds__(256) void k3_grad_sum(const double* __restrict__ cand, const int32_t* __restrict__ perm,
                                                   int n, double C, const double* __restrict__ part, int nslices,
                                                   int npad, double* __restrict__ F)
{
  constexpr int NACC = grad3_nacc<SHEAR>();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double acc[NACC];
  #pragma unroll
  for (int q = 0; q < NACC; ++q) acc[q] = 0.0;
  for (int sli = 0; sli < nslices; ++sli)
    #pragma unroll
    for (int q = 0; q < NACC; ++q) acc[q] += part[((size_t)sli * NACC + q) * npad + k];
  #pragma unroll
  for (int q = 0; q < NACC; ++q) acc[q] *= C;
  const double* me = cand + (size_t)k * kCandG3;
  const int id = perm[k];
  auto put = [&](int c, double v) { F[(size_t)c * npad + id] = v; };
  const double sla = me[19], isla = 1.0 / sla;
  #pragma unroll
  for (int d = 0; d < 3; ++d) {
    put(CCAKE_F_GRADP0 + d, sla * (acc[3 + d] + me[4] * acc[d]));
    put(CCAKE_F_GRADE0 + d, (acc[6 + d] - me[5] * acc[d]) * isla);
    put(CCAKE_F_GRADBULK0 + d, sla * (acc[9 + d] + me[6] * acc[d]));
    #pragma unroll
    for (int e = 0; e < 3; ++e) put(CCAKE_F_GRADV0 + 3 * d + e, (acc[12 + 3 * d + e] - me[7 + e] * acc[d]) * isla);
  }
  if (SHEAR) {
    const double acv = acc[SHEAR ? 21 : 0];
    #pragma unroll
    for (int e = 0; e < 3; ++e) put(CCAKE_F_GRADSHEAR0 + e, sla * (acc[SHEAR ? 22 + e : 0] + me[10 + e] * acv));
    // pi_a^{ji} a1_i with pi_a = (xx, xy, x eta, yy, y eta, eta eta) at me[13..18]
    put(CCAKE_F_DIVSHEAR0 + 0, sla * (acc[SHEAR ? 25 : 0] + me[13] * acc[0] + me[14] * acc[1] + me[15] * acc[2]));
    put(CCAKE_F_DIVSHEAR0 + 1, sla * (acc[SHEAR ? 26 : 0] + me[14] * acc[0] + me[16] * acc[1] + me[17] * acc[2]));
    put(CCAKE_F_DIVSHEAR0 + 2, sla * (acc[SHEAR ? 27 : 0] + me[15] * acc[0] + me[17] * acc[1] + me[18] * acc[2]));
  } else {
    for (int e = 0; e < 3; ++e) { put(CCAKE_F_GRADSHEAR0 + e, 0.0); put(CCAKE_F_DIVSHEAR0 + e, 0.0); }
  }
}

// ---------------------------------------------------------------------------
// Equations of motion, D = 3 (diffusion off).  Milne helpers
// (include/milne.hpp): a D-vector is lowered by (-1, -1, -tau^2) and raised by
// (-1, -1, -1/tau^2); Matrix<4,4>::make_covariant(slot) flips slots 1, 2 and
// multiplies slot 3 by -tau^2; Matrix<D,D>::make_contravariant(0) flips rows
// 0, 1 and multiplies row 2 by -1/tau^2.
struct Col3 {
  double* F; int npad; int i;
  __device__ __forceinline__ double& operator()(int c) const { return F[(size_t)c * npad + i]; }
};

__device__ __forceinline__ void mat4_cov(double M[4][4], int slot, double t2)
{
  if (slot == 0) {
    #pragma unroll
    for (int j = 0; j < 4; ++j) { M[1][j] = -M[1][j]; M[2][j] = -M[2][j]; M[3][j] *= -t2; }
  } else {
    #pragma unroll
    for (int i = 0; i < 4; ++i) { M[i][1] = -M[i][1]; M[i][2] = -M[i][2]; M[i][3] *= -t2; }
  }
}

// milne::inverse of a 3x3 matrix (cofactors over the rule-of-Sarrus determinant).
__device__ __forceinline__ void inverse3(const double A[3][3], double R[3][3])
{
  const double det = A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1]
                   - A[0][2] * A[1][1] * A[2][0] - A[0][1] * A[1][0] * A[2][2] - A[0][0] * A[1][2] * A[2][1];
  R[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
  R[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
  R[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
  R[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
  R[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
  R[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
  R[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
  R[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
  R[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
}

template <bool SHEAR>
__device__ __forceinline__ void eom_d3(const Col3& col, double t, double use_vort, double du_dt[3], double* d_s,
                                       double dN[3], double* d_b, double dshv[5], double* theta_out,
                                       double* snu_out, double* dez_out)
{
  const double t2 = t * t;
  const double cmd[4] = { 1.0, -1.0, -1.0, -1.0 / t / t };   // get_contra_metric_diagonal<3>
  const double dl[3] = { 0.0, 0.0, 1.0 };                     // delta_i_eta<3>
  const double gamma = col(CCAKE_F_GAMMA), sigma = col(CCAKE_F_SIGMA), bulk = col(CCAKE_F_BULK);
  const double T = col(CCAKE_F_T), w = col(CCAKE_F_W), s = col(CCAKE_F_S);
  const double dwds = col(CCAKE_F_DWDS);
  const double mu[3] = { col(CCAKE_F_MUB), col(CCAKE_F_MUS), col(CCAKE_F_MUQ) };
  const double rho[3] = { col(CCAKE_F_RHOB), col(CCAKE_F_RHOS), col(CCAKE_F_RHOQ) };
  const double dwdr[3] = { col(CCAKE_F_DWDB), col(CCAKE_F_DWDS_S), col(CCAKE_F_DWDQ) };
  const double zeta = col(CCAKE_F_ZETA_PI), tau_Pi = col(CCAKE_F_TAU_BULK);
  const double delta_PiPi = col(CCAKE_F_DELTA_BB), lambda_Pipi = col(CCAKE_F_LAMBDA_BPI);
  const double phi1 = col(CCAKE_F_PHI1), phi3 = col(CCAKE_F_PHI3);

  const double u[3] = { col(CCAKE_F_UX), col(CCAKE_F_UY), col(CCAKE_F_UETA) };
  const double v[3] = { col(CCAKE_F_VX), col(CCAKE_F_VY), col(CCAKE_F_VETA) };
  const double uc[3] = { -u[0], -u[1], u[2] * -t2 };            // u_cov (also the fixed-size u_cov)
  double gradV[3][3], divV = 0.0;
  #pragma unroll
  for (int a = 0; a < 3; ++a) {
    divV += col(CCAKE_F_GRADV0 + 4 * a);
    #pragma unroll
    for (int b = 0; b < 3; ++b) gradV[a][b] = col(CCAKE_F_GRADV0 + 3 * a + b);
  }
  const double gf = dl[2] * u[2] * u[2] * t / gamma;          // geometric_factor
  const double expf = gf + gamma / t + gamma * divV;
  double S[4][4], H[4][4], C[4][4];
  #pragma unroll
  for (int a = 0; a < 4; ++a)
    #pragma unroll
    for (int b = 0; b < 4; ++b) S[a][b] = col(CCAKE_F_SHV0 + 4 * a + b);
  #pragma unroll
  for (int a = 0; a < 4; ++a)
    #pragma unroll
    for (int b = 0; b < 4; ++b) H[a][b] = S[a][b];
  mat4_cov(H, 1, t2);
  #pragma unroll
  for (int a = 0; a < 4; ++a)
    #pragma unroll
    for (int b = 0; b < 4; ++b) C[a][b] = H[a][b];
  mat4_cov(C, 0, t2);
  double shv_shvcov = 0.0;
  #pragma unroll
  for (int a = 0; a < 4; ++a)
    #pragma unroll
    for (int b = 0; b < 4; ++b) shv_shvcov += S[a][b] * C[a][b];

  // ---------------- calculate_MRF_shear (src/eom_default.cpp:882-1412) ----------------
  double M_snu[3] = { 0, 0, 0 }, F_snu = 0.0;
  double F_0i[3] = { 0, 0, 0 }, M_0i[3][3] = { { 0 } };
  double F_es[2][3] = { { 0 } }, M_es[2][3][3] = { { { 0 } } };
  if (SHEAR) {
    const double eta_pi = col(CCAKE_F_ETA_PI), tau_pi = col(CCAKE_F_TAU_PI);
    const double delta_pipi = col(CCAKE_F_DELTA_PIPI), lambda_piPi = col(CCAKE_F_LAMBDA_PIB);
    const double phi6 = col(CCAKE_F_PHI6), phi7 = col(CCAKE_F_PHI7), tau_pipi = col(CCAKE_F_TAU_PIPI);
    double guj[3][3], gu0[3];
    #pragma unroll
    for (int a = 0; a < 3; ++a)
      #pragma unroll
      for (int b = 0; b < 3; ++b) {
        double x = gamma * gradV[a][b];
        #pragma unroll
        for (int k = 0; k < 3; ++k) x += -gamma * u[b] * uc[k] * gradV[a][k];
        guj[a][b] = x;
      }
    #pragma unroll
    for (int a = 0; a < 3; ++a) {
      double c = 0.0;
      #pragma unroll
      for (int b = 0; b < 3; ++b) c += guj[a][b] * uc[b];
      gu0[a] = -1. * c / gamma;
    }
    const double gu0c[3] = { -gu0[0], -gu0[1], gu0[2] * (-1.0 / t2) };
    double cg[3][3];                                            // contra_grad_uj
    #pragma unroll
    for (int b = 0; b < 3; ++b) { cg[0][b] = -guj[0][b]; cg[1][b] = -guj[1][b]; cg[2][b] = guj[2][b] * (-1.0 / t2); }
    double guv[3] = { 0, 0, 0 };                                // contract(grad_uj, v, FirstIndex)
    #pragma unroll
    for (int b = 0; b < 3; ++b)
      #pragma unroll
      for (int a = 0; a < 3; ++a) guv[b] += guj[a][b] * v[a];
    double gu0v = 0.0;
    #pragma unroll
    for (int a = 0; a < 3; ++a) gu0v += gu0[a] * v[a];
    double vgs[3] = { 0, 0, 0 }, gugs[3] = { 0, 0, 0 }, sgui[3] = { 0, 0, 0 }, vsgu[3] = { 0, 0, 0 };
    #pragma unroll
    for (int a = 0; a < 3; ++a) {
      #pragma unroll
      for (int b = 0; b < 3; ++b) vgs[a] += guv[b] * H[a + 1][b + 1];
      vgs[a] += gu0v * H[a + 1][0];
      #pragma unroll
      for (int b = 0; b < 3; ++b) gugs[a] += cg[a][b] * H[0][b + 1];
      gugs[a] += cmd[a + 1] * gu0[a] * H[0][0];
      #pragma unroll
      for (int b = 0; b < 3; ++b) sgui[a] += S[b + 1][0] * guj[b][a] + S[b + 1][a + 1] * gu0[b];
      #pragma unroll
      for (int b = 0; b < 3; ++b) vsgu[a] += v[b] * S[0][0] * guj[b][a] + v[b] * S[0][a + 1] * gu0[b];
    }
    double sgu = 0.0;                                           // shv_grad_u
    #pragma unroll
    for (int a = 0; a < 3; ++a) {
      #pragma unroll
      for (int b = 0; b < 3; ++b) sgu += (H[a + 1][b + 1] - v[a] * H[0][b + 1]) * guj[a][b];
      sgu += (H[a + 1][0] - v[a] * H[0][0]) * gu0[a];
    }
    double sl0l[3] = { 0, 0, 0 };
    #pragma unroll
    for (int a = 0; a < 3; ++a)
      #pragma unroll
      for (int b = 0; b < 4; ++b) sl0l[a] += (S[0][b] * H[a + 1][b] + S[a + 1][b] * H[0][b]) / 2.;
    const double uD = u[2];
    #pragma unroll
    for (int a = 0; a < 3; ++a) {
      double Fsig = -(guv[a] - gu0c[a]) / 2. + u[a] * gamma * (gf + gamma / t + gamma * divV) / 3.
                  - gamma * gf * u[a] / 2.;
      double FD = gamma * (u[a] * H[0][0] + gamma * H[a + 1][0]) * gf + t * uD * S[3][a + 1] * dl[a];
      double Fds = -(vgs[a] - gugs[a] - 4. * gamma * u[a] * sgu / 3. - sgui[a] + vsgu[a]) / 4.
                 + (H[a + 1][0] - gamma * gamma * H[a + 1][0] - gamma * u[a] * H[0][0] + S[0][a + 1]
                    + 4. * u[a] * gamma * H[0][0] / 3.) * gf / 4.
                 + (gamma * cmd[a + 1] * H[0][3] * dl[a] / t + gamma * S[3][0] * dl[a] / t
                    + 4. * gamma * u[a] * (gamma * H[3][3] / t) / 3.) / 4.
                 - S[0][a + 1] * (gamma / t + gf + gamma * divV) / 3.;
      double Fdo = -(vgs[a] - gugs[a] + sgui[a] - vsgu[a]) / 4.
                 + (H[a + 1][0] - gamma * gamma * H[a + 1][0] - gamma * u[a] * H[0][0] - S[0][a + 1]) * gf / 4.
                 + (gamma * cmd[a + 1] * H[0][3] * dl[a] / t - gamma * S[3][0] * dl[a] / t) / 4.;
      const double Fdd = S[a + 1][0] * expf;
      double Msig[3], MD[3], Mds[3], Mdo[3], Mdd[3];
      #pragma unroll
      for (int b = 0; b < 3; ++b) {
        Msig[b] = u[a] * uc[b] / (2. * 3.);
        MD[b] = gamma * u[a] * (H[0][b + 1] - H[0][0] * uc[b] / gamma)
              + gamma * gamma * (H[a + 1][b + 1] - H[a + 1][0] * uc[b] / gamma);
        Mds[b] = (H[a + 1][b + 1] - H[a + 1][0] * uc[b] / gamma - gamma * gamma * H[a + 1][b + 1]
                  - gamma * u[a] * H[0][b + 1] + uc[b] * gamma * H[a + 1][0] + uc[b] * u[a] * H[0][0]
                  + 4. * gamma * u[a] * (H[0][b + 1] - H[0][0] * uc[b] / gamma) / 3.
                  + (-S[0][a + 1] + 4. * S[0][a + 1] / 3.) * uc[b] / gamma) / 4.;
        Mdo[b] = (H[a + 1][b + 1] - H[a + 1][0] * uc[b] / gamma - gamma * gamma * H[a + 1][b + 1]
                  - gamma * u[a] * H[0][b + 1] + uc[b] * gamma * H[a + 1][0] + uc[b] * u[a] * H[0][0]
                  + S[0][a + 1] * uc[b] / gamma) / 4.;
        Mdd[b] = -S[a + 1][0] * uc[b] / gamma;
      }
      Msig[a] += (1. - gamma * gamma) / 2.;
      Mdo[a] += -S[0][0] / 4.;
      Mds[a] += +S[0][0] / 4.;
      // metric-only terms
      Fsig += dl[a] * (-(-uD * t - uD / t + gamma * uD * uD * t + 2. * gamma * gamma * uD / t) / 2.);
      FD += dl[a] * (uD * S[0][0] / t + gamma * S[3][0] / t);
      Fds += (t * uD * dl[a] * cmd[a + 1] * H[0][0] + uD * H[a + 1][3] * gamma / t / 4. + H[a + 1][3] * t * uD
              + S[0][0] * dl[a] * uD / t + t * uD * uD * (gamma * H[a + 1][0] + u[a] * H[0][0])
              + 2. * gamma * uD * (gamma * H[a + 1][3] + u[a] * S[0][3]) / t) / 4.
           + gamma * u[a] * (H[3][0] * uD * t + uD * H[0][3] / t) / 3.;
      Fdo += (t * uD * dl[a] * cmd[a + 1] * H[0][0] + uD * H[a + 1][3] * gamma / t - H[3][a + 1] * t * uD
              - S[0][0] * dl[a] * uD / t + t * uD * uD * (gamma * H[a + 1][0] + u[a] * H[0][0])
              + 2. * gamma * uD * (gamma * H[a + 1][3] + u[a] * S[0][3]) / t) / 4.;
      F_0i[a] = (-S[a + 1][0] - tau_pi * FD - delta_pipi * Fdd + 2. * tau_pi * Fdo * use_vort - tau_pipi * Fds
                 + (2. * eta_pi + lambda_piPi * bulk) * Fsig + phi6 * bulk * S[a + 1][0]
                 + phi7 * (sl0l[a] + gamma * u[a] * shv_shvcov / 3.)) / (tau_pi * gamma);
      #pragma unroll
      for (int b = 0; b < 3; ++b)
        M_0i[a][b] = (-tau_pi * MD[b] - delta_pipi * Mdd[b] + 2. * tau_pi * Mdo[b] * use_vort - tau_pipi * Mds[b]
                      + (2. * eta_pi + lambda_piPi * bulk) * Msig[b]) / (tau_pi * gamma);
    }
    // shv_nabla_u = pi^{ij} nabla_i u_j
    #pragma unroll
    for (int a = 0; a < 3; ++a) {
      M_snu[a] = H[0][a + 1] - H[0][0] * uc[a] / gamma;
      #pragma unroll
      for (int b = 0; b < 3; ++b) F_snu += (H[a + 1][b + 1] - v[a] * H[0][b + 1]) * guj[a][b];
      F_snu += (H[a + 1][0] - v[a] * H[0][0]) * gu0[a];
    }
    F_snu += +H[3][3] * gamma / t + H[0][0] * gf + (t * H[3][0] + H[0][3] / t) * u[2] * dl[2];

    // extensive-shear MRFs, components (a, b) = (0,0), (0,1), (0,2), (1,1), (1,2)
    const double tilde_delta = delta_pipi - tau_pi;
    #pragma unroll
    for (int a = 0; a < 2; ++a) {
      double FDm[3] = { 0, 0, 0 }, Fsm[3] = { 0, 0, 0 }, Fdl[3] = { 0, 0, 0 }, Fdom[3] = { 0, 0, 0 },
             Fdsm[3] = { 0, 0, 0 };
      double MDm[3][3] = { { 0 } }, Msm[3][3] = { { 0 } }, Mdl[3][3] = { { 0 } }, Mdom[3][3] = { { 0 } },
             Mdsm[3][3] = { { 0 } };
      double sls[3] = { 0, 0, 0 };
      #pragma unroll
      for (int b = a; b < 3; ++b) {
        double gugj = 0.0, sgu2 = 0.0, vsg = 0.0;
        #pragma unroll
        for (int k = 0; k < 3; ++k) {
          gugj += cg[a][k] * H[b + 1][k + 1] + cg[b][k] * H[a + 1][k + 1];
          sgu2 += S[k + 1][a + 1] * guj[k][b] + S[k + 1][b + 1] * guj[k][a];
          vsg += v[k] * S[0][a + 1] * guj[k][b] + v[k] * S[0][b + 1] * guj[k][a];
        }
        #pragma unroll
        for (int k = 0; k < 4; ++k) sls[b] += (S[a + 1][k] * H[b + 1][k] + S[b + 1][k] * H[a + 1][k]) / 2.;
        FDm[b] = gamma * gf * (u[a] * H[b + 1][0] + u[b] * H[a + 1][0]);
        Fsm[b] = cg[a][b] / 2. + cg[b][a] / 2. + u[a] * u[b] * expf / 3.;
        Fdl[b] = S[a + 1][b + 1] * expf;
        const double metric_terms =
            t * u[2] * (cmd[a + 1] * H[b + 1][0] * dl[a] + cmd[b + 1] * H[a + 1][0] * dl[b])
          + gamma * (cmd[a + 1] * H[b + 1][3] * dl[a] + cmd[b + 1] * H[a + 1][3] * dl[b]) / t
          + u[2] * u[2] * t * (u[a] * H[b + 1][0] + u[b] * H[a + 1][0])
          + 2. * gamma * u[2] * (u[a] * H[b + 1][3] + u[b] * H[a + 1][3]) / t;
        Fdom[b] = -(-gugj + sgu2 - vsg) / 4.
                + (-u[a] * gamma * H[b + 1][0] - u[b] * gamma * H[a + 1][0]) * gf / 4.
                + (metric_terms - gamma * (S[3][a + 1] * dl[b] + S[3][b + 1] * dl[a]) / t
                   - u[2] * (S[0][a + 1] * dl[b] + S[0][b + 1] * dl[a]) / t) / 4.;
        Fdsm[b] = -(-gugj - sgu2 + vsg - 4. * (u[a] * u[b]) * sgu / 3.) / 4.
                + (-u[a] * gamma * H[b + 1][0] - u[b] * gamma * H[a + 1][0] + 4. * u[a] * u[b] * H[0][0] / 3.) * gf / 4.
                + (metric_terms + gamma * (S[3][a + 1] * dl[b] + S[3][b + 1] * dl[a]) / t
                   + u[2] * (S[0][a + 1] * dl[b] + S[0][b + 1] * dl[a]) / t
                   + 4. * u[a] * u[b] * (H[3][0] * u[2] * t + gamma * H[3][3] / t + u[2] * H[0][3] / t) / 3.) / 4.
                - S[a + 1][b + 1] * (gamma / t + gf + gamma * divV) / 3.;
        #pragma unroll
        for (int k = 0; k < 3; ++k) {
          MDm[b][k] = gamma * (u[a] * H[b + 1][k + 1] - u[a] * H[b + 1][0] * uc[k] / gamma
                               + u[b] * H[a + 1][k + 1] - u[b] * H[a + 1][0] * uc[k] / gamma);
          Msm[b][k] = -u[a] * u[b] * uc[k] / (3. * gamma);
          Mdl[b][k] = -S[a + 1][b + 1] * uc[k] / gamma;
          Mdom[b][k] = (-u[a] * H[b + 1][k + 1] * gamma - u[b] * H[a + 1][k + 1] * gamma
                        + uc[k] * u[a] * H[b + 1][0] + uc[k] * u[b] * H[a + 1][0]) / 4.;
          Mdsm[b][k] = (-u[a] * H[b + 1][k + 1] * gamma - u[b] * H[a + 1][k + 1] * gamma
                        + uc[k] * u[a] * H[b + 1][0] + uc[k] * u[b] * H[a + 1][0]
                        + 4. * u[a] * u[b] * (H[0][k + 1] - H[0][0] * uc[k] / gamma) / 3.
                        + 4. * S[a + 1][b + 1] * uc[k] / 3. / gamma) / 4.;
        }
        Mdsm[b][a] += S[0][b + 1] / 4.;
        Mdom[b][a] += -S[0][b + 1] / 4.;
        Msm[b][a] += -gamma * u[b] / 2.;
      }
      #pragma unroll
      for (int b = a; b < 3; ++b) {
        Msm[b][b] += -gamma * u[a] / 2.;
        Mdsm[b][b] += S[0][a + 1] / 4.;
        Mdom[b][b] += -S[0][a + 1] / 4.;
      }
      #pragma unroll
      for (int k = 0; k < 3; ++k) {
        Msm[a][k] += cmd[a + 1] * uc[k] / (gamma * 3.);
        Mdsm[a][k] += -cmd[a + 1] * (H[0][k + 1] - H[0][0] * uc[k] / gamma) / 3.;
      }
      const double ueta = u[2];
      FDm[2] += gamma * S[3][a + 1] / t + ueta * S[0][a + 1] / t;
      Fsm[a] += -cmd[a + 1] * expf / 3.;
      Fsm[2] += -u[a] * ueta * ueta * t / 2. - u[a] * ueta * gamma / t;
      Fdsm[a] += -cmd[a + 1] * sgu / 3. - cmd[a + 1] * H[0][0] * gf / 3.
                 - cmd[a + 1] * (H[3][0] * u[2] * t + gamma * H[3][3] / t + u[2] * H[0][3] / t) / 3.;
      #pragma unroll
      for (int b = a; b < 3; ++b) {
        F_es[a][b] = (-S[a + 1][b + 1] - tilde_delta * Fdl[b] - tau_pi * FDm[b] + 2. * tau_pi * Fdom[b] * use_vort
                      - tau_pipi * Fdsm[b] + (2. * eta_pi + lambda_piPi * bulk) * Fsm[b] + phi6 * bulk * S[a + 1][b + 1]
                      + phi7 * (sls[b] + u[a] * u[b] * shv_shvcov / 3.)) / (sigma * tau_pi * gamma);
        #pragma unroll
        for (int k = 0; k < 3; ++k)
          M_es[a][b][k] = (-tau_pi * MDm[b][k] - tilde_delta * Mdl[b][k] + 2. * tau_pi * Mdom[b][k] * use_vort
                           - tau_pipi * Mdsm[b][k] + (2. * eta_pi + lambda_piPi * bulk) * Msm[b][k])
                        / (sigma * tau_pi * gamma);
      }
      F_es[a][a] += phi7 * (-cmd[a + 1] * shv_shvcov) / (3. * tau_pi * gamma * sigma);
    }
  }

  // ---------------- fill_auxiliary_variables (230-362) ----------------
  const double zeta_tilde = zeta + (delta_PiPi - tau_Pi) * bulk;
  double M_eb[3], M_S[3];
  #pragma unroll
  for (int a = 0; a < 3; ++a) {
    M_eb[a] = (zeta_tilde * uc[a] / gamma) / (sigma * gamma * tau_Pi);
    M_S[a] = (bulk * uc[a] / gamma) / (sigma * gamma * T);
  }
  double R_S[3];
  #pragma unroll
  for (int c = 0; c < 3; ++c) R_S[c] = -mu[c] / T;
  double F_eb = -(zeta_tilde * (gamma * divV + gamma / t + gf) + bulk - phi1 * bulk * bulk) / (sigma * gamma * tau_Pi);
  double F_S = (-bulk * (gamma * divV + gamma / t + gf)) / (sigma * gamma * T);
  if (SHEAR) {
    F_S += F_snu / (sigma * gamma * T);
    F_eb += -(lambda_Pipi * F_snu - phi3 * shv_shvcov) / (sigma * gamma * tau_Pi);
    #pragma unroll
    for (int a = 0; a < 3; ++a) {
      M_eb[a] += -lambda_Pipi * M_snu[a] / (sigma * gamma * tau_Pi);
      M_S[a] += M_snu[a] / (sigma * gamma * T);
    }
  }

  // ---------------- compute_charge_derivative (369-453), diffusion off ----------------
  // R_N = gamma sigma I, F_N = 0 (no sources), M_N = 0: dN/dt = 0.

  // ---------------- compute_velocity_derivative (459-618) ----------------
  const double gP[3] = { -col(CCAKE_F_GRADP0), -col(CCAKE_F_GRADP0 + 1), col(CCAKE_F_GRADP0 + 2) * (-1.0 / t2) };
  const double gB[3] = { -col(CCAKE_F_GRADBULK0), -col(CCAKE_F_GRADBULK0 + 1),
                         col(CCAKE_F_GRADBULK0 + 2) * (-1.0 / t2) };
  const double rdw = rho[0] * dwdr[0] + rho[1] * dwdr[1] + rho[2] * dwdr[2];
  double M_w[3], R_w[3];
  #pragma unroll
  for (int a = 0; a < 3; ++a) M_w[a] = (s * dwds / gamma + rdw / gamma) * uc[a] / gamma + sigma * dwds * M_S[a];
  #pragma unroll
  for (int c = 0; c < 3; ++c) R_w[c] = sigma * dwdr[c] + sigma * dwds * R_S[c];
  (void)R_w;   // R_u R_N^{-1} multiplies M_N = 0 and F_N = 0 without diffusion
  double F_S_ex = F_S;
  const double F_w = -(s * dwds / gamma + rdw / gamma) * (gf + gamma / t + gamma * divV) + sigma * dwds * F_S_ex;
  const double F_bulk = sigma * F_eb - bulk * (gamma * divV + gamma / t + gf) / gamma;
  double M_u[3][3], F_u[3];
  #pragma unroll
  for (int a = 0; a < 3; ++a) {
    #pragma unroll
    for (int b = 0; b < 3; ++b) M_u[a][b] = u[a] * gamma * (M_eb[b] + M_w[b]) - u[a] * uc[b] * (w + bulk) / gamma;
    M_u[a][a] += gamma * (w + bulk);
  }
  #pragma unroll
  for (int a = 0; a < 3; ++a)
    F_u[a] = (gP[a] + gB[a]) - gamma * (F_w + F_bulk) * u[a] - (w + bulk) * (gamma * divV + gamma / t + gf) * u[a]
           - 2. * (w + bulk) * gamma * u[2] * dl[a] / t;
  if (SHEAR) {
    #pragma unroll
    for (int a = 0; a < 3; ++a) {
      F_u[a] += -S[a + 1][0] / t - 2. * dl[a] * S[0][3] / t - col(CCAKE_F_DIVSHEAR0 + a) + col(CCAKE_F_GRADSHEAR0 + a)
              - F_0i[a];
      #pragma unroll
      for (int b = 0; b < 3; ++b) M_u[a][b] += M_0i[a][b];
    }
  }
  double MI[3][3];
  inverse3(M_u, MI);
  #pragma unroll
  for (int a = 0; a < 3; ++a) {
    double x = 0.0;
    #pragma unroll
    for (int b = 0; b < 3; ++b) x += MI[a][b] * (F_u[b] + 0.0);
    du_dt[a] = x;
  }

  // ---------------- compute_derivatives (623-825) ----------------
  dN[0] = dN[1] = dN[2] = 0.0;
  double c1 = 0.0, c2 = 0.0, c3 = 0.0, ucdu = 0.0;
  #pragma unroll
  for (int a = 0; a < 3; ++a) {
    c1 += M_S[a] * du_dt[a];
    c2 += M_eb[a] * du_dt[a];
    c3 += M_snu[a] * du_dt[a];
    ucdu += uc[a] * du_dt[a];
  }
  const double rs = R_S[0] * dN[0] + R_S[1] * dN[1] + R_S[2] * dN[2];
  *d_s = c1 + F_S + rs;
  *d_b = c2 + F_eb + 0.0;
  *snu_out = c3 + F_snu;
  *theta_out = -1. * ucdu / gamma + gamma * divV + gamma / t + gf;
  if (SHEAR) {
    int q = 0;
    #pragma unroll
    for (int a = 0; a < 2; ++a)
      #pragma unroll
      for (int b = a; b < 3; ++b) {
        double m = 0.0;
        #pragma unroll
        for (int c = 0; c < 3; ++c) m += M_es[a][b][c] * du_dt[c];
        dshv[q++] = m + F_es[a][b];
      }
  }
  // ---------------- compute_Ez_derivative (828-862) ----------------
  const double p = col(CCAKE_F_P), e = col(CCAKE_F_E);
  *dez_out = ((e + p + bulk) * u[2] * u[2] + (p + bulk) / t2 + S[3][3]) * col(CCAKE_F_MASS_S) * t2
             / col(CCAKE_F_SIGMA_LAB);
}

__device__ __forceinline__ void cache_state3(const Col3& col)
{
  for (int q = 0; q < 5; ++q) col(P_C_SHV0 + q) = col(CCAKE_F_EXT_SHV0 + q);
  col(P_C_UX) = col(CCAKE_F_UX); col(P_C_UY) = col(CCAKE_F_UY); col(P_C_UETA) = col(CCAKE_F_UETA);
  col(P_C_X) = col(CCAKE_F_X); col(P_C_Y) = col(CCAKE_F_Y); col(P_C_ETA) = col(CCAKE_F_ETA);
  col(P_C_S) = col(CCAKE_F_EXT_S); col(P_C_BULK) = col(CCAKE_F_EXT_BULK);
  col(P_C_B) = col(CCAKE_F_EXT_B); col(P_C_SS) = col(CCAKE_F_EXT_SS); col(P_C_Q) = col(CCAKE_F_EXT_Q);
  col(P_C_EZ) = col(CCAKE_F_EZ);
}

// y = y0 + dt f with the negative-entropy rule (Evolver::step_rk,
// src/evolver.cpp:363-484), from the derivatives stored in F.
__device__ __forceinline__ void rk_update3(const Col3& col, int freeze, double dt, int32_t* __restrict__ err)
{
  double es = col(P_C_S) + dt * col(CCAKE_F_DEXT_S);
  if (es < 0.0 && freeze > 3) es = 1.e-3;
  else if (es < 0.0) atomicOr(err, DE_NEG_ENTROPY);
  col(CCAKE_F_EXT_S) = es;
  col(CCAKE_F_EXT_BULK) = col(P_C_BULK) + dt * col(CCAKE_F_DEXT_BULK);
  col(CCAKE_F_EXT_B) = col(P_C_B) + dt * col(CCAKE_F_DEXT_B);
  col(CCAKE_F_EXT_SS) = col(P_C_SS) + dt * col(CCAKE_F_DEXT_SS);
  col(CCAKE_F_EXT_Q) = col(P_C_Q) + dt * col(CCAKE_F_DEXT_Q);
  col(CCAKE_F_EZ) = col(P_C_EZ) + dt * col(CCAKE_F_DEZ);
  col(CCAKE_F_X) = col(P_C_X) + dt * col(CCAKE_F_VX);
  col(CCAKE_F_Y) = col(P_C_Y) + dt * col(CCAKE_F_VY);
  col(CCAKE_F_ETA) = col(P_C_ETA) + dt * col(CCAKE_F_VETA);
  col(CCAKE_F_UX) = col(P_C_UX) + dt * col(CCAKE_F_DUDT0 + 0);
  col(CCAKE_F_UY) = col(P_C_UY) + dt * col(CCAKE_F_DUDT0 + 1);
  col(CCAKE_F_UETA) = col(P_C_UETA) + dt * col(CCAKE_F_DUDT0 + 2);
  for (int q = 0; q < 5; ++q) col(CCAKE_F_EXT_SHV0 + q) = col(P_C_SHV0 + q) + dt * col(CCAKE_F_DEXT_SHV0 + q);
}

// RK: 0 derivatives only, 1 and the RK update, 2 the RK cache first, then
// the update (the first stage of a step).
template <bool SHEAR>
__global__ __launch_bounds__(256) void k3_eom(double* __restrict__ F, int npad, int n,
                                              const int32_t* __restrict__ freeze, double t, double use_vort,
                                              double dt, int32_t* __restrict__ err, int rk)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Col3 col{ F, npad, i };
  double du[3], ds, dN[3], db, dshv[5] = { 0, 0, 0, 0, 0 }, theta, snu, dez;
  eom_d3<SHEAR>(col, t, use_vort, du, &ds, dN, &db, dshv, &theta, &snu, &dez);
  for (int a = 0; a < 3; ++a) col(CCAKE_F_DUDT0 + a) = du[a];
  col(CCAKE_F_DEXT_S) = ds;
  col(CCAKE_F_DEXT_B) = dN[0]; col(CCAKE_F_DEXT_SS) = dN[1]; col(CCAKE_F_DEXT_Q) = dN[2];
  col(CCAKE_F_DEXT_BULK) = db;
  if (SHEAR) for (int q = 0; q < 5; ++q) col(CCAKE_F_DEXT_SHV0 + q) = dshv[q];
  col(CCAKE_F_THETA) = theta;
  col(CCAKE_F_SHV_NABLA_U) = snu;
  col(CCAKE_F_DEZ) = dez;
  if (rk == 0) return;
  if (rk == 2) cache_state3(col);
  rk_update3(col, freeze[i], dt, err);
}

__global__ __launch_bounds__(256) void k3_rk(double* __restrict__ F, int npad, int n,
                                             const int32_t* __restrict__ freeze, double dt, int32_t* __restrict__ err)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rk_update3(Col3{ F, npad, i }, freeze[i], dt, err);
}

__global__ __launch_bounds__(256) void k3_cache(double* __restrict__ F, int npad, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cache_state3(Col3{ F, npad, i });
}

// ---------------------------------------------------------------------------
// Launchers.
int launch3_stage_begin(ccake_event* ev, int do_gamma)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k3_stage_begin, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n, ev->bbox,
                     do_gamma, ev->t * ev->t);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch3_stage_gather(ccake_event* ev)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k3_stage_gather, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n, ev->perm,
                     ev->cand_s, (unsigned long long*)ev->pair_count, ev->t * ev->t);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

double smooth_norm2_event(const ccake_event* ev)
{
  if (ev->p.dim == 3) return kern3(ev->p).norm2;
  return smooth_norm2(ev->p.hT);
}

int launch3_smooth_pairs(ccake_event* ev, int mode, bool count_pairs)
{
  const int n = ev->n;
  ev->smooth_sum_slices = ev->nslices;
  const double rc = 2.0 * ev->p.hT;
  const double rc2 = rc * rc;
  const double rcs = search_radius(ev);
  const int ntiles = (n + kTile - 1) / kTile;
  hipStream_t st = ev->stream;
  if (mode == 2) {
    int32_t* bt = ev->I + (size_t)CCAKE_I_BTRACK * ev->npad;
    unsigned long long* pc = count_pairs ? (unsigned long long*)ev->pair_count : nullptr;
    hipLaunchKernelGGL(k3_walk<2>, dim3(ntiles), dim3(kTile), 0, st, ev->cand_s, ev->skey, ev->perm, ev->cell_off,
                       ev->grid, n, rc2, rcs, bt, pc, nullptr, nullptr, nullptr);
    return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
  }
  const Kern3 K = kern3(ev->p);
  const dim3 grid(ntiles * ev->nslices), blk(kWave);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, blk, 0, st, ev->cand_s, ev->skey, ev->cell_off, ev->grid, n, rc2, rcs, K.hinv2,
                       K.heinv2, ev->nslices, ev->part, ev->npad);
  };
  if (ev->p.factorized_kernel) {
    if (mode == 0) go(k3_smooth_slice<0, true>); else go(k3_smooth_slice<1, true>);
  } else {
    if (mode == 0) go(k3_smooth_slice<0, false>); else go(k3_smooth_slice<1, false>);
  }
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch3_export_pairs(ccake_event* ev, int32_t* pi, int32_t* pj, size_t cap, size_t* n_out)
{
  const int n = ev->n;
  hipStream_t st = ev->stream;
  std::vector<int32_t> saved_bt(n), cnt(n);
  int32_t* bt = ev->I + (size_t)CCAKE_I_BTRACK * ev->npad;
  CCAKE_HIP(hipMemcpyAsync(saved_bt.data(), bt, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipMemsetAsync(bt, 0, n * sizeof(int32_t), st));
  const int rc_count = launch3_smooth_pairs(ev, 2, false);
  if (rc_count) { (void)hipStreamSynchronize(st); return rc_count; }
  CCAKE_HIP(hipMemcpyAsync(cnt.data(), bt, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipStreamSynchronize(st));
  CCAKE_HIP(hipMemcpyAsync(bt, saved_bt.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  CCAKE_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> off(n + 1, 0);
  for (int i = 0; i < n; ++i) off[i + 1] = off[i] + cnt[i];
  const size_t total = (size_t)off[n];
  *n_out = total;
  if (total > cap) return CCAKE_ERR_CAPACITY;
  struct DevBuf {
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
  hipLaunchKernelGGL(k3_walk<3>, dim3((n + kTile - 1) / kTile), dim3(kTile), 0, st, ev->cand_s, ev->skey, ev->perm,
                     ev->cell_off, ev->grid, n, rc * rc, search_radius(ev), bt, nullptr, d_off, d_i, d_j);
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess) { (void)hipStreamSynchronize(st); return hip_fail(le, "k3_walk"); }
  CCAKE_HIP(hipMemcpyAsync(pi, d_i, total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipMemcpyAsync(pj, d_j, total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < n; ++i) std::sort(pj + off[i], pj + off[i + 1]);
  return 0;
}

int launch3_grad(ccake_event* ev)
{
  const int n = ev->n;
  const double rc = 2.0 * ev->p.hT;
  const double rcs = search_radius(ev);
  const Kern3 K = kern3(ev->p);
  const int ntiles = (n + kTile - 1) / kTile;
  hipStream_t st = ev->stream;
  const dim3 grid(ntiles * ev->nslices), blk(kWave);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, blk, 0, st, ev->cand_g, ev->skey, ev->cell_off, ev->grid, n, rc * rc, rcs,
                       K.hinv2, K.heinv2, K.kfac, ev->nslices, ev->part, ev->npad);
  };
  const bool shear = ev->p.using_shear != 0;
  if (shear) {
    if (ev->p.factorized_kernel) go(k3_grad_slice<true, true>); else go(k3_grad_slice<true, false>);
  } else {
    if (ev->p.factorized_kernel) go(k3_grad_slice<false, true>); else go(k3_grad_slice<false, false>);
  }
  const dim3 g2((n + 255) / 256), b2(256);
  if (shear)
    hipLaunchKernelGGL(k3_grad_sum<true>, g2, b2, 0, st, ev->cand_g, ev->perm, n, K.C, ev->part, ev->nslices, ev->npad,
                       ev->F);
  else
    hipLaunchKernelGGL(k3_grad_sum<false>, g2, b2, 0, st, ev->cand_g, ev->perm, n, K.C, ev->part, ev->nslices,
                       ev->npad, ev->F);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch3_eom(ccake_event* ev, int fuse_rk, double dt)
{
  const int n = ev->n;
  const dim3 g((n + 255) / 256), b(256);
  const int32_t* fz = ev->I + (size_t)CCAKE_I_FREEZE * ev->npad;
  const double uv = ev->p.using_vorticity ? 1.0 : 0.0;
  if (ev->p.using_shear)
    hipLaunchKernelGGL(k3_eom<true>, g, b, 0, ev->stream, ev->F, ev->npad, n, fz, ev->t, uv, dt, ev->err, fuse_rk);
  else
    hipLaunchKernelGGL(k3_eom<false>, g, b, 0, ev->stream, ev->F, ev->npad, n, fz, ev->t, uv, dt, ev->err, fuse_rk);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch3_rk(ccake_event* ev, double dt)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k3_rk, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n,
                     ev->I + (size_t)CCAKE_I_FREEZE * ev->npad, dt, ev->err);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

int launch3_cache(ccake_event* ev)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k3_cache, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

}  // namespace ccake
