// Internal declarations of the MI355X SPH library (host + device).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>
#include "../../include/ccake_amd.h"
#include "preinv_device.h"

namespace ccake {

constexpr double kHbarcMeVfm = 197.3269804;    // include/constants.h:16
constexpr double kPi = 3.1415926535897932384626433;
constexpr int kWave = 64;                       // CDNA wavefront width
constexpr int kTile = 64;                       // targets per pair-kernel workgroup

// Private per-particle columns stored after the public fields.
enum PrivateField {
  P_C_UX = CCAKE_NUM_FIELDS, P_C_UY, P_C_UETA,     // RK cache: u0
  P_C_X, P_C_Y, P_C_ETA,                           // r0
  P_C_S, P_C_B, P_C_SS, P_C_Q, P_C_BULK, P_C_EZ,   // extensive s, B, S, Q, Pi; Ez
  P_C_SHV0,                                        // 5 extensive shear
  P_FB_E0 = P_C_SHV0 + 5,                          // fback..fback4 energy (4 columns)
  P_FB_G0 = P_FB_E0 + 4,                           // fback..fback4 gradP_x (4 columns)
  P_FRZ1_G = P_FB_G0 + 4, P_FRZ2_G,                // frz1 / frz2 gradP_x
  P_C_Q0,                                          // RK cache: 9 extensive diffusion
  NUM_COLUMNS = P_C_Q0 + 9
};
// Multi-column members own disjoint column ranges: every enumerator after a
// block of k columns starts k past it (an overlap would let one member's
// writes land in another's column, or past the last column of the buffer).
static_assert(P_C_X == P_C_UX + 3 && P_C_S == P_C_X + 3 && P_C_SHV0 == P_C_S + 6, "RK cache columns overlap");
static_assert(P_FB_E0 == P_C_SHV0 + 5 && P_FB_G0 == P_FB_E0 + 4 && P_FRZ1_G == P_FB_G0 + 4,
              "freeze-out fallback columns overlap");
static_assert(P_FRZ2_G == P_FRZ1_G + 1 && P_C_Q0 == P_FRZ2_G + 1 && NUM_COLUMNS == P_C_Q0 + 9,
              "private columns overlap");

// Uniform cell grid of one neighbour build (written on the device).
// Cell grid of one event.  A batch of events (ccake_batch_create) keeps one
// Grid per event; their cells share one key space of rows of `stride`
// cells, event e owning the key rows [row0, row0 + nc[1]), so the cell sort
// orders particles event by event and no pair walk crosses events.  For a
// single event stride = nc[0] and row0 = 0.  A 3+1D event (dim 3) has one
// key row per (y, eta) cell pair, row = c_eta * nc[1] + c_y, so the key is
// (c_eta * nc[1] + c_y) * nc[0] + c_x; a 2+1D grid has nc[2] = 1.
struct Grid {
  double lo[3];       // grid origin
  double cell;        // cell edge along y (and eta): one key row per y (eta) cell
  double inv_cell;
  double cellx;       // cell edge along x: cell / x_subcells, so a row's candidate
  double inv_cellx;   // runs are cut close to the cutoff disc without adding rows
  int nc[3];          // cells per axis
  int ncells;         // cells of the whole key space (cell_off holds ncells + 1 entries)
  int stride;         // cells per key row
  int row0;           // first key row of this event
  int nev;            // events sharing the key space
  int dim;            // 2 or 3
};

// Batches: at most kMaxBatchEvents events, each with at most
// kBatchAxisCells cells per axis, so the key space (at most 2^24 cells)
// sorts in three 8-bit radix passes.
constexpr int kMaxBatchEvents = 1024;
constexpr int kBatchAxisCells = 128;
constexpr int kBatchCellsPerEvent = kBatchAxisCells * kBatchAxisCells;

// Device error flags (sticky until ccake_event_check).
enum DevErr { DE_NEG_ENTROPY = 1, DE_EOS_FAIL = 2, DE_GRID_OVERFLOW = 4 };

// EoS chain on the device.
struct EosMember {
  int kind;                 // 0 table, 1 tanh, 2 conformal, 3 conformal_diagonal
  int pad;
  double c, T0, muB0, muS0, muQ0, Tc, Ts;
  double mins[4], maxs[4];
  int dims[4];
  double gmin[4], gdel[4];
  // table: smallest and largest node value of (s, B, S, Q, e), the bounds of
  // the multilinear interpolant (dmin > dmax for members without them)
  double dmin[5], dmax[5];
};
constexpr int kMaxMembers = 4;
struct EosChain {
  int n;
  int kind;
  EosMember m[kMaxMembers];
  const double* tab;        // node-major, 17 fields per node (device)
  const double* tab4;       // node-major (s, B, S, Q, e) packed per node for the solver
};

struct Profile {
  bool enabled = false;
  hipEvent_t ev0[CCAKE_NUM_KERNELS][2];
  double total_ms[CCAKE_NUM_KERNELS] = {};
  int64_t launches[CCAKE_NUM_KERNELS] = {};
};

}  // namespace ccake

struct ccake_dev {
  int gpu = 0;
  int simds = 1024;          // SIMDs of the device (4 per CU)
  void* scratch = nullptr;   // this device's address of the static pair-export scratch array
  std::mutex mu;             // guards the memory accounting and live counts below
  size_t budget = 0;
  size_t used = 0;
  int live_eos = 0;          // handles created on this device and not yet destroyed
  int live_events = 0;
  ccake::Profile prof;
};

struct ccake_eos {
  ccake_dev* dev = nullptr;
  ccake::EosChain chain{};
  double* d_tab = nullptr;
  double* d_tab4 = nullptr;
  size_t bytes = 0;
  int live_events = 0;       // events created on this EoS and not yet destroyed
  // host copy for host-side evaluation (efreeze / full_thermo)
  double* h_tab = nullptr;
  // pre-inverted flat (s, nB) table (CCAKE_EOS_PREINVERTED): device view, its
  // buffers (variables, then the s axis) and a host copy for efreeze_at_T
  ccake::Preinv pre{};
  double* d_pre = nullptr;
  std::vector<double> h_pre;
};

struct ccake_event {
  ccake_dev* dev = nullptr;
  ccake_eos* eos = nullptr;
  hipStream_t stream = nullptr;  // every launch and copy of the event (default priority)
  bool counted = false;      // included in the device's and EoS's live counts
  ccake_params p{};
  int n = 0;
  int npad = 0;
  double t = 0.0;
  int frzc = 0;
  // device buffers
  double* F = nullptr;          // NUM_COLUMNS x npad
  int32_t* I = nullptr;         // CCAKE_NUM_IFIELDS x npad
  // neighbour build
  int cap_cells = 0;
  int sort_passes = 2;          // 8-bit radix passes over the cell keys
  // a batch of events (ccake_batch_create): event e owns the particle ids
  // [off[e], off[e+1]) and grid[e]; a single event has nev = 1
  int nev = 1;
  std::vector<int> h_off;       // host copy of the offsets (nev + 1)
  int32_t* ev_off = nullptr;    // device offsets (nev + 1)
  int32_t* eid = nullptr;       // event of each particle id (batches only)
  double* red_ev = nullptr;     // per-event reduction results (nev)
  ccake::Grid* grid = nullptr;  // device
  double* bbox = nullptr;       // device partials
  int32_t* key = nullptr;       // cell key per particle (id order)
  int32_t* key_a = nullptr, *key_b = nullptr;   // radix ping-pong keys
  int32_t* val_a = nullptr, *val_b = nullptr;   // radix ping-pong ids
  int32_t* hist = nullptr;      // radix block histograms
  int32_t* scan_tmp = nullptr;  // device-wide scan partials
  int32_t* cell_off = nullptr;  // cap_cells + 1
  int32_t* perm = nullptr;      // sorted position -> particle id (alias of the final val buffer)
  int32_t* skey = nullptr;      // sorted keys (alias)
  int32_t* rank = nullptr;      // particle id -> sorted position
  int64_t* pair_count = nullptr;  // device accumulators: [0] last stage, [1] since the last reset
  int radix_blocks = 0;
  // reuse of the cell sort: bound on |x - x_sorted| of every particle
  bool sort_valid = false;
  bool no_charges = false;      // every particle's charges are zero (set at initialize, cleared by uploads)
  bool no_bulk = false;         // zeta = 0 and every particle's bulk pressure zero (likewise)
  double sort_disp = 0.0;
  double skin_frac = 0.04;      // re-sort when the bound would exceed skin_frac * rc (measured best at 40k)
  // freeze-out hypersurface (particlization on): 6 snapshot sets x 33 columns
  // x npad, the appended rows (n x CCAKE_FO_ROW) and their count
  double* snaps = nullptr;
  double* fo_rows = nullptr;
  int* fo_count = nullptr;
  double taup = 0.0, taupp = 0.0;   // times of the last two freeze-out snapshots
  int fo_step = 0;                  // steps since initialize
  int init_phase = 0;               // initialisation phases done (1, 2, 3)
  // pair-kernel partial sums: nslices x (up to 17) x npad, sorted order
  int nslices = 1;
  int nslices_cap = 1;          // slices the partial-sum buffer holds
  // EoS stage: particles whose inversion needs the full cascade (positions),
  // two alternating counters and which one the next stage appends to
  double* k4 = nullptr;         // RK4 accumulator (allocated at the first fourth-order step)
  int32_t* hard = nullptr;
  int32_t* hard_count = nullptr;
  int hard_parity = 0;
  double* part = nullptr;
  // gathered candidates (sorted order)
  double* cand_s = nullptr;     // 8 doubles per particle
  double* cand_g = nullptr;     // 16 doubles per particle
  double* cand_d = nullptr;     // diffusion pair records, 16 doubles per particle (diffusion on)
  // mixed precision (pair_precision 1): FP32 candidate records next to the
  // FP64 ones, kCandS32 and kCandG32 dwords per particle
  uint32_t* cand_s32 = nullptr;
  uint32_t* cand_g32 = nullptr;
  int snap_cols = 0;            // columns per freeze-out snapshot (33, or 45 with diffusion)
  // reductions
  double* red = nullptr;
  int32_t* err = nullptr;       // device error flags
  // Retired events of a batch: an event whose continue flag has turned
  // false (every particle frozen out, or tau past max_tau) has ended, as
  // BSQHydro::run ends for an event alone.  Its particles get the key past
  // every cell, so the cell sort places them after the live events'; the
  // time step's sorted-order kernels stop at n_live and the per-id kernels
  // that change state skip them, so an ended event keeps its final state.
  int32_t* live = nullptr;      // device, per event: 1 live, 0 retired (batches only)
  std::vector<int> h_live;
  int n_live = 0;               // particles of the live events (n for a single event)
  // Slices of the partial sums the sum passes read: nslices, or 1 when the
  // pair kernel added a tile's slices in LDS (set by each pair launch).
  int smooth_sum_slices = 1;
  int grad_sum_slices = 1;
  bool lds_reduce = true;       // 2+1D pair kernels reduce a tile's slices in LDS
  int32_t* eos_done = nullptr;  // per sorted position: the EoS stage's first attempt decided it (split stage)
  int64_t last_pairs = 0;
  size_t bytes = 0;
};

namespace ccake {

// error reporting
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
#define CCAKE_HIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return ccake::hip_fail(e_, #x); } while (0)

// profiling helpers (HIP events on the event's stream)
void prof_begin(ccake_event* ev, int k);
void prof_end(ccake_event* ev, int k);

inline double* col(ccake_event* ev, int f) { return ev->F + (size_t)f * ev->npad; }

// The live flags of a batch's events for the per-id kernels; empty for a
// single event.
struct Live {
  const int32_t* flag = nullptr;
  const int32_t* eid = nullptr;
  __device__ __forceinline__ bool retired(int i) const { return flag != nullptr && flag[eid[i]] == 0; }
};
inline Live live_of(const ccake_event* ev)
{
  Live l;
  if (ev->live && ev->n_live < ev->n) { l.flag = ev->live; l.eid = ev->eid; }
  return l;
}

// Search radius of the pair walks: rc with a rounding margin, widened by the
// bound on how far particles have moved since the cell sort (a candidate's
// sorted cell then lies within the radius of every query it can reach).
inline double search_radius(const ccake_event* ev)
{
  const double rc = 2.0 * ev->p.hT;
  double r = rc * (1.0 + 1e-12) + 1e-12;
  if (ev->sort_disp > 0.0) r += ev->sort_disp * (1.0 + 1e-12) + 1e-12;
  return r;
}
inline int32_t* icol(ccake_event* ev, int f) { return ev->I + (size_t)f * ev->npad; }

// Waves per workgroup of the 2+1D pair kernels that add a tile's slices in
// LDS (one wave per slice), or 0 for one single-wave workgroup per (tile,
// slice) with the slices' sums through HBM.
inline int lds_slices(const ccake_event* ev)
{
  if (ev->p.dim != 2 || !ev->lds_reduce) return 0;
  return ev->nslices >= 8 ? 8 : (ev->nslices >= 4 ? 4 : 0);
}

// ---- launchers (one per .hip unit) ----
// neighbours
int launch_stage_begin(ccake_event* ev);            // gamma/v, bbox
int launch_bbox(ccake_event* ev);                   // bbox only
int launch_neighbor_sort(ccake_event* ev);          // grid, keys, radix sort, cell offsets, smoothing records
int launch_export_pairs(ccake_event* ev, int32_t* pi, int32_t* pj, size_t cap, size_t* n);
int launch_export_pairs_noalloc(ccake_event* ev, int32_t* pi, int32_t* pj, size_t cap, size_t* n);
int scratch_address(void** out);   // the calling thread's current device's address of the scratch array
// smoothing
int launch_gather_smooth(ccake_event* ev);
int launch_smooth(ccake_event* ev, int mode, bool count_pairs = true);  // mode 0 fields, 1 initial sigma, 2 counts
int launch_smooth_pairs(ccake_event* ev, int mode, bool count_pairs);   // pair part (mode 2: whole count walk)
int launch_smooth_sum(ccake_event* ev, int mode, bool count_pairs);
int launch_stage_gather(ccake_event* ev);   // kept sort: gamma, v, smoothing records
int launch_charge_check(ccake_event* ev, bool* any_charge, bool* any_bulk);     // slice sums + fields (modes 0, 1)
// EoS + per-particle physics
// fuse_post: 0 EoS only, 1 with the stage's per-particle tail, 2 also with
// the smoothing sum pass folded in front (mode 0 only)
int launch_eos(ccake_event* ev, int mode, int fuse_post);
double smooth_norm2(double hT);   // fuse_post: reset_pi + transport + gradient gather
int launch_reset_pi(ccake_event* ev);
int launch_viscosities(ccake_event* ev);
int launch_gather_grad(ccake_event* ev);
int host_eos_full(const ccake_eos* eos, int member, const double* x, double* r17);
// gradients
int launch_grad(ccake_event* ev);
// BSQ diffusion (k_diff.hip): prep what = 1 q^mu + regulator, 2 kappa_q / tau_q,
// 4 pair records (perm: sorted order); the diffusion sums of the gradient pass
struct DiffParams {
  int mode = 0, relax = 0;
  double k2[9] = {};
  double C_B = 0.0;
  int reg = 0;
  double chi0 = 0, e0 = 0, xi0 = 1, rq_max = 0, fermi_offset = 0;
};
DiffParams make_diff_params(const ccake_params& p);
int launch_diff_prep(ccake_event* ev, int what, const int32_t* perm);
int launch_diff_grad(ccake_event* ev);
int launch_grad_pairs(ccake_event* ev);
int launch_grad_sum(ccake_event* ev);
// EoM + RK
// fold_grad_sum: the gradient pass's sum runs inside the EoM kernel (in a
// time step, after k_grad_slice), which then walks particles in cell order
// begin_next (with the fold and an RK update): also begin the next stage,
// which must keep the cell sort (gamma, v, smoothing records)
int launch_eom(ccake_event* ev, int fuse_rk, double dt, double t0, bool fold_grad_sum = false,
               bool begin_next = false);
double grad_sum_scale(double hT);
int launch_cache(ccake_event* ev);
int launch_rk4_update_k(ccake_event* ev, double w);   // k += w f (Evolver::update_k)
int launch_rk4_update(ccake_event* ev, double dt);    // y = y0 + dt k (Evolver::update_rk4)
int rk4_columns();
int launch_rk(ccake_event* ev, double dt, double t0);
// mixed-precision pair passes (pair_precision 1), in place of the FP64 ones
int launch_smooth_pairs32(ccake_event* ev);
int launch_grad_pairs32(ccake_event* ev);
// 3+1D Milne path (k_milne3.hip, dim 3): the stages whose data layout or
// geometry differ from 2+1D
int launch3_stage_begin(ccake_event* ev, int do_gamma);     // gamma/v (tau^2 metric), bbox in (x, y, eta)
int launch3_stage_gather(ccake_event* ev);                  // kept sort: gamma, v, smoothing records
int launch3_smooth_pairs(ccake_event* ev, int mode, bool count_pairs);
int launch3_export_pairs(ccake_event* ev, int32_t* pi, int32_t* pj, size_t cap, size_t* n);
int launch3_post(ccake_event* ev, int what);                // 1 reset_pi, 2 transport, 4 gradient records
int launch3_grad(ccake_event* ev);                          // gradient pairs + sums
int launch3_eom(ccake_event* ev, int fuse_rk, double dt);   // EoM (+ RK update, + cache first)
int launch3_rk(ccake_event* ev, double dt);
int launch3_cache(ccake_event* ev);
double smooth_norm2_event(const ccake_event* ev);           // W = norm2 w(q) of the event's kernel
// misc
int launch_gamma_velocity(ccake_event* ev);
int launch_init_particles(ccake_event* ev, int phase);
int launch_freeze(ccake_event* ev, int phase);
int launch_reduce(ccake_event* ev, int which, double* out);
int launch_reduce_events(ccake_event* ev, int which, double* out);   // one result per event of a batch

// Smoothing candidate record of particle i (8 doubles): x, y, m_s,
// m_s*ext_s, m_B*ext_B, m_S*ext_S, m_Q*ext_Q, 0.
constexpr int kCandS = 8;
__device__ __forceinline__ void smooth_record(const double* __restrict__ F, int npad, int i, double* __restrict__ o)
{
  auto f = [&](int c) { return F[(size_t)c * npad + i]; };
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

// 3+1D candidate records.  Smoothing (8 doubles): x, y, eta, m_s,
// m_s*ext_s, m_B*ext_B, m_S*ext_S, m_Q*ext_Q.  Gradients (24): x, y, eta,
// m_s, p/sl^2, e, Pi/sl^2, v^x, v^y, v^eta, pi^{0i}/sl^2 (3), the spatial
// pi^{ij}/sl^2 (xx, xy, x eta, yy, y eta, eta eta), sl, 4 unused.
constexpr int kCandG3 = 24;
__device__ __forceinline__ void smooth_record3(const double* __restrict__ F, int npad, int i, double* __restrict__ o)
{
  auto f = [&](int c) { return F[(size_t)c * npad + i]; };
  const double ms = f(CCAKE_F_MASS_S);
  o[0] = f(CCAKE_F_X);
  o[1] = f(CCAKE_F_Y);
  o[2] = f(CCAKE_F_ETA);
  o[3] = ms;
  o[4] = ms * f(CCAKE_F_EXT_S);
  o[5] = f(CCAKE_F_MASS_B) * f(CCAKE_F_EXT_B);
  o[6] = f(CCAKE_F_MASS_SS) * f(CCAKE_F_EXT_SS);
  o[7] = f(CCAKE_F_MASS_Q) * f(CCAKE_F_EXT_Q);
}

// FP32 candidate records of the mixed-precision pair kernels: the position
// stays FP64 (the pair predicate is evaluated in FP64 on it, so the pair set
// is the FP64 path's), every other entry is rounded to FP32.  Dwords:
//   smoothing (12): x, y (2 each), m_s, m_s*ext_s, m_B*ext_B, m_S*ext_S,
//                   m_Q*ext_Q, 3 unused;
//   gradients (16): x, y (2 each), m_s, p/sl^2, e, Pi/sl^2, v^x, v^y,
//                   pi^{0x}/sl^2, pi^{0y}/sl^2, pi^{xx}/sl^2, pi^{xy}/sl^2,
//                   pi^{yy}/sl^2, 1 unused.
constexpr int kCandS32 = 12;
constexpr int kCandG32 = 16;
__device__ __forceinline__ void put_f64(uint32_t* o, double v)
{
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  o[0] = (uint32_t)b;
  o[1] = (uint32_t)(b >> 32);
}
__device__ __forceinline__ void put_f32(uint32_t* o, double v) { o[0] = __float_as_uint((float)v); }
__device__ __forceinline__ void smooth_record32(const double* __restrict__ o64, uint32_t* __restrict__ o)
{
  put_f64(o, o64[0]);
  put_f64(o + 2, o64[1]);
  for (int q = 2; q < 7; ++q) put_f32(o + 2 + q, o64[q]);
  o[9] = o[10] = o[11] = 0u;
}
__device__ __forceinline__ void grad_record32(const double* __restrict__ o64, uint32_t* __restrict__ o)
{
  put_f64(o, o64[0]);
  put_f64(o + 2, o64[1]);
  for (int q = 2; q < 13; ++q) put_f32(o + 2 + q, o64[q]);
  o[15] = 0u;
}

}  // namespace ccake
