// C-ABI of the MI355X SPH library: lifetime, copies, phase entry points,
// the RK2 step and the initialisation sequence.  See include/ccake_amd.h.
#include "ccake_internal.h"
#include "eos_device.h"
#include "pair_common.h"
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <mutex>
#include <vector>

namespace ccake {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int hip_fail(hipError_t e, const char* what)
{
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return CCAKE_ERR_HIP;
}

// ---- profiling: HIP events on the library stream --------------------------
struct PendingProf { int k; hipEvent_t a, b; };
static std::mutex g_prof_mu;
static std::vector<PendingProf> g_pending;
static std::vector<hipEvent_t> g_free_events;
static thread_local hipEvent_t t_open[CCAKE_NUM_KERNELS];

static hipEvent_t get_event()
{
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (!g_free_events.empty()) { hipEvent_t e = g_free_events.back(); g_free_events.pop_back(); return e; }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
void prof_begin(ccake_event* ev, int k)
{
  if (!ev->dev->prof.enabled) return;
  hipEvent_t e = get_event();
  if (e && hipEventRecord(e, ev->stream) == hipSuccess) t_open[k] = e;
}
void prof_end(ccake_event* ev, int k)
{
  if (!ev->dev->prof.enabled || !t_open[k]) return;
  hipEvent_t e = get_event();
  if (e && hipEventRecord(e, ev->stream) == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_pending.push_back({ k, t_open[k], e });
  }
  t_open[k] = nullptr;
}
static void prof_drain(ccake_dev* d)
{
  if (hipSetDevice(d->gpu) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (auto& p : g_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      d->prof.total_ms[p.k] += ms;
      d->prof.launches[p.k] += 1;
    }
    g_free_events.push_back(p.a);
    g_free_events.push_back(p.b);
  }
  g_pending.clear();
}

static int check_launch(int rc) { return rc; }

// ---- reuse of the cell sort ------------------------------------------------
// Each RK update moves a particle by less than c * dt (|v| = |u| / gamma < 1).
// A sort taken at positions x_s therefore stays usable for the pair walks
// while a host-side bound B >= |x - x_s| is added to their search radius;
// the exact predicate still decides every pair, so the pair set is the one a
// fresh sort gives.  B is 0 right after a sort, grows by dt/2 into stage 2
// and by dt per completed step; a stage sorts again when B would exceed
// skin_frac * rc (default 0.04; CCAKE_SORT_SKIN overrides it for tuning).

static int fresh_sort(ccake_event* ev)
{
  int rc;
  if ((rc = launch_bbox(ev))) return rc;
  if ((rc = launch_neighbor_sort(ev))) return rc;
  ev->sort_valid = true;
  ev->sort_disp = 0.0;
  return 0;
}

// Walking entry points of the phase-by-phase C-ABI sort first when the
// positions changed by an unknown amount since the last sort.
static int ensure_sorted(ccake_event* ev)
{
  return ev->sort_valid ? 0 : fresh_sort(ev);
}

// ---- one RK stage: SPHWorkstation::get_time_derivatives ------------------
// may_reuse: the stage may keep the last cell sort (ev->sort_disp bounds the
// displacement since); *sorted reports whether it sorted.
// Which sum passes run inside the next kernel (bit 1: smoothing sum in the
// EoS kernel, bit 2: gradient sum in the EoM kernel).  The defaults follow
// measurement: the smoothing fold pays off with every EoS chain; the
// gradient fold pays off with the light ideal-fluid EoM kernel and costs
// with the register-bound shear one (32k lumpy event: 0.837 ms per step
// with bit 1 only, 0.860 with bit 2 only).  CCAKE_FOLD_SUMS overrides them
// for tuning.
static int fold_sums(const ccake_event* ev)
{
  if (ev->p.dim == 3) return 0;                 // the 3+1D stage runs its passes as kernels of their own
  static const int env = [] { const char* e = getenv("CCAKE_FOLD_SUMS"); return e ? atoi(e) : -1; }();
  if (env >= 0) return env;
  // the gradient sums fold into the EoM kernel unless the shear walk's sums
  // arrive in many slices (their loads would crowd the EoM kernel)
  return 1 | ((!ev->p.using_shear || lds_slices(ev)) ? 2 : 0);
}

// Within one stepping call, a stage that keeps the cell sort can be begun by
// the previous stage's EoM kernel (`begun`).  next_disp: the sort bound B at
// the next stage's start if this stage keeps the sort ([0]) or sorts ([1]);
// negative when no stage follows in this call.  *begun_next reports whether
// this stage's EoM kernel began the next one.
struct NextStage {
  bool begun = false;
  double next_disp[2] = { -1.0, -1.0 };
  bool begun_next = false;
};

static int time_derivatives(ccake_event* ev, int fuse_rk, double dt_stage, double t0, bool may_reuse,
                            bool* sorted = nullptr, NextStage* ns = nullptr)
{
  int rc;
  const int fold = fold_sums(ev);
  const double bound = ev->skin_frac * 2.0 * ev->p.hT;
  const bool reuse = may_reuse && ev->sort_valid && ev->sort_disp <= bound;
  if (reuse && ns && ns->begun) {
    // gamma, v and the smoothing records were written by the last EoM kernel
  } else if (reuse) {
    prof_begin(ev, CCAKE_K_SORT);
    if ((rc = launch_stage_gather(ev))) return rc;              // 02 gamma, v; candidate records at current positions
    prof_end(ev, CCAKE_K_SORT);
  } else {
    prof_begin(ev, CCAKE_K_BIN);
    if ((rc = launch_stage_begin(ev))) return rc;               // 02 gamma, v (+ bbox)
    prof_end(ev, CCAKE_K_BIN);
    prof_begin(ev, CCAKE_K_SORT);
    if ((rc = launch_neighbor_sort(ev))) return rc;             // 01 cell sort + candidate records
    prof_end(ev, CCAKE_K_SORT);
    ev->sort_valid = true;
    ev->sort_disp = 0.0;
  }
  if (sorted) *sorted = !reuse;
  prof_begin(ev, CCAKE_K_SMOOTH);
  if ((rc = launch_smooth_pairs(ev, 0, true))) return rc;       // 03 pair sums per slice
  prof_end(ev, CCAKE_K_SMOOTH);
  if (!(fold & 1)) {
    prof_begin(ev, CCAKE_K_SMOOTH_SUM);
    if ((rc = launch_smooth_sum(ev, 0, true))) return rc;       // 03 fields (+ btrack)
    prof_end(ev, CCAKE_K_SMOOTH_SUM);
  }
  prof_begin(ev, CCAKE_K_EOS);
  if (ev->p.dim == 3) {
    if ((rc = launch_eos(ev, 0, 0))) return rc;                 // 04
    if ((rc = launch3_post(ev, 1 | 2 | 4))) return rc;          // 05 + 08 + gradient records
    prof_end(ev, CCAKE_K_EOS);
    prof_begin(ev, CCAKE_K_GRAD);
    if ((rc = launch_grad(ev))) return rc;                      // 09
    prof_end(ev, CCAKE_K_GRAD);
    prof_begin(ev, CCAKE_K_EOM);
    if ((rc = launch_eom(ev, fuse_rk, dt_stage, t0))) return rc;  // 10 (+ RK)
    prof_end(ev, CCAKE_K_EOM);
    if (ns) ns->begun_next = false;
    return 0;
  }
  if ((rc = launch_eos(ev, 0, (fold & 1) ? 2 : 1))) return rc;  // (03 fields,) 04 + 05 + 08 + gather
  if (ev->p.using_diffusion && (rc = launch_diff_prep(ev, 15, ev->perm))) return rc;  // q, 06, kappa/tau_q
  prof_end(ev, CCAKE_K_EOS);
  prof_begin(ev, CCAKE_K_GRAD);
  if ((rc = launch_grad_pairs(ev))) return rc;                  // 09 pair sums per slice
  prof_end(ev, CCAKE_K_GRAD);
  // the diffusion pair pass reuses the partial-sum buffer, so with diffusion
  // the gradient sums run first; otherwise they run inside the EoM kernel
  const bool diff = ev->p.using_diffusion != 0;
  const bool fold_grad = !diff && (fold & 2);
  if (!fold_grad) {
    prof_begin(ev, CCAKE_K_GRAD_SUM);
    if ((rc = launch_grad_sum(ev))) return rc;                  // 09 gradients
    if (diff && (rc = launch_diff_grad(ev))) return rc;         // 09 diffusion sums
    prof_end(ev, CCAKE_K_GRAD_SUM);
  }
  bool begin_next = false;
  if (ns) {
    const double nd = ns->next_disp[reuse ? 0 : 1];
    begin_next = fold_grad && fuse_rk && may_reuse && nd >= 0.0 && nd <= bound;
    ns->begun_next = begin_next;
  }
  prof_begin(ev, CCAKE_K_EOM);
  if ((rc = launch_eom(ev, fuse_rk, dt_stage, t0, fold_grad, begin_next))) return rc;  // (09,) 10 (+ RK, + next stage)
  prof_end(ev, CCAKE_K_EOM);
  return 0;
}

// SPHWorkstation::freeze_out_particles (src/sph_workstation.cpp:1475-1485):
// status check and bsqsvfreezeout with its snapshot times
// (include/freeze_out.h:227-281, 549-636): taupp is set at the first step,
// taup at the second, and from the third on they move by one step after the
// interpolation.
static int freeze_step(ccake_event* ev)
{
  const int phase = ev->frzc;
  ev->fo_step++;
  if (phase == 0) ev->taupp = ev->t;
  else if (phase == 1) ev->taup = ev->t;
  const int rc = launch_freeze(ev, phase);
  if (rc) return rc;
  if (phase >= 2) { ev->taupp = ev->taup; ev->taup = ev->t; }
  if (ev->frzc < 2) ev->frzc++;
  return 0;
}

static int rk2_step(ccake_event* ev, double dt, bool more = false, bool* begun = nullptr)
{
  int rc;
  // the RK cache: taken inside the first stage's EoM kernel, or up front when
  // the diffusion regulator may rewrite the evolved currents within the stage
  const bool sep_cache = ev->p.using_diffusion != 0;
  if (sep_cache) {
    prof_begin(ev, CCAKE_K_RK);
    if ((rc = launch_cache(ev))) return rc;
    prof_end(ev, CCAKE_K_RK);
  }
  const double t0 = ev->t;
  const double adt = fabs(dt);
  bool sorted = false;
  NextStage s1;
  s1.begun = begun && *begun;
  s1.next_disp[0] = ev->sort_disp + 0.5 * adt;     // the bound update after stage 1, below
  s1.next_disp[1] = 0.5 * adt;
  if ((rc = time_derivatives(ev, sep_cache ? 1 : 2, 0.5 * dt, t0, true, &sorted, &s1))) return rc;
  const double d0 = ev->sort_disp;                 // bound at the step's start positions x0
  ev->sort_disp = d0 + 0.5 * adt;                  // x0 + dt/2 v
  ev->t = t0 + 0.5 * dt;
  NextStage s2;
  s2.begun = s1.begun_next;
  if (more) {                                      // the bound update at the end of the step, below
    s2.next_disp[0] = d0 + adt;
    s2.next_disp[1] = 1.5 * adt;
  }
  if ((rc = time_derivatives(ev, 1, dt, t0, true, &sorted, &s2))) return rc;
  // end of step: x0 + dt v_mid, within dt of x0 and 1.5 dt of x0 + dt/2 v
  ev->sort_disp = sorted ? 1.5 * adt : d0 + adt;
  ev->t = t0 + dt;
  if (begun) *begun = s2.begun_next;
  if (ev->p.particlization) {
    prof_begin(ev, CCAKE_K_MISC);
    if ((rc = freeze_step(ev))) return rc;
    prof_end(ev, CCAKE_K_MISC);
  }
  return 0;
}

// Bytes added past every event buffer (the host emulation build sets 0 so
// that any access past a buffer's end is caught).
#ifndef CCAKE_ALLOC_PAD
#define CCAKE_ALLOC_PAD 256
#endif

template <class T>
static int dalloc(ccake_event* ev, T** p, size_t count);

// Evolver::allocate_k_values (src/evolver.cpp:519-526): the RK4
// accumulator, allocated at an event's first fourth-order step and cleared
// at every one (with a fill kernel: the step may be captured).
__global__ void k_fill_f64(double* __restrict__ a, size_t n, double v)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
static int rk4_allocate_k(ccake_event* ev)
{
  if (ev->p.dim == 3) { set_error("the fourth-order step is built for D = 2"); return CCAKE_ERR_UNSUPPORTED; }
  const size_t cnt = (size_t)rk4_columns() * ev->npad;
  if (!ev->k4) {
    const int rc = dalloc(ev, &ev->k4, cnt);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, ev->stream, ev->k4, cnt, 0.0);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

// Evolver::advance_timestep_rk4 (src/evolver.cpp:519-561) as the reference
// runs it (not the classical method): the first stage's derivatives are not
// kept; k = k2/6 + 2 k3/6 + 2 k4/6 + 2 k5/6 with k5 the derivatives at
// t0 + dt after the fourth stage; y = y0 + dt k.  Every stage sorts afresh
// (the stages' points lie up to dt apart), and none begins the next.
static int rk4_step(ccake_event* ev, double dt)
{
  int rc;
  if ((rc = launch_cache(ev))) return rc;                   // set_current_timestep_quantities
  if ((rc = rk4_allocate_k(ev))) return rc;                 // allocate_k_values
  const double t0 = ev->t;
  if ((rc = time_derivatives(ev, 1, 0.5 * dt, t0, false))) return rc;   // step 1
  ev->t = t0 + 0.5 * dt;
  if ((rc = time_derivatives(ev, 1, 0.5 * dt, t0, false))) return rc;   // step 2
  if ((rc = launch_rk4_update_k(ev, 1.0 / 6.0))) return rc;
  if ((rc = time_derivatives(ev, 1, 0.5 * dt, t0, false))) return rc;   // step 3
  if ((rc = launch_rk4_update_k(ev, 2.0 / 6.0))) return rc;
  if ((rc = time_derivatives(ev, 1, dt, t0, false))) return rc;         // step 4
  ev->t = t0 + dt;
  if ((rc = launch_rk4_update_k(ev, 2.0 / 6.0))) return rc;
  if ((rc = time_derivatives(ev, 0, 0.0, ev->t, false))) return rc;    // the last functional call
  if ((rc = launch_rk4_update_k(ev, 2.0 / 6.0))) return rc;
  if ((rc = launch_rk4_update(ev, dt))) return rc;                      // update_rk4
  ev->sort_valid = false;
  if (ev->p.particlization) {
    prof_begin(ev, CCAKE_K_MISC);
    if ((rc = freeze_step(ev))) return rc;
    prof_end(ev, CCAKE_K_MISC);
  }
  return 0;
}

template <class T>
static int dalloc(ccake_event* ev, T** p, size_t count)
{
  const size_t bytes = count * sizeof(T) + CCAKE_ALLOC_PAD;
  ccake_dev* d = ev->dev;
  {
    std::lock_guard<std::mutex> lk(d->mu);
    if (d->budget && d->used + bytes > d->budget) { set_error("HBM budget exceeded"); return CCAKE_ERR_OOM; }
    d->used += bytes;                 // reserved before the allocation, so concurrent creators see it
  }
  if (hipMalloc((void**)p, bytes) != hipSuccess) {
    std::lock_guard<std::mutex> lk(d->mu);
    d->used -= bytes;
    set_error("hipMalloc failed");
    return CCAKE_ERR_OOM;
  }
  ev->bytes += bytes;
  return 0;
}

}  // namespace ccake

using namespace ccake;

extern "C" {

const char* ccake_last_error(void) { return g_err.c_str(); }

int ccake_dev_create(int gpu, size_t budget, ccake_dev** out)
{
  if (!out) return CCAKE_ERR_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_error("no HIP device"); return CCAKE_ERR_HIP; }
  if (gpu < 0 || gpu >= ndev) { set_error("bad device index"); return CCAKE_ERR_INVALID; }
  CCAKE_HIP(hipSetDevice(gpu));
  ccake_dev* d = new ccake_dev();
  d->gpu = gpu;
  d->budget = budget;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, gpu) == hipSuccess && prop.multiProcessorCount > 0)
    d->simds = 4 * prop.multiProcessorCount;
  // this device's address of the static scratch array of the
  // allocation-free pair export (the device is current here)
  if (const int rc = scratch_address(&d->scratch)) { delete d; return rc; }
  *out = d;
  return 0;
}

int ccake_dev_destroy(ccake_dev* d)
{
  if (!d) return CCAKE_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(d->mu);
    if (d->live_eos || d->live_events) {
      set_error("ccake_dev_destroy: destroy the device's events and EoS objects first");
      return CCAKE_ERR_INVALID;
    }
  }
  prof_drain(d);
  delete d;
  return 0;
}

int ccake_dev_sync(ccake_dev* d)
{
  if (!d) return CCAKE_ERR_INVALID;
  CCAKE_HIP(hipSetDevice(d->gpu));
  CCAKE_HIP(hipDeviceSynchronize());
  return 0;
}

int ccake_dev_memory(ccake_dev* d, size_t* used, size_t* budget)
{
  if (!d) return CCAKE_ERR_INVALID;
  std::lock_guard<std::mutex> lk(d->mu);
  if (used) *used = d->used;
  if (budget) *budget = d->budget;
  return 0;
}

// ---------------------------------------------------------------- EoS
static void init_analytic(EosMember& m, int kind, double c, double T0, double muB0, double muQ0, double muS0,
                          double Tc, double Ts)
{
  std::memset(&m, 0, sizeof(m));
  m.kind = kind; m.c = c; m.T0 = T0; m.muB0 = muB0; m.muQ0 = muQ0; m.muS0 = muS0; m.Tc = Tc; m.Ts = Ts;
  m.mins[0] = 0.0; m.maxs[0] = 1e10;
  for (int a = 1; a < 4; ++a) { m.mins[a] = -1e10; m.maxs[a] = 1e10; }
}

// EoS_Interpolator::load_flat_table (src/eos_interpolator.cpp:291-527): the
// flat (s, nB) table in MeV units -> fm units; log10-uniform axes (checked to
// 1e-6); dw/ds and dw/dB from the baryon-only Schur complement of the
// susceptibilities (dw/dS = dw/dQ = 0).  Fills e->pre on the device and the
// host copy e->h_pre (variables, then the s axis).
static int make_preinverted(ccake_dev* dev, const ccake_eos_desc* desc, ccake_eos* e)
{
  const int Ns = desc->dims[0], NB = desc->dims[1];
  if (Ns < 2 || NB < 2 || !desc->table) { set_error("pre-inverted table needs dims[0], dims[1] >= 2 and a table"); return CCAKE_ERR_INVALID; }
  const size_t nt = (size_t)Ns * NB;
  const double* sB = desc->table;                 // s, nB, T, muB, muS, muQ, e, p, cs2, chiTT, chiTB, chiBB
  auto in = [&](int k, size_t i) { return sB[(size_t)k * nt + i]; };
  const double h1 = 1.0 / kHbarcMeVfm, h2 = h1 * h1, h3 = h2 * h1, h4 = h2 * h2;
  std::vector<double> sax(Ns), nax(NB);
  for (int is = 0; is < Ns; ++is) sax[is] = in(0, (size_t)is * NB) * h3;
  for (int iB = 0; iB < NB; ++iB) nax[iB] = in(1, (size_t)iB) * h3;
  auto axis = [](const std::vector<double>& a, double* l0, double* step) {
    for (double x : a) if (!(x > 0.0)) return false;
    *l0 = std::log10(a.front());
    *step = (std::log10(a.back()) - *l0) / (a.size() - 1);
    double dev_max = 0.0;
    for (size_t i = 1; i + 1 < a.size(); ++i)
      dev_max = std::max(dev_max, std::fabs(std::log10(a[i + 1]) - std::log10(a[i]) - *step));
    return dev_max <= 1e-6;
  };
  Preinv& p = e->pre;
  p.Ns = Ns; p.NB = NB;
  if (!axis(sax, &p.ls0, &p.dls) || !axis(nax, &p.lb0, &p.dlb)) {
    set_error("pre-inverted table: s and nB axes must be positive and uniform in log10");
    return CCAKE_ERR_INVALID;
  }
  p.s_first = sax.front(); p.s_last = sax.back(); p.nB_first = nax.front();
  e->h_pre.assign((size_t)PI_NUM * nt + Ns, 0.0);
  double* v = e->h_pre.data();
  const double EPS = 1e-30, TINY = 1e-25;
  for (int is = 0; is < Ns; ++is)
    for (int iB = 0; iB < NB; ++iB) {
      const size_t i = (size_t)is * NB + iB;
      const double T = in(2, i) * h1, muB = in(3, i) * h1;
      const double cTT = in(9, i) * h2, cTB = in(10, i) * h2, cBB = in(11, i) * h2;
      const double s_c = sax[is], nB_c = nax[iB];
      const double dentr_dt = cTT - (cTB * cTB) / (cBB + EPS);
      const double dentr_dmub = cTB - (cTT * cBB) / (cTB + EPS);
      const double db_dt = cTB - (cBB * cTT) / (cTB + EPS);
      const double db_dmub = cBB - (cTB * cTB) / (cTT + EPS);
      double dwds = T + s_c / dentr_dt;
      if (std::fabs(nB_c) > TINY) dwds += nB_c / dentr_dmub;
      double dwdB = muB;
      if (std::fabs(nB_c) > TINY) dwdB += s_c / db_dt + nB_c / db_dmub;
      v[PI_T * nt + i] = T; v[PI_MUB * nt + i] = muB;
      v[PI_MUS * nt + i] = in(4, i) * h1; v[PI_MUQ * nt + i] = in(5, i) * h1;
      v[PI_E * nt + i] = in(6, i) * h4; v[PI_P * nt + i] = in(7, i) * h4;
      v[PI_CS2 * nt + i] = in(8, i);
      v[PI_DWDS * nt + i] = dwds; v[PI_DWDB * nt + i] = dwdB;
      v[PI_DWDQ * nt + i] = 0.0; v[PI_DWDS_S * nt + i] = 0.0;
    }
  std::copy(sax.begin(), sax.end(), v + (size_t)PI_NUM * nt);
  const size_t bytes = e->h_pre.size() * sizeof(double);
  {
    std::lock_guard<std::mutex> lk(dev->mu);
    if (dev->budget && dev->used + bytes > dev->budget) { set_error("HBM budget exceeded"); return CCAKE_ERR_OOM; }
    dev->used += bytes;
  }
  e->bytes = bytes;
  if (hipMalloc((void**)&e->d_pre, bytes) != hipSuccess) {
    std::lock_guard<std::mutex> lk(dev->mu);
    dev->used -= bytes;
    e->bytes = 0;
    set_error("hipMalloc failed");
    return CCAKE_ERR_OOM;
  }
  CCAKE_HIP(hipMemcpy(e->d_pre, v, bytes, hipMemcpyHostToDevice));
  p.v = e->d_pre;
  p.s_axis = e->d_pre + (size_t)PI_NUM * nt;
  return 0;
}

int ccake_eos_create(ccake_dev* dev, const ccake_eos_desc* desc, ccake_eos** out)
{
  if (!dev || !desc || !out) return CCAKE_ERR_INVALID;
  CCAKE_HIP(hipSetDevice(dev->gpu));
  ccake_eos* e = new ccake_eos();
  e->dev = dev;
  EosChain& ch = e->chain;
  ch.kind = desc->kind;
  const double Nc = 3.0, Nf = 2.5;
  const double c0 = kPi * kPi * (2.0 * (Nc * Nc - 1.0) + (7.0 / 2.0) * Nc * Nf) / 90.0;
  if (desc->kind == CCAKE_EOS_CONFORMAL) {
    ch.n = 2;
    init_analytic(ch.m[0], M_CONF, c0, 1, 1, 1, 1, 0, 0);
    init_analytic(ch.m[1], M_DIAG, c0, 1, 1, 1, 1, 0, 0);
    std::lock_guard<std::mutex> lk(dev->mu);
    dev->live_eos++;
    *out = e;
    return 0;
  }
  if (desc->kind == CCAKE_EOS_PREINVERTED) {
    const int rc = make_preinverted(dev, desc, e);
    if (rc) { if (e->d_pre) (void)hipFree(e->d_pre); delete e; return rc; }
    std::lock_guard<std::mutex> lk(dev->mu);
    dev->live_eos++;
    *out = e;
    return 0;
  }
  if (desc->kind != CCAKE_EOS_TABLE || !desc->table) { delete e; set_error("bad EoS description"); return CCAKE_ERR_INVALID; }
  ch.n = 4;
  EosMember& t = ch.m[0];
  std::memset(&t, 0, sizeof(t));
  t.kind = M_TABLE;
  size_t nn = 1;
  for (int a = 0; a < 4; ++a) {
    t.dims[a] = desc->dims[a];
    nn *= (size_t)desc->dims[a];
    t.gmin[a] = desc->mins[a];
    t.gdel[a] = desc->dims[a] <= 1 ? 0.0 : (desc->maxs[a] - desc->mins[a]) / (desc->dims[a] - 1.0);
    t.mins[a] = desc->mins[a]; t.maxs[a] = desc->maxs[a];
  }
  e->h_tab = new double[nn * 17];
  std::memcpy(e->h_tab, desc->table, nn * 17 * sizeof(double));
  std::vector<double> t5(nn * 5);
  for (int q = 0; q < 5; ++q) { t.dmin[q] = INFINITY; t.dmax[q] = -INFINITY; }
  bool finite = true;
  for (size_t k = 0; k < nn; ++k) {
    const double* c = desc->table + k * 17;
    t5[k * 5 + 0] = c[X_S]; t5[k * 5 + 1] = c[X_B]; t5[k * 5 + 2] = c[X_SS]; t5[k * 5 + 3] = c[X_Q];
    t5[k * 5 + 4] = c[X_E];
    for (int q = 0; q < 5; ++q) {
      const double v = t5[k * 5 + q];
      if (!std::isfinite(v)) finite = false;
      t.dmin[q] = std::min(t.dmin[q], v);
      t.dmax[q] = std::max(t.dmax[q], v);
    }
  }
  if (!finite) for (int q = 0; q < 5; ++q) { t.dmin[q] = 1.0; t.dmax[q] = 0.0; }   // no bounds: never skipped
  e->bytes = nn * 22 * sizeof(double);
  {
    std::lock_guard<std::mutex> lk(dev->mu);
    if (dev->budget && dev->used + e->bytes > dev->budget) {
      delete[] e->h_tab; delete e; set_error("HBM budget exceeded"); return CCAKE_ERR_OOM;
    }
    dev->used += e->bytes;
  }
  {
    // on any failure: free what was allocated and give the budget back
    auto fail = [&](hipError_t err, const char* what) {
      if (e->d_tab) (void)hipFree(e->d_tab);
      if (e->d_tab4) (void)hipFree(e->d_tab4);
      {
        std::lock_guard<std::mutex> lk(dev->mu);
        dev->used -= e->bytes;
      }
      delete[] e->h_tab;
      delete e;
      return hip_fail(err, what);
    };
    hipError_t err;
    if ((err = hipMalloc((void**)&e->d_tab, nn * 17 * sizeof(double))) != hipSuccess) return fail(err, "hipMalloc(table)");
    if ((err = hipMalloc((void**)&e->d_tab4, nn * 5 * sizeof(double))) != hipSuccess) return fail(err, "hipMalloc(table)");
    if ((err = hipMemcpy(e->d_tab, desc->table, nn * 17 * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
      return fail(err, "hipMemcpy(table)");
    if ((err = hipMemcpy(e->d_tab4, t5.data(), nn * 5 * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
      return fail(err, "hipMemcpy(table)");
  }
  ch.tab = e->d_tab;
  ch.tab4 = e->d_tab4;
  // matched fallbacks at Tmax = 1.1 T_FO (src/eos_initialization.cpp:270-488)
  const double Tmax = 1.1 * desc->T_freeze_out;
  double r[17], pt[4] = { Tmax, 0, 0, 0 };
  member_fields(t, e->h_tab, pt, r);
  const double pT = r[X_P];
  const double c = pT / (Tmax * Tmax * Tmax * Tmax);
  double mu0[4] = { 1, 1, 1, 1 }, mud[4] = { 1, 1, 1, 1 };
  for (int a = 1; a < 4; ++a) {
    const double mmax = t.maxs[a];
    if (std::fabs(mmax) < 1e-30) continue;
    double q[4] = { Tmax, 0, 0, 0 };
    q[a] = mmax;
    member_fields(t, e->h_tab, q, r);
    mu0[a] = pow(c, 0.25) * mmax / sqrt(sqrt(r[X_P]) - sqrt(pT));
    mud[a] = pow(c / (r[X_P] - pT), 0.25) * mmax;
  }
  const double Tc = 220.0 / kHbarcMeVfm, Ts = 120.0 / kHbarcMeVfm;
  init_analytic(ch.m[1], M_TANH, c, 1.0, mu0[1], mu0[2], mu0[3], Tc, Ts);
  init_analytic(ch.m[2], M_CONF, c, 1.0, mu0[1], mu0[2], mu0[3], 0, 0);
  init_analytic(ch.m[3], M_DIAG, c, 1.0, mud[1], mud[2], mud[3], 0, 0);
  std::lock_guard<std::mutex> lk(dev->mu);
  dev->live_eos++;
  *out = e;
  return 0;
}

int ccake_eos_destroy(ccake_eos* e)
{
  if (!e) return CCAKE_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(e->dev->mu);
    if (e->live_events) {
      set_error("ccake_eos_destroy: destroy the events that use this EoS first");
      return CCAKE_ERR_INVALID;
    }
    e->dev->live_eos--;
    e->dev->used -= (e->dev->used >= e->bytes ? e->bytes : e->dev->used);
  }
  if (e->d_tab) (void)hipFree(e->d_tab);
  if (e->d_tab4) (void)hipFree(e->d_tab4);
  if (e->d_pre) (void)hipFree(e->d_pre);
  delete[] e->h_tab;
  delete e;
  return 0;
}

int ccake_eos_efreeze(ccake_eos* e, double T_fo, double* e_out)
{
  if (!e || !e_out) return CCAKE_ERR_INVALID;
  if (e->chain.kind == CCAKE_EOS_PREINVERTED) {
    // EoS_Interpolator::efreeze_at_T (src/eos_interpolator.cpp:865-892): T along
    // the smallest-nB column, bisection, linear in the bracketing cell
    const Preinv& p = e->pre;
    const size_t nt = (size_t)p.Ns * p.NB;
    const double* T = e->h_pre.data() + (size_t)PI_T * nt;
    const double* E = e->h_pre.data() + (size_t)PI_E * nt;
    auto at = [&](const double* a, int is) { return a[(size_t)is * p.NB]; };
    if (T_fo <= at(T, 0)) { *e_out = at(E, 0); return 0; }
    if (T_fo >= at(T, p.Ns - 1)) { *e_out = at(E, p.Ns - 1); return 0; }
    int lo = 0, hi = p.Ns - 1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (at(T, mid) <= T_fo) lo = mid; else hi = mid;
    }
    const double Tl = at(T, lo), Th = at(T, hi);
    const double alpha = (Th > Tl) ? (T_fo - Tl) / (Th - Tl) : 0.0;
    *e_out = (1.0 - alpha) * at(E, lo) + alpha * at(E, hi);
    return 0;
  }
  double r[17], x[4] = { T_fo, 0, 0, 0 };
  member_fields(e->chain.m[0], e->h_tab, x, r);
  *e_out = r[X_E];
  return 0;
}

int ccake_eos_full_thermo(ccake_eos* e, int member, const double tbqs[4], double out17[17])
{
  if (!e || !tbqs || !out17) return CCAKE_ERR_INVALID;
  if (e->chain.kind == CCAKE_EOS_PREINVERTED) {
    set_error("a pre-inverted EoS has no (T, mu) evaluation");
    return CCAKE_ERR_UNSUPPORTED;
  }
  return host_eos_full(e, member, tbqs, out17);
}

// ---------------------------------------------------------------- events
static int event_create(ccake_dev* dev, const ccake_params* p, ccake_eos* eos, int nev, const int* counts,
                        ccake_event** out)
{
  if (!dev || !p || !eos || !out || nev <= 0 || !counts) return CCAKE_ERR_INVALID;
  long long ntot = 0;
  for (int e = 0; e < nev; ++e) {
    if (counts[e] <= 0) { set_error("every event of a batch needs particles"); return CCAKE_ERR_INVALID; }
    ntot += counts[e];
  }
  if (ntot > INT_MAX / 32) { set_error("too many particles"); return CCAKE_ERR_INVALID; }
  if (nev > kMaxBatchEvents) { set_error("too many events in one batch"); return CCAKE_ERR_INVALID; }
  const int n = (int)ntot;
  if (p->dim != 2 && p->dim != 3) { set_error("D = 2 (2+1D) and D = 3 (3+1D Milne) are built"); return CCAKE_ERR_UNSUPPORTED; }
  if (p->dim == 3) {
    // the 3+1D path: one event per handle, FP64 pair sums, no diffusion, RK2
    if (nev != 1) { set_error("batches are built for D = 2"); return CCAKE_ERR_UNSUPPORTED; }
    if (p->using_diffusion) { set_error("BSQ diffusion is built for D = 2"); return CCAKE_ERR_UNSUPPORTED; }
    if (p->pair_precision != 0) { set_error("mixed-precision pair sums are built for D = 2"); return CCAKE_ERR_UNSUPPORTED; }
    if (p->factorized_kernel && !(p->hEta > 0.0)) { set_error("the factorized kernel needs hEta > 0"); return CCAKE_ERR_INVALID; }
  }
  if (p->using_diffusion && (p->diffusion_mode == 1 || p->diffusion_mode > 2 || p->diffusion_relax_mode > 1)) {
    set_error("DNMR diffusion coefficients are not built (the reference returns zero matrices for them)");
    return CCAKE_ERR_UNSUPPORTED;
  }
  if (p->pair_precision != 0 && p->pair_precision != 1) {
    set_error("pair_precision must be 0 (FP64) or 1 (FP32 with compensated sums)");
    return CCAKE_ERR_INVALID;
  }
  CCAKE_HIP(hipSetDevice(dev->gpu));
  ccake_event* ev = new ccake_event();
  ev->dev = dev; ev->eos = eos; ev->p = *p; ev->n = n;
  ev->npad = (n + 255) / 256 * 256;
  ev->t = p->t0;
  ev->nev = nev;
  ev->h_off.assign(nev + 1, 0);
  for (int e = 0; e < nev; ++e) ev->h_off[e + 1] = ev->h_off[e] + counts[e];
  // one event: up to 65,536 cells (16-bit keys, two radix passes); a batch:
  // up to 128 x 128 cells per event in one key space (three passes)
  ev->cap_cells = nev == 1 ? 65536 : nev * kBatchCellsPerEvent;
  if (p->dim == 3) ev->cap_cells = 1 << 24;                   // (x, y, eta) cells: three passes
  {
    // the largest key: ncells - 1 for one event; a batch adds the retired
    // events' key, ncells (at most cap_cells)
    const long long kmax = nev > 1 ? (long long)ev->cap_cells : (long long)ev->cap_cells - 1;
    ev->sort_passes = kmax < (1LL << 16) ? 2 : (kmax < (1LL << 24) ? 3 : 4);
  }
  ev->n_live = n;
  ev->radix_blocks = (n + 1023) / 1024;
  {
    const hipError_t se = hipStreamCreateWithFlags(&ev->stream, hipStreamNonBlocking);
    if (se != hipSuccess) { ev->stream = nullptr; delete ev; return hip_fail(se, "hipStreamCreateWithFlags"); }
  }
  int rc = 0;
  const size_t np = (size_t)ev->npad;
  const int nscan = 256 * ev->radix_blocks;
  const int nscan_blocks = (nscan + 2047) / 2048 + 1;
  rc |= dalloc(ev, &ev->F, (size_t)NUM_COLUMNS * np);
  rc |= dalloc(ev, &ev->I, (size_t)CCAKE_NUM_IFIELDS * np);
  rc |= dalloc(ev, &ev->grid, (size_t)nev);
  rc |= dalloc(ev, &ev->bbox, 6 * std::max(np / 256 + 1, (size_t)nev));
  if (nev > 1) {
    rc |= dalloc(ev, &ev->ev_off, (size_t)nev + 1);
    rc |= dalloc(ev, &ev->eid, np);
    rc |= dalloc(ev, &ev->red_ev, (size_t)nev);
    rc |= dalloc(ev, &ev->live, (size_t)nev);
  }
  rc |= dalloc(ev, &ev->key, np);
  rc |= dalloc(ev, &ev->key_a, np);
  rc |= dalloc(ev, &ev->key_b, np);
  rc |= dalloc(ev, &ev->val_a, np);
  rc |= dalloc(ev, &ev->val_b, np);
  rc |= dalloc(ev, &ev->hist, (size_t)nscan);
  rc |= dalloc(ev, &ev->scan_tmp, (size_t)nscan_blocks);
  rc |= dalloc(ev, &ev->cell_off, (size_t)ev->cap_cells + 1);
  rc |= dalloc(ev, &ev->pair_count, 2);
  ev->nslices = choose_slices((n + kTile - 1) / kTile, dev->simds);
  if (const char* ov = std::getenv("CCAKE_PAIR_SLICES")) {     // tuning override, 1..16
    const int v = std::atoi(ov);
    if (v >= 1 && v <= kMaxSlices) ev->nslices = v;
  }
  if (const char* ov = std::getenv("CCAKE_LDS_SLICES"))        // tuning override: 0 sums slices through HBM
    ev->lds_reduce = std::atoi(ov) != 0;
  if (const char* ov = std::getenv("CCAKE_SORT_SKIN")) {       // tuning override, fraction of rc
    const double v = std::atof(ov);
    if (v >= 0.0 && v <= 1.0) ev->skin_frac = v;
  }
  ev->nslices_cap = ev->nslices;
  // partial sums: 17 per slice in 2+1D, 28 in 3+1D (the 3D gradient sums)
  rc |= dalloc(ev, &ev->part, (size_t)ev->nslices_cap * (p->dim == 3 ? 28 : kMaxAcc) * np);
  rc |= dalloc(ev, &ev->hard, np);
  {
    // the EoS stage in two kernels (table chains): CCAKE_EOS_SPLIT=0 keeps one
    const char* es = std::getenv("CCAKE_EOS_SPLIT");
    if (!(es && std::atoi(es) == 0)) rc |= dalloc(ev, &ev->eos_done, np);
  }
  rc |= dalloc(ev, &ev->hard_count, 2);
  rc |= dalloc(ev, &ev->cand_s, 8 * np);
  rc |= dalloc(ev, &ev->cand_g, (size_t)(p->dim == 3 ? kCandG3 : 16) * np);
  if (p->using_diffusion) rc |= dalloc(ev, &ev->cand_d, 16 * np);
  if (p->pair_precision == 1) {
    rc |= dalloc(ev, &ev->cand_s32, (size_t)kCandS32 * np);
    rc |= dalloc(ev, &ev->cand_g32, (size_t)kCandG32 * np);
  }
  ev->snap_cols = p->using_diffusion ? 45 : 33;
  rc |= dalloc(ev, &ev->red, 1024);
  rc |= dalloc(ev, &ev->err, 2);          // [0] error flags, [1] charge-check scratch
  if (p->particlization && p->dim == 2) {     // 3+1D runs the status machine without hypersurface rows
    rc |= dalloc(ev, &ev->snaps, (size_t)6 * ev->snap_cols * np);
    rc |= dalloc(ev, &ev->fo_rows, (size_t)n * CCAKE_FO_ROW);
    rc |= dalloc(ev, &ev->fo_count, 1);
  }
  if (rc) { ccake_event_destroy(ev); return CCAKE_ERR_OOM; }
  hipStream_t st = ev->stream;
  CCAKE_HIP(hipMemsetAsync(ev->F, 0, (size_t)NUM_COLUMNS * np * sizeof(double), st));
  CCAKE_HIP(hipMemsetAsync(ev->I, 0, (size_t)CCAKE_NUM_IFIELDS * np * sizeof(int32_t), st));
  CCAKE_HIP(hipMemsetAsync(ev->err, 0, sizeof(int32_t), st));
  CCAKE_HIP(hipMemsetAsync(ev->pair_count, 0, 2 * sizeof(int64_t), st));
  CCAKE_HIP(hipMemsetAsync(ev->hard_count, 0, 2 * sizeof(int32_t), st));
  if (ev->snaps) {
    CCAKE_HIP(hipMemsetAsync(ev->snaps, 0, (size_t)6 * ev->snap_cols * np * sizeof(double), st));
    CCAKE_HIP(hipMemsetAsync(ev->fo_count, 0, sizeof(int), st));
  }
  // sph_mass = 1 and gamma = 1 until the caller overrides them
  std::vector<double> ones(n, 1.0);
  for (int f : { (int)CCAKE_F_MASS_S, (int)CCAKE_F_MASS_B, (int)CCAKE_F_MASS_SS, (int)CCAKE_F_MASS_Q, (int)CCAKE_F_GAMMA })
    CCAKE_HIP(hipMemcpyAsync(col(ev, f), ones.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
  if (nev > 1) {
    std::vector<int32_t> eid(n);
    for (int e = 0; e < nev; ++e)
      for (int i = ev->h_off[e]; i < ev->h_off[e + 1]; ++i) eid[i] = e;
    CCAKE_HIP(hipMemcpyAsync(ev->ev_off, ev->h_off.data(), (nev + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    CCAKE_HIP(hipMemcpyAsync(ev->eid, eid.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    ev->h_live.assign(nev, 1);
    CCAKE_HIP(hipMemcpyAsync(ev->live, ev->h_live.data(), nev * sizeof(int32_t), hipMemcpyHostToDevice, st));
    CCAKE_HIP(hipStreamSynchronize(st));
  }
  CCAKE_HIP(hipStreamSynchronize(st));
  ev->perm = ev->val_a;
  ev->skey = ev->key_a;
  std::lock_guard<std::mutex> lk(dev->mu);
  dev->live_events++;
  eos->live_events++;
  ev->counted = true;
  *out = ev;
  return 0;
}

int ccake_event_create(ccake_dev* dev, const ccake_params* p, ccake_eos* eos, int n, ccake_event** out)
{
  if (n <= 0) return CCAKE_ERR_INVALID;
  return event_create(dev, p, eos, 1, &n, out);
}

int ccake_batch_create(ccake_dev* dev, const ccake_params* p, ccake_eos* eos, int n_events, const int* counts,
                       ccake_event** out)
{
  return event_create(dev, p, eos, n_events, counts, out);
}

int ccake_batch_events(ccake_event* ev, int* n_events, int* offsets)
{
  if (!ev || !n_events) return CCAKE_ERR_INVALID;
  *n_events = ev->nev;
  if (offsets) std::copy(ev->h_off.begin(), ev->h_off.end(), offsets);
  return 0;
}

int ccake_batch_reduce(ccake_event* ev, int which, double* out)
{
  if (!ev || !out || which < 0 || which >= CCAKE_NUM_REDUCTIONS) return CCAKE_ERR_INVALID;
  if (ev->nev == 1) return launch_reduce(ev, which, out);
  return launch_reduce_events(ev, which, out);
}

// The live flags back to the device (every event live again after
// process_initial_conditions; an ended event retired by the continue test).
static int upload_live(ccake_event* ev)
{
  if (!ev->live) return 0;
  int nl = 0;
  for (int e = 0; e < ev->nev; ++e)
    if (ev->h_live[e]) nl += ev->h_off[e + 1] - ev->h_off[e];
  ev->n_live = nl;
  ev->sort_valid = false;               // the retired particles move behind the live ones at the next sort
  CCAKE_HIP(hipMemcpyAsync(ev->live, ev->h_live.data(), ev->nev * sizeof(int32_t), hipMemcpyHostToDevice,
                           ev->stream));
  CCAKE_HIP(hipStreamSynchronize(ev->stream));
  return 0;
}

int ccake_batch_continue_evolution(ccake_event* ev, double max_tau, int* keep_going)
{
  if (!ev || !keep_going) return CCAKE_ERR_INVALID;
  std::vector<double> frozen(ev->nev, 0.0);
  if (ev->p.particlization) {
    const int rc = ccake_batch_reduce(ev, CCAKE_R_FROZEN, frozen.data());
    if (rc) return rc;
  }
  bool retire = false;
  for (int e = 0; e < ev->nev; ++e) {
    const int ne = ev->h_off[e + 1] - ev->h_off[e];
    const int keep = ev->p.particlization ? ((int)(frozen[e] + 0.5) < ne) : 1;
    keep_going[e] = keep && (ev->t <= max_tau);
    // an event that has ended stays ended: its frozen count cannot fall and
    // tau only grows
    if (ev->live && !keep_going[e] && ev->h_live[e]) { ev->h_live[e] = 0; retire = true; }
    if (ev->live && !ev->h_live[e]) keep_going[e] = 0;
  }
  return retire ? upload_live(ev) : 0;
}

int ccake_event_destroy(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  if (ev->stream) (void)hipStreamSynchronize(ev->stream);
  void* bufs[] = { ev->F, ev->I, ev->grid, ev->bbox, ev->key, ev->key_a, ev->key_b, ev->val_a, ev->val_b, ev->hist,
                   ev->scan_tmp, ev->cell_off, ev->pair_count, ev->part, ev->hard, ev->hard_count, ev->k4, ev->cand_s, ev->cand_g, ev->red,
                   ev->err, ev->snaps, ev->fo_rows, ev->fo_count, ev->cand_d, ev->cand_s32, ev->cand_g32,
                   ev->ev_off, ev->eid, ev->red_ev, ev->live, ev->eos_done };
  for (void* b : bufs) if (b) (void)hipFree(b);
  if (ev->stream) (void)hipStreamDestroy(ev->stream);
  {
    std::lock_guard<std::mutex> lk(ev->dev->mu);
    ev->dev->used -= (ev->dev->used >= ev->bytes ? ev->bytes : ev->dev->used);
    if (ev->counted) { ev->dev->live_events--; ev->eos->live_events--; }
  }
  delete ev;
  return 0;
}

int ccake_event_size(ccake_event* ev, int* n) { if (!ev || !n) return CCAKE_ERR_INVALID; *n = ev->n; return 0; }
int ccake_event_set_time(ccake_event* ev, double t) { if (!ev) return CCAKE_ERR_INVALID; ev->t = t; return 0; }
int ccake_event_get_time(ccake_event* ev, double* t) { if (!ev || !t) return CCAKE_ERR_INVALID; *t = ev->t; return 0; }

int ccake_event_upload(ccake_event* ev, int field, const double* host, size_t n)
{
  if (!ev || !host || field < 0 || field >= CCAKE_NUM_FIELDS || n != (size_t)ev->n) return CCAKE_ERR_INVALID;
  ev->no_charges = false;               // re-established by the next initialize
  ev->no_bulk = false;
  if (field == CCAKE_F_X || field == CCAKE_F_Y || field == CCAKE_F_ETA) {
    for (size_t i = 0; i < n; ++i)
      if (!std::isfinite(host[i])) { set_error("non-finite particle position"); return CCAKE_ERR_INVALID; }
    ev->sort_valid = false;             // the cell sort no longer matches
  }
  CCAKE_HIP(hipMemcpyAsync(col(ev, field), host, n * sizeof(double), hipMemcpyHostToDevice, ev->stream));
  CCAKE_HIP(hipStreamSynchronize(ev->stream));
  return 0;
}

int ccake_event_download(ccake_event* ev, int field, double* host, size_t n)
{
  if (!ev || !host || field < 0 || field >= CCAKE_NUM_FIELDS || n != (size_t)ev->n) return CCAKE_ERR_INVALID;
  CCAKE_HIP(hipMemcpyAsync(host, col(ev, field), n * sizeof(double), hipMemcpyDeviceToHost, ev->stream));
  CCAKE_HIP(hipStreamSynchronize(ev->stream));
  return 0;
}

int ccake_event_upload_int(ccake_event* ev, int f, const int32_t* host, size_t n)
{
  if (!ev || !host || f < 0 || f >= CCAKE_NUM_IFIELDS || n != (size_t)ev->n) return CCAKE_ERR_INVALID;
  CCAKE_HIP(hipMemcpyAsync(icol(ev, f), host, n * sizeof(int32_t), hipMemcpyHostToDevice, ev->stream));
  CCAKE_HIP(hipStreamSynchronize(ev->stream));
  return 0;
}

int ccake_event_download_int(ccake_event* ev, int f, int32_t* host, size_t n)
{
  if (!ev || !host || f < 0 || f >= CCAKE_NUM_IFIELDS || n != (size_t)ev->n) return CCAKE_ERR_INVALID;
  CCAKE_HIP(hipMemcpyAsync(host, icol(ev, f), n * sizeof(int32_t), hipMemcpyDeviceToHost, ev->stream));
  CCAKE_HIP(hipStreamSynchronize(ev->stream));
  return 0;
}

// ---------------------------------------------------------------- phases
int ccake_build_neighbors(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  int rc;
  if ((rc = fresh_sort(ev))) return rc;
  return launch_smooth(ev, 2);
}

int ccake_export_pairs(ccake_event* ev, int32_t* pi, int32_t* pj, size_t cap, size_t* n_pairs)
{
  if (!ev || !n_pairs) return CCAKE_ERR_INVALID;
  int rc = ensure_sorted(ev);
  if (!rc) rc = launch_gather_smooth(ev);
  if (rc) return rc;
  return launch_export_pairs(ev, pi, pj, cap, n_pairs);
}

int ccake_export_pairs_noalloc(ccake_event* ev, int32_t* pi, int32_t* pj, size_t cap, size_t* n_pairs)
{
  if (!ev || !n_pairs) return CCAKE_ERR_INVALID;
  if (ev->p.dim == 3) { set_error("the allocation-free export is built for D = 2"); return CCAKE_ERR_UNSUPPORTED; }
  int rc = ensure_sorted(ev);
  if (!rc) rc = launch_gather_smooth(ev);
  if (rc) return rc;
  return launch_export_pairs_noalloc(ev, pi, pj, cap, n_pairs);
}

int ccake_gamma_velocity(ccake_event* ev) { return ev ? launch_stage_begin(ev) : CCAKE_ERR_INVALID; }

int ccake_smooth_fields(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  int rc = ensure_sorted(ev);
  if (!rc) rc = launch_gather_smooth(ev);
  return rc ? rc : launch_smooth(ev, 0);
}

int ccake_initial_sigma(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  int rc = ensure_sorted(ev);
  if (!rc) rc = launch_gather_smooth(ev);
  return rc ? rc : launch_smooth(ev, 1);
}

int ccake_eos_invert(ccake_event* ev, int mode)
{
  if (!ev || (mode != 0 && mode != 1)) return CCAKE_ERR_INVALID;
  int rc = launch_eos(ev, mode == 1 ? 1 : 2, 0);
  if (rc) return rc;
  return ccake_event_check(ev);
}

int ccake_reset_pi_tensor(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  int rc = launch_reset_pi(ev);
  // with diffusion: q^mu from the extensive diffusion, then the regulator
  if (!rc && ev->p.using_diffusion) rc = launch_diff_prep(ev, 1 | 8, nullptr);
  return rc;
}

int ccake_viscosities(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  int rc = launch_viscosities(ev);
  if (!rc && ev->p.using_diffusion) rc = launch_diff_prep(ev, 2, nullptr);
  return rc;
}

int ccake_smooth_gradients(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  int rc = ensure_sorted(ev);
  if (!rc) rc = launch_gather_grad(ev);
  if (!rc && ev->p.using_diffusion) rc = launch_diff_prep(ev, 4, ev->perm);
  if (!rc) rc = launch_grad(ev);
  if (!rc && ev->p.using_diffusion) rc = launch_diff_grad(ev);
  return rc;
}

int ccake_eom(ccake_event* ev) { return ev ? launch_eom(ev, 0, 0.0, ev->t) : CCAKE_ERR_INVALID; }

int ccake_get_time_derivatives(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  int rc = time_derivatives(ev, 0, 0.0, ev->t, false);
  if (rc) return rc;
  return ccake_event_check(ev);
}

int ccake_cache_state(ccake_event* ev) { return ev ? launch_cache(ev) : CCAKE_ERR_INVALID; }

int ccake_rk_update(ccake_event* ev, double dt, double t0)
{
  if (!ev) return CCAKE_ERR_INVALID;
  int rc = launch_rk(ev, dt, t0);
  if (rc) return rc;
  ev->sort_valid = false;               // positions moved from outside a step
  ev->t = t0 + dt;
  return 0;
}

int ccake_advance_timestep(ccake_event* ev, double dt, int rk_order)
{
  if (!ev) return CCAKE_ERR_INVALID;
  if (rk_order != 2 && rk_order != 4) { set_error("Invalid Runge-Kutta order"); return CCAKE_ERR_INVALID; }
  int rc = rk_order == 4 ? rk4_step(ev, dt) : rk2_step(ev, dt);
  if (rc) return rc;
  return ccake_event_check(ev);
}

int ccake_allocate_k_values(ccake_event* ev) { return ev ? rk4_allocate_k(ev) : CCAKE_ERR_INVALID; }

int ccake_update_k(ccake_event* ev, int n)
{
  if (!ev || n < 1 || n > 4) return CCAKE_ERR_INVALID;
  if (!ev->k4) { set_error("allocate_k_values first"); return CCAKE_ERR_INVALID; }
  static const double w[5] = { 0.0, 1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0 };
  return launch_rk4_update_k(ev, w[n]);
}

int ccake_update_rk4(ccake_event* ev, double dt)
{
  if (!ev) return CCAKE_ERR_INVALID;
  if (!ev->k4) { set_error("allocate_k_values first"); return CCAKE_ERR_INVALID; }
  const int rc = launch_rk4_update(ev, dt);
  if (rc) return rc;
  ev->sort_valid = false;               // positions moved from outside a step
  return ccake_event_check(ev);
}

int ccake_run_steps(ccake_event* ev, double dt, int rk_order, int n_steps, int* steps_done)
{
  if (!ev) return CCAKE_ERR_INVALID;
  if (rk_order != 2 && rk_order != 4) { set_error("Invalid Runge-Kutta order"); return CCAKE_ERR_INVALID; }
  int done = 0;
  bool begun = false;                  // the next stage was begun by the last EoM kernel
  for (int k = 0; k < n_steps; ++k) {
    int rc = rk_order == 4 ? rk4_step(ev, dt) : rk2_step(ev, dt, k + 1 < n_steps, &begun);
    if (rc) { if (steps_done) *steps_done = done; return rc; }
    ++done;
  }
  if (steps_done) *steps_done = done;
  return ccake_event_check(ev);
}

int ccake_process_initial_conditions(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  ev->t = ev->p.t0;
  ev->frzc = 0;
  ev->fo_step = 0;
  ev->taup = ev->taupp = 0.0;
  ev->init_phase = 1;
  if (ev->live) {                                           // a new run: every event live
    ev->h_live.assign(ev->nev, 1);
    const int rc = upload_live(ev);
    if (rc) return rc;
  }
  if (ev->snaps) {                                          // a new surface
    CCAKE_HIP(hipMemsetAsync(ev->snaps, 0, (size_t)6 * ev->snap_cols * ev->npad * sizeof(double), ev->stream));
    CCAKE_HIP(hipMemsetAsync(ev->fo_count, 0, sizeof(int), ev->stream));
  }
  return launch_init_particles(ev, 0);
}

int ccake_initialize_entropy_and_charge_densities(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  if (ev->init_phase < 1) { set_error("process_initial_conditions first"); return CCAKE_ERR_INVALID; }
  int rc;
  if ((rc = ensure_sorted(ev))) return rc;                  // initialize_linklist, if the caller has not
  if ((rc = launch_eos(ev, 1, 0))) return rc;               // e -> s inversion
  if ((rc = launch_init_particles(ev, 1))) return rc;       // freeze below e_freeze
  if ((rc = ccake_initial_sigma(ev))) return rc;            // calculate_intial_sigma
  if ((rc = launch_init_particles(ev, 2))) return rc;       // init_particles
  ev->init_phase = 2;
  return 0;
}

int ccake_initial_smoothing(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  if (ev->init_phase < 2) { set_error("initialize_entropy_and_charge_densities first"); return CCAKE_ERR_INVALID; }
  int rc;
  if ((rc = launch_stage_begin(ev))) return rc;             // calculate_gamma_and_velocities
  if ((rc = ccake_smooth_fields(ev))) return rc;            // smooth_all_particle_fields
  if ((rc = launch_init_particles(ev, 3))) return rc;       // extensive shear, bulk, set_diffusion
  if ((rc = launch_reset_pi(ev))) return rc;
  if (ev->p.using_diffusion && (rc = launch_diff_prep(ev, 1, nullptr))) return rc;   // q^mu
  if ((rc = launch_eos(ev, 2, 0))) return rc;               // update_all_particle_thermodynamics
  if ((rc = launch_freeze(ev, -1))) return rc;              // check_freeze_out_status
  // charges stay zero without diffusion (dN/dt = 0), and the bulk pressure
  // stays zero with zeta = 0 and no shear (dPi/dt = 0 at Pi = 0; with shear
  // the lambda_Pipi and phi3 couplings can source it): the pair passes can
  // then skip those sums
  bool any_charge = true, any_bulk = true;
  if ((rc = launch_charge_check(ev, &any_charge, &any_bulk))) return rc;
  ev->no_charges = !ev->p.using_diffusion && !any_charge;
  ev->no_bulk = ev->p.zeta_mode == CCAKE_ZETA_CONSTANT && ev->p.constant_zeta_over_s == 0.0 && !ev->p.using_shear
                && !any_bulk;
  ev->init_phase = 3;
  return ccake_event_check(ev);
}

int ccake_initialize(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  int rc;
  if ((rc = ccake_process_initial_conditions(ev))) return rc;
  if ((rc = ccake_build_neighbors(ev))) return rc;          // initialize_linklist
  if ((rc = ccake_initialize_entropy_and_charge_densities(ev))) return rc;
  return ccake_initial_smoothing(ev);                       // begins with calculate_gamma_and_velocities
}

int ccake_freeze_out_particles(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  return freeze_step(ev);
}

int ccake_continue_evolution(ccake_event* ev, double max_tau, int* keep_going)
{
  if (!ev || !keep_going) return CCAKE_ERR_INVALID;
  int keep = 1;
  if (ev->p.particlization) {
    double frozen = 0.0;
    int rc = launch_reduce(ev, CCAKE_R_FROZEN, &frozen);
    if (rc) return rc;
    keep = (int)(frozen + 0.5) < ev->n;
  }
  *keep_going = keep && (ev->t <= max_tau);
  return 0;
}

int ccake_event_check(ccake_event* ev)
{
  if (!ev) return CCAKE_ERR_INVALID;
  int32_t e = 0;
  hipStream_t st = ev->stream;
  CCAKE_HIP(hipMemcpyAsync(&e, ev->err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipStreamSynchronize(st));
  if (e) CCAKE_HIP(hipMemsetAsync(ev->err, 0, sizeof(int32_t), st));
  if (e & DE_EOS_FAIL) { set_error("no EoS solution"); return CCAKE_ERR_EOS_NO_SOLUTION; }
  if (e & DE_NEG_ENTROPY) { set_error("negative entropy density"); return CCAKE_ERR_NEG_ENTROPY; }
  if (e & DE_GRID_OVERFLOW) { set_error("non-finite positions"); return CCAKE_ERR_INVALID; }
  return 0;
}

int ccake_reduce(ccake_event* ev, int which, double* out)
{
  if (!ev || !out || which < 0 || which >= CCAKE_NUM_REDUCTIONS) return CCAKE_ERR_INVALID;
  return launch_reduce(ev, which, out);
}

int ccake_freeze_out_surface(ccake_event* ev, double* rows, size_t cap_rows, size_t* n_rows)
{
  if (!ev || !n_rows) return CCAKE_ERR_INVALID;
  if (ev->p.dim == 3) { set_error("freeze-out hypersurface rows are built for D = 2"); return CCAKE_ERR_UNSUPPORTED; }
  if (!ev->fo_rows) { set_error("freeze-out surface needs particlization at event creation"); return CCAKE_ERR_INVALID; }
  int cnt = 0;
  CCAKE_HIP(hipMemcpyAsync(&cnt, ev->fo_count, sizeof(int), hipMemcpyDeviceToHost, ev->stream));
  CCAKE_HIP(hipStreamSynchronize(ev->stream));
  const size_t n = (size_t)(cnt < ev->n ? cnt : ev->n);    // each particle appends at most once
  *n_rows = n;
  if (!rows) return 0;
  if (cap_rows < n) { set_error("row buffer too small"); return CCAKE_ERR_CAPACITY; }
  std::vector<double> tmp(n * CCAKE_FO_ROW);
  if (n) {
    CCAKE_HIP(hipMemcpyAsync(tmp.data(), ev->fo_rows, n * CCAKE_FO_ROW * sizeof(double), hipMemcpyDeviceToHost,
                             ev->stream));
    CCAKE_HIP(hipStreamSynchronize(ev->stream));
  }
  // the device appends in arrival order; order by (step, particle id)
  std::vector<size_t> idx(n);
  for (size_t k = 0; k < n; ++k) idx[k] = k;
  std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    const double* ra = &tmp[a * CCAKE_FO_ROW];
    const double* rb = &tmp[b * CCAKE_FO_ROW];
    return ra[1] != rb[1] ? ra[1] < rb[1] : ra[0] < rb[0];
  });
  for (size_t k = 0; k < n; ++k)
    std::memcpy(rows + k * CCAKE_FO_ROW, &tmp[idx[k] * CCAKE_FO_ROW], CCAKE_FO_ROW * sizeof(double));
  return 0;
}

int ccake_profile_enable(ccake_dev* d, int enable)
{
  if (!d) return CCAKE_ERR_INVALID;
  d->prof.enabled = enable != 0;
  return 0;
}

int ccake_profile_read(ccake_dev* d, int k, double* total_ms, int64_t* launches)
{
  if (!d || k < 0 || k >= CCAKE_NUM_KERNELS) return CCAKE_ERR_INVALID;
  prof_drain(d);
  if (total_ms) *total_ms = d->prof.total_ms[k];
  if (launches) *launches = d->prof.launches[k];
  return 0;
}

int ccake_profile_reset(ccake_dev* d)
{
  if (!d) return CCAKE_ERR_INVALID;
  prof_drain(d);
  for (int k = 0; k < CCAKE_NUM_KERNELS; ++k) { d->prof.total_ms[k] = 0; d->prof.launches[k] = 0; }
  return 0;
}

int ccake_last_pair_count(ccake_event* ev, int64_t* n_pairs)
{
  if (!ev || !n_pairs) return CCAKE_ERR_INVALID;
  int64_t v = 0;
  CCAKE_HIP(hipMemcpyAsync(&v, ev->pair_count, sizeof(int64_t), hipMemcpyDeviceToHost, ev->stream));
  CCAKE_HIP(hipStreamSynchronize(ev->stream));
  *n_pairs = v;
  return 0;
}

int ccake_event_pair_slices(ccake_event* ev, int set, int* nslices)
{
  if (!ev) return CCAKE_ERR_INVALID;
  if (set > 0) {
    if (set > ev->nslices_cap) { set_error("slice count above the event's partial-sum capacity"); return CCAKE_ERR_INVALID; }
    ev->nslices = set;
  }
  if (nslices) *nslices = ev->nslices;
  return 0;
}

int ccake_pair_total(ccake_event* ev, int reset, int64_t* n_pairs)
{
  if (!ev || !n_pairs) return CCAKE_ERR_INVALID;
  int64_t v = 0;
  CCAKE_HIP(hipMemcpyAsync(&v, ev->pair_count + 1, sizeof(int64_t), hipMemcpyDeviceToHost, ev->stream));
  CCAKE_HIP(hipStreamSynchronize(ev->stream));
  *n_pairs = v;
  if (reset) {
    CCAKE_HIP(hipMemsetAsync(ev->pair_count + 1, 0, sizeof(int64_t), ev->stream));
    CCAKE_HIP(hipStreamSynchronize(ev->stream));
  }
  return 0;
}

}  // extern "C"
