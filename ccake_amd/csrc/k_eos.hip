// Per-particle physics after the smoothing pass, fused in one kernel:
//   04 EoS inversion        SPHWorkstation::update_all_particle_thermodynamics
//                           (src/sph_workstation.cpp:1713-1772) -- on the device,
//                           where the reference runs a serial host loop
//   05 reset_pi_tensor      EoM_default::reset_pi_tensor (src/eom_default.cpp:56-172)
//   08 viscosities          SPHWorkstation::update_all_particle_viscosities (1622-1699)
//   gather of the gradient-pass candidate records in cell-sorted order.
// Threads walk the cell-sorted order, so the lanes of a wave hold nearby
// particles with similar (T, mu) and share table cells.
#include "eos_device.h"
#include "preinv_device.h"
#include <vector>
#include <cstring>

namespace ccake {

constexpr int kCandG = 16;   // doubles per gradient candidate record

struct Transport {
  int eta_mode, shear_relax_mode, zeta_mode, modulate;
  double eta_over_s, zeta_over_s, zA, zp, gA, gT0, gW;
  int delta_pipi, tau_pipi, lambda_piPi, phi7, delta_PiPi, lambda_Pipi;
};

__device__ __forceinline__ double tc_eta(const Transport& P, double T, double w)
{
  if (P.eta_mode == 1) return (T <= 0.0) ? 0.0 : P.eta_over_s * (w / T);
  return 0.20 * (w / T);
}
__device__ __forceinline__ double tc_zeta(const Transport& P, double T, double w, double cs2)
{
  switch (P.zeta_mode) {
    case 1: return P.zeta_over_s * (w / T);
    case 2: {
      double factor = 1.0;
      if (P.modulate) {
        const double Ttr = 150.0 / kHbarcMeVfm, Tsc = 10.0 / kHbarcMeVfm;
        const double thx = tanh((T - Ttr) / Tsc);
        const double xp = pow(T / Ttr, P.zp);
        factor = 0.5 * (1.0 + xp) + 0.5 * (1.0 - xp) * thx;
      }
      return P.zA * factor * pow((1.0 / 3.0) - fmin(cs2, 1.0), P.zp) * w / T;
    }
    case 3: {
      const double d = T - P.gT0;
      return P.gA * exp(-d * d / (2.0 * P.gW * P.gW)) * w / T;
    }
    default: return 0.005 * (w / T);
  }
}

// transport_coefficients dispatchers (src/transport_coefficients.cpp:284-912)
__device__ __forceinline__ void transport(const Transport& P, double T, double w, double p, double cs2, double* o)
{
  const double eta = tc_eta(P, T, w);
  const double x = 5.0 * eta / w;
  const double tau_pi = (P.shear_relax_mode == 2) ? x : fmax(x, P.shear_relax_mode == 1 ? 0.001 : 0.005);
  const double zeta = tc_zeta(P, T, w, cs2);
  const double cc = cs2 > 1 ? 1 : cs2;
  const double pw = pow(1.0 / 3.0 - cc, P.zp);
  const double tau_Pi = (pw < 1e-10) ? 1e10 : fmax(5.0 * zeta / (pw * w), 0.1);
  o[0] = eta;
  o[1] = tau_pi;
  o[2] = zeta;
  o[3] = tau_Pi;
  o[4] = P.delta_PiPi == 1 ? 0.0 : (P.delta_PiPi == 2 ? tau_Pi : 2. * tau_Pi / 3.0);
  o[5] = P.delta_pipi == 1 ? 0.0 : (P.delta_pipi == 2 ? tau_pi : 4. * tau_pi / 3.0);
  o[6] = P.lambda_piPi == 1 ? 0.0 : 6. / 5. * tau_pi;
  o[7] = P.lambda_Pipi == 1 ? 0.0 : 8. * (1. / 3. - cc) * tau_Pi / 5.;
  o[8] = 0.0;
  o[9] = 0.0;
  o[10] = 0.0;
  o[11] = P.phi7 == 1 ? 0.0 : 9. / (70. * p);
  o[12] = P.tau_pipi == 1 ? 0.0 : 10. * tau_pi / 7.0;
}

// MODE 0: s-inversion + reset_pi + transport + gradient gather (sorted order).
// MODE 1: e-inversion of the inputs (id order), writes input s.
// MODE 2: s-inversion only (id order).
// reset_pi_tensor, transport coefficients and the gradient candidate record
// of particle i (sorted position k) after its thermodynamics are known: the
// tail of the fused EoS stage.
__device__ __forceinline__ void stage_post(double* __restrict__ F, int npad, int i, int k, double t,
                                           const Transport& P, double T, double w, double pr, double cs2,
                                           double e, double* __restrict__ cand)
{
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
  // ---- 05 reset_pi_tensor, D = 2 (fixed-size u has u^eta = 0) ----
  const double gamma = col(CCAKE_F_GAMMA), sigma = col(CCAKE_F_SIGMA);
  const double t2 = t * t;
  const double ux = col(CCAKE_F_UX), uy = col(CCAKE_F_UY);
  const double fc0 = -ux, fc1 = -uy;
  const double bulk = col(CCAKE_F_EXT_BULK) * sigma;
  col(CCAKE_F_BULK) = bulk;
  const double s11 = col(CCAKE_F_EXT_SHV0 + 0) * sigma, s12 = col(CCAKE_F_EXT_SHV0 + 1) * sigma;
  const double s13 = col(CCAKE_F_EXT_SHV0 + 2) * sigma, s22 = col(CCAKE_F_EXT_SHV0 + 3) * sigma;
  const double s23 = col(CCAKE_F_EXT_SHV0 + 4) * sigma;
  const double s00 = (fc0 * fc0 * s11 + fc1 * fc1 * s22 + 2. * fc1 * fc0 * s12) / (gamma * gamma);
  const double s33 = (s00 - s11 - s22) / t2;
  const double s01 = -s11 * fc0 / gamma + -s12 * fc1 / gamma;
  const double s02 = -s12 * fc0 / gamma + -s22 * fc1 / gamma;
  const double s03 = -s13 * fc0 / gamma + -s23 * fc1 / gamma;
  const double S[16] = { s00, s01, s02, s03, s01, s11, s12, s13, s02, s12, s22, s23, s03, s13, s23, s33 };
  for (int q = 0; q < 16; ++q) col(CCAKE_F_SHV0 + q) = S[q];

  // ---- 08 transport coefficients ----
  double tc[13];
  transport(P, T, w, pr, cs2, tc);
  for (int q = 0; q < 13; ++q) col(CCAKE_F_ETA_PI + q) = tc[q];

  // ---- gradient candidate record: x, y, m_s, p/sl^2, e, Pi/sl^2, v, pi^{0i}/sl^2, pi^{ij}/sl^2
  const double sl = col(CCAKE_F_SIGMA_LAB);
  const double isl2 = 1.0 / (sl * sl);
  double* o = cand + (size_t)k * kCandG;
  o[0] = col(CCAKE_F_X); o[1] = col(CCAKE_F_Y); o[2] = col(CCAKE_F_MASS_S);
  o[3] = pr * isl2; o[4] = e; o[5] = bulk * isl2;
  o[6] = col(CCAKE_F_VX); o[7] = col(CCAKE_F_VY);
  o[8] = s01 * isl2; o[9] = s02 * isl2;
  o[10] = s11 * isl2; o[11] = s12 * isl2; o[12] = s22 * isl2;
  o[13] = sl; o[14] = 0.0; o[15] = 0.0;
}

// The smoothing sum pass folded into the stage's EoS kernel (MODE 3): the
// slices' partial sums of sorted position k in slice order, the smoothed
// fields, btrack and the stage's pair total, as k_smooth_sum<0, S> writes
// them, then the EoS stage of MODE 0 on the fields just written.
struct SmoothSum {
  const double* part = nullptr;
  int nslices = 0;
  double norm2 = 0.0;
  unsigned long long* pair_total = nullptr;
};

__device__ __forceinline__ void smooth_sum_into(const SmoothSum& sm, double* __restrict__ F, int npad, int n,
                                                int k, int id, double t, int32_t* __restrict__ btrack)
{
  constexpr int NACC = 6;
  const bool act = k < n;
  double acc[NACC];
  #pragma unroll
  for (int f = 0; f < NACC; ++f) acc[f] = 0.0;
  if (act) {
    // blocks of 8 slices, every load of a block issued before its adds (a
    // clamped slice index past the last slice adds an exact 0.0)
    const int ns = sm.nslices;
    for (int s0 = 0; s0 < ns; s0 += 8) {
      double v[8][NACC];
      #pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int sli = s0 + u < ns ? s0 + u : ns - 1;
        #pragma unroll
        for (int f = 0; f < NACC; ++f) v[u][f] = sm.part[((size_t)sli * NACC + f) * npad + k];
      }
      #pragma unroll
      for (int u = 0; u < 8; ++u)
        #pragma unroll
        for (int f = 0; f < NACC; ++f) acc[f] += (s0 + u < ns) ? v[u][f] : 0.0;
    }
  }
  const int cnt = act ? (int)acc[1] - 1 : 0;    // neighbours, self excluded
  if (sm.pair_total) {
    // one pair of atomics per wave
    unsigned long long tot = (unsigned long long)cnt;
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
    if ((threadIdx.x & (kWave - 1)) == 0 && tot) { atomicAdd(sm.pair_total, tot); atomicAdd(sm.pair_total + 1, tot); }
  }
  if (!act) return;
  const int b = btrack[id];
  btrack[id] = (b == -1) ? -1 : cnt;
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + id]; };
  const double sl = acc[0] * sm.norm2;
  const double gam = col(CCAKE_F_GAMMA);
  col(CCAKE_F_SIGMA_LAB) = sl;
  col(CCAKE_F_SIGMA) = sl / gam / t;            // get_LRF (src/eom_default.cpp:198-203)
  const double ss = acc[2] * sm.norm2, sB = acc[3] * sm.norm2;
  const double sS = acc[4] * sm.norm2, sQ = acc[5] * sm.norm2;
  col(CCAKE_F_SM_S) = ss; col(CCAKE_F_SM_B) = sB; col(CCAKE_F_SM_SS) = sS; col(CCAKE_F_SM_Q) = sQ;
  col(CCAKE_F_S) = ss / gam / t;
  col(CCAKE_F_RHOB) = sB / gam / t;
  col(CCAKE_F_RHOS) = sS / gam / t;
  col(CCAKE_F_RHOQ) = sQ / gam / t;
}

// The EoS stage runs in two kernels.  k_eos (one thread per particle) tries
// the first attempt of the inversion -- the starting member, the first seed,
// find_root's attempt 0 -- with a Newton budget; a particle for which that
// attempt decides the result (almost every particle of a time step, which
// starts at last step's solution) is finished there.  The others are
// appended to the event's hard list, and k_eos_hard runs the whole
// inversion for each of them with one wave per particle (invert_wave), so a
// particle that needs the full seed cascade no longer holds 63 others.
// Both give invert_point's result, so the split changes no value.
constexpr int kFastBudget = 12;     // Newton iterations of the first attempt in k_eos

struct HardList {
  int32_t* idx = nullptr;           // positions (sorted for MODE 0, 3; id for 1, 2)
  int32_t* count = nullptr;         // this stage's counter
  int32_t* next = nullptr;          // the next stage's counter, cleared by k_eos_hard
};

// KM: the member kinds the chain holds (kAllMembers, or kConformalChain for
// the analytic conformal chain, which then compiles without the table code).
// MODE 0: the stage (EoS, reset_pi_tensor, transport, gradient records);
// 1: the initial e inversion; 2: EoS only; 3: MODE 0 after the smoothing sum.
#ifndef CCAKE_EOS_WAVES_PER_SIMD
#define CCAKE_EOS_WAVES_PER_SIMD 1
#endif
// PART 0: the whole first attempt and tail in one kernel.  PART 1 and 2
// split it where the live registers are fewest: PART 1 runs the smoothing
// sum and the Newton attempt and stores the root (T, mu) and member with a
// per-position flag `done`; PART 2 (after k_eos_hard) evaluates the
// thermodynamics and the tail for the flagged positions.  The values are
// those of PART 0: the same operations on the same operands.
template <int MODE, int KM, int PART>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CCAKE_EOS_WAVES_PER_SIMD)))
void k_eos(EosChain ch, double* __restrict__ F, int npad, int n,
                                             int32_t* __restrict__ I, const int32_t* __restrict__ perm,
                                             double t, Transport P, double* __restrict__ cand,
                                             int32_t* __restrict__ err, SmoothSum sm, HardList hl,
                                             int32_t* __restrict__ done)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = k < n ? k : n - 1;
  const int i = (MODE == 0 || MODE == 3) ? perm[kk] : kk;
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
  int32_t* eos_id = I + (size_t)CCAKE_I_EOS_ID * npad;
  // the Newton seed and member are read before the smoothing sum's stores,
  // so their loads overlap with the partial sums' loads
  double x[4] = { col(CCAKE_F_T), col(CCAKE_F_MUB), col(CCAKE_F_MUQ), col(CCAKE_F_MUS) };
  int member = eos_id[i];
  if (PART == 2) {
    if (k >= n || !done[k]) return;
  } else {
    if (MODE == 3) smooth_sum_into(sm, F, npad, n, k, i, t, I + (size_t)CCAKE_I_BTRACK * npad);
    if (k >= n) return;
  }
  if (PART != 2) {
  double tgt[4];
  if (MODE == 1) {
    tgt[0] = col(CCAKE_F_IN_E); tgt[1] = col(CCAKE_F_IN_B); tgt[2] = col(CCAKE_F_IN_SS); tgt[3] = col(CCAKE_F_IN_Q);
  } else {
    tgt[0] = col(CCAKE_F_S); tgt[1] = col(CCAKE_F_RHOB); tgt[2] = col(CCAKE_F_RHOS); tgt[3] = col(CCAKE_F_RHOQ);
  }
  if (invert_first<KM>(ch, ch.tab, ch.tab4, MODE == 1 ? 1 : 0, tgt, x, &member, kFastBudget) != 1) {
    hl.idx[atomicAdd(hl.count, 1)] = k;      // the whole inversion in k_eos_hard
    if (PART == 1) done[k] = 0;
    return;
  }
  if (PART == 1) {
    col(CCAKE_F_T) = x[0]; col(CCAKE_F_MUB) = x[1]; col(CCAKE_F_MUQ) = x[2]; col(CCAKE_F_MUS) = x[3];
    eos_id[i] = member;
    done[k] = 1;
    return;
  }
  }
  double th[31];
  set_thermo<KM>(ch, ch.tab, member, x, th);
  for (int q = 0; q < 31; ++q) col(CCAKE_F_T + q) = th[q];
  eos_id[i] = member;
  if (MODE == 1) { col(CCAKE_F_IN_S) = th[5]; return; }
  if (MODE == 2) return;

  stage_post(F, npad, i, k, t, P, th[0], th[11], th[9], th[10], th[4], cand);
}

// The whole inversion (invert_point, with the initial inversion's freeze and
// zero-charge retry, src/sph_workstation.cpp:1825-1848) of every particle on
// the hard list, one wave per particle, then the same tail as k_eos.  Waves
// stride over the list; every lane of a wave takes part in invert_wave.
template <int MODE, int KM>
__global__ __launch_bounds__(kHardLanes) void k_eos_hard(EosChain ch, double* __restrict__ F, int npad, int n,
                                                         int32_t* __restrict__ I, const int32_t* __restrict__ perm,
                                                         double t, Transport P, double* __restrict__ cand,
                                                         int32_t* __restrict__ err, HardList hl)
{
  const int lane = threadIdx.x;
  const int cnt = *hl.count;
  if (blockIdx.x == 0 && lane == 0) *hl.next = 0;   // the next stage's list starts empty
  int32_t* eos_id = I + (size_t)CCAKE_I_EOS_ID * npad;
  int32_t* freeze = I + (size_t)CCAKE_I_FREEZE * npad;
  for (int h = blockIdx.x; h < cnt; h += gridDim.x) {
    const int k = hl.idx[h];
    const int i = (MODE == 0 || MODE == 3) ? perm[k] : k;
    auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
    double x[4] = { col(CCAKE_F_T), col(CCAKE_F_MUB), col(CCAKE_F_MUQ), col(CCAKE_F_MUS) };
    int member = eos_id[i];
    double tgt[4];
    bool ok;
    if (MODE == 1) {
      tgt[0] = col(CCAKE_F_IN_E); tgt[1] = col(CCAKE_F_IN_B); tgt[2] = col(CCAKE_F_IN_SS); tgt[3] = col(CCAKE_F_IN_Q);
      ok = invert_wave<KM>(ch, ch.tab, ch.tab4, 1, tgt, x, &member);
      if (!ok) {
        // src/sph_workstation.cpp:1825-1848: freeze and retry at zero charge
        if (lane == 0) freeze[i] = 4;
        const double z[4] = { tgt[0], 0.0, 0.0, 0.0 };
        ok = invert_wave<KM>(ch, ch.tab, ch.tab4, 1, z, x, &member);
      }
    } else {
      tgt[0] = col(CCAKE_F_S); tgt[1] = col(CCAKE_F_RHOB); tgt[2] = col(CCAKE_F_RHOS); tgt[3] = col(CCAKE_F_RHOQ);
      ok = invert_wave<KM>(ch, ch.tab, ch.tab4, 0, tgt, x, &member);
    }
    if (lane != 0) continue;
    if (!ok) { atomicOr(err, DE_EOS_FAIL); continue; }
    double th[31];
    set_thermo<KM>(ch, ch.tab, member, x, th);
    for (int q = 0; q < 31; ++q) col(CCAKE_F_T + q) = th[q];
    eos_id[i] = member;
    if (MODE == 1) col(CCAKE_F_IN_S) = th[5];
    else if (MODE == 0 || MODE == 3) stage_post(F, npad, i, k, t, P, th[0], th[11], th[9], th[10], th[4], cand);
  }
}

// The EoS stage with the pre-inverted table (online_inverter_enabled =
// false): MODE 0 fill_thermodynamics and the fused tail, MODE 2 the lookup
// only (SPHWorkstation::update_all_particle_thermodynamics,
// src/sph_workstation.cpp:1730-1734), MODE 1 the initial e -> s inversion
// (invert_e_to_s_host and query_thermo_host, src/sph_workstation.cpp:290-305).
// Only the fields the reference's lookups write change.
template <int MODE>
__global__ __launch_bounds__(256) void k_preinv(Preinv pre, double* __restrict__ F, int npad, int n,
                                                const int32_t* __restrict__ perm, double t, Transport P,
                                                double* __restrict__ cand)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = (MODE == 0) ? perm[k] : k;
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
  if (MODE == 1) {
    const double rhoB = col(CCAKE_F_IN_B);
    const double s = preinv_e_to_s(pre, col(CCAKE_F_IN_E), rhoB);
    double o[10];
    preinv_query(pre, s, rhoB, o);
    col(CCAKE_F_T) = o[0]; col(CCAKE_F_MUB) = o[1]; col(CCAKE_F_MUS) = o[2]; col(CCAKE_F_MUQ) = o[3];
    col(CCAKE_F_E) = o[4]; col(CCAKE_F_P) = o[5]; col(CCAKE_F_CS2) = o[6];
    col(CCAKE_F_S) = s; col(CCAKE_F_RHOB) = rhoB; col(CCAKE_F_RHOS) = 0.0; col(CCAKE_F_RHOQ) = 0.0;
    col(CCAKE_F_W) = o[9];
    col(CCAKE_F_DWDS) = o[7]; col(CCAKE_F_DWDB) = o[8]; col(CCAKE_F_DWDS_S) = 0.0; col(CCAKE_F_DWDQ) = 0.0;
    col(CCAKE_F_EOS_IS_TABLE) = 1.0;
    col(CCAKE_F_IN_S) = s;
    return;
  }
  double o[12];
  preinv_fill(pre, col(CCAKE_F_S), col(CCAKE_F_RHOB), o);
  col(CCAKE_F_T) = o[0]; col(CCAKE_F_MUB) = o[1]; col(CCAKE_F_MUS) = o[2]; col(CCAKE_F_MUQ) = o[3];
  col(CCAKE_F_E) = o[4]; col(CCAKE_F_P) = o[5]; col(CCAKE_F_CS2) = o[6];
  col(CCAKE_F_DWDS) = o[7]; col(CCAKE_F_DWDB) = o[8]; col(CCAKE_F_DWDS_S) = o[9]; col(CCAKE_F_DWDQ) = o[10];
  col(CCAKE_F_W) = o[11];
  if (MODE == 2) return;
  stage_post(F, npad, i, k, t, P, o[0], o[11], o[5], o[6], o[4], cand);
}

static Transport make_transport(const ccake_params& p)
{
  Transport T;
  T.eta_mode = p.eta_mode; T.shear_relax_mode = p.shear_relax_mode; T.zeta_mode = p.zeta_mode;
  T.modulate = p.modulate_zeta_with_tanh;
  T.eta_over_s = p.constant_eta_over_s; T.zeta_over_s = p.constant_zeta_over_s;
  T.zA = p.cs2_dependent_zeta_A; T.zp = p.cs2_dependent_zeta_p;
  T.gA = p.zeta_gauss_A; T.gT0 = p.zeta_gauss_T0; T.gW = p.zeta_gauss_width;
  T.delta_pipi = p.delta_pipi_mode; T.tau_pipi = p.tau_pipi_mode; T.lambda_piPi = p.lambda_piPi_mode;
  T.phi7 = p.phi7_mode; T.delta_PiPi = p.delta_PiPi_mode; T.lambda_Pipi = p.lambda_Pipi_mode;
  return T;
}

// Workgroup size of the EoS kernel.  It runs at one wave per SIMD (its
// register footprint), so small workgroups spread the waves over every CU
// (64 measured best at 32k-40k particles, 256 and 128 slower).
#ifndef CCAKE_EOS_THREADS
#define CCAKE_EOS_THREADS 64
#endif
constexpr int kEosThreads = CCAKE_EOS_THREADS;
constexpr int kHardWaves = 512;     // waves of k_eos_hard, striding over the hard list
constexpr int kConformalChain = (1 << M_CONF) | (1 << M_DIAG);

static int chain_kinds(const EosChain& ch)
{
  int km = 0;
  for (int k = 0; k < ch.n; ++k) km |= 1 << ch.m[k].kind;
  return km;
}

int launch_eos(ccake_event* ev, int mode, int fuse_post)
{
  const int n = ev->n_live;          // sorted positions of the live events' particles
  const dim3 g((n + kEosThreads - 1) / kEosThreads), b(kEosThreads);
  hipStream_t st = ev->stream;
  const Transport P = make_transport(ev->p);
  if (ev->eos->chain.kind == CCAKE_EOS_PREINVERTED) {
    if (fuse_post == 2) {                 // the lookup kernel keeps the sum pass separate
      const int rc = launch_smooth_sum(ev, 0, true);
      if (rc) return rc;
      fuse_post = 1;
    }
    const dim3 g2((n + 255) / 256), b2(256);
    const Preinv& pr = ev->eos->pre;
    if (mode == 1)
      hipLaunchKernelGGL(k_preinv<1>, g2, b2, 0, st, pr, ev->F, ev->npad, n, ev->perm, ev->t, P, ev->cand_g);
    else if (fuse_post)
      hipLaunchKernelGGL(k_preinv<0>, g2, b2, 0, st, pr, ev->F, ev->npad, n, ev->perm, ev->t, P, ev->cand_g);
    else
      hipLaunchKernelGGL(k_preinv<2>, g2, b2, 0, st, pr, ev->F, ev->npad, n, ev->perm, ev->t, P, ev->cand_g);
    return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
  }
  const int md = (mode == 1) ? 1 : (fuse_post == 2 ? 3 : (fuse_post ? 0 : 2));
  SmoothSum sm;
  if (md == 3) {
    sm.part = ev->part;
    sm.nslices = ev->smooth_sum_slices;
    sm.norm2 = smooth_norm2(ev->p.hT);
    sm.pair_total = (unsigned long long*)ev->pair_count;
  }
  // the hard list's counters alternate between stages: this stage's k_eos
  // appends to one, its k_eos_hard clears the other for the next stage
  HardList hl;
  hl.idx = ev->hard;
  hl.count = ev->hard_count + ev->hard_parity;
  hl.next = ev->hard_count + (1 - ev->hard_parity);
  ev->hard_parity = 1 - ev->hard_parity;
  auto launch = [&](auto kern, int32_t* done) {
    hipLaunchKernelGGL(kern, g, b, 0, st, ev->eos->chain, ev->F, ev->npad, n, ev->I, ev->perm, ev->t, P,
                       ev->cand_g, ev->err, sm, hl, done);
  };
  auto hard = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(kHardWaves), dim3(kHardLanes), 0, st, ev->eos->chain, ev->F, ev->npad, n, ev->I,
                       ev->perm, ev->t, P, ev->cand_g, ev->err, hl);
  };
  // The time-step modes with the table chain run split (eos_split): the
  // Newton part, the hard list, then the thermodynamics and tail.
  const bool split = ev->eos_done && (md == 0 || md == 3);
  if (chain_kinds(ev->eos->chain) == kConformalChain) {
    if (md == 1) { launch(k_eos<1, kConformalChain, 0>, nullptr); hard(k_eos_hard<1, kConformalChain>); }
    else if (md == 0) { launch(k_eos<0, kConformalChain, 0>, nullptr); hard(k_eos_hard<0, kConformalChain>); }
    else if (md == 3) { launch(k_eos<3, kConformalChain, 0>, nullptr); hard(k_eos_hard<0, kConformalChain>); }
    else { launch(k_eos<2, kConformalChain, 0>, nullptr); hard(k_eos_hard<2, kConformalChain>); }
  } else if (split) {
    if (md == 0) launch(k_eos<0, kAllMembers, 1>, ev->eos_done);
    else launch(k_eos<3, kAllMembers, 1>, ev->eos_done);
    hard(k_eos_hard<0, kAllMembers>);
    launch(k_eos<0, kAllMembers, 2>, ev->eos_done);
  } else {
    if (md == 1) { launch(k_eos<1, kAllMembers, 0>, nullptr); hard(k_eos_hard<1, kAllMembers>); }
    else if (md == 0) { launch(k_eos<0, kAllMembers, 0>, nullptr); hard(k_eos_hard<0, kAllMembers>); }
    else if (md == 3) { launch(k_eos<3, kAllMembers, 0>, nullptr); hard(k_eos_hard<0, kAllMembers>); }
    else { launch(k_eos<2, kAllMembers, 0>, nullptr); hard(k_eos_hard<2, kAllMembers>); }
  }
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

// Stand-alone reset_pi_tensor / viscosities / gather (same code paths as the
// fused kernel, for the phase-by-phase C-ABI).
__global__ __launch_bounds__(256) void k_post(double* __restrict__ F, int npad, int n,
                                              const int32_t* __restrict__ perm, double t, Transport P,
                                              double* __restrict__ cand, int what)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm ? perm[k] : k;
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
  if (what == 0) {
    const double gamma = col(CCAKE_F_GAMMA), sigma = col(CCAKE_F_SIGMA);
    const double t2 = t * t;
    const double fc0 = -col(CCAKE_F_UX), fc1 = -col(CCAKE_F_UY);
    col(CCAKE_F_BULK) = col(CCAKE_F_EXT_BULK) * sigma;
    const double s11 = col(CCAKE_F_EXT_SHV0 + 0) * sigma, s12 = col(CCAKE_F_EXT_SHV0 + 1) * sigma;
    const double s13 = col(CCAKE_F_EXT_SHV0 + 2) * sigma, s22 = col(CCAKE_F_EXT_SHV0 + 3) * sigma;
    const double s23 = col(CCAKE_F_EXT_SHV0 + 4) * sigma;
    const double s00 = (fc0 * fc0 * s11 + fc1 * fc1 * s22 + 2. * fc1 * fc0 * s12) / (gamma * gamma);
    const double s33 = (s00 - s11 - s22) / t2;
    const double s01 = -s11 * fc0 / gamma + -s12 * fc1 / gamma;
    const double s02 = -s12 * fc0 / gamma + -s22 * fc1 / gamma;
    const double s03 = -s13 * fc0 / gamma + -s23 * fc1 / gamma;
    const double S[16] = { s00, s01, s02, s03, s01, s11, s12, s13, s02, s12, s22, s23, s03, s13, s23, s33 };
    for (int q = 0; q < 16; ++q) col(CCAKE_F_SHV0 + q) = S[q];
  } else if (what == 1) {
    double tc[13];
    transport(P, col(CCAKE_F_T), col(CCAKE_F_W), col(CCAKE_F_P), col(CCAKE_F_CS2), tc);
    for (int q = 0; q < 13; ++q) col(CCAKE_F_ETA_PI + q) = tc[q];
  } else {
    const double sl = col(CCAKE_F_SIGMA_LAB);
    const double isl2 = 1.0 / (sl * sl);
    double* o = cand + (size_t)k * kCandG;
    o[0] = col(CCAKE_F_X); o[1] = col(CCAKE_F_Y); o[2] = col(CCAKE_F_MASS_S);
    o[3] = col(CCAKE_F_P) * isl2; o[4] = col(CCAKE_F_E); o[5] = col(CCAKE_F_BULK) * isl2;
    o[6] = col(CCAKE_F_VX); o[7] = col(CCAKE_F_VY);
    o[8] = col(CCAKE_F_SHV0 + 1) * isl2; o[9] = col(CCAKE_F_SHV0 + 2) * isl2;
    o[10] = col(CCAKE_F_SHV0 + 5) * isl2; o[11] = col(CCAKE_F_SHV0 + 6) * isl2;
    o[12] = col(CCAKE_F_SHV0 + 10) * isl2;
    o[13] = sl; o[14] = 0.0; o[15] = 0.0;
  }
}

// The per-particle tail of a 3+1D EoS stage (after k_eos in its EoS-only
// mode): `what` bit 1 reset_pi_tensor (EoM_default::reset_pi_tensor,
// src/eom_default.cpp:56-172, with the tau^2 metric and u^eta), bit 2 the
// transport coefficients, bit 4 the gradient candidate record (kCandG3
// layout) at sorted position k.  With perm, thread k works on perm[k].
__global__ __launch_bounds__(256) void k3_post(double* __restrict__ F, int npad, int n,
                                               const int32_t* __restrict__ perm, double t, Transport P,
                                               double* __restrict__ cand, int what)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm ? perm[k] : k;
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
  if (what & 1) {
    const double gamma = col(CCAKE_F_GAMMA), sigma = col(CCAKE_F_SIGMA);
    const double t2 = t * t;
    const double fc[3] = { -col(CCAKE_F_UX), -col(CCAKE_F_UY), col(CCAKE_F_UETA) * -t2 };
    col(CCAKE_F_BULK) = col(CCAKE_F_EXT_BULK) * sigma;
    double S[4][4];
    for (int a = 0; a < 4; ++a) for (int b = 0; b < 4; ++b) S[a][b] = 0.0;
    int q = 0;
    for (int a = 0; a < 2; ++a) for (int b = a; b < 3; ++b) S[a + 1][b + 1] = col(CCAKE_F_EXT_SHV0 + q++) * sigma;
    for (int a = 1; a < 4; ++a) for (int b = a + 1; b < 4; ++b) S[b][a] = S[a][b];
    S[0][0] = (fc[0] * fc[0] * S[1][1] + fc[1] * fc[1] * S[2][2] + 2. * fc[1] * fc[0] * S[2][1]
               + 2. * fc[2] * fc[0] * S[3][1] + 2. * fc[2] * fc[1] * S[3][2]
               - fc[2] * fc[2] * (S[1][1] + S[2][2]) / t2) / (gamma * gamma - fc[2] * fc[2] / t2);
    S[3][3] = (S[0][0] - S[1][1] - S[2][2]) / t2;
    for (int a = 0; a < 3; ++a) {
      double p0 = 0.0;
      for (int b = 0; b < 3; ++b) p0 += -S[a + 1][b + 1] * fc[b] / gamma;
      S[0][a + 1] = p0;
    }
    for (int a = 1; a < 4; ++a) S[a][0] = S[0][a];
    for (int a = 0; a < 4; ++a) for (int b = 0; b < 4; ++b) col(CCAKE_F_SHV0 + 4 * a + b) = S[a][b];
  }
  if (what & 2) {
    double tc[13];
    transport(P, col(CCAKE_F_T), col(CCAKE_F_W), col(CCAKE_F_P), col(CCAKE_F_CS2), tc);
    for (int q = 0; q < 13; ++q) col(CCAKE_F_ETA_PI + q) = tc[q];
  }
  if (what & 4) {
    const double sl = col(CCAKE_F_SIGMA_LAB);
    const double isl2 = 1.0 / (sl * sl);
    double* o = cand + (size_t)k * kCandG3;
    o[0] = col(CCAKE_F_X); o[1] = col(CCAKE_F_Y); o[2] = col(CCAKE_F_ETA); o[3] = col(CCAKE_F_MASS_S);
    o[4] = col(CCAKE_F_P) * isl2; o[5] = col(CCAKE_F_E); o[6] = col(CCAKE_F_BULK) * isl2;
    o[7] = col(CCAKE_F_VX); o[8] = col(CCAKE_F_VY); o[9] = col(CCAKE_F_VETA);
    o[10] = col(CCAKE_F_SHV0 + 1) * isl2; o[11] = col(CCAKE_F_SHV0 + 2) * isl2; o[12] = col(CCAKE_F_SHV0 + 3) * isl2;
    o[13] = col(CCAKE_F_SHV0 + 5) * isl2; o[14] = col(CCAKE_F_SHV0 + 6) * isl2; o[15] = col(CCAKE_F_SHV0 + 7) * isl2;
    o[16] = col(CCAKE_F_SHV0 + 10) * isl2; o[17] = col(CCAKE_F_SHV0 + 11) * isl2;
    o[18] = col(CCAKE_F_SHV0 + 15) * isl2;
    o[19] = sl; o[20] = 0.0; o[21] = 0.0; o[22] = 0.0; o[23] = 0.0;
  }
}

int launch3_post(ccake_event* ev, int what)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k3_post, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n,
                     (what & 4) ? ev->perm : nullptr, ev->t, make_transport(ev->p), ev->cand_g, what);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

static int post(ccake_event* ev, int what, const int32_t* perm)
{
  if (ev->p.dim == 3) return launch3_post(ev, what == 0 ? 1 : (what == 1 ? 2 : 4));
  const int n = perm ? ev->n_live : ev->n;   // sorted positions of the live events, or every particle id
  hipLaunchKernelGGL(k_post, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n, perm, ev->t,
                     make_transport(ev->p), ev->cand_g, what);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}
int launch_reset_pi(ccake_event* ev) { return post(ev, 0, nullptr); }
int launch_viscosities(ccake_event* ev) { return post(ev, 1, nullptr); }
int launch_gather_grad(ccake_event* ev) { return post(ev, 2, ev->perm); }

int host_eos_full(const ccake_eos* eos, int member, const double* x, double* r17)
{
  if (member < 0 || member >= eos->chain.n) return CCAKE_ERR_INVALID;
  member_fields(eos->chain.m[member], eos->h_tab, x, r17);
  return 0;
}

}  // namespace ccake
