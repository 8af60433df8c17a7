// Initialisation, freeze-out status and reductions.
#include "ccake_internal.h"

namespace ccake {

// phase 0: SPHWorkstation::process_initial_conditions (src/sph_workstation.cpp:1228-1340)
// phase 1: freeze flag after the e inversion (src/sph_workstation.cpp:283-287 / 314-318)
// phase 2: init_particles (src/sph_workstation.cpp:338-364): sph_mass = 1, extensive = input/sigma
// phase 3: calculate_extensive_shv (1392-1432) + set_bulk_Pi (1349-1384)
__global__ __launch_bounds__(256) void k_init(double* __restrict__ F, int npad, int n, int32_t* __restrict__ I,
                                              int phase, double t0, double efcheck, int input_shear,
                                              int bulk_from_trace, int diffusion, int dim)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
  int32_t* freeze = I + (size_t)CCAKE_I_FREEZE * npad;
  if (phase == 0) {
    const double ux = col(CCAKE_F_UX), uy = col(CCAKE_F_UY);
    if (dim == 3) {
      // gamma_calc(u, t0^2): u_cov = (-u^x, -u^y, -t0^2 u^eta)
      const double ue = col(CCAKE_F_UETA);
      double c = 0.0;
      c += (-ux) * ux;
      c += (-uy) * uy;
      c += (ue * -(t0 * t0)) * ue;
      col(CCAKE_F_GAMMA) = sqrt(1.0 - c);
    } else {
      col(CCAKE_F_GAMMA) = sqrt(1.0 - ((-ux) * ux + (-uy) * uy));
    }
    col(CCAKE_F_EXT_S) = col(CCAKE_F_EXT_B) = col(CCAKE_F_EXT_SS) = col(CCAKE_F_EXT_Q) = 1.0;
    col(CCAKE_F_MASS_S) = col(CCAKE_F_MASS_B) = col(CCAKE_F_MASS_SS) = col(CCAKE_F_MASS_Q) = 1.0;
    col(CCAKE_F_T) = 580.0 / kHbarcMeVfm;
    col(CCAKE_F_MUB) = col(CCAKE_F_MUS) = col(CCAKE_F_MUQ) = 0.0;
    I[(size_t)CCAKE_I_EOS_ID * npad + i] = 0;
    I[(size_t)CCAKE_I_BTRACK * npad + i] = 0;
    freeze[i] = 0;
    if (dim != 3) { col(CCAKE_F_ETA) = 0.0; col(CCAKE_F_UETA) = 0.0; }
    col(CCAKE_F_EZ) = 0.0;
    col(CCAKE_F_FRZ1_E) = col(CCAKE_F_FRZ2_E) = 0.0;
    for (int q = 0; q < 4; ++q) { col(P_FB_E0 + q) = 0.0; col(P_FB_G0 + q) = 0.0; }
    col(P_FRZ1_G) = col(P_FRZ2_G) = 0.0;
    for (int q = 0; q < 5; ++q) col(CCAKE_F_DEXT_SHV0 + q) = 0.0;
  } else if (phase == 1) {
    if (col(CCAKE_F_IN_E) > efcheck && freeze[i] != 4) freeze[i] = 0;
    else freeze[i] = 4;
  } else if (phase == 2) {
    const double sig = col(CCAKE_F_SIGMA);
    col(CCAKE_F_MASS_S) = col(CCAKE_F_MASS_B) = col(CCAKE_F_MASS_SS) = col(CCAKE_F_MASS_Q) = 1.0;
    col(CCAKE_F_EXT_S) = col(CCAKE_F_IN_S) / sig;
    col(CCAKE_F_EXT_B) = col(CCAKE_F_IN_B) / sig;
    col(CCAKE_F_EXT_SS) = col(CCAKE_F_IN_SS) / sig;
    col(CCAKE_F_EXT_Q) = col(CCAKE_F_IN_Q) / sig;
    if (col(CCAKE_F_IN_S) < 0.0) freeze[i] = 4;
  } else {
    const double sig = col(CCAKE_F_SIGMA);
    int q = 0;
    for (int a = 0; a < 2; ++a)
      for (int b = a; b < 3; ++b, ++q)
        col(CCAKE_F_EXT_SHV0 + q) = input_shear ? col(CCAKE_F_SHV0 + 4 * (a + 1) + (b + 1)) / sig : 0.0;
    if (!input_shear) for (int c = 0; c < 16; ++c) col(CCAKE_F_SHV0 + c) = 0.0;
    if (bulk_from_trace) {
      const double sl = col(CCAKE_F_SIGMA_LAB), g = col(CCAKE_F_GAMMA);
      col(CCAKE_F_EXT_BULK) = (col(CCAKE_F_BULK) - col(CCAKE_F_P)) / (sl / g / t0);
    } else {
      col(CCAKE_F_EXT_BULK) = 0.0;
      col(CCAKE_F_BULK) = 0.0;
    }
    // set_diffusion (src/sph_workstation.cpp:1441-1459): extensive q^i = q^i / sigma
    if (diffusion)
      for (int d = 0; d < 3; ++d)
        for (int c = 0; c < 3; ++c) col(CCAKE_F_EXT_Q0 + 3 * c + d) = col(CCAKE_F_Q0 + 4 * c + 1 + d) / sig;
  }
}

int launch_init_particles(ccake_event* ev, int phase)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k_init, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n, ev->I, phase,
                     ev->p.t0, ev->p.e_freeze, ev->p.input_initial_shear, ev->p.bulk_from_trace,
                     ev->p.using_diffusion, ev->p.dim);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

// ---- freeze-out hypersurface -------------------------------------------
// Snapshot of one particle (FreezeOut's freeze_array members for D = 2,
// include/freeze_out.h:29-35): the F column of each of the kSnap entries,
// then, with diffusion on, the kSnapDiff components of q^mu for B, S, Q
// (the diffusion members stay zero while diffusion is off and are not stored).
constexpr int kSnap = 33;
constexpr int kSnapDiff = 12;
enum SnapIdx {
  SN_SHEAR = 0,                 // 3x3 pi^{ab}, a, b = 0..2, row-major
  SN_R = 9, SN_U = 11, SN_GRADP = 13, SN_GRADE = 15,
  SN_S = 17, SN_E, SN_RHOB, SN_RHOS, SN_RHOQ, SN_T, SN_MUB, SN_MUS, SN_MUQ,
  SN_THETA, SN_BULK, SN_SIGMA_LAB, SN_SHEAR33, SN_SHV_NABLA_U, SN_W, SN_CS2
};
__device__ __forceinline__ int snap_col(int q)
{
  switch (q) {
    case 0: return CCAKE_F_SHV0 + 0;  case 1: return CCAKE_F_SHV0 + 1;  case 2: return CCAKE_F_SHV0 + 2;
    case 3: return CCAKE_F_SHV0 + 4;  case 4: return CCAKE_F_SHV0 + 5;  case 5: return CCAKE_F_SHV0 + 6;
    case 6: return CCAKE_F_SHV0 + 8;  case 7: return CCAKE_F_SHV0 + 9;  case 8: return CCAKE_F_SHV0 + 10;
    case 9: return CCAKE_F_X;         case 10: return CCAKE_F_Y;
    case 11: return CCAKE_F_UX;       case 12: return CCAKE_F_UY;
    case 13: return CCAKE_F_GRADP0;   case 14: return CCAKE_F_GRADP0 + 1;
    case 15: return CCAKE_F_GRADE0;   case 16: return CCAKE_F_GRADE0 + 1;
    case 17: return CCAKE_F_S;        case 18: return CCAKE_F_E;
    case 19: return CCAKE_F_RHOB;     case 20: return CCAKE_F_RHOS;     case 21: return CCAKE_F_RHOQ;
    case 22: return CCAKE_F_T;        case 23: return CCAKE_F_MUB;      case 24: return CCAKE_F_MUS;
    case 25: return CCAKE_F_MUQ;      case 26: return CCAKE_F_THETA;    case 27: return CCAKE_F_BULK;
    case 28: return CCAKE_F_SIGMA_LAB; case 29: return CCAKE_F_SHV0 + 15;
    case 30: return CCAKE_F_SHV_NABLA_U; case 31: return CCAKE_F_W;
    case 32: return CCAKE_F_CS2;
    default: return CCAKE_F_Q0 + (q - kSnap);
  }
}
// Snapshot sets: frz1, frz2, fback, fback2, fback3, fback4.
enum SnapSet { SS_FRZ1 = 0, SS_FRZ2, SS_FB1, SS_FB2, SS_FB3, SS_FB4, SS_NUM };

struct Snaps {
  double* p; int npad; int i; int nc;    // nc: kSnap, or kSnap + kSnapDiff with diffusion
  __device__ __forceinline__ double& at(int set, int q) const
  {
    return p[((size_t)set * nc + q) * npad + i];
  }
  __device__ __forceinline__ void copy(int dst, int src) const
  {
    #pragma unroll
    for (int q = 0; q < kSnap; ++q) at(dst, q) = at(src, q);
    if (nc > kSnap)
      #pragma unroll
      for (int q = kSnap; q < kSnap + kSnapDiff; ++q) at(dst, q) = at(src, q);
  }
  __device__ __forceinline__ void take(int dst, const double* F) const
  {
    #pragma unroll
    for (int q = 0; q < kSnap; ++q) at(dst, q) = F[(size_t)snap_col(q) * npad + i];
    if (nc > kSnap)
      #pragma unroll
      for (int q = kSnap; q < kSnap + kSnapDiff; ++q) at(dst, q) = F[(size_t)snap_col(q) * npad + i];
  }
};

// FreezeOut::bsqsvinterpolate for one particle with status 3
// (include/freeze_out.h:638-930, D = 2, hyperbolic coordinates): pick the
// snapshot whose energy is closer to the freeze-out energy, derive the
// surface quantities, and append the row in the column order of
// Output<2, EoM_default>::print_freeze_out (src/output.cpp:502-560), after
// the particle id and the step number.
__device__ void interpolate_row(const Snaps& sn, const double* __restrict__ F, int npad, int i, int bt, double efc,
                                double taup, double taupp, double dt, int step, double* __restrict__ rows,
                                int* __restrict__ count, int cap)
{
  constexpr double kHbarcGeVfm = 0.1973269804;        // include/constants.h:15
  const double d1 = fabs(sn.at(SS_FRZ1, SN_E) - efc);
  const double d2 = fabs(sn.at(SS_FRZ2, SN_E) - efc);
  const int set = (d1 < d2) ? SS_FRZ1 : SS_FRZ2;
  const double tau_s = (set == SS_FRZ1) ? taup : taupp;
  const double tlist = (bt != -1) ? tau_s : tau_s - dt;
  double u[2], gP[2], gE[2];
  for (int a = 0; a < 2; ++a) {
    u[a] = sn.at(set, SN_U + a); gP[a] = sn.at(set, SN_GRADP + a); gE[a] = sn.at(set, SN_GRADE + a);
  }
  // D = 2 Milne vectors: lowering and raising flip the sign (include/milne.hpp:353-364)
  const double u_cov[2] = { -u[0], -u[1] };
  const double gP_con[2] = { -gP[0], -gP[1] };
  const double gE_con[2] = { -gE[0], -gE[1] };
  const double norm2u = u[0] * u_cov[0] + u[1] * u_cov[1];
  const double gsub = sqrt(-norm2u + 1.0);
  const double sigsub = sn.at(set, SN_SIGMA_LAB) / (gsub * tlist);
  const double swsub = F[(size_t)CCAKE_F_MASS_S * npad + i] / sigsub;
  double sFO = sn.at(set, SN_S);
  double divT[2], divE[2], divT_con[2];
  for (int a = 0; a < 2; ++a) {
    divT[a] = (1.0 / sFO) * gP[a];
    divE[a] = gE[a];
    divT_con[a] = (1.0 / sFO) * gP_con[a];
  }
  const double cs2 = sn.at(set, SN_CS2), w = sn.at(set, SN_W);
  const double bulk = sn.at(set, SN_BULK), theta = sn.at(set, SN_THETA);
  const double snu = sn.at(set, SN_SHV_NABLA_U);
  double inner = u[0] * gP[0] + u[1] * gP[1];
  const double divPpress = -(1.0 / gsub) * (cs2 * (w + bulk) * theta - cs2 * snu + inner);
  double divTtemp = divPpress / sFO;
  inner = u[0] * gE[0] + u[1] * gE[1];
  double divEener = -(1.0 / gsub) * ((w + bulk) * theta - snu + inner);
  double n2 = divT[0] * divT_con[0] + divT[1] * divT_con[1];
  const double insub = divTtemp * divTtemp + n2;
  double norm = -sqrt(fabs(insub));
  norm = fabs(norm) > 1.E-14 ? norm : 1.E-14;
  divTtemp /= norm;
  divT[0] /= norm; divT[1] /= norm;
  n2 = divE[0] * gE_con[0] + divE[1] * gE_con[1];
  const double insubE = divEener * divEener + n2;
  double normE = -sqrt(fabs(insubE));
  normE = fabs(normE) > 1.E-14 ? normE : 1.E-14;
  divEener /= normE;
  divE[0] /= normE; divE[1] /= normE;
  const double Tfluc0 = sn.at(set, SN_T);
  const double Tk = Tfluc0 * kHbarcGeVfm;
  sFO *= Tk * Tk * Tk;
  const int slot = atomicAdd(count, 1);
  if (slot >= cap) return;
  double* r = rows + (size_t)slot * CCAKE_FO_ROW;
  r[0] = (double)i; r[1] = (double)step;
  double* c = r + 2;
  c[0] = divEener; c[1] = divE[0]; c[2] = divE[1];
  c[3] = gsub; c[4] = u[0]; c[5] = u[1];
  c[6] = swsub; c[7] = bulk;
  c[8] = sn.at(set, SN_SHEAR + 0); c[9] = sn.at(set, SN_SHEAR + 4); c[10] = sn.at(set, SN_SHEAR + 8);
  c[11] = sn.at(set, SN_SHEAR33); c[12] = sn.at(set, SN_SHEAR + 5);
  c[13] = tlist; c[14] = sn.at(set, SN_R); c[15] = sn.at(set, SN_R + 1);
  c[16] = sFO; c[17] = sn.at(set, SN_E); c[18] = Tk;
  c[19] = sn.at(set, SN_MUB); c[20] = sn.at(set, SN_MUS); c[21] = sn.at(set, SN_MUQ);
  c[22] = w; c[23] = cs2;
  c[24] = sn.at(set, SN_RHOB); c[25] = sn.at(set, SN_RHOS); c[26] = sn.at(set, SN_RHOQ);
  for (int q = 0; q < kSnapDiff; ++q)                 // diffout(charge, mu); zero with diffusion off
    c[27 + q] = sn.nc > kSnap ? sn.at(set, kSnap + q) : 0.0;
  c[39] = F[(size_t)CCAKE_F_EOS_IS_TABLE * npad + i];
}

// FreezeOut::check_freeze_out_status (include/freeze_out.h:455-542) and
// bsqsvfreezeout (549-636): the status machine, the snapshot shifts of
// freeze_out_evo (322-374) including the no-neighbour freeze, and the
// hypersurface rows of particles at status 3 (bsqsvinterpolate), which then
// move to 4.  The energy / gradP_x snapshots that drive the status machine
// live in F; with `snaps` the full snapshots follow the same shifts.
// phase -1: status check only (initial smoothing); 0, 1, 2: the frzc stage.
__global__ __launch_bounds__(256) void k_freeze(double* __restrict__ F, int npad, int n, int32_t* __restrict__ I,
                                                double efcheck, int phase, double* __restrict__ snaps, int snap_cols,
                                                double* __restrict__ rows, int* __restrict__ count, int cap,
                                                double taup, double taupp, double dt, int step)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
  int32_t& fz = I[(size_t)CCAKE_I_FREEZE * npad + i];
  int32_t& bt = I[(size_t)CCAKE_I_BTRACK * npad + i];
  const bool full = snaps != nullptr;
  const Snaps sn{ snaps, npad, i, snap_cols };
  const double e = col(CCAKE_F_E);
  switch (fz) {
    case 0: if (e <= efcheck) fz = 1; break;
    case 1:
      if (bt == -1) fz = 3;
      else if (e > col(CCAKE_F_FRZ1_E)) fz = 1;
      else if (e <= efcheck) fz = 3;
      else fz = 0;
      break;
    default: break;
  }
  if (phase < 0) return;
  const double g0 = col(CCAKE_F_GRADP0);
  if (phase == 0) {
    col(CCAKE_F_FRZ2_E) = e; col(P_FRZ2_G) = g0;
    if (full) sn.take(SS_FRZ2, F);
    return;
  }
  if (phase == 1) {
    col(CCAKE_F_FRZ1_E) = e; col(P_FRZ1_G) = g0;
    if (full) sn.take(SS_FRZ1, F);
    if (fz == 3) {
      if (full) interpolate_row(sn, F, npad, i, bt, efcheck, taup, taupp, dt, step, rows, count, cap);
      fz = 4;
    }
    return;
  }
  // freeze_out_evo: snapshot shifts for few-neighbour particles
  if (fz < 4 && bt <= 3 && bt > 0) {
    col(P_FB_E0 + 3) = col(P_FB_E0 + 1); col(P_FB_G0 + 3) = col(P_FB_G0 + 1);
    col(P_FB_E0 + 2) = col(P_FB_E0 + 0); col(P_FB_G0 + 2) = col(P_FB_G0 + 0);
    col(P_FB_E0 + 1) = col(CCAKE_F_FRZ2_E); col(P_FB_G0 + 1) = col(P_FRZ2_G);
    col(P_FB_E0 + 0) = col(CCAKE_F_FRZ1_E); col(P_FB_G0 + 0) = col(P_FRZ1_G);
    if (full) { sn.copy(SS_FB4, SS_FB2); sn.copy(SS_FB3, SS_FB1); sn.copy(SS_FB2, SS_FRZ2); sn.copy(SS_FB1, SS_FRZ1); }
  }
  if (fz < 4 && bt == 0 && col(P_FB_G0 + 0) != 0) {
    col(CCAKE_F_FRZ2_E) = col(P_FB_E0 + 1); col(P_FRZ2_G) = col(P_FB_G0 + 1);
    col(CCAKE_F_FRZ1_E) = col(P_FB_E0 + 0); col(P_FRZ1_G) = col(P_FB_G0 + 0);
    if (full) { sn.copy(SS_FRZ2, SS_FB2); sn.copy(SS_FRZ1, SS_FB1); }
  }
  if (fz < 4 && bt == 0 && col(P_FB_G0 + 0) == 0) {
    col(CCAKE_F_FRZ2_E) = col(P_FB_E0 + 3); col(P_FRZ2_G) = col(P_FB_G0 + 3);
    col(CCAKE_F_FRZ1_E) = col(P_FB_E0 + 2); col(P_FRZ1_G) = col(P_FB_G0 + 2);
    if (full) { sn.copy(SS_FRZ2, SS_FB4); sn.copy(SS_FRZ1, SS_FB3); }
  }
  if (fz < 4 && bt == 0) { fz = 4; bt = -1; }
  if (fz == 3) {
    if (full) interpolate_row(sn, F, npad, i, bt, efcheck, taup, taupp, dt, step, rows, count, cap);
    fz = 4;
  }
  if (fz < 4) {
    col(CCAKE_F_FRZ2_E) = col(CCAKE_F_FRZ1_E); col(P_FRZ2_G) = col(P_FRZ1_G);
    if (full) sn.copy(SS_FRZ2, SS_FRZ1);
  }
  col(CCAKE_F_FRZ1_E) = e; col(P_FRZ1_G) = g0;
  if (full) sn.take(SS_FRZ1, F);
}

int launch_freeze(ccake_event* ev, int phase)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k_freeze, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n, ev->I,
                     ev->p.e_freeze, phase, ev->snaps, ev->snap_cols, ev->fo_rows, ev->fo_count, n, ev->taup, ev->taupp,
                     ev->p.dt, ev->fo_step);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

// Deterministic two-pass reductions (SystemState::conservation_*,
// src/system_state.cpp:738-885; frozen count, src/sph_workstation.cpp:2011-2017).
// One particle's term of reduction `which`.
__device__ __forceinline__ double reduce_term(const double* __restrict__ F, int npad, const int32_t* __restrict__ I,
                                              int which, double t, int i)
{
  auto col = [&](int c) { return F[(size_t)c * npad + i]; };
  switch (which) {
    case CCAKE_R_ENTROPY: return col(CCAKE_F_EXT_S) * col(CCAKE_F_MASS_S);
    case CCAKE_R_BARYON: return col(CCAKE_F_EXT_B) * col(CCAKE_F_MASS_B);
    case CCAKE_R_STRANGE: return col(CCAKE_F_EXT_SS) * col(CCAKE_F_MASS_SS);
    case CCAKE_R_CHARGE: return col(CCAKE_F_EXT_Q) * col(CCAKE_F_MASS_Q);
    case CCAKE_R_ENERGY: {
      const double w = col(CCAKE_F_W), p = col(CCAKE_F_P), bulk = col(CCAKE_F_BULK), g = col(CCAKE_F_GAMMA);
      return ((w + bulk) * g * g - p - bulk + col(CCAKE_F_SHV0)) * col(CCAKE_F_MASS_S) * t / col(CCAKE_F_SIGMA_LAB);
    }
    case CCAKE_R_EZ: return col(CCAKE_F_EZ);
    case CCAKE_R_FROZEN: return (I[(size_t)CCAKE_I_FREEZE * npad + i] >= 4) ? 1.0 : 0.0;
    default: return 0.0;
  }
}

__global__ __launch_bounds__(256) void k_reduce(const double* __restrict__ F, int npad, int n,
                                                const int32_t* __restrict__ I, int which, double t,
                                                double* __restrict__ part)
{
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    acc += reduce_term(F, npad, I, which, t, i);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  __shared__ double s[4];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ __launch_bounds__(256) void k_reduce_final(double* __restrict__ part, int np)
{
  double acc = 0.0;
  for (int k = threadIdx.x; k < np; k += blockDim.x) acc += part[k];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  __shared__ double s[4];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[np] = s[0] + s[1] + s[2] + s[3];
}

// A batch: workgroup e reduces the ids [off[e], off[e+1]) of event e.
__global__ __launch_bounds__(256) void k_reduce_events(const double* __restrict__ F, int npad,
                                                       const int32_t* __restrict__ I, const int32_t* __restrict__ off,
                                                       int which, double t, double* __restrict__ out)
{
  const int e = blockIdx.x;
  double acc = 0.0;
  for (int i = off[e] + threadIdx.x; i < off[e + 1]; i += blockDim.x) acc += reduce_term(F, npad, I, which, t, i);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  __shared__ double s[4];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[e] = s[0] + s[1] + s[2] + s[3];
}

constexpr int kRedBlocks = 512;

int launch_reduce(ccake_event* ev, int which, double* out)
{
  hipStream_t st = ev->stream;
  hipLaunchKernelGGL(k_reduce, dim3(kRedBlocks), dim3(256), 0, st, ev->F, ev->npad, ev->n, ev->I, which, ev->t, ev->red);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, ev->red, kRedBlocks);
  CCAKE_HIP(hipGetLastError());
  CCAKE_HIP(hipMemcpyAsync(out, ev->red + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipStreamSynchronize(st));
  return 0;
}

int launch_reduce_events(ccake_event* ev, int which, double* out)
{
  hipStream_t st = ev->stream;
  hipLaunchKernelGGL(k_reduce_events, dim3(ev->nev), dim3(256), 0, st, ev->F, ev->npad, ev->I, ev->ev_off, which,
                     ev->t, ev->red_ev);
  CCAKE_HIP(hipGetLastError());
  CCAKE_HIP(hipMemcpyAsync(out, ev->red_ev, ev->nev * sizeof(double), hipMemcpyDeviceToHost, st));
  CCAKE_HIP(hipStreamSynchronize(st));
  return 0;
}

}  // namespace ccake
