// Initialisation, freeze-out status and reductions.
#include "ccake_internal.h"

namespace ccake {

// phase 0: SPHWorkstation::process_initial_conditions (src/sph_workstation.cpp:1228-1340)
// phase 1: freeze flag after the e inversion (src/sph_workstation.cpp:283-287 / 314-318)
// phase 2: init_particles (src/sph_workstation.cpp:338-364): sph_mass = 1, extensive = input/sigma
// phase 3: calculate_extensive_shv (1392-1432) + set_bulk_Pi (1349-1384)
__global__ __launch_bounds__(256) void k_init(double* __restrict__ F, int npad, int n, int32_t* __restrict__ I,
                                              int phase, double t0, double efcheck, int input_shear,
                                              int bulk_from_trace)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
  int32_t* freeze = I + (size_t)CCAKE_I_FREEZE * npad;
  if (phase == 0) {
    const double ux = col(CCAKE_F_UX), uy = col(CCAKE_F_UY);
    col(CCAKE_F_GAMMA) = sqrt(1.0 - ((-ux) * ux + (-uy) * uy));
    col(CCAKE_F_EXT_S) = col(CCAKE_F_EXT_B) = col(CCAKE_F_EXT_SS) = col(CCAKE_F_EXT_Q) = 1.0;
    col(CCAKE_F_MASS_S) = col(CCAKE_F_MASS_B) = col(CCAKE_F_MASS_SS) = col(CCAKE_F_MASS_Q) = 1.0;
    col(CCAKE_F_T) = 580.0 / kHbarcMeVfm;
    col(CCAKE_F_MUB) = col(CCAKE_F_MUS) = col(CCAKE_F_MUQ) = 0.0;
    I[(size_t)CCAKE_I_EOS_ID * npad + i] = 0;
    I[(size_t)CCAKE_I_BTRACK * npad + i] = 0;
    freeze[i] = 0;
    col(CCAKE_F_ETA) = 0.0;
    col(CCAKE_F_UETA) = 0.0;
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
  }
}

int launch_init_particles(ccake_event* ev, int phase)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k_init, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n, ev->I, phase,
                     ev->p.t0, ev->p.e_freeze, ev->p.input_initial_shear, ev->p.bulk_from_trace);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

// FreezeOut::check_freeze_out_status (include/freeze_out.h:455-542) and the
// status side of bsqsvfreezeout (549-626): snapshot energies / gradP_x that
// drive the machine, the no-neighbour freeze of freeze_out_evo (322-374),
// and 3 -> 4 at interpolation.  phase -1: status check only (initial
// smoothing); 0, 1, 2: the frzc stage of the step.
__global__ __launch_bounds__(256) void k_freeze(double* __restrict__ F, int npad, int n, int32_t* __restrict__ I,
                                                double efcheck, int phase)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
  int32_t& fz = I[(size_t)CCAKE_I_FREEZE * npad + i];
  int32_t& bt = I[(size_t)CCAKE_I_BTRACK * npad + i];
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
  if (phase == 0) { col(CCAKE_F_FRZ2_E) = e; col(P_FRZ2_G) = g0; return; }
  if (phase == 1) {
    col(CCAKE_F_FRZ1_E) = e; col(P_FRZ1_G) = g0;
    if (fz == 3) fz = 4;
    return;
  }
  // freeze_out_evo: snapshot shifts for few-neighbour particles
  if (fz < 4 && bt <= 3 && bt > 0) {
    col(P_FB_E0 + 3) = col(P_FB_E0 + 1); col(P_FB_G0 + 3) = col(P_FB_G0 + 1);
    col(P_FB_E0 + 2) = col(P_FB_E0 + 0); col(P_FB_G0 + 2) = col(P_FB_G0 + 0);
    col(P_FB_E0 + 1) = col(CCAKE_F_FRZ2_E); col(P_FB_G0 + 1) = col(P_FRZ2_G);
    col(P_FB_E0 + 0) = col(CCAKE_F_FRZ1_E); col(P_FB_G0 + 0) = col(P_FRZ1_G);
  }
  if (fz < 4 && bt == 0 && col(P_FB_G0 + 0) != 0) {
    col(CCAKE_F_FRZ2_E) = col(P_FB_E0 + 1); col(P_FRZ2_G) = col(P_FB_G0 + 1);
    col(CCAKE_F_FRZ1_E) = col(P_FB_E0 + 0); col(P_FRZ1_G) = col(P_FB_G0 + 0);
  }
  if (fz < 4 && bt == 0 && col(P_FB_G0 + 0) == 0) {
    col(CCAKE_F_FRZ2_E) = col(P_FB_E0 + 3); col(P_FRZ2_G) = col(P_FB_G0 + 3);
    col(CCAKE_F_FRZ1_E) = col(P_FB_E0 + 2); col(P_FRZ1_G) = col(P_FB_G0 + 2);
  }
  if (fz < 4 && bt == 0) { fz = 4; bt = -1; }
  if (fz == 3) fz = 4;
  if (fz < 4) { col(CCAKE_F_FRZ2_E) = col(CCAKE_F_FRZ1_E); col(P_FRZ2_G) = col(P_FRZ1_G); }
  col(CCAKE_F_FRZ1_E) = e; col(P_FRZ1_G) = g0;
}

int launch_freeze(ccake_event* ev, int phase)
{
  const int n = ev->n;
  hipLaunchKernelGGL(k_freeze, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n, ev->I,
                     ev->p.e_freeze, phase);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

// Deterministic two-pass reductions (SystemState::conservation_*,
// src/system_state.cpp:738-885; frozen count, src/sph_workstation.cpp:2011-2017).
__global__ __launch_bounds__(256) void k_reduce(const double* __restrict__ F, int npad, int n,
                                                const int32_t* __restrict__ I, int which, double t,
                                                double* __restrict__ part)
{
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    auto col = [&](int c) { return F[(size_t)c * npad + i]; };
    switch (which) {
      case CCAKE_R_ENTROPY: acc += col(CCAKE_F_EXT_S) * col(CCAKE_F_MASS_S); break;
      case CCAKE_R_BARYON: acc += col(CCAKE_F_EXT_B) * col(CCAKE_F_MASS_B); break;
      case CCAKE_R_STRANGE: acc += col(CCAKE_F_EXT_SS) * col(CCAKE_F_MASS_SS); break;
      case CCAKE_R_CHARGE: acc += col(CCAKE_F_EXT_Q) * col(CCAKE_F_MASS_Q); break;
      case CCAKE_R_ENERGY: {
        const double w = col(CCAKE_F_W), p = col(CCAKE_F_P), bulk = col(CCAKE_F_BULK), g = col(CCAKE_F_GAMMA);
        acc += ((w + bulk) * g * g - p - bulk + col(CCAKE_F_SHV0)) * col(CCAKE_F_MASS_S) * t / col(CCAKE_F_SIGMA_LAB);
        break;
      }
      case CCAKE_R_EZ: acc += col(CCAKE_F_EZ); break;
      case CCAKE_R_FROZEN: acc += (I[(size_t)CCAKE_I_FREEZE * npad + i] >= 4) ? 1.0 : 0.0; break;
      default: break;
    }
  }
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

}  // namespace ccake
