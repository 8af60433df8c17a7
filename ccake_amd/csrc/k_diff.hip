// BSQ diffusion on gfx950, D = 2; launched only when diffusion is on.
//  - k_diff_prep, per particle after the EoS stage: q^mu from the extensive
//    diffusion (EoM_default::reset_pi_tensor, src/eom_default.cpp:139-163),
//    the regulator (SPHWorkstation::regulate_diffusion,
//    src/sph_workstation.cpp:133-218), kappa_q and tau_q
//    (src/transport_coefficients.cpp:571-620, 917-1025), and the neighbour
//    record of the diffusion pair pass.
//  - k_diff_slice / k_diff_sum: the diffusion sums of the gradient pass
//    (src/sph_workstation.cpp:909-967), in the slice layout of the gradient
//    kernels.  With K_ab = m_b sl_a grad W_ab the reference sums per b
//      grad_qa_c   += (K_ab . v_a)(q0_b/sl_b^2 + q0_a/sl_a^2)
//      div_qa_c    += sum_i K_ab,i (q^i_b/sl_b^2 + q^i_a/sl_a^2)
//      grad_alpha  += (m_b/sl_a)(alpha_b - alpha_a) grad W_ab;
//    the owner-side factors are constant over b, so the pair kernel keeps
//    17 neighbour-side sums of w = m_b grad W_ab and the sum pass applies
//    the owner terms (and v_a) once.
#include "pair_common.h"

namespace ccake {

constexpr int kCandD = 16;   // x, y, m_s, q0_c/sl^2 (3), qx_c/sl^2 (3), qy_c/sl^2 (3), alpha_c (3), sl

// kappa_q and tau_q (3x3, row-major) of one particle.
__device__ __forceinline__ void diff_transport(const DiffParams& P, double T, double muB, double nB, double w,
                                               double* kq, double* tq)
{
  constexpr double kTiny = 1e-25;                       // include/eos_header.h:16
  for (int q = 0; q < 9; ++q) { kq[q] = 0.0; tq[q] = 0.0; }
  if (P.mode == 2) {                                    // kappaB_coth
    const double x = (T > kTiny) ? muB / T : 0.0;
    if (T > kTiny && fabs(w) > kTiny && fabs(x) > 1.0e-2) {
      const double kB = (P.C_B / T) * nB * ((1.0 / 3.0) / tanh(x) - nB * T / w);
      if (isfinite(kB) && kB > 0.0) kq[0] = kB;
    }
    if (P.relax == 1) {                                 // tauqB_only
      if (T > kTiny) for (int a = 0; a < 3; ++a) tq[4 * a] = P.C_B / T;
      return;
    }
  } else {                                              // default_kappa
    for (int q = 0; q < 9; ++q) kq[q] = 1.0 * P.k2[q] * (T * T);
  }
  for (int a = 0; a < 3; ++a) tq[4 * a] += 0.2 / T;     // default_tauq
}

// what: 1 rebuild q^mu, 8 the regulator (when on), 2 kappa_q / tau_q,
// 4 neighbour record (sorted position k; particle perm[k], or k without perm).
__global__ __launch_bounds__(256) void k_diff_prep(double* __restrict__ F, int npad, int n,
                                                   const int32_t* __restrict__ perm, double t, DiffParams P,
                                                   double* __restrict__ cand, int what)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm ? perm[k] : k;
  auto col = [&](int c) -> double& { return F[(size_t)c * npad + i]; };
  const double sigma = col(CCAKE_F_SIGMA);
  if (what & 1) {
    const double gamma = col(CCAKE_F_GAMMA);
    const double t2 = t * t;
    const double fc[3] = { -col(CCAKE_F_UX), -col(CCAKE_F_UY), 0.0 * -t2 };
    for (int c = 0; c < 3; ++c) {
      double qv[3];
      for (int d = 0; d < 3; ++d) qv[d] = col(CCAKE_F_EXT_Q0 + 3 * c + d) * sigma;
      double q0 = 0.0;
      for (int d = 0; d < 3; ++d) q0 += -qv[d] * fc[d] / gamma;
      col(CCAKE_F_Q0 + 4 * c) = q0;
      for (int d = 0; d < 3; ++d) col(CCAKE_F_Q0 + 4 * c + 1 + d) = qv[d];
    }
  }
  if ((what & 8) && P.reg) {
    const double t2 = t * t;
    {
      const double e = col(CCAKE_F_E);
      const double ferm = 1.0 / (exp(-(e - P.e0) / P.xi0) + 1.0);
      const double f_str = P.chi0 * (ferm - P.fermi_offset);
      const double rho[3] = { col(CCAKE_F_RHOB), col(CCAKE_F_RHOS), col(CCAKE_F_RHOQ) };
      for (int c = 0; c < 3; ++c) {
        const double n_abs = fabs(rho[c]);
        if (n_abs < 1.0e-15) continue;
        const double q0 = col(CCAKE_F_Q0 + 4 * c), q1 = col(CCAKE_F_Q0 + 4 * c + 1);
        const double q2 = col(CCAKE_F_Q0 + 4 * c + 2), q3 = col(CCAKE_F_Q0 + 4 * c + 3);
        const double q_sq = q1 * q1 + q2 * q2 + t2 * q3 * q3 - q0 * q0;
        if (q_sq <= 0.0) continue;
        const double q_mag = sqrt(q_sq);
        const double threshold = P.rq_max * n_abs * f_str;
        if (q_mag <= threshold) continue;
        const double scale = threshold / q_mag;
        col(CCAKE_F_Q0 + 4 * c) = q0 * scale; col(CCAKE_F_Q0 + 4 * c + 1) = q1 * scale;
        col(CCAKE_F_Q0 + 4 * c + 2) = q2 * scale; col(CCAKE_F_Q0 + 4 * c + 3) = q3 * scale;
        for (int d = 0; d < 3; ++d) col(CCAKE_F_EXT_Q0 + 3 * c + d) = col(CCAKE_F_Q0 + 4 * c + 1 + d) / sigma;
      }
    }
  }
  if (what & 2) {
    double kq[9], tq[9];
    diff_transport(P, col(CCAKE_F_T), col(CCAKE_F_MUB), col(CCAKE_F_RHOB), col(CCAKE_F_W), kq, tq);
    for (int q = 0; q < 9; ++q) { col(CCAKE_F_KAPPA_Q0 + q) = kq[q]; col(CCAKE_F_TAU_Q0 + q) = tq[q]; }
  }
  if (what & 4) {
    const double sl = col(CCAKE_F_SIGMA_LAB);
    const double isl2 = 1.0 / (sl * sl);
    const double T = col(CCAKE_F_T);
    const double al[3] = { col(CCAKE_F_MUB) / T, col(CCAKE_F_MUS) / T, col(CCAKE_F_MUQ) / T };
    double* o = cand + (size_t)k * kCandD;
    o[0] = col(CCAKE_F_X); o[1] = col(CCAKE_F_Y); o[2] = col(CCAKE_F_MASS_S);
    for (int c = 0; c < 3; ++c) {
      o[3 + c] = col(CCAKE_F_Q0 + 4 * c) * isl2;
      o[6 + c] = col(CCAKE_F_Q0 + 4 * c + 1) * isl2;
      o[9 + c] = col(CCAKE_F_Q0 + 4 * c + 2) * isl2;
      o[12 + c] = al[c];
    }
    o[15] = sl;
  }
}

constexpr int kDiffAcc = 17;

__global__ __launch_bounds__(kWave) void k_diff_slice(const double* __restrict__ cand,
                                                      const int32_t* __restrict__ skey,
                                                      const int32_t* __restrict__ cell_off,
                                                      const Grid* __restrict__ gridp, int n,
                                                      double rc2, double rcs, double hinv2, int nslices,
                                                      double* __restrict__ part, int npad)
{
  const Grid g0 = *gridp;
  const int lane = threadIdx.x;
  const int unit = xcd_tile(blockIdx.x, gridDim.x);
  const int tile = unit / nslices;
  const int slice = unit - tile * nslices;
  const int k = tile * kTile + lane;
  const bool act = k < n;
  const int kk = act ? k : 0;
  const double* me = cand + (size_t)kk * kCandD;
  const double xa = me[0], ya = me[1];
  const int myrow = act ? skey[kk] / g0.stride : INT_MAX;

  double acc[kDiffAcc];
  for (int q = 0; q < kDiffAcc; ++q) acc[q] = 0.0;

  for_segments(act, myrow, xa, ya, [&](const Seg& seg, bool in) {
    const Grid& g = gridp[event_index(gridp, g0, seg.row)];
    for_spans_slice(g, cell_off, seg, rcs, slice, nslices, in, [&](int j0, int j1) {
      walk_records<15, kCandD, 1>(cand, xa, ya, j0, j1, [&](const double* c, double dx, double dy, int j, bool valid) {
        const double r2 = dist2_nocontract(dx, dy);
        if (kSkipEmpty && wave_none(valid && r2 <= rc2)) return;
        double q, iq;
        q_and_inverse(r2, hinv2, q, iq);
        const double gm = valid ? c[2] * spline_g(q, iq) : 0.0;
        const double wx = gm * dx, wy = gm * dy;
        acc[0] += wx; acc[1] += wy;
        #pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          acc[2 + ch] += c[3 + ch] * wx; acc[5 + ch] += c[3 + ch] * wy;      // q0_b / sl_b^2
          acc[8 + ch] += c[6 + ch] * wx + c[9 + ch] * wy;                    // q^i_b / sl_b^2
          acc[11 + ch] += c[12 + ch] * wx; acc[14 + ch] += c[12 + ch] * wy;  // alpha_b
        }
      });
    });
  });

  double* p = part + (size_t)slice * kDiffAcc * npad + k;
  for (int q = 0; q < kDiffAcc; ++q) p[(size_t)q * npad] = acc[q];
}

template <int S>
__global__ __launch_bounds__(256) void k_diff_sum(const double* __restrict__ cand,
                                                  const int32_t* __restrict__ perm, int n, double C,
                                                  const double* __restrict__ part, int npad,
                                                  double* __restrict__ F)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double acc[kDiffAcc];
  #pragma unroll
  for (int q = 0; q < kDiffAcc; ++q) acc[q] = 0.0;
  #pragma unroll
  for (int sli = 0; sli < S; ++sli)
    #pragma unroll
    for (int q = 0; q < kDiffAcc; ++q) acc[q] += part[((size_t)sli * kDiffAcc + q) * npad + k];
  #pragma unroll
  for (int q = 0; q < kDiffAcc; ++q) acc[q] *= C;
  const double* me = cand + (size_t)k * kCandD;
  const int id = perm[k];
  auto put = [&](int c, double v) { F[(size_t)c * npad + id] = v; };
  const double vax = F[(size_t)CCAKE_F_VX * npad + id], vay = F[(size_t)CCAKE_F_VY * npad + id];
  const double sla = me[15], isla = 1.0 / sla;
  const double a1x = acc[0], a1y = acc[1];
  const double av = vax * a1x + vay * a1y;
  #pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    put(CCAKE_F_GRAD_QA0 + ch, sla * (vax * acc[2 + ch] + vay * acc[5 + ch] + me[3 + ch] * av));
    put(CCAKE_F_DIV_QA0 + ch, sla * (acc[8 + ch] + me[6 + ch] * a1x + me[9 + ch] * a1y));
    put(CCAKE_F_GRAD_ALPHA0 + ch, (acc[11 + ch] - me[12 + ch] * a1x) * isla);
    put(CCAKE_F_GRAD_ALPHA0 + 3 + ch, (acc[14 + ch] - me[12 + ch] * a1y) * isla);
  }
}

DiffParams make_diff_params(const ccake_params& p)
{
  DiffParams P;
  P.mode = p.diffusion_mode;
  P.relax = p.diffusion_relax_mode;
  for (int q = 0; q < 9; ++q) P.k2[q] = p.kappa_over_T2[q];
  P.C_B = p.C_B;
  P.reg = p.diffusion_regulator;
  const double hbarc_GeVfm = 0.1973269804;
  P.chi0 = p.regulator_chi0;
  P.e0 = p.regulator_e0 / hbarc_GeVfm;
  P.xi0 = p.regulator_xi0 / hbarc_GeVfm;
  P.rq_max = p.regulator_rq_max;
  P.fermi_offset = 1.0 / (exp(P.e0 / P.xi0) + 1.0);
  return P;
}

int launch_diff_prep(ccake_event* ev, int what, const int32_t* perm)
{
  const int n = perm ? ev->n_live : ev->n;   // sorted positions of the live events, or every particle id
  hipLaunchKernelGGL(k_diff_prep, dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->F, ev->npad, n, perm, ev->t,
                     make_diff_params(ev->p), ev->cand_d, what);
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

template <int S = 1>
static int diff_sum_dispatch(ccake_event* ev, double C)
{
  if constexpr (S > kMaxSlices) {
    set_error("slice count out of range");
    return CCAKE_ERR_INVALID;
  } else {
    if (ev->nslices != S) return diff_sum_dispatch<S + 1>(ev, C);
    const int n = ev->n_live;
    hipLaunchKernelGGL((k_diff_sum<S>), dim3((n + 255) / 256), dim3(256), 0, ev->stream, ev->cand_d, ev->perm, n, C,
                       ev->part, ev->npad, ev->F);
    return 0;
  }
}

int launch_diff_grad(ccake_event* ev)
{
  const int n = ev->n_live;          // sorted positions of the live events' particles
  const double h = ev->p.hT;
  const double rc = 2.0 * h;
  const int ntiles = (n + kTile - 1) / kTile;
  hipLaunchKernelGGL(k_diff_slice, dim3(ntiles * ev->nslices), dim3(kWave), 0, ev->stream, ev->cand_d, ev->skey,
                     ev->cell_off, ev->grid, n, rc * rc, search_radius(ev), 1.0 / (h * h), ev->nslices, ev->part,
                     ev->npad);
  if (hipGetLastError() != hipSuccess) return CCAKE_ERR_HIP;
  const double C = -0.5 * (15.0 / (7.0 * kPi)) / pow(h, 4);
  const int rc2 = diff_sum_dispatch(ev, C);
  if (rc2) return rc2;
  return hipGetLastError() == hipSuccess ? 0 : CCAKE_ERR_HIP;
}

}  // namespace ccake
