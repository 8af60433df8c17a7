/*
 * ccake_oracle.h -- CPU restatement of CCAKE's per-timestep SPH path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the parity checker and the timed
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker -- never as the
 * thing measured or shipped.
 *
 * Plain C99 + OpenMP, FP64, written from the reference's algorithm
 * (the-nuclear-confectionery/CCAKE); each function cites the file:line it
 * follows.  Field numbering mirrors the product's field ids so that a test
 * can compare field k of both sides directly.
 *
 * Pinned by tests/test_oracle.py against the reference's known answers,
 * brute force, closed forms, finite differences and the analytic ideal
 * Gubser solution, and by tests/test_gubser_viscous.py against the
 * reference's semi-analytic Israel-Stewart Gubser profiles (shear terms,
 * Gubser-mode tau_pi).  BSQ diffusion is pinned by tests/test_diffusion.py
 * (at-rest Israel-Stewart relaxation, continuity and transverse-momentum
 * conservation of a uniform moving fluid, SPH gradient identities, the
 * regulator cap, transport special values).  tests/test_oracle_pins.py pins
 * the bulk-pressure terms (0+1D Bjorken interior against an ODE solve), the
 * six gradient sums (linear profiles) and the fourth-order step (a closed
 * form); the d alpha / d(s, rho) fields and the Knudsen numbers with flow
 * are pinned in tests/test_oracle.py.  Parity unpinned:
 *   - which root the reference's GSL hybrids solver returns where several
 *     exist (the inversion here is Newton with the same absolute stop);
 *   - diffusion MRF terms multiplied by the geometric factor or u^eta (zero
 *     for D = 2).
 */
#ifndef CCAKE_ORACLE_H
#define CCAKE_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum orc_field {
  ORC_F_X = 0, ORC_F_Y, ORC_F_ETA,
  ORC_F_UX, ORC_F_UY, ORC_F_UETA,
  ORC_F_GAMMA, ORC_F_VX, ORC_F_VY, ORC_F_VETA,
  ORC_F_MASS_S, ORC_F_MASS_B, ORC_F_MASS_SS, ORC_F_MASS_Q,
  ORC_F_EXT_S, ORC_F_EXT_B, ORC_F_EXT_SS, ORC_F_EXT_Q,
  ORC_F_IN_E, ORC_F_IN_S, ORC_F_IN_B, ORC_F_IN_SS, ORC_F_IN_Q,
  ORC_F_SM_S, ORC_F_SM_B, ORC_F_SM_SS, ORC_F_SM_Q,
  ORC_F_SIGMA_LAB, ORC_F_SIGMA,
  ORC_F_T, ORC_F_MUB, ORC_F_MUS, ORC_F_MUQ,
  ORC_F_E, ORC_F_S, ORC_F_RHOB, ORC_F_RHOS, ORC_F_RHOQ,
  ORC_F_P, ORC_F_CS2, ORC_F_W,
  ORC_F_DWDS, ORC_F_DWDB, ORC_F_DWDS_S, ORC_F_DWDQ,
  ORC_F_DALPHA0,
  ORC_F_CHIBB = ORC_F_DALPHA0 + 12, ORC_F_CHIBB0, ORC_F_EOS_IS_TABLE,
  ORC_F_BULK, ORC_F_EXT_BULK,
  ORC_F_EXT_SHV0,
  ORC_F_SHV0 = ORC_F_EXT_SHV0 + 5,
  ORC_F_ETA_PI = ORC_F_SHV0 + 16, ORC_F_TAU_PI, ORC_F_ZETA_PI,
  ORC_F_TAU_BULK, ORC_F_DELTA_BB, ORC_F_DELTA_PIPI, ORC_F_LAMBDA_PIB,
  ORC_F_LAMBDA_BPI, ORC_F_PHI1, ORC_F_PHI3, ORC_F_PHI6, ORC_F_PHI7,
  ORC_F_TAU_PIPI,
  ORC_F_GRADP0, ORC_F_GRADE0 = ORC_F_GRADP0 + 3,
  ORC_F_GRADBULK0 = ORC_F_GRADE0 + 3, ORC_F_GRADSHEAR0 = ORC_F_GRADBULK0 + 3,
  ORC_F_DIVSHEAR0 = ORC_F_GRADSHEAR0 + 3, ORC_F_GRADV0 = ORC_F_DIVSHEAR0 + 3,
  ORC_F_DUDT0 = ORC_F_GRADV0 + 9,
  ORC_F_DEXT_S = ORC_F_DUDT0 + 3, ORC_F_DEXT_B, ORC_F_DEXT_SS, ORC_F_DEXT_Q,
  ORC_F_DEXT_BULK, ORC_F_DEXT_SHV0,
  ORC_F_THETA = ORC_F_DEXT_SHV0 + 5, ORC_F_SHV_NABLA_U,
  ORC_F_EZ, ORC_F_DEZ,
  ORC_F_FRZ1_E, ORC_F_FRZ2_E,
  ORC_F_SHEAR_KNUDSEN, ORC_F_BULK_KNUDSEN, ORC_F_INV_REYNOLDS_SHEAR, ORC_F_INV_REYNOLDS_BULK,
  /* BSQ diffusion: q^mu of B, S, Q (4 each); extensive q^i and its time derivative
     (3 per charge); kappa_q and tau_q (3x3); the gradient-pass sums grad q^0, div q
     (per charge) and grad alpha (direction-major, 2 x 3); Knudsen and inverse
     Reynolds numbers per charge */
  ORC_F_Q0,
  ORC_F_EXT_Q0 = ORC_F_Q0 + 12,
  ORC_F_DEXT_Q0 = ORC_F_EXT_Q0 + 9,
  ORC_F_KAPPA_Q0 = ORC_F_DEXT_Q0 + 9,
  ORC_F_TAU_Q0 = ORC_F_KAPPA_Q0 + 9,
  ORC_F_GRAD_QA0 = ORC_F_TAU_Q0 + 9,
  ORC_F_DIV_QA0 = ORC_F_GRAD_QA0 + 3,
  ORC_F_GRAD_ALPHA0 = ORC_F_DIV_QA0 + 3,
  ORC_F_KN_DIFF0 = ORC_F_GRAD_ALPHA0 + 6,
  ORC_F_RE_DIFF0 = ORC_F_KN_DIFF0 + 3,
  ORC_NUM_FIELDS = ORC_F_RE_DIFF0 + 3
};

enum orc_ifield { ORC_I_FREEZE = 0, ORC_I_BTRACK, ORC_I_EOS_ID, ORC_NUM_IFIELDS };

enum { ORC_EOS_CONFORMAL = 0, ORC_EOS_TABLE = 1, ORC_EOS_PREINVERTED = 2 };
/* ORC_EOS_PREINVERTED: the flat (s, nB) pre-inverted table.  dims[0] = Ns,
 * dims[1] = NB; table = 12 arrays of Ns*NB doubles (s slowest): s, nB, T,
 * muB, muS, muQ, e, p, cs2, chiTT, chiTB, chiBB, in the MeV units of the
 * reference's HDF5 file.  orc_eos_create returns NULL for a table it rejects. */

/* Same layout as the product's ccake_params. */
typedef struct orc_params {
  int    dim;
  int    rk_order;
  double hT, hEta;
  double t0;
  double dt;
  double e_freeze;
  int    particlization;
  int    using_shear;
  int    using_vorticity;
  int    input_initial_shear;
  int    bulk_from_trace;
  int    prohibit_unstable_cs2, prohibit_acausal_cs2;
  int    eta_mode;            double constant_eta_over_s;
  int    shear_relax_mode;
  int    zeta_mode;           double constant_zeta_over_s;
  double cs2_dependent_zeta_A, cs2_dependent_zeta_p;
  int    modulate_zeta_with_tanh;
  double zeta_gauss_A, zeta_gauss_T0, zeta_gauss_width;
  int    delta_pipi_mode, tau_pipi_mode, lambda_piPi_mode, phi6_mode, phi7_mode;
  int    delta_PiPi_mode, lambda_Pipi_mode, phi1_mode, phi3_mode;
  int    cells_per_rc;
  int    using_diffusion;
  int    diffusion_mode;
  int    diffusion_relax_mode;
  double kappa_over_T2[9];
  double C_B;
  int    diffusion_regulator;
  double regulator_chi0, regulator_e0, regulator_xi0, regulator_rq_max;
  int    pair_precision;      /* the product's switch; the oracle is FP64 and ignores it */
  int    factorized_kernel;   /* D = 3: W_2D(r_perp, hT) W_1D(|d eta|, hEta) (settings.h:114) */
  int    x_subcells;          /* the product's neighbour-search tuning; the oracle ignores it */
} orc_params;

typedef struct orc_eos_desc {
  int    kind;
  double T_freeze_out;
  int    dims[4];
  double mins[4], maxs[4];
  const double* table;
} orc_eos_desc;

typedef struct orc_eos orc_eos;
typedef struct orc_event orc_event;

/* status codes match the product's */
enum { ORC_OK = 0, ORC_ERR_INVALID = 1, ORC_ERR_NEG_ENTROPY = 4,
       ORC_ERR_EOS_NO_SOLUTION = 5, ORC_ERR_EOS_OUT_OF_RANGE = 6 };

/* kernel (src/kernel.cpp) */
double orc_kernel(int D, double r, double h);
void   orc_grad_kernel(int D, const double* rel, double r, double h, double* grad);

/* EoS */
orc_eos* orc_eos_create(const orc_eos_desc* desc);
void     orc_eos_destroy(orc_eos* e);
int      orc_eos_members(orc_eos* e);
void     orc_eos_full_thermo(orc_eos* e, int member, const double tbqs[4], double out[17]);
double   orc_eos_efreeze(orc_eos* e, double T_fo);
/* pre-inverted table point queries: which = 0 fill_thermodynamics
 * (out: T, muB, muS, muQ, e, p, cs2, dwds, dwdB, dwdS, dwdQ, w), 1
 * query_thermo_host (T, muB, muS, muQ, e, p, cs2, dwds, dwdB, w), 2
 * invert_e_to_s_host (out[0] = s); (a, b) = (s or e, rhoB). */
void     orc_preinv_point(orc_eos* e, int which, double a, double b, double out[12]);
/* invert one point: mode 0 -> target (s,B,S,Q), mode 1 -> (e,B,S,Q).
 * seed/out tbqs = (T, muB, muQ, muS); *member in: start member, out: member used.
 * thermo31: the 31 thermodynamic_info doubles. Returns 1 on success. */
int      orc_eos_invert_point(orc_eos* e, int mode, const double target[4],
                              double tbqs[4], int* member, double thermo31[31]);

/* events */
orc_event* orc_event_create(const orc_params* p, orc_eos* eos, int n);
void     orc_event_destroy(orc_event* ev);
double*  orc_field(orc_event* ev, int f);
int32_t* orc_ifield(orc_event* ev, int f);
void     orc_set_time(orc_event* ev, double t);
double   orc_get_time(orc_event* ev);

int      orc_build_neighbors(orc_event* ev, int brute_force);
int64_t  orc_pair_count(orc_event* ev);
int      orc_export_pairs(orc_event* ev, int32_t* pi, int32_t* pj, int64_t cap);
void     orc_gamma_velocity(orc_event* ev);
void     orc_smooth_fields(orc_event* ev);
void     orc_initial_sigma(orc_event* ev);
int      orc_eos_invert(orc_event* ev, int mode);
void     orc_reset_pi_tensor(orc_event* ev);
void     orc_viscosities(orc_event* ev);
void     orc_smooth_gradients(orc_event* ev);
void     orc_eom(orc_event* ev);
int      orc_get_time_derivatives(orc_event* ev);
void     orc_cache_state(orc_event* ev);
int      orc_rk_update(orc_event* ev, double dt, double t0);
int      orc_advance_timestep(orc_event* ev, double dt, int rk_order);
int      orc_initialize(orc_event* ev);
void     orc_freeze_out_particles(orc_event* ev);
/* Freeze-out hypersurface rows accumulated since orc_initialize (events with
 * particlization): ORC_FO_ROW doubles each -- particle id, step number, then
 * the 40 columns of Output<2,EoM_default>::print_freeze_out
 * (src/output.cpp:502-560) -- in the order the reference writes them (step,
 * then particle index).  Returns the row count; copies min(count, cap). */
#define ORC_FO_ROW 42
int      orc_freeze_out_surface(orc_event* ev, double* rows, int cap);
int      orc_continue_evolution(orc_event* ev, double max_tau);
double   orc_reduce(orc_event* ev, int which);
/* thread count used by the OpenMP loops (0 = runtime default) */
void     orc_set_threads(int n);
int      orc_get_threads(void);
/* Timing variants of the CPU baseline: the EoS inversion on one thread (the
 * reference's serial host loop), and wall time per phase (index k = the
 * reference's phase 0k of get_time_derivatives; 11 the RK cache and
 * updates; 12 the freeze-out step).                                       */
#define ORC_NUM_PHASES 13
void     orc_set_serial_eos(int on);
void     orc_phase_reset(void);
void     orc_phase_read(double* seconds, long long* calls);

#ifdef __cplusplus
}
#endif
#endif
