/*
 * ccake_amd.h -- C-ABI of the MI355X-native SPH time-step library.
 *
 * Everything crossing this boundary is plain C: opaque handles, plain
 * pointers, sizes and int status codes.  No C++ or device types.  The
 * library owns every device buffer and the EoS tables; the caller owns
 * host arrays.  Every event owns a HIP stream (default priority): the calls
 * on one event are ordered on it and asynchronous except *_upload,
 * *_download, *_reduce, *_check, *_sync and the query calls.  Different
 * events may be driven from different host threads at once (their work
 * overlaps on the device); one event is driven by one thread at a time.
 *
 * Each entry point names the reference interface it replaces
 * (the-nuclear-confectionery/CCAKE, file:line).  Units are the reference's
 * internal units: lengths in fm, energy/pressure densities in fm^-4,
 * entropy and charge densities in fm^-3, T and mu in fm^-1.
 */
#ifndef CCAKE_AMD_H
#define CCAKE_AMD_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define CCAKE_API __attribute__((visibility("default")))
#else
#define CCAKE_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes.  The library never exits or aborts; the C++ host layer
 * maps these to the reference behaviour:
 *   CCAKE_ERR_NEG_ENTROPY     -> exit(EXIT_FAILURE)   src/evolver.cpp:437-450
 *   CCAKE_ERR_EOS_NO_SOLUTION -> exit(404)            src/sph_workstation.cpp:1752-1760
 *   CCAKE_ERR_EOS_OUT_OF_RANGE-> exit(8)              src/eos.cpp:154-159          */
enum ccake_status {
  CCAKE_OK = 0,
  CCAKE_ERR_INVALID = 1,          /* bad argument / handle                      */
  CCAKE_ERR_OOM = 2,              /* HBM budget or hipMalloc failure            */
  CCAKE_ERR_HIP = 3,              /* HIP runtime error                          */
  CCAKE_ERR_NEG_ENTROPY = 4,      /* s~ < 0 on a non-frozen particle            */
  CCAKE_ERR_EOS_NO_SOLUTION = 5,  /* no EoS in the chain inverted the point     */
  CCAKE_ERR_EOS_OUT_OF_RANGE = 6, /* seed point outside every EoS range         */
  CCAKE_ERR_CAPACITY = 7,         /* output buffer too small (export_pairs)     */
  CCAKE_ERR_UNSUPPORTED = 8       /* configuration not built in this library    */
};

/* ------------------------------------------------------------------ */
/* Per-particle fields addressable by upload/download.  Vector fields use
 * three consecutive ids (x, y, eta); D x D matrices nine (row major);
 * pi^{mu nu} sixteen (row major, index 0 = tau).  The names follow the
 * reference data model: include/particle.h:54-136,
 * include/hydrodynamic_info.h:23-310, include/thermodynamic_info.h:18-70. */
enum ccake_field {
  CCAKE_F_X = 0, CCAKE_F_Y, CCAKE_F_ETA,
  CCAKE_F_UX, CCAKE_F_UY, CCAKE_F_UETA,
  CCAKE_F_GAMMA, CCAKE_F_VX, CCAKE_F_VY, CCAKE_F_VETA,
  CCAKE_F_MASS_S, CCAKE_F_MASS_B, CCAKE_F_MASS_SS, CCAKE_F_MASS_Q,   /* sph_mass   */
  CCAKE_F_EXT_S, CCAKE_F_EXT_B, CCAKE_F_EXT_SS, CCAKE_F_EXT_Q,       /* extensive  */
  CCAKE_F_IN_E, CCAKE_F_IN_S, CCAKE_F_IN_B, CCAKE_F_IN_SS, CCAKE_F_IN_Q, /* input */
  CCAKE_F_SM_S, CCAKE_F_SM_B, CCAKE_F_SM_SS, CCAKE_F_SM_Q,           /* smoothed   */
  CCAKE_F_SIGMA_LAB, CCAKE_F_SIGMA,
  /* thermodynamic_info, in the reference's thermo_info order */
  CCAKE_F_T, CCAKE_F_MUB, CCAKE_F_MUS, CCAKE_F_MUQ,
  CCAKE_F_E, CCAKE_F_S, CCAKE_F_RHOB, CCAKE_F_RHOS, CCAKE_F_RHOQ,
  CCAKE_F_P, CCAKE_F_CS2, CCAKE_F_W,
  CCAKE_F_DWDS, CCAKE_F_DWDB, CCAKE_F_DWDS_S, CCAKE_F_DWDQ,
  CCAKE_F_DALPHA0,                       /* 12 dalpha_{B,S,Q}d{s,B,S,Q}      */
  CCAKE_F_CHIBB = CCAKE_F_DALPHA0 + 12, CCAKE_F_CHIBB0, CCAKE_F_EOS_IS_TABLE,
  /* viscous state */
  CCAKE_F_BULK, CCAKE_F_EXT_BULK,
  CCAKE_F_EXT_SHV0,                      /* 5: (11,12,13,22,23) extensive pi */
  CCAKE_F_SHV0 = CCAKE_F_EXT_SHV0 + 5,   /* 16: pi^{mu nu}                   */
  /* transport coefficients (src/sph_workstation.cpp:1631-1697 order) */
  CCAKE_F_ETA_PI = CCAKE_F_SHV0 + 16, CCAKE_F_TAU_PI, CCAKE_F_ZETA_PI,
  CCAKE_F_TAU_BULK, CCAKE_F_DELTA_BB, CCAKE_F_DELTA_PIPI, CCAKE_F_LAMBDA_PIB,
  CCAKE_F_LAMBDA_BPI, CCAKE_F_PHI1, CCAKE_F_PHI3, CCAKE_F_PHI6, CCAKE_F_PHI7,
  CCAKE_F_TAU_PIPI,
  /* gradients (src/sph_workstation.cpp:796-986) */
  CCAKE_F_GRADP0, CCAKE_F_GRADE0 = CCAKE_F_GRADP0 + 3,
  CCAKE_F_GRADBULK0 = CCAKE_F_GRADE0 + 3, CCAKE_F_GRADSHEAR0 = CCAKE_F_GRADBULK0 + 3,
  CCAKE_F_DIVSHEAR0 = CCAKE_F_GRADSHEAR0 + 3, CCAKE_F_GRADV0 = CCAKE_F_DIVSHEAR0 + 3,
  /* time derivatives and EoM outputs (src/eom_default.cpp:211-869) */
  CCAKE_F_DUDT0 = CCAKE_F_GRADV0 + 9,
  CCAKE_F_DEXT_S = CCAKE_F_DUDT0 + 3, CCAKE_F_DEXT_B, CCAKE_F_DEXT_SS, CCAKE_F_DEXT_Q,
  CCAKE_F_DEXT_BULK, CCAKE_F_DEXT_SHV0,
  CCAKE_F_THETA = CCAKE_F_DEXT_SHV0 + 5, CCAKE_F_SHV_NABLA_U,
  CCAKE_F_EZ, CCAKE_F_DEZ,               /* contribution_to_total_Ez / dEz   */
  CCAKE_F_FRZ1_E, CCAKE_F_FRZ2_E,        /* freeze-out energy snapshots      */
  /* Knudsen and inverse Reynolds numbers, written by the EoM when shear is on
     (EoM_default::compute_hydro_numbers, src/eom_default.cpp:1831-1995) */
  CCAKE_F_SHEAR_KNUDSEN, CCAKE_F_BULK_KNUDSEN, CCAKE_F_INV_REYNOLDS_SHEAR, CCAKE_F_INV_REYNOLDS_BULK,
  /* BSQ diffusion: q^mu of B, S, Q (4 each); extensive q^i and its time derivative
     (3 per charge); kappa_q and tau_q (3x3); the gradient-pass sums grad q^0, div q
     (per charge) and grad alpha (direction-major, 2 x 3); Knudsen and inverse
     Reynolds numbers per charge */
  CCAKE_F_Q0,
  CCAKE_F_EXT_Q0 = CCAKE_F_Q0 + 12,
  CCAKE_F_DEXT_Q0 = CCAKE_F_EXT_Q0 + 9,
  CCAKE_F_KAPPA_Q0 = CCAKE_F_DEXT_Q0 + 9,
  CCAKE_F_TAU_Q0 = CCAKE_F_KAPPA_Q0 + 9,
  CCAKE_F_GRAD_QA0 = CCAKE_F_TAU_Q0 + 9,
  CCAKE_F_DIV_QA0 = CCAKE_F_GRAD_QA0 + 3,
  CCAKE_F_GRAD_ALPHA0 = CCAKE_F_DIV_QA0 + 3,
  CCAKE_F_KN_DIFF0 = CCAKE_F_GRAD_ALPHA0 + 6,
  CCAKE_F_RE_DIFF0 = CCAKE_F_KN_DIFF0 + 3,
  CCAKE_NUM_FIELDS = CCAKE_F_RE_DIFF0 + 3
};

/* Integer per-particle fields (upload_int / download_int). */
enum ccake_ifield {
  CCAKE_I_FREEZE = 0,   /* Freeze status 0,1,3,4 (include/freeze_out.h:455-542) */
  CCAKE_I_BTRACK,       /* neighbour count, sticky -1 (src/system_state.cpp:712-715) */
  CCAKE_I_EOS_ID,       /* index of the EoS in the chain this particle last used */
  CCAKE_NUM_IFIELDS
};

/* EoS kinds; a chain is built as in src/eos_initialization.cpp:216-501. */
enum ccake_eos_kind {
  CCAKE_EOS_CONFORMAL = 0,   /* chain: conformal -> conformal_diagonal            */
  CCAKE_EOS_TABLE = 1,       /* chain: table -> tanh_conformal -> conformal -> conformal_diagonal */
  CCAKE_EOS_PREINVERTED = 2  /* pre-inverted flat (s, nB) table, online_inverter_enabled = false
                                (EoS_Interpolator, include/eos_interpolator.h) */
};

/* Transport-coefficient modes (src/transport_coefficients.cpp:9-283). */
enum ccake_eta_mode   { CCAKE_ETA_DEFAULT = 0, CCAKE_ETA_CONSTANT = 1 };
enum ccake_taupi_mode { CCAKE_TAUPI_DEFAULT = 0, CCAKE_TAUPI_MINVAL = 1, CCAKE_TAUPI_GUBSER = 2 };
enum ccake_zeta_mode  { CCAKE_ZETA_DEFAULT = 0, CCAKE_ZETA_CONSTANT = 1,
                        CCAKE_ZETA_CS2_DEPENDENT = 2,
                        CCAKE_ZETA_GAUSSIAN = 3 /* zeta/s = A exp(-(T-T0)^2/(2 w^2)); not in the reference */ };
enum ccake_onoff_mode { CCAKE_MODE_DEFAULT = 0, CCAKE_MODE_DISABLED = 1, CCAKE_MODE_IS = 2 };
/* diffusion_mode / diffusion_relax_mode (src/transport_coefficients.cpp:82-115) */
enum ccake_diff_mode  { CCAKE_DIFF_KAPPA_DEFAULT = 0, CCAKE_DIFF_KAPPA_DNMR = 1, CCAKE_DIFF_KAPPA_B_COTH = 2 };
enum ccake_tauq_mode  { CCAKE_TAUQ_DEFAULT = 0, CCAKE_TAUQ_CB_OVER_T = 1 };

/* Settings consumed by the time-step path (a POD restatement of the fields
 * of include/settings.h the path reads). */
typedef struct ccake_params {
  int    dim;                 /* spatial dimension D; 2 is built            */
  int    rk_order;            /* 2 or 4 (the reference's fourth-order step)   */
  double hT, hEta;            /* smoothing lengths, fm                       */
  double t0;                  /* tau_0, fm/c                                 */
  double dt;                  /* fm/c                                        */
  double e_freeze;            /* efcheck, fm^-4                              */
  int    particlization;      /* freeze-out status machine on/off            */
  int    using_shear;         /* shear evolution (eta/s > 1e-3)              */
  int    using_vorticity;
  int    input_initial_shear;
  int    bulk_from_trace;
  int    prohibit_unstable_cs2, prohibit_acausal_cs2;
  /* transport */
  int    eta_mode;            double constant_eta_over_s;
  int    shear_relax_mode;
  int    zeta_mode;           double constant_zeta_over_s;
  double cs2_dependent_zeta_A, cs2_dependent_zeta_p;
  int    modulate_zeta_with_tanh;
  double zeta_gauss_A, zeta_gauss_T0, zeta_gauss_width;   /* fm^-1 */
  int    delta_pipi_mode, tau_pipi_mode, lambda_piPi_mode, phi6_mode, phi7_mode;
  int    delta_PiPi_mode, lambda_Pipi_mode, phi1_mode, phi3_mode;
  /* neighbour search tuning: cells per interaction radius (1 = reference) */
  int    cells_per_rc;
  /* BSQ diffusion (transport: src/transport_coefficients.cpp:571-620, 917-1059;
     regulator: src/sph_workstation.cpp:108-218, PhysRevC.98.034916 Eqs. C6-C8) */
  int    using_diffusion;
  int    diffusion_mode;          /* 0 kappa = (kappa/T^2)_ij T^2, 2 baryon coth form (1, DNMR: unsupported) */
  int    diffusion_relax_mode;    /* 0 tau_q = 0.2/T, 1 tau_q = C_B/T (with mode 2) */
  double kappa_over_T2[9];        /* 3x3 row-major, charges (B, S, Q) */
  double C_B;
  int    diffusion_regulator;     /* regulator on/off */
  double regulator_chi0, regulator_e0, regulator_xi0, regulator_rq_max;   /* e0, xi0 in GeV/fm^3 */
  /* Precision of the time step's smoothing and gradient pair sums (no
     counterpart in the reference, which is FP64 throughout):
     0 FP64 (the reference's arithmetic);
     1 FP32 pair terms with compensated FP32 accumulation.  Positions, the
       pair predicate (so the pair set), the slices' sums and everything
       outside the pair loops (EoS inversion, EoM, RK accumulation, shear
       update, initialisation, diffusion sums) stay FP64. */
  int    pair_precision;
  /* D = 3 only: the factorized kernel W_2D(r_perp, hT) W_1D(|d eta|, hEta)
     (SPHkernel<3>::kernel / gradKernel, src/kernel.cpp:129-170; selected by
     Settings::factorized_kernel, include/settings.h:114); 0 is the
     isotropic 3D spline in (x, y, eta) with hT. */
  int    factorized_kernel;
  /* neighbour search tuning: x cells per cell of a key row (the x cell edge
     is rc / cells_per_rc / x_subcells; 0 or 1 = square cells).  Finer x
     cells cut each row's candidate run closer to the cutoff disc without
     adding rows to walk; the pair set does not depend on it. */
  int    x_subcells;
} ccake_params;

/* Description of an equation of state.  For CCAKE_EOS_TABLE the table is a
 * uniform 4D grid over (T, muB, muQ, muS), T slowest, 17 fields per node in
 * the order p, s, B, S, Q, e, cs2, chiBB, chiQQ, chiSS, chiBQ, chiBS, chiQS,
 * chiTB, chiTQ, chiTS, chiTT (src/eos_table.cpp:30-34), already converted to
 * fm units (no 1/T^n normalisation).
 * For CCAKE_EOS_PREINVERTED (EoS_Interpolator::load_flat_table,
 * src/eos_interpolator.cpp:291-527) dims[0] = Ns, dims[1] = NB and the table
 * is the content of the reference's flat HDF5 file: 12 arrays of Ns*NB
 * doubles (s slowest) in the order s, nB, T, muB, muS, muQ, e, p, cs2, chiTT,
 * chiTB, chiBB, in its MeV units (MeV^3, MeV, MeV^4, MeV^2; cs2 plain).  The
 * s and nB axes (row 0 / column 0 of the s and nB arrays) must be positive
 * and uniform in log10 to 1e-6; mins, maxs and dims[2..3] are not read.
 * Per step the lookup replaces the Newton solve; initialisation inverts e
 * by bisection on the table; ccake_eos_efreeze reads T(s, smallest nB). */
typedef struct ccake_eos_desc {
  int    kind;
  double T_freeze_out;        /* fm^-1; sets the fallback matching scale     */
  int    dims[4];             /* nodes per axis (T, muB, muQ, muS)            */
  double mins[4], maxs[4];    /* axis ranges, fm^-1                           */
  const double* table;        /* dims[0]*dims[1]*dims[2]*dims[3]*17 doubles   */
} ccake_eos_desc;

typedef struct ccake_dev   ccake_dev;
typedef struct ccake_eos   ccake_eos;
typedef struct ccake_event ccake_event;

/* ---------------------------- lifetime ---------------------------- */
/* Device context: one HIP device and an HBM budget (0 = all) that event and
 * EoS creation is admitted against (CCAKE_ERR_OOM beyond it).
 * Replaces the Kokkos DeviceType/ExecutionSpace typedefs
 * (include/utilities.h:42-55). */
CCAKE_API int  ccake_dev_create(int gpu, size_t hbm_budget_bytes, ccake_dev** out);
/* Handles are destroyed children first: events, then EoS objects, then the
 * device; destroying a parent that still has live children is refused
 * (CCAKE_ERR_INVALID) and leaves everything valid. */
CCAKE_API int  ccake_dev_destroy(ccake_dev* dev);
/* Waits for the work of every event on the device. */
CCAKE_API int  ccake_dev_sync(ccake_dev* dev);
/* Bytes of HBM the library holds on this device, and its budget. */
CCAKE_API int  ccake_dev_memory(ccake_dev* dev, size_t* used, size_t* budget);
CCAKE_API const char* ccake_last_error(void);

/* EoS chain on the device.  Replaces EquationOfState::init /
 * set_up_chosen_EOSs (src/eos_initialization.cpp:216-501). */
CCAKE_API int  ccake_eos_create(ccake_dev* dev, const ccake_eos_desc* desc, ccake_eos** out);
CCAKE_API int  ccake_eos_destroy(ccake_eos* eos);
/* e(T_FO, mu=0) of the chain's default EoS: EquationOfState::efreeze
 * (src/eos.cpp, used at src/sph_workstation.cpp:71). */
CCAKE_API int  ccake_eos_efreeze(ccake_eos* eos, double T_freeze_out, double* e_out);
/* Host evaluation of the 17 forward fields of chain member `member` at
 * (T, muB, muQ, muS): EoS_base::get_full_thermo (include/eos_base.h). */
CCAKE_API int  ccake_eos_full_thermo(ccake_eos* eos, int member, const double tbqs[4], double out17[17]);

/* One event (one SystemState).  Replaces SystemState::allocate_cabana_particles
 * (src/system_state.cpp:57-63) and Evolver::allocate_cache (include/evolver.h). */
CCAKE_API int  ccake_event_create(ccake_dev* dev, const ccake_params* params, ccake_eos* eos,
                        int n_particles, ccake_event** out);
CCAKE_API int  ccake_event_destroy(ccake_event* ev);
CCAKE_API int  ccake_event_size(ccake_event* ev, int* n_particles);
CCAKE_API int  ccake_event_set_time(ccake_event* ev, double t);        /* SystemState::t */
CCAKE_API int  ccake_event_get_time(ccake_event* ev, double* t);

/* Host <-> device copies of one field: SystemState::copy_host_to_device /
 * copy_device_to_host (src/system_state.cpp:71-628), field by field. */
CCAKE_API int  ccake_event_upload(ccake_event* ev, int field, const double* host, size_t n);
CCAKE_API int  ccake_event_download(ccake_event* ev, int field, double* host, size_t n);
CCAKE_API int  ccake_event_upload_int(ccake_event* ev, int ifield, const int32_t* host, size_t n);
CCAKE_API int  ccake_event_download_int(ccake_event* ev, int ifield, int32_t* host, size_t n);

/* ---------------------------- phases ------------------------------ */
/* 01: SystemState::reset_neighbour_list (src/system_state.cpp:650-719).
 * Cell sort of the current positions; btrack = neighbour count. */
CCAKE_API int  ccake_build_neighbors(ccake_event* ev);
/* Materialise the pair set (i, j), i != j, |r_i - r_j|^2 <= (2 hT)^2, rows
 * ordered by i then j ascending (parity checks only; never on the hot path).
 * In a batch, the pairs of the live events (an ended event has none).
 * Replaces reading Cabana::NeighborList (include/utilities.h:57-62). */
CCAKE_API int  ccake_export_pairs(ccake_event* ev, int32_t* pi, int32_t* pj, size_t cap, size_t* n_pairs);
/* ccake_export_pairs without any allocation, host or device: no counterpart
 * in the reference.  For callers that cannot allocate memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it).  The
 * device-side counts, row offsets and rows live in a static device array of
 * 128 MiB per device (held until the process ends), so the event may hold
 * at most 1,048,576 particles and 15,204,320 pairs; larger ones return
 * CCAKE_ERR_CAPACITY with a message.  Same results as ccake_export_pairs;
 * calls that overlap in time on one device race on that array. */
CCAKE_API int  ccake_export_pairs_noalloc(ccake_event* ev, int32_t* pi, int32_t* pj, size_t cap, size_t* n_pairs);
/* 02: EoM_default::update_velocity (src/eom_default.cpp:17-38). */
CCAKE_API int  ccake_gamma_velocity(ccake_event* ev);
/* 03: SPHWorkstation::smooth_all_particle_fields (src/sph_workstation.cpp:494-578). */
CCAKE_API int  ccake_smooth_fields(ccake_event* ev);
/* SPHWorkstation::calculate_intial_sigma (src/sph_workstation.cpp:582-640). */
CCAKE_API int  ccake_initial_sigma(ccake_event* ev);
/* 04: SPHWorkstation::update_all_particle_thermodynamics (1713-1772):
 * mode 0 inverts (s, rhoB, rhoS, rhoQ) from the thermo fields;
 * mode 1 inverts (e, rhoB, rhoS, rhoQ) from the input fields and writes
 * input.s (locate_phase_diagram_point_eBSQ, 1791-1851). */
CCAKE_API int  ccake_eos_invert(ccake_event* ev, int mode);
/* 05: EoM_default::reset_pi_tensor (src/eom_default.cpp:56-172). */
CCAKE_API int  ccake_reset_pi_tensor(ccake_event* ev);
/* 08: SPHWorkstation::update_all_particle_viscosities (1622-1699). */
CCAKE_API int  ccake_viscosities(ccake_event* ev);
/* 09: SPHWorkstation::smooth_all_particle_gradients (796-986). */
CCAKE_API int  ccake_smooth_gradients(ccake_event* ev);
/* 10: EoM_default::evaluate_time_derivatives (src/eom_default.cpp:211-869). */
CCAKE_API int  ccake_eom(ccake_event* ev);
/* SPHWorkstation::get_time_derivatives (src/sph_workstation.cpp:1509-1612):
 * phases 01-10 in order. */
CCAKE_API int  ccake_get_time_derivatives(ccake_event* ev);
/* Evolver::set_current_timestep_quantities (src/evolver.cpp:124-168). */
CCAKE_API int  ccake_cache_state(ccake_event* ev);
/* Evolver::step_rk update part (src/evolver.cpp:363-484): y = y0 + dt*f(y),
 * t = t0 + dt.  Flags a negative extensive entropy on a non-frozen
 * particle; the status is returned by ccake_event_check. */
CCAKE_API int  ccake_rk_update(ccake_event* ev, double dt, double t0);
/* SPHWorkstation::advance_timestep (src/sph_workstation.cpp:2042-2073) with
 * Evolver::advance_timestep_rk2 (src/evolver.cpp:493-517), or with rk_order
 * 4 advance_timestep_rk4 (519-561), and, when
 * particlization is on, freeze_out_particles (1474-1485). */
CCAKE_API int  ccake_advance_timestep(ccake_event* ev, double dt, int rk_order);
/* Evolver::allocate_k_values, update_k(n) and update_rk4(dt)
 * (src/evolver.cpp:171-341, 519-561): the pieces of the reference's
 * fourth-order step for a caller that drives it phase by phase.
 * update_k(n) adds w_n (1/6, 2/6, 2/6, 1/6 for n = 1..4) times the last
 * derivatives to k; update_rk4 sets y = y0 + dt k from the RK cache.
 * ccake_advance_timestep(ev, dt, 4) runs the whole step. */
CCAKE_API int  ccake_allocate_k_values(ccake_event* ev);
CCAKE_API int  ccake_update_k(ccake_event* ev, int n);
CCAKE_API int  ccake_update_rk4(ccake_event* ev, double dt);
/* Run `n_steps` timesteps; stops early when continue_evolution is false. */
CCAKE_API int  ccake_run_steps(ccake_event* ev, double dt, int rk_order, int n_steps, int* steps_done);
/* The initialisation phases, in the order BSQHydro::initialize_hydrodynamics
 * calls them (src/BSQHydro.cpp:576-602), after upload of positions, u, input
 * e and charges, pi and Pi:
 * SPHWorkstation::process_initial_conditions (src/sph_workstation.cpp:1228-1340):
 * gamma from u, extensive densities and SPH masses 1, the Newton seed
 * (580 MeV, mu = 0), Freeze and btrack 0; the event's time to t0. */
CCAKE_API int  ccake_process_initial_conditions(ccake_event* ev);
/* SPHWorkstation::initialize_entropy_and_charge_densities (252-375): the
 * e -> s inversion of the inputs (freezing particles below e_freeze, and
 * those without a solution at their charges), calculate_intial_sigma and
 * the extensive densities input / sigma.  Needs the neighbour build. */
CCAKE_API int  ccake_initialize_entropy_and_charge_densities(ccake_event* ev);
/* SPHWorkstation::initial_smoothing (394-457): gamma and v, the smoothing
 * pass, calculate_extensive_shv, set_bulk_Pi, set_diffusion,
 * reset_pi_tensor, update_all_particle_thermodynamics and the first
 * check_freeze_out_status. */
CCAKE_API int  ccake_initial_smoothing(ccake_event* ev);
/* All of initialisation in one call: process_initial_conditions,
 * the neighbour build, initialize_entropy_and_charge_densities,
 * calculate_gamma_and_velocities and initial_smoothing
 * (BSQHydro::initialize_hydrodynamics, src/BSQHydro.cpp:568-646). */
CCAKE_API int  ccake_initialize(ccake_event* ev);
/* FreezeOut::check_freeze_out_status + status transitions of bsqsvfreezeout
 * (include/freeze_out.h:322-542, 549-626). */
CCAKE_API int  ccake_freeze_out_particles(ccake_event* ev);
/* SPHWorkstation::continue_evolution (src/sph_workstation.cpp:1992-2024).
 * For a batch: true while any of its events continues. */
CCAKE_API int  ccake_continue_evolution(ccake_event* ev, double max_tau, int* keep_going);

/* ----------------------------- batches ------------------------------ */
/* Many independent events in one handle, stepped together: every stage
 * launches each kernel once for all of them.  No counterpart in the
 * reference, which runs one event per process (BSQHydro::run,
 * src/BSQHydro.cpp:658-866); each event evolves exactly as it would alone
 * (same pair sets, same arithmetic; only the grouping of its pair sums into
 * slices can differ, by rounding).  Event e owns the particle ids
 * [offsets[e], offsets[e+1]) of the batch, counts[e] of them: upload and
 * download the concatenated columns with the ordinary event calls.  All
 * events share the parameters, the EoS and the time (they start at t0
 * together).  At most 1024 events; each event's cell grid has at most 128
 * cells per axis (coarser cells for very wide events).  Freeze-out rows
 * carry batch particle ids. */
CCAKE_API int  ccake_batch_create(ccake_dev* dev, const ccake_params* p, ccake_eos* eos, int n_events,
                                  const int* counts, ccake_event** out);
/* Number of events (1 for an ordinary event) and, when offsets != NULL,
 * the n_events + 1 offsets. */
CCAKE_API int  ccake_batch_events(ccake_event* ev, int* n_events, int* offsets);
/* ccake_reduce per event: out[e] for each of the n_events events. */
CCAKE_API int  ccake_batch_reduce(ccake_event* ev, int which, double* out);
/* continue_evolution per event: keep_going[e] for each event.  An event
 * whose flag is 0 has ended (as BSQHydro::run ends for it alone): later
 * steps of the batch leave it unchanged and cost it nothing, until the next
 * process_initial_conditions (or initialize) makes every event live again. */
CCAKE_API int  ccake_batch_continue_evolution(ccake_event* ev, double max_tau, int* keep_going);
/* Sticky device-side error flags raised by kernels since the last call
 * (negative entropy, EoS failure); clears them. */
CCAKE_API int  ccake_event_check(ccake_event* ev);

/* Device reductions: SystemState::conservation_entropy / _BSQ / _energy
 * (src/system_state.cpp:738-885) and the frozen-out count. */
enum ccake_reduction { CCAKE_R_ENTROPY = 0, CCAKE_R_BARYON, CCAKE_R_STRANGE, CCAKE_R_CHARGE,
                       CCAKE_R_ENERGY, CCAKE_R_EZ, CCAKE_R_FROZEN, CCAKE_NUM_REDUCTIONS };
CCAKE_API int  ccake_reduce(ccake_event* ev, int which, double* out);

/* ------------------------ freeze-out surface ------------------------ */
/* The freeze-out hypersurface accumulated since ccake_initialize, for events
 * created with particlization on: one row of CCAKE_FO_ROW doubles for every
 * particle that froze out through the interpolation (status 3 -> 4): the
 * particle id, the step number (1 = the first step after initialize), then
 * the 40 columns of freeze_out.dat for D = 2 in the order of
 * Output<2, EoM_default>::print_freeze_out (src/output.cpp:502-560):
 * divEener, divE[2], gsub, uout[2], swsub, bulksub, shear 00, 11, 22, 33, 12,
 * tlist, rsub[2], sFO, E, T (GeV), muB, muS, muQ, w, cs2, rhoB, rhoS, rhoQ,
 * the 12 diffusion components q^mu of B, S, Q (zero with diffusion off) and
 * the table-EoS flag.  Rows are
 * ordered by step, then particle id, the order the reference writes them.
 * Replaces FreezeOut::bsqsvinterpolate's results with
 * Output::print_freeze_out / flush_freeze_out (include/freeze_out.h:638-930,
 * src/output.cpp:412-498).  rows == NULL or cap_rows too small: *n_rows is
 * set and CCAKE_ERR_CAPACITY returned (rows NULL: 0 returned). */
#define CCAKE_FO_ROW 42
CCAKE_API int  ccake_freeze_out_surface(ccake_event* ev, double* rows, size_t cap_rows, size_t* n_rows);

/* --------------------------- profiling ----------------------------- */
/* Per-kernel HIP-event timing on the events' streams (off by default).
 * SMOOTH and GRAD time the pair kernels alone; SMOOTH_SUM and GRAD_SUM the
 * passes that add their per-slice partial sums and write the fields. */
enum ccake_kernel_id { CCAKE_K_BIN = 0, CCAKE_K_SORT, CCAKE_K_SMOOTH, CCAKE_K_EOS,
                       CCAKE_K_GRAD, CCAKE_K_EOM, CCAKE_K_RK, CCAKE_K_MISC,
                       CCAKE_K_SMOOTH_SUM, CCAKE_K_GRAD_SUM, CCAKE_NUM_KERNELS };
CCAKE_API int  ccake_profile_enable(ccake_dev* dev, int enable);
CCAKE_API int  ccake_profile_read(ccake_dev* dev, int kernel_id, double* total_ms, int64_t* launches);
CCAKE_API int  ccake_profile_reset(ccake_dev* dev);
/* Pair count of the last neighbour build (directed pairs, i != j). */
CCAKE_API int  ccake_last_pair_count(ccake_event* ev, int64_t* n_pairs);
/* Directed pairs summed over every smoothing pass since the event was
 * created or since the last call with reset != 0 (the work count behind the
 * pair kernels' throughput).  Diagnostic; no counterpart in the reference. */
CCAKE_API int  ccake_pair_total(ccake_event* ev, int reset, int64_t* n_pairs);
/* Slices per tile of the pair kernels (the work split of the smoothing and
 * gradient passes).  set > 0 chooses a new count, at most the count the
 * event was created with (its partial-sum buffer is sized for that); the
 * current count is returned in *nslices.  Results do not depend on it beyond
 * summation order.  Tuning; no counterpart in the reference. */
CCAKE_API int  ccake_event_pair_slices(ccake_event* ev, int set, int* nslices);

#ifdef __cplusplus
}
#endif
#endif /* CCAKE_AMD_H */
