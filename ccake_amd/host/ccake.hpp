// C++ host-side call surface of the per-timestep SPH path, over the C-ABI of
// libccake_amd.so (include/ccake_amd.h).  The class and method names are the
// reference's (include/system_state.h:34-201, include/evolver.h:57-116,
// include/sph_workstation.h:54-159), so that a caller such as
// BSQHydro::initialize_hydrodynamics / BSQHydro::run (src/BSQHydro.cpp:568-866)
// keeps its sequence of calls; every call lands on the gfx950 library, which
// owns the particle state in HBM.  The host keeps a mirror
// (`SystemState::particles`) that `copy_device_to_host()` fills field by
// field, as the reference does for its host AoS (src/system_state.cpp:334-555).
//
// Errors: a library status other than CCAKE_OK throws ccake::Error.  With
// Settings::exit_on_error the reference's exit codes are used instead:
// negative entropy exits with EXIT_FAILURE (src/evolver.cpp:437-450), a
// failed EoS inversion with 404 (src/sph_workstation.cpp:1752-1760).
#pragma once

#include <cmath>
#ifdef CCAKE_WITH_HDF5
#include "ccake_hdf5.hpp"
#endif
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ccake_amd.h"

namespace ccake {

struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string& what)
    : std::runtime_error(what + ": status " + std::to_string(st) + " (" + ccake_last_error() + ")"), status(st) {}
};

// The settings the path reads (include/settings.h, include/defaults.h:14-66),
// under the reference's names.  Transport modes use the C-ABI enums.
struct Settings {
  int gpu = 0;
  size_t hbm_budget = 0;              // 0: no limit
  std::string eos_type = "conformal"; // "conformal" or "table"
  // table EoS (eos_type == "table"): 17 fields per node, T slowest
  std::vector<double> eos_table;
  int eos_dims[4] = { 0, 0, 0, 0 };
  double eos_mins[4] = { 0, 0, 0, 0 }, eos_maxs[4] = { 0, 0, 0, 0 };
  double Freeze_Out_Temperature = 150.0 / 197.3269804;   // fm^-1
  double t0 = 1.0, dt = 0.005, max_tau = 50.0;
  double hT = 0.3, hEta = 0.3;
  int rk_order = 2;
  bool particlization_enabled = false;
  bool input_initial_shear = false, bulk_from_trace = false;
  bool using_shear = false, using_vorticity = false;
  bool prohibit_unstable_cs2 = true, prohibit_acausal_cs2 = true;
  int eta_mode = CCAKE_ETA_CONSTANT;           double constant_eta_over_s = 0.0;
  int shear_relax_mode = CCAKE_TAUPI_DEFAULT;
  int zeta_mode = CCAKE_ZETA_CONSTANT;         double constant_zeta_over_s = 0.0;
  double cs2_dependent_zeta_A = 1.67552, cs2_dependent_zeta_p = 2.0;
  bool modulate_zeta_with_tanh = true;
  double zeta_gauss_A = 0.0, zeta_gauss_T0 = 0.0, zeta_gauss_width = 1.0;
  int delta_pipi_mode = CCAKE_MODE_IS, tau_pipi_mode = CCAKE_MODE_DISABLED, lambda_piPi_mode = CCAKE_MODE_DISABLED;
  int phi6_mode = CCAKE_MODE_DISABLED, phi7_mode = CCAKE_MODE_DISABLED;
  int delta_PiPi_mode = CCAKE_MODE_IS, lambda_Pipi_mode = CCAKE_MODE_DISABLED;
  int phi1_mode = CCAKE_MODE_DISABLED, phi3_mode = CCAKE_MODE_DISABLED;
  // BSQ diffusion (include/settings.h:195-215, src/input.cpp:570-640):
  // on unless diffusionMode is "disabled"; modes "constant_over_T2" and
  // "kappaB_coth" ("DNMR" is rejected); relaxation "constant_over_T" or
  // "cb_over_T", anything else falls back to the former with a warning.
  std::string diffusionMode = "disabled";
  std::string diffusionRelaxMode = "default";
  double constant_kappa_over_T2[3][3] = {};
  double C_B = 0.0;
  bool critical_scaling_diffusion = false;     // needs the critical-point table inputs: rejected
  bool diffusion_regulator_enabled = false;
  double diffusion_regulator_chi0 = 10.0, diffusion_regulator_e0 = 0.1;   // e0, xi0 in GeV/fm^3
  double diffusion_regulator_xi0 = 0.01, diffusion_regulator_rq_max = 1.0;
  int cells_per_rc = 3;
  int x_subcells = 8;     // x cells per row cell (ccake_params::x_subcells)
  // 0: FP64 pair sums; 1: FP32 pair terms with compensated accumulation
  // (ccake_params::pair_precision; no counterpart in the reference)
  int pair_precision = 0;
  // D = 3: W_2D(r_perp, hT) W_1D(|d eta|, hEta) (include/settings.h:114)
  bool factorized_kernel = false;
  bool print_conservation_status = false;
  bool exit_on_error = false;
  // initial conditions and outputs (include/settings.h, include/defaults.h:25-72)
  std::string IC_type = "ccake";      // "ccake" or "ICCING"
  std::string IC_file;
  double e_cutoff = 0.15;             // GeV/fm^3; points at or below it are dropped
  bool input_as_entropy = false;
  bool input_initial_diffusion = false;
  bool baryon_charge_enabled = true, strange_charge_enabled = true, electric_charge_enabled = true;
  double stepx = 0.0, stepy = 0.0, stepEta = 0.0, xmin = 0.0, ymin = 0.0, etamin = 0.0;
  std::vector<std::string> headers;
  std::string results_directory = ".";
  bool txt_evolution = false, hdf_evolution = false;
};

inline void check(int st, const char* what, const Settings* s = nullptr)
{
  if (st == CCAKE_OK) return;
  if (s && s->exit_on_error) {
    std::fprintf(stderr, "%s: %s\n", what, ccake_last_error());
    if (st == CCAKE_ERR_NEG_ENTROPY) std::exit(EXIT_FAILURE);
    if (st == CCAKE_ERR_EOS_NO_SOLUTION) std::exit(404);
    std::exit(EXIT_FAILURE);
  }
  throw Error(st, what);
}

// Host record of one particle: the fields of the reference's Particle that
// initial conditions set and outputs read (include/particle.h,
// include/hydrodynamic_info.h, include/thermodynamic_info.h).
template <unsigned int D>
struct Particle {
  int ID = 0;
  double r[D] = {};
  double u[D] = {};
  double input_e = 0, input_rhoB = 0, input_rhoS = 0, input_rhoQ = 0;
  // filled by copy_device_to_host()
  double sigma = 0, sigma_lab = 0, gamma = 1, v[D] = {};
  double s_ext = 0, rhoB_ext = 0, rhoS_ext = 0, rhoQ_ext = 0, Bulk_ext = 0;
  double T_ = 0, muB = 0, muS = 0, muQ = 0, e_ = 0, s = 0, rhoB = 0, rhoS = 0, rhoQ = 0, p_ = 0, cs2 = 0, w = 0;
  double Bulk = 0;
  double shv[4][4] = {};              // also the initial pi^{mu nu} an IC reader sets
  double diffusion[3][4] = {};        // q^mu of B, S, Q: the initial currents, then the evolved ones
  // transport coefficients and diagnostics (include/hydrodynamic_info.h:29-62)
  double eta_pi = 0, zeta_Pi = 0, tau_Pi = 0, tau_pi = 0, theta = 0;
  double shear_knudsen = 0, bulk_knudsen = 0, inverse_reynolds_shear = 0, inverse_reynolds_bulk = 0;
  double knudsen_diffusion[3] = {}, inverse_reynolds_diffusion[3] = {};   // zero with diffusion off
  int causality = 0;                  // check_causality is not on this path
  std::string eos_name = "default";   // the host name the reference prints (never set from the device)
  int Freeze = 0, btrack = 0, eos_id = 0;
  double T() const { return T_; }
  double e() const { return e_; }
  double p() const { return p_; }
  std::string get_current_eos_name() const { return eos_name; }
};

// EoM_default's static helpers (src/eom_default.cpp:17-38, 180-203): the
// equations themselves run on the device.
template <unsigned int D>
struct EoM_default {
  static double gamma_calc(const double* u, double /*time_squared*/)
  {
    double c = 0.0;
    for (unsigned a = 0; a < D; ++a) c += u[a] * u[a];
    return std::sqrt(1.0 + c);
  }
  static double get_LRF(double lab, double gamma, double t) { return lab / gamma / t; }
};

template <unsigned int D>
class SystemState {
 public:
  explicit SystemState(std::shared_ptr<Settings> s) : settingsPtr(s)
  {
    static_assert(D == 2, "the gfx950 library is built for 2+1D");
    check(ccake_dev_create(s->gpu, s->hbm_budget, &dev), "ccake_dev_create", s.get());
    ccake_eos_desc d{};
    d.T_freeze_out = s->Freeze_Out_Temperature;
    if (s->eos_type == "table") {
      d.kind = CCAKE_EOS_TABLE;
      for (int a = 0; a < 4; ++a) { d.dims[a] = s->eos_dims[a]; d.mins[a] = s->eos_mins[a]; d.maxs[a] = s->eos_maxs[a]; }
      d.table = s->eos_table.data();
    } else {
      d.kind = CCAKE_EOS_CONFORMAL;
    }
    check(ccake_eos_create(dev, &d, &eos), "ccake_eos_create", s.get());
  }
  SystemState(const SystemState&) = delete;
  SystemState& operator=(const SystemState&) = delete;
  ~SystemState()
  {
    if (ev) ccake_event_destroy(ev);
    if (eos) ccake_eos_destroy(eos);
    if (dev) ccake_dev_destroy(dev);
  }

  double t = 0.0;
  double S = 0.0, S0 = 0.0, Btotal = 0.0, Btotal0 = 0.0, Stotal = 0.0, Stotal0 = 0.0, Qtotal = 0.0, Qtotal0 = 0.0;
  double E = 0.0, Ez = 0.0, E0 = 0.0, Etot = 0.0;
  double efcheck = 0.0;
  int number_part_fo = 0;
  int n_particles = 0;
  int number_of_elapsed_timesteps = 0;
  std::vector<Particle<D>> particles;

  void add_particle(Particle<D> p)
  {
    p.ID = (int)particles.size();
    particles.push_back(p);
    n_particles = (int)particles.size();
  }
  int n() const { return n_particles; }

  // SystemState::allocate_cabana_particles (src/system_state.cpp:58-63):
  // the device event (the reference's Cabana AoSoA) for the particles added
  // so far, filled from them (copy_host_to_device).  An existing event of
  // the same size and settings is kept and refilled.  If
  // SPHWorkstation::process_initial_conditions ran on the host particles,
  // its device part runs here, so the device holds the processed particles
  // as the reference's copy does.
  void allocate_cabana_particles()
  {
    const ccake_params p = params();
    if (ev && (ev_n != n_particles || std::memcmp(&ev_params, &p, sizeof p) != 0)) {
      ccake_event_destroy(ev);
      ev = nullptr;
    }
    if (!ev) {
      check(ccake_event_create(dev, &p, eos, n_particles, &ev), "ccake_event_create", settingsPtr.get());
      ev_n = n_particles;
      ev_params = p;
    }
    copy_host_to_device();
    if (ic_processed)
      check(ccake_process_initial_conditions(ev), "ccake_process_initial_conditions", settingsPtr.get());
  }
  // SystemState::initialize (src/system_state.cpp:40-53)
  void initialize()
  {
    number_of_elapsed_timesteps = 0;
    t = settingsPtr->t0;
  }
  // set by SPHWorkstation::process_initial_conditions
  bool ic_processed = false;

  // Initial-condition fields host -> device (the path's inputs).
  void copy_host_to_device()
  {
    std::vector<double> a(n_particles);
    auto up = [&](int f, auto get) {
      for (int i = 0; i < n_particles; ++i) a[i] = get(particles[i]);
      check(ccake_event_upload(ev, f, a.data(), a.size()), "ccake_event_upload", settingsPtr.get());
    };
    up(CCAKE_F_X, [](const Particle<D>& q) { return q.r[0]; });
    up(CCAKE_F_Y, [](const Particle<D>& q) { return q.r[1]; });
    up(CCAKE_F_UX, [](const Particle<D>& q) { return q.u[0]; });
    up(CCAKE_F_UY, [](const Particle<D>& q) { return q.u[1]; });
    up(CCAKE_F_IN_E, [](const Particle<D>& q) { return q.input_e; });
    up(CCAKE_F_IN_B, [](const Particle<D>& q) { return q.input_rhoB; });
    up(CCAKE_F_IN_SS, [](const Particle<D>& q) { return q.input_rhoS; });
    up(CCAKE_F_IN_Q, [](const Particle<D>& q) { return q.input_rhoQ; });
    // initial Pi and pi^{mu nu}: read by initialize when bulk_from_trace /
    // input_initial_shear are set, zeroed there otherwise
    up(CCAKE_F_BULK, [](const Particle<D>& q) { return q.Bulk; });
    for (int c = 0; c < 16; ++c)
      up(CCAKE_F_SHV0 + c, [c](const Particle<D>& q) { return q.shv[c / 4][c % 4]; });
    // initial q^mu (zero unless an IC reader set it): initialize derives the
    // extensive currents from it when diffusion is on
    for (int c = 0; c < 12; ++c)
      up(CCAKE_F_Q0 + c, [c](const Particle<D>& q) { return q.diffusion[c / 4][c % 4]; });
    check(ccake_event_set_time(ev, t), "ccake_event_set_time", settingsPtr.get());
  }

  // Device -> host mirror, field by field (src/system_state.cpp:334-555).
  void copy_device_to_host()
  {
    std::vector<double> a(n_particles);
    std::vector<int32_t> b(n_particles);
    auto dn = [&](int f, auto set) {
      check(ccake_event_download(ev, f, a.data(), a.size()), "ccake_event_download", settingsPtr.get());
      for (int i = 0; i < n_particles; ++i) set(particles[i], a[i]);
    };
    dn(CCAKE_F_X, [](Particle<D>& q, double v) { q.r[0] = v; });
    dn(CCAKE_F_Y, [](Particle<D>& q, double v) { q.r[1] = v; });
    dn(CCAKE_F_UX, [](Particle<D>& q, double v) { q.u[0] = v; });
    dn(CCAKE_F_UY, [](Particle<D>& q, double v) { q.u[1] = v; });
    dn(CCAKE_F_VX, [](Particle<D>& q, double v) { q.v[0] = v; });
    dn(CCAKE_F_VY, [](Particle<D>& q, double v) { q.v[1] = v; });
    dn(CCAKE_F_GAMMA, [](Particle<D>& q, double v) { q.gamma = v; });
    dn(CCAKE_F_SIGMA, [](Particle<D>& q, double v) { q.sigma = v; });
    dn(CCAKE_F_SIGMA_LAB, [](Particle<D>& q, double v) { q.sigma_lab = v; });
    dn(CCAKE_F_EXT_S, [](Particle<D>& q, double v) { q.s_ext = v; });
    dn(CCAKE_F_EXT_B, [](Particle<D>& q, double v) { q.rhoB_ext = v; });
    dn(CCAKE_F_EXT_SS, [](Particle<D>& q, double v) { q.rhoS_ext = v; });
    dn(CCAKE_F_EXT_Q, [](Particle<D>& q, double v) { q.rhoQ_ext = v; });
    dn(CCAKE_F_EXT_BULK, [](Particle<D>& q, double v) { q.Bulk_ext = v; });
    dn(CCAKE_F_T, [](Particle<D>& q, double v) { q.T_ = v; });
    dn(CCAKE_F_MUB, [](Particle<D>& q, double v) { q.muB = v; });
    dn(CCAKE_F_MUS, [](Particle<D>& q, double v) { q.muS = v; });
    dn(CCAKE_F_MUQ, [](Particle<D>& q, double v) { q.muQ = v; });
    dn(CCAKE_F_E, [](Particle<D>& q, double v) { q.e_ = v; });
    dn(CCAKE_F_S, [](Particle<D>& q, double v) { q.s = v; });
    dn(CCAKE_F_RHOB, [](Particle<D>& q, double v) { q.rhoB = v; });
    dn(CCAKE_F_RHOS, [](Particle<D>& q, double v) { q.rhoS = v; });
    dn(CCAKE_F_RHOQ, [](Particle<D>& q, double v) { q.rhoQ = v; });
    dn(CCAKE_F_P, [](Particle<D>& q, double v) { q.p_ = v; });
    dn(CCAKE_F_CS2, [](Particle<D>& q, double v) { q.cs2 = v; });
    dn(CCAKE_F_W, [](Particle<D>& q, double v) { q.w = v; });
    dn(CCAKE_F_BULK, [](Particle<D>& q, double v) { q.Bulk = v; });
    dn(CCAKE_F_ETA_PI, [](Particle<D>& q, double v) { q.eta_pi = v; });
    dn(CCAKE_F_ZETA_PI, [](Particle<D>& q, double v) { q.zeta_Pi = v; });
    dn(CCAKE_F_TAU_BULK, [](Particle<D>& q, double v) { q.tau_Pi = v; });
    dn(CCAKE_F_TAU_PI, [](Particle<D>& q, double v) { q.tau_pi = v; });
    dn(CCAKE_F_THETA, [](Particle<D>& q, double v) { q.theta = v; });
    dn(CCAKE_F_SHEAR_KNUDSEN, [](Particle<D>& q, double v) { q.shear_knudsen = v; });
    dn(CCAKE_F_BULK_KNUDSEN, [](Particle<D>& q, double v) { q.bulk_knudsen = v; });
    dn(CCAKE_F_INV_REYNOLDS_SHEAR, [](Particle<D>& q, double v) { q.inverse_reynolds_shear = v; });
    dn(CCAKE_F_INV_REYNOLDS_BULK, [](Particle<D>& q, double v) { q.inverse_reynolds_bulk = v; });
    for (int c = 0; c < 16; ++c) {
      check(ccake_event_download(ev, CCAKE_F_SHV0 + c, a.data(), a.size()), "ccake_event_download", settingsPtr.get());
      for (int i = 0; i < n_particles; ++i) particles[i].shv[c / 4][c % 4] = a[i];
    }
    if (settingsPtr->diffusionMode != "disabled") {
      for (int c = 0; c < 12; ++c)
        dn(CCAKE_F_Q0 + c, [c](Particle<D>& q, double v) { q.diffusion[c / 4][c % 4] = v; });
      for (int c = 0; c < 3; ++c) {
        dn(CCAKE_F_KN_DIFF0 + c, [c](Particle<D>& q, double v) { q.knudsen_diffusion[c] = v; });
        dn(CCAKE_F_RE_DIFF0 + c, [c](Particle<D>& q, double v) { q.inverse_reynolds_diffusion[c] = v; });
      }
    }
    auto dni = [&](int f, auto set) {
      check(ccake_event_download_int(ev, f, b.data(), b.size()), "ccake_event_download_int", settingsPtr.get());
      for (int i = 0; i < n_particles; ++i) set(particles[i], b[i]);
    };
    dni(CCAKE_I_FREEZE, [](Particle<D>& q, int v) { q.Freeze = v; });
    dni(CCAKE_I_BTRACK, [](Particle<D>& q, int v) { q.btrack = v; });
    dni(CCAKE_I_EOS_ID, [](Particle<D>& q, int v) { q.eos_id = v; });
    check(ccake_event_get_time(ev, &t), "ccake_event_get_time", settingsPtr.get());
  }

  // SystemState::initialize_linklist / reset_neighbour_list
  // (src/system_state.cpp:650-733): cell sort plus neighbour counts (btrack).
  void initialize_linklist() { reset_neighbour_list(); }
  void reset_neighbour_list() { check(ccake_build_neighbors(ev), "ccake_build_neighbors", settingsPtr.get()); }

  // SystemState::conservation_* (src/system_state.cpp:738-885)
  void conservation_entropy(bool first_iteration = false)
  {
    S = reduce(CCAKE_R_ENTROPY);
    if (first_iteration) S0 = S;
  }
  void conservation_BSQ(bool first_iteration = false)
  {
    Btotal = reduce(CCAKE_R_BARYON); Stotal = reduce(CCAKE_R_STRANGE); Qtotal = reduce(CCAKE_R_CHARGE);
    if (first_iteration) { Btotal0 = Btotal; Stotal0 = Stotal; Qtotal0 = Qtotal; }
  }
  void conservation_energy(bool first_iteration = false, double /*t*/ = 1.0)
  {
    E = reduce(CCAKE_R_ENERGY);
    Ez = reduce(CCAKE_R_EZ);
    if (first_iteration) E0 = E;
    Etot = E + Ez;
  }

  // handles for the workstation and evolver
  ccake_event* event() const { return ev; }
  ccake_eos* equation_of_state() const { return eos; }
  const Settings& settings() const { return *settingsPtr; }

 private:
  double reduce(int which)
  {
    double v = 0.0;
    check(ccake_reduce(ev, which, &v), "ccake_reduce", settingsPtr.get());
    return v;
  }
  ccake_params params() const
  {
    const Settings& s = *settingsPtr;
    ccake_params p{};
    p.dim = (int)D; p.rk_order = s.rk_order; p.hT = s.hT; p.hEta = s.hEta; p.t0 = s.t0; p.dt = s.dt;
    p.e_freeze = efcheck; p.particlization = s.particlization_enabled; p.using_shear = s.using_shear;
    p.using_vorticity = s.using_vorticity; p.input_initial_shear = s.input_initial_shear;
    p.bulk_from_trace = s.bulk_from_trace; p.prohibit_unstable_cs2 = s.prohibit_unstable_cs2;
    p.prohibit_acausal_cs2 = s.prohibit_acausal_cs2; p.eta_mode = s.eta_mode;
    p.constant_eta_over_s = s.constant_eta_over_s; p.shear_relax_mode = s.shear_relax_mode;
    p.zeta_mode = s.zeta_mode; p.constant_zeta_over_s = s.constant_zeta_over_s;
    p.cs2_dependent_zeta_A = s.cs2_dependent_zeta_A; p.cs2_dependent_zeta_p = s.cs2_dependent_zeta_p;
    p.modulate_zeta_with_tanh = s.modulate_zeta_with_tanh; p.zeta_gauss_A = s.zeta_gauss_A;
    p.zeta_gauss_T0 = s.zeta_gauss_T0; p.zeta_gauss_width = s.zeta_gauss_width;
    p.delta_pipi_mode = s.delta_pipi_mode; p.tau_pipi_mode = s.tau_pipi_mode; p.lambda_piPi_mode = s.lambda_piPi_mode;
    p.phi6_mode = s.phi6_mode; p.phi7_mode = s.phi7_mode; p.delta_PiPi_mode = s.delta_PiPi_mode;
    p.lambda_Pipi_mode = s.lambda_Pipi_mode; p.phi1_mode = s.phi1_mode; p.phi3_mode = s.phi3_mode;
    p.cells_per_rc = s.cells_per_rc;
    // diffusion (src/transport_coefficients.cpp:82-115, src/input.cpp:577-640)
    p.using_diffusion = s.diffusionMode != "disabled";
    if (p.using_diffusion) {
      if (s.diffusionMode == "DNMR" || s.critical_scaling_diffusion)
        throw Error(CCAKE_ERR_UNSUPPORTED, "diffusion: DNMR kappa and critical scaling are not on this path");
      if (s.diffusionMode == "kappaB_coth") {
        if (!(s.baryon_charge_enabled && !s.strange_charge_enabled && !s.electric_charge_enabled))
          throw Error(CCAKE_ERR_INVALID, "diffusion mode 'kappaB_coth' requires a baryon-only charge configuration");
        p.diffusion_mode = CCAKE_DIFF_KAPPA_B_COTH;
      } else {
        if (s.diffusionMode != "constant_over_T2")
          std::cout << "WARNING: Unknown diffusion mode: " << s.diffusionMode << "\nDefaulting to constant_over_T2.\n";
        p.diffusion_mode = CCAKE_DIFF_KAPPA_DEFAULT;
      }
      if (s.diffusionRelaxMode == "cb_over_T") p.diffusion_relax_mode = CCAKE_TAUQ_CB_OVER_T;
      else {
        if (s.diffusionRelaxMode != "constant_over_T")
          std::cout << "WARNING: Unknown diffusion relaxation mode: " << s.diffusionRelaxMode
                    << "\nDefaulting to constant_over_T (tau_q = 0.2/T).\n";
        p.diffusion_relax_mode = CCAKE_TAUQ_DEFAULT;
      }
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) p.kappa_over_T2[3 * a + b] = s.constant_kappa_over_T2[a][b];
      p.C_B = s.C_B;
      p.diffusion_regulator = s.diffusion_regulator_enabled;
      p.regulator_chi0 = s.diffusion_regulator_chi0; p.regulator_e0 = s.diffusion_regulator_e0;
      p.regulator_xi0 = s.diffusion_regulator_xi0; p.regulator_rq_max = s.diffusion_regulator_rq_max;
    }
    p.pair_precision = s.pair_precision;
    p.factorized_kernel = s.factorized_kernel;
    p.x_subcells = s.x_subcells;
    return p;
  }

  std::shared_ptr<Settings> settingsPtr;
  ccake_dev* dev = nullptr;
  ccake_eos* eos = nullptr;
  ccake_event* ev = nullptr;
  int ev_n = 0;                 // size and settings the event was created with
  ccake_params ev_params{};
};

// Evolver<D> (src/evolver.cpp:97-517): the RK cache lives on the device.
template <unsigned int D>
class Evolver {
 public:
  Evolver(std::shared_ptr<Settings> s, std::shared_ptr<SystemState<D>> sys) : settingsPtr(s), systemPtr(sys) {}

  void allocate_cache() {}   // the event owns its cache columns
  void set_current_timestep_quantities()
  {
    check(ccake_cache_state(systemPtr->event()), "ccake_cache_state", settingsPtr.get());
  }
  // y = y0 + dt f(y, t), t = t0 + dt, after the caller's derivatives
  void step_rk(double dt, double t0, std::function<void(void)> time_derivatives_functional)
  {
    time_derivatives_functional();
    check(ccake_rk_update(systemPtr->event(), dt, t0), "ccake_rk_update", settingsPtr.get());
    systemPtr->t = t0 + dt;
  }
  // Evolver::allocate_k_values / update_k / update_rk4 (src/evolver.cpp:171-341)
  void allocate_k_values() { check(ccake_allocate_k_values(systemPtr->event()), "ccake_allocate_k_values", settingsPtr.get()); }
  void update_k(int n) { check(ccake_update_k(systemPtr->event(), n), "ccake_update_k", settingsPtr.get()); }
  void update_rk4(double dt) { check(ccake_update_rk4(systemPtr->event(), dt), "ccake_update_rk4", settingsPtr.get()); }

  // Evolver::advance_timestep_rk2 (src/evolver.cpp:493-517)
  void advance_timestep_rk2(double dt, std::function<void(void)> time_derivatives_functional)
  {
    set_current_timestep_quantities();
    const double t0 = systemPtr->t;
    step_rk(0.5 * dt, t0, time_derivatives_functional);
    step_rk(dt, t0, time_derivatives_functional);
  }
  // Evolver::advance_timestep_rk4 (src/evolver.cpp:519-561), the reference's
  // sequence of calls as it is
  void advance_timestep_rk4(double dt, std::function<void(void)> time_derivatives_functional)
  {
    set_current_timestep_quantities();
    allocate_k_values();
    const double t0 = systemPtr->t;
    step_rk(0.5 * dt, t0, time_derivatives_functional);
    step_rk(0.5 * dt, t0, time_derivatives_functional);
    update_k(1);
    step_rk(0.5 * dt, t0, time_derivatives_functional);
    update_k(2);
    step_rk(dt, t0, time_derivatives_functional);
    update_k(3);
    time_derivatives_functional();
    update_k(3);
    update_rk4(dt);
  }
  // Evolver::execute_timestep (src/evolver.cpp:97-117)
  void execute_timestep(double dt, int rk_order, std::function<void(void)> time_derivatives_functional)
  {
    if (rk_order == 2) advance_timestep_rk2(dt, time_derivatives_functional);
    else if (rk_order == 4) advance_timestep_rk4(dt, time_derivatives_functional);
    else check(CCAKE_ERR_INVALID, "Evolver::execute_timestep: Invalid Runge-Kutta order", settingsPtr.get());
  }

 private:
  std::shared_ptr<Settings> settingsPtr;
  std::shared_ptr<SystemState<D>> systemPtr;
};

// FreezeOut<D> (include/freeze_out.h:136-930): the snapshots, the status
// machine and the interpolation run on the device inside every step; this
// handle reads the hypersurface they have produced.
template <unsigned int D>
class FreezeOut {
 public:
  explicit FreezeOut(std::shared_ptr<SystemState<D>> sys) : systemPtr(std::move(sys)) {}
  // Rows of CCAKE_FO_ROW doubles (particle id, step, then the columns of
  // freeze_out.dat), ordered by (step, particle id).
  std::vector<double> surface(size_t* n_rows) const
  {
    ccake_event* ev = systemPtr->event();
    size_t n = 0;
    check(ccake_freeze_out_surface(ev, nullptr, 0, &n), "ccake_freeze_out_surface", &systemPtr->settings());
    std::vector<double> rows(n * CCAKE_FO_ROW);
    if (n) check(ccake_freeze_out_surface(ev, rows.data(), n, &n), "ccake_freeze_out_surface", &systemPtr->settings());
    *n_rows = n;
    return rows;
  }

 private:
  std::shared_ptr<SystemState<D>> systemPtr;
};

// Output<D, TEOM> (src/output.cpp): the outputs of the run.
//  - print_system_state (63-83) downloads the particles and writes
//    system_state_<k>.dat (print_system_state_to_txt, 102-172) when
//    txt_evolution is set and a frame of output_directory/system_state.h5
//    (print_system_state_to_HDF, 213-266, ccake_hdf5.hpp) when
//    hdf_evolution is set; k counts the calls.  HDF5 output needs a program
//    built with CCAKE_WITH_HDF5 and linked to libhdf5; without it asking for
//    hdf_evolution fails at construction.
//  - freeze-out rows are buffered as text at every print_freeze_out and
//    written to output_directory/freeze_out.dat by flush_freeze_out
//    (412-498), after a header line "# B S Q" with the totals passed at the
//    first call.
// Numbers are formatted as the reference formats them: the default stream
// format (6 significant digits) for freeze_out.dat and the first two columns
// of the first system-state line, std::scientific with 6 digits after it.
template <unsigned int D, template <unsigned int> class TEOM>
class Output {
 public:
  explicit Output(std::string dir) : output_directory(std::move(dir)) {}
  Output(std::shared_ptr<Settings> s, std::shared_ptr<SystemState<D>> sys)
    : output_directory(s->results_directory), settingsPtr(std::move(s)), systemPtr(std::move(sys))
  {
    if (settingsPtr->hdf_evolution) {
#ifdef CCAKE_WITH_HDF5
      // the run's attributes (src/output.cpp:33-47)
      hdf5_file.initialize(output_directory + "/system_state.h5", { settingsPtr->hT, settingsPtr->e_cutoff },
                           { "h", "e_cutoff" });
#else
      check(CCAKE_ERR_UNSUPPORTED, "Output (hdf_evolution: this program is built without HDF5)", settingsPtr.get());
#endif
    }
  }

  void print_system_state()
  {
    systemPtr->copy_device_to_host();
    if (settingsPtr->txt_evolution) print_system_state_to_txt();
#ifdef CCAKE_WITH_HDF5
    if (settingsPtr->hdf_evolution) print_system_state_to_HDF();
#endif
    n_timesteps_output++;
  }

#ifdef CCAKE_WITH_HDF5
  // One frame: x, y (fm), T, mu (MeV), e (MeV/fm^3), s, B, S, Q (1/fm^3) and
  // each particle's EoS tag, by particle id (src/output.cpp:213-266).
  void print_system_state_to_HDF()
  {
    constexpr double hc = 197.3269804;        // include/constants.h:16
    const int width = (int)std::ceil(std::log10(std::ceil(settingsPtr->max_tau / settingsPtr->dt)));
    static const std::vector<std::string> names = { "x", "y", "T", "muB", "muS", "muQ", "e", "s", "B", "S", "Q" };
    static const std::vector<std::string> units = { "fm", "fm", "MeV", "MeV", "MeV", "MeV", "MeV/fm^3", "1/fm^3",
                                                    "1/fm^3", "1/fm^3", "1/fm^3" };
    const SystemState<D>& sys = *systemPtr;
    const size_t n = sys.particles.size();
    std::vector<std::vector<double>> data(names.size(), std::vector<double>(n));
    std::vector<int> labels(n);
    for (size_t id = 0; id < n; ++id) {
      const Particle<D>& p = sys.particles[id];
      data[0][id] = p.r[0];
      data[1][id] = p.r[1];
      data[2][id] = p.T() * hc;
      data[3][id] = p.muB * hc;
      data[4][id] = p.muS * hc;
      data[5][id] = p.muQ * hc;
      data[6][id] = p.e() * hc;
      data[7][id] = p.s;
      data[8][id] = p.rhoB;
      data[9][id] = p.rhoS;
      data[10][id] = p.rhoQ;
      // the reference's map of EoS names to tags; a name outside it is
      // inserted with tag 0 by std::map::operator[]
      const std::string nm = p.get_current_eos_name();
      labels[id] = nm == "tanh_conformal" ? 1 : nm == "conformal" ? 2 : nm == "conformal_diagonal" ? 3 : 0;
    }
    hdf5_file.output_dataset(names, units, data, width, n_timesteps_output, labels, { sys.t, -10.0, -10.0 },
                             { "Time", "e_2_X", "e_2_P" });
  }
#endif

  void print_system_state_to_txt()
  {
    constexpr double hc = 0.1973269804;      // include/constants.h:15
    const std::string name = output_directory + "/system_state_" + std::to_string(n_timesteps_output) + ".dat";
    std::ofstream out(name);
    if (!out) throw std::runtime_error("Could not open " + name);
    const SystemState<D>& sys = *systemPtr;
    out << sys.t << "\n";
    int id = 0;
    for (const Particle<D>& p : sys.particles) {
      out << id++ << " " << sys.t << " " << std::setprecision(6) << std::scientific;
      for (unsigned a = 0; a < D; ++a) out << p.r[a] << " ";
      for (unsigned a = D; a < 3; ++a) out << 0.0 << " ";
      out << p.p() * hc << " " << p.T() * hc << " " << p.muB * hc << " " << p.muS * hc << " " << p.muQ * hc << " "
          << p.e() * hc << " " << p.rhoB << " " << p.rhoS << " " << p.rhoQ << " " << p.s << " "
          << p.eta_pi << " " << p.zeta_Pi << " " << p.tau_Pi << " " << p.tau_pi << " " << p.theta << " "
          << p.inverse_reynolds_shear << " " << p.inverse_reynolds_bulk << " "
          << p.shear_knudsen << " " << p.bulk_knudsen << " "
          << p.shv[0][0] << " " << p.shv[1][1] << " " << p.shv[2][2] << " " << p.shv[0][1] << " "
          << p.shv[0][2] << " " << p.shv[1][2] << " " << p.shv[1][3] << " " << p.shv[2][3] << " "
          << p.shv[3][3] << " ";
      for (unsigned a = 0; a < D; ++a) out << p.u[a] << " ";
      for (unsigned a = D; a < 3; ++a) out << 0.0 << " ";
      out << p.gamma << " " << p.Freeze << " " << p.get_current_eos_name() << " " << p.causality << " ";
      for (int c = 0; c < 3; ++c) out << p.knudsen_diffusion[c] << " ";
      for (int c = 0; c < 3; ++c) out << p.inverse_reynolds_diffusion[c] << " ";
      out << "\n";
    }
    out << std::flush;
  }

  void print_freeze_out(std::shared_ptr<FreezeOut<D>> freeze_out, double B, double Q, double S)
  {
    if (fo_header.empty()) {
      std::ostringstream hdr;
      hdr << "# " << B << " " << S << " " << Q << "\n";
      fo_header = hdr.str();
    }
    size_t n = 0;
    const std::vector<double> rows = freeze_out->surface(&n);
    int count = 0;
    for (size_t k = rows_buffered; k < n; ++k) {      // rows of the steps since the last call
      const double* c = &rows[k * CCAKE_FO_ROW + 2];
      for (int q = 0; q < CCAKE_FO_ROW - 3; ++q) fo_buffer << c[q] << " ";
      fo_buffer << c[CCAKE_FO_ROW - 3] << "\n";
      ++count;
    }
    rows_buffered = n;
    fo_particles_buffered += count;
  }

  void flush_freeze_out()
  {
    const std::string name = output_directory + "/freeze_out.dat";
    std::ofstream FO(name, std::ios::out | std::ios::trunc);
    if (!FO) throw std::runtime_error("Could not open " + name);
    FO << fo_header;
    FO << fo_buffer.str();
  }

  int particles_buffered() const { return fo_particles_buffered; }
  std::string output_directory;
  int n_timesteps_output = 0;

 private:
  std::shared_ptr<Settings> settingsPtr;
  std::shared_ptr<SystemState<D>> systemPtr;
  std::string fo_header;
  std::ostringstream fo_buffer;
  size_t rows_buffered = 0;
  int fo_particles_buffered = 0;
#ifdef CCAKE_WITH_HDF5
  Hdf5SystemState hdf5_file;
#endif
};

// SPHWorkstation<D, TEOM> (src/sph_workstation.cpp).
template <unsigned int D, template <unsigned int> class TEOM>
class SPHWorkstation {
 public:
  SPHWorkstation(std::shared_ptr<Settings> s, std::shared_ptr<SystemState<D>> sys)
    : evolver(s, sys), freezePtr(std::make_shared<FreezeOut<D>>(sys)), settingsPtr(s), systemPtr(sys) {}

  // SPHWorkstation::initialize (src/sph_workstation.cpp:29-110):
  // SystemState::initialize and e_freeze from the EoS at T_FO.
  void initialize()
  {
    systemPtr->initialize();
    double ef = 0.0;
    check(ccake_eos_efreeze(systemPtr->equation_of_state(), settingsPtr->Freeze_Out_Temperature, &ef),
          "ccake_eos_efreeze", settingsPtr.get());
    systemPtr->efcheck = ef;
  }
  void setup_freeze_out() {}   // the freeze-out state lives in the event (snapshots, status)

  // SPHWorkstation::process_initial_conditions (src/sph_workstation.cpp:1228-1340)
  // on the host particles: gamma from u, extensive densities and SPH masses 1,
  // the Newton seed T = 580 MeV, mu = 0, IDs.  The device event gets the same
  // values when allocate_cabana_particles fills it.
  void process_initial_conditions()
  {
    const double t2 = settingsPtr->t0 * settingsPtr->t0;
    int id = 0;
    for (Particle<D>& p : systemPtr->particles) {
      p.gamma = TEOM<D>::gamma_calc(p.u, t2);
      p.s_ext = p.rhoB_ext = p.rhoS_ext = p.rhoQ_ext = 1.0;
      p.T_ = 580.0 / 197.3269804;
      p.muB = p.muS = p.muQ = 0.0;
      p.eos_name = "default";
      p.ID = id++;
    }
    systemPtr->n_particles = (int)systemPtr->particles.size();
    systemPtr->ic_processed = true;
    if (systemPtr->event())       // an event already allocated: process it too
      call(ccake_process_initial_conditions(ev()), "process_initial_conditions");
  }
  // SPHWorkstation::initialize_entropy_and_charge_densities (252-375)
  void initialize_entropy_and_charge_densities()
  {
    call(ccake_initialize_entropy_and_charge_densities(ev()), "initialize_entropy_and_charge_densities");
  }
  // SPHWorkstation::initial_smoothing (394-457), with its initial energy
  // conservation record
  void initial_smoothing()
  {
    call(ccake_initial_smoothing(ev()), "initial_smoothing");
    check(ccake_event_get_time(ev(), &systemPtr->t), "ccake_event_get_time", settingsPtr.get());
    systemPtr->conservation_energy(true, settingsPtr->t0);
  }

  // The whole of BSQHydro::initialize_hydrodynamics (src/BSQHydro.cpp:568-640)
  // in one device call (allocating and filling the event first if the caller
  // has not): process_initial_conditions, initialize_linklist,
  // initialize_entropy_and_charge_densities, calculate_gamma_and_velocities,
  // initial_smoothing (calculate_extensive_shv, set_bulk_Pi, set_diffusion,
  // reset_pi_tensor, update_all_particle_thermodynamics, the first
  // freeze-out check), then the initial conservation records.
  void initialize_hydrodynamics()
  {
    if (!systemPtr->event()) systemPtr->allocate_cabana_particles();
    check(ccake_initialize(ev()), "ccake_initialize", settingsPtr.get());
    check(ccake_event_get_time(ev(), &systemPtr->t), "ccake_event_get_time", settingsPtr.get());
    systemPtr->conservation_entropy(true);
    systemPtr->conservation_BSQ(true);
    systemPtr->conservation_energy(true, systemPtr->t);
  }

  // the phases, one device call each
  void calculate_gamma_and_velocities() { call(ccake_gamma_velocity(ev()), "calculate_gamma_and_velocities"); }
  void smooth_all_particle_fields(double /*time_squared*/) { call(ccake_smooth_fields(ev()), "smooth_all_particle_fields"); }
  void calculate_intial_sigma(double /*time_squared*/) { call(ccake_initial_sigma(ev()), "calculate_intial_sigma"); }
  void smooth_all_particle_gradients(double /*time_squared*/) { call(ccake_smooth_gradients(ev()), "smooth_all_particle_gradients"); }
  void update_all_particle_thermodynamics() { call(ccake_eos_invert(ev(), 0), "update_all_particle_thermodynamics"); }
  void update_all_particle_viscosities() { call(ccake_viscosities(ev()), "update_all_particle_viscosities"); }
  void reset_pi_tensor(double /*time_squared*/) { call(ccake_reset_pi_tensor(ev()), "reset_pi_tensor"); }
  // SPHWorkstation::get_time_derivatives (src/sph_workstation.cpp:1509-1612)
  void get_time_derivatives(double /*dt*/) { call(ccake_get_time_derivatives(ev()), "get_time_derivatives"); }
  void freeze_out_particles() { call(ccake_freeze_out_particles(ev()), "freeze_out_particles"); }

  // SPHWorkstation::advance_timestep (src/sph_workstation.cpp:2042-2073): the
  // RK2 step with its freeze-out check, fused on the device.
  void advance_timestep(double dt, int rk_order)
  {
    call(ccake_advance_timestep(ev(), dt, rk_order), "advance_timestep");
    check(ccake_event_get_time(ev(), &systemPtr->t), "ccake_event_get_time", settingsPtr.get());
    systemPtr->number_of_elapsed_timesteps++;
    if (settingsPtr->print_conservation_status) {
      systemPtr->conservation_entropy();
      systemPtr->conservation_BSQ();
      systemPtr->conservation_energy(false, systemPtr->t);
    }
  }
  // the same step through the Evolver and an arbitrary derivative functional
  void advance_timestep_with_evolver(double dt, int rk_order)
  {
    evolver.execute_timestep(dt, rk_order, [this, dt] { this->get_time_derivatives(dt); });
    if (settingsPtr->particlization_enabled) freeze_out_particles();
    systemPtr->number_of_elapsed_timesteps++;
  }

  // SPHWorkstation::continue_evolution (src/sph_workstation.cpp:1992-2024)
  bool continue_evolution()
  {
    int keep = 0;
    call(ccake_continue_evolution(ev(), settingsPtr->max_tau, &keep), "continue_evolution");
    return keep != 0;
  }

  Evolver<D> evolver;
  std::shared_ptr<FreezeOut<D>> freezePtr;

 private:
  ccake_event* ev() const { return systemPtr->event(); }
  void call(int st, const char* what) { check(st, what, settingsPtr.get()); }
  std::shared_ptr<Settings> settingsPtr;
  std::shared_ptr<SystemState<D>> systemPtr;
};

}  // namespace ccake
