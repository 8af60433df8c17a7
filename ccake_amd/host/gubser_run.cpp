// Ideal 2+1D Gubser flow through the C++ call surface (SystemState,
// SPHWorkstation, Evolver) on the gfx950 library, in the order
// BSQHydro::run uses (src/BSQHydro.cpp:568-866).  Prints one JSON line:
// the energy density against the analytic solution, entropy and energy
// conservation, and (with --evolver-steps K) the largest difference between
// K steps taken by the fused device step and by Evolver::execute_timestep
// with the workstation's get_time_derivatives as the functional.  With
// --freeze-out-T (MeV) particlization is on, the hypersurface is buffered
// after every step (Output::print_freeze_out, as BSQHydro::run does,
// src/BSQHydro.cpp:797) and written to OUTPUT/freeze_out.dat at the end (859-860).
// With --ic FILE the particles come from a ccake-format initial-condition
// file (BSQHydro::read_ccake) instead of the built-in lattice; with
// --hdf-evolution K a frame of OUTPUT/system_state.h5 every K steps (and at the
//                  start; with --txt-evolution the text files share its stride)
// --txt-evolution K the system state is written to OUTPUT/system_state_<k>.dat
// after initialisation and every K steps (src/BSQHydro.cpp:689-691, 791-796).
//
// With --rk 4 every step is the reference's fourth-order step
// (Evolver::advance_timestep_rk4, src/evolver.cpp:519-561).
//
//   gubser_run [--spacing 0.05] [--steps 120] [--dt 0.005] [--evolver-steps 0] [--rk 2]
//              [--freeze-out-T MeV] [--ic FILE --e-cutoff GeV/fm^3]
//              [--txt-evolution K] [--hdf-evolution K] [--output DIR]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "ccake.hpp"
#include "ccake_io.hpp"

namespace {

constexpr double kHbarc = 0.1973269804;   // GeV fm

// Ideal Gubser energy density (fm^-4) and radial four-velocity, q = 1/fm,
// eps0 = 80 GeV/fm^3 (utilities/ideal_gubser_ic_generator.py:23-38).
double gubser_e(double tau, double r, double q = 1.0, double eps0 = 80.0)
{
  const double den = 1.0 + 2.0 * q * q * (tau * tau + r * r) + std::pow(q, 4) * std::pow(tau * tau - r * r, 2);
  return (eps0 / kHbarc) * std::pow(2.0 * q, 8.0 / 3.0) / (std::pow(tau, 4.0 / 3.0) * std::pow(den, 4.0 / 3.0));
}
double gubser_ur(double tau, double r, double q = 1.0)
{
  return std::sinh(std::atanh(2.0 * q * q * tau * r / (1.0 + q * q * tau * tau + q * q * r * r)));
}

// lattice coordinate -L + a i, rounded to 12 decimals as read from a text file
double coord(double L, double a, int i)
{
  const double v = -L + a * i;
  return std::nearbyint(v * 1e12) / 1e12 + 0.0;
}

std::shared_ptr<ccake::SystemState<2>> make_system(std::shared_ptr<ccake::Settings> s, double a)
{
  auto sys = std::make_shared<ccake::SystemState<2>>(s);
  const double L = 5.0;
  const int n = (int)std::lround(2 * L / a) + 1;
  for (int iy = 0; iy < n; ++iy)
    for (int ix = 0; ix < n; ++ix) {
      ccake::Particle<2> p;
      p.r[0] = coord(L, a, ix);
      p.r[1] = coord(L, a, iy);
      const double r = std::hypot(p.r[0], p.r[1]);
      p.input_e = gubser_e(s->t0, r);
      const double ur = gubser_ur(s->t0, r);
      p.u[0] = r > 0 ? ur * p.r[0] / r : 0.0;
      p.u[1] = r > 0 ? ur * p.r[1] / r : 0.0;
      sys->add_particle(p);
    }
  return sys;
}

}  // namespace

int main(int argc, char** argv)
{
  double a = 0.05;
  int steps = 120, evolver_steps = 0, rk = 2;
  double dt = 0.005, t_fo_mev = 0.0, e_cutoff = 0.15;
  int txt_stride = 0, hdf_stride = 0;
  std::string out_dir = ".", ic_file;
  for (int k = 1; k + 1 < argc; k += 2) {
    if (!std::strcmp(argv[k], "--spacing")) a = std::atof(argv[k + 1]);
    else if (!std::strcmp(argv[k], "--steps")) steps = std::atoi(argv[k + 1]);
    else if (!std::strcmp(argv[k], "--dt")) dt = std::atof(argv[k + 1]);
    else if (!std::strcmp(argv[k], "--evolver-steps")) evolver_steps = std::atoi(argv[k + 1]);
    else if (!std::strcmp(argv[k], "--freeze-out-T")) t_fo_mev = std::atof(argv[k + 1]);
    else if (!std::strcmp(argv[k], "--output")) out_dir = argv[k + 1];
    else if (!std::strcmp(argv[k], "--ic")) ic_file = argv[k + 1];
    else if (!std::strcmp(argv[k], "--e-cutoff")) e_cutoff = std::atof(argv[k + 1]);
    else if (!std::strcmp(argv[k], "--txt-evolution")) txt_stride = std::atoi(argv[k + 1]);
    else if (!std::strcmp(argv[k], "--hdf-evolution")) hdf_stride = std::atoi(argv[k + 1]);
    else if (!std::strcmp(argv[k], "--rk")) rk = std::atoi(argv[k + 1]);
  }
  auto s = std::make_shared<ccake::Settings>();
  s->eos_type = "conformal";
  s->t0 = 1.0; s->dt = dt; s->hT = 0.3; s->rk_order = rk;
  s->particlization_enabled = t_fo_mev > 0.0;
  if (s->particlization_enabled) s->Freeze_Out_Temperature = t_fo_mev / 197.3269804;
  s->using_shear = false;
  s->eta_mode = CCAKE_ETA_CONSTANT; s->constant_eta_over_s = 0.0;
  s->zeta_mode = CCAKE_ZETA_CONSTANT; s->constant_zeta_over_s = 0.0;
  s->results_directory = out_dir;
  s->txt_evolution = txt_stride > 0;
  s->hdf_evolution = hdf_stride > 0;
  const int out_stride = txt_stride > 0 ? txt_stride : hdf_stride;
  s->IC_file = ic_file;
  s->e_cutoff = e_cutoff;

  try {
    std::shared_ptr<ccake::SystemState<2>> sys;
    if (ic_file.empty()) {
      sys = make_system(s, a);
    } else {
      sys = std::make_shared<ccake::SystemState<2>>(s);
      ccake::read_in_initial_conditions(*s, *sys);
    }
    ccake::SPHWorkstation<2, ccake::EoM_default> ws(s, sys);
    ws.initialize();
    ws.initialize_hydrodynamics();
    ccake::Output<2, ccake::EoM_default> out(s, sys);
    if (out_stride > 0) out.print_system_state();
    for (int k = 0; k < steps; ++k) {
      ws.advance_timestep(s->dt, s->rk_order);
      if (out_stride > 0 && (k + 1) % out_stride == 0) out.print_system_state();
      if (s->particlization_enabled)
        out.print_freeze_out(ws.freezePtr, sys->Btotal0, sys->Qtotal0, sys->Stotal0);
    }
    if (s->particlization_enabled) out.flush_freeze_out();
    sys->conservation_entropy();
    sys->conservation_energy();
    sys->copy_device_to_host();
    double worst = 0.0;
    for (const auto& p : sys->particles) {
      const double r = std::hypot(p.r[0], p.r[1]);
      if (r >= 2.5) continue;
      const double ea = gubser_e(sys->t, r);
      worst = std::max(worst, std::fabs(p.e() - ea) / ea);
    }

    double evolver_diff = -1.0;
    if (evolver_steps > 0) {
      auto s1 = make_system(s, a), s2 = make_system(s, a);
      ccake::SPHWorkstation<2, ccake::EoM_default> w1(s, s1), w2(s, s2);
      w1.initialize(); w1.initialize_hydrodynamics();
      w2.initialize(); w2.initialize_hydrodynamics();
      for (int k = 0; k < evolver_steps; ++k) {
        w1.advance_timestep(s->dt, s->rk_order);
        w2.advance_timestep_with_evolver(s->dt, s->rk_order);
      }
      s1->copy_device_to_host();
      s2->copy_device_to_host();
      evolver_diff = std::fabs(s1->t - s2->t);
      for (int i = 0; i < s1->n(); ++i) {
        const auto &p = s1->particles[i], &q = s2->particles[i];
        evolver_diff = std::max({ evolver_diff, std::fabs(p.e() - q.e()), std::fabs(p.u[0] - q.u[0]),
                                  std::fabs(p.u[1] - q.u[1]), std::fabs(p.r[0] - q.r[0]), std::fabs(p.s_ext - q.s_ext) });
      }
    }
    std::printf("{\"particles\": %d, \"steps\": %d, \"t\": %.15g, \"max_rel_err_e_r_lt_2.5\": %.6g, "
                "\"entropy_rel_change\": %.6g, \"energy_rel_change\": %.6g, \"evolver_vs_fused_max_abs_diff\": %.6g, "
                "\"freeze_out_rows\": %d}\n",
                sys->n(), steps, sys->t, worst, sys->S / sys->S0 - 1.0, (sys->Etot) / (sys->E0) - 1.0, evolver_diff,
                out.particles_buffered());
  } catch (const ccake::Error& e) {
    std::fprintf(stderr, "gubser_run: %s\n", e.what());
    return 2;
  }
  return 0;
}
