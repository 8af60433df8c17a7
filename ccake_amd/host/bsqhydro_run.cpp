// The reference's driver sequence, BSQHydro::initialize_hydrodynamics and
// the loop of BSQHydro::run (src/BSQHydro.cpp:576-602, 745-797), written out
// call for call against the C++ call surface (ccake.hpp), on a ccake-format
// initial condition and the viscous BSQ settings of the lumpy workload.
//
// The event is run twice: once through the reference's phase-by-phase
// sequence (initialize, process_initial_conditions,
// allocate_cabana_particles, initialize_linklist,
// initialize_entropy_and_charge_densities, setup_freeze_out,
// calculate_gamma_and_velocities, initial_smoothing), once through the fused
// SPHWorkstation::initialize_hydrodynamics; both then step with
// continue_evolution / advance_timestep / print_freeze_out until every
// particle has frozen out or --max-steps (the CCAKE_MAX_STEPS cap of
// src/BSQHydro.cpp:740-766) is reached.  Prints one JSON line saying whether
// every field, every integer field and the freeze-out rows of the two runs
// are equal bit for bit.
//
//   bsqhydro_run --ic FILE [--table FILE] [--e-cutoff 0.15] [--max-steps N]
//                [--output DIR]
// --table: binary file of int32 dims[4], float64 mins[4], float64 maxs[4],
// then dims[0]*dims[1]*dims[2]*dims[3]*17 float64 (the 17 EoS fields per
// node, T slowest); without it the conformal EoS is used.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "ccake.hpp"
#include "ccake_io.hpp"

namespace {

using WS = ccake::SPHWorkstation<2, ccake::EoM_default>;
using SYS = ccake::SystemState<2>;

struct Run {
  std::shared_ptr<SYS> sys;
  std::unique_ptr<WS> ws;
  std::unique_ptr<ccake::Output<2, ccake::EoM_default>> out;
  int steps = 0;
};

bool load_table(const std::string& path, ccake::Settings& s)
{
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  int32_t dims[4];
  f.read(reinterpret_cast<char*>(dims), sizeof dims);
  f.read(reinterpret_cast<char*>(s.eos_mins), sizeof s.eos_mins);
  f.read(reinterpret_cast<char*>(s.eos_maxs), sizeof s.eos_maxs);
  size_t nn = 1;
  for (int a = 0; a < 4; ++a) { s.eos_dims[a] = dims[a]; nn *= (size_t)dims[a]; }
  s.eos_table.resize(nn * 17);
  f.read(reinterpret_cast<char*>(s.eos_table.data()), (std::streamsize)(s.eos_table.size() * sizeof(double)));
  s.eos_type = "table";
  return (bool)f;
}

// BSQHydro::run's loop (src/BSQHydro.cpp:745-797), the outputs it keeps
void evolve(Run& r, const ccake::Settings& s, int max_steps)
{
  while (r.ws->continue_evolution()) {
    r.ws->advance_timestep(s.dt, s.rk_order);
    ++r.steps;
    if (max_steps > 0 && r.steps >= max_steps) break;
    if (s.particlization_enabled)
      r.out->print_freeze_out(r.ws->freezePtr, r.sys->Btotal0, r.sys->Qtotal0, r.sys->Stotal0);
  }
}

Run make_run(std::shared_ptr<ccake::Settings> s)
{
  Run r;
  r.sys = std::make_shared<SYS>(s);
  ccake::read_in_initial_conditions(*s, *r.sys);
  r.ws = std::make_unique<WS>(s, r.sys);
  r.out = std::make_unique<ccake::Output<2, ccake::EoM_default>>(s, r.sys);
  return r;
}

}  // namespace

int main(int argc, char** argv)
{
  auto s = std::make_shared<ccake::Settings>();
  std::string table;
  int max_steps = 0;
  s->IC_type = "ccake";
  for (int k = 1; k + 1 < argc; k += 2) {
    if (!std::strcmp(argv[k], "--ic")) s->IC_file = argv[k + 1];
    else if (!std::strcmp(argv[k], "--table")) table = argv[k + 1];
    else if (!std::strcmp(argv[k], "--e-cutoff")) s->e_cutoff = std::atof(argv[k + 1]);
    else if (!std::strcmp(argv[k], "--max-steps")) max_steps = std::atoi(argv[k + 1]);
    else if (!std::strcmp(argv[k], "--output")) s->results_directory = argv[k + 1];
  }
  if (!table.empty() && !load_table(table, *s)) {
    std::fprintf(stderr, "bsqhydro_run: cannot read table %s\n", table.c_str());
    return 2;
  }
  // viscous BSQ settings of the lumpy workload
  const double hc = 0.1973269804;
  s->t0 = 0.6; s->dt = 0.02; s->hT = 0.3; s->rk_order = 2; s->max_tau = 30.0;
  s->particlization_enabled = true;
  s->Freeze_Out_Temperature = 150.0 / 197.3269804;
  s->using_shear = true;
  s->eta_mode = CCAKE_ETA_CONSTANT; s->constant_eta_over_s = 0.08;
  s->zeta_mode = CCAKE_ZETA_GAUSSIAN;
  s->zeta_gauss_A = 0.1; s->zeta_gauss_T0 = 0.170 / hc; s->zeta_gauss_width = 0.020 / hc;

  try {
    // (a) BSQHydro::initialize_hydrodynamics, src/BSQHydro.cpp:576-602
    Run a = make_run(s);
    a.ws->initialize();
    a.ws->process_initial_conditions();
    a.sys->allocate_cabana_particles();
    a.sys->initialize_linklist();
    a.ws->initialize_entropy_and_charge_densities();
    a.ws->setup_freeze_out();
    a.ws->calculate_gamma_and_velocities();
    a.ws->initial_smoothing();
    a.sys->conservation_entropy(true);
    a.sys->conservation_BSQ(true);
    evolve(a, *s, max_steps);

    // (b) the fused initialisation
    Run b = make_run(s);
    b.ws->initialize();
    b.ws->initialize_hydrodynamics();
    evolve(b, *s, max_steps);

    // compare every column of the two events bit for bit
    const int n = a.sys->n();
    std::vector<double> fa(n), fb(n);
    int fields_differ = 0, ints_differ = 0;
    for (int f = 0; f < CCAKE_NUM_FIELDS; ++f) {
      ccake::check(ccake_event_download(a.sys->event(), f, fa.data(), n), "download");
      ccake::check(ccake_event_download(b.sys->event(), f, fb.data(), n), "download");
      if (std::memcmp(fa.data(), fb.data(), n * sizeof(double)) != 0) {
        if (fields_differ < 5) std::fprintf(stderr, "field %d differs\n", f);
        ++fields_differ;
      }
    }
    std::vector<int32_t> ia(n), ib(n);
    for (int f = 0; f < CCAKE_NUM_IFIELDS; ++f) {
      ccake::check(ccake_event_download_int(a.sys->event(), f, ia.data(), n), "download_int");
      ccake::check(ccake_event_download_int(b.sys->event(), f, ib.data(), n), "download_int");
      if (ia != ib) ++ints_differ;
    }
    size_t na = 0, nb = 0;
    const std::vector<double> ra = a.ws->freezePtr->surface(&na), rb = b.ws->freezePtr->surface(&nb);
    const bool rows_equal = na == nb && ra == rb;
    int frozen = 0;
    ccake::check(ccake_event_download_int(a.sys->event(), CCAKE_I_FREEZE, ia.data(), n), "download_int");
    for (int i = 0; i < n; ++i) frozen += ia[i] >= 4;
    std::printf("{\"particles\": %d, \"steps_sequence\": %d, \"steps_fused\": %d, \"t_sequence\": %.17g, "
                "\"t_fused\": %.17g, \"fields_differing\": %d, \"int_fields_differing\": %d, \"rows\": %zu, "
                "\"rows_equal\": %s, \"frozen_out\": %d, \"freeze_out_lines_buffered\": %d}\n",
                n, a.steps, b.steps, a.sys->t, b.sys->t, fields_differ, ints_differ, na, rows_equal ? "true" : "false",
                frozen, a.out->particles_buffered());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "bsqhydro_run: %s\n", e.what());
    return 2;
  }
  return 0;
}
