// Reads an initial-condition file with the C++ readers (ccake_io.hpp) and
// prints what they produced, without touching a device: a line with the
// header's grid values, the particle count, then one line per particle:
// x y e[fm^-4] rhoB rhoS rhoQ ux uy Pi[fm^-4], the 16 pi^{mu nu} [fm^-4]
// and the 12 diffusion components, each with 17 significant digits.
//
//   ic_dump --type ccake|ICCING --file F [--e-cutoff 0.15] [--no-baryon]
//           [--no-strange] [--no-electric] [--diffusion]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ccake.hpp"
#include "ccake_io.hpp"

namespace {
struct Collect {
  std::vector<ccake::Particle<2>> v;
  void add_particle(const ccake::Particle<2>& p) { v.push_back(p); }
};
}  // namespace

int main(int argc, char** argv)
{
  ccake::Settings s;
  for (int k = 1; k < argc; ++k) {
    if (!std::strcmp(argv[k], "--type") && k + 1 < argc) s.IC_type = argv[++k];
    else if (!std::strcmp(argv[k], "--file") && k + 1 < argc) s.IC_file = argv[++k];
    else if (!std::strcmp(argv[k], "--e-cutoff") && k + 1 < argc) s.e_cutoff = std::atof(argv[++k]);
    else if (!std::strcmp(argv[k], "--no-baryon")) s.baryon_charge_enabled = false;
    else if (!std::strcmp(argv[k], "--no-strange")) s.strange_charge_enabled = false;
    else if (!std::strcmp(argv[k], "--no-electric")) s.electric_charge_enabled = false;
    else if (!std::strcmp(argv[k], "--diffusion")) s.input_initial_diffusion = true;
  }
  Collect c;
  try {
    ccake::read_in_initial_conditions(s, c);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ic_dump: %s\n", e.what());
    return 2;
  }
  std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", s.stepx, s.stepy, s.stepEta, s.xmin, s.ymin, s.etamin);
  std::printf("%zu\n", c.v.size());
  for (const auto& p : c.v) {
    std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g", p.r[0], p.r[1], p.input_e, p.input_rhoB,
                p.input_rhoS, p.input_rhoQ, p.u[0], p.u[1], p.Bulk);
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) std::printf(" %.17g", p.shv[a][b]);
    for (int q = 0; q < 3; ++q)
      for (int k = 0; k < 4; ++k) std::printf(" %.17g", p.diffusion[q][k]);
    std::printf("\n");
  }
  return 0;
}
