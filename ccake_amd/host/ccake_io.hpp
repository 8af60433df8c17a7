// Initial-condition readers of the C++ call surface, for D = 2:
// BSQHydro::read_ccake (src/BSQHydro.cpp:160-425) and
// BSQHydro<2, EoM_default>::read_ICCING (83-139), dispatched by
// read_in_initial_conditions (37-71).  Every accepted point becomes a
// Particle<2> through sink.add_particle(p); the sink is a SystemState<2> or
// any collector with that member, so a file can be read without a device.
//
// Numbers are taken with std::istringstream extraction, as the reference
// takes them, and the per-line variables live across lines, so short or
// malformed lines behave as they do there: fields missing at the end of a
// line keep the previous line's values (the in-place unit conversions of
// the shear and bulk columns then apply again), a token that does not parse
// reads as 0 and ends the line, and the energy column starts from 0 on
// every line.
#pragma once

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "ccake.hpp"

namespace ccake {

namespace io_detail {

constexpr double kHbarcGeVfm = 0.1973269804;     // include/constants.h:15

[[noreturn]] inline void open_failed(const Settings& s, const std::string& file)
{
  const std::string msg = "File " + file + " could not be opened!";
  if (s.exit_on_error) { std::cerr << msg << std::endl; std::abort(); }
  throw std::runtime_error(msg);
}

}  // namespace io_detail

// ccake format: a header line "#<ignored> dx dy deta <ignored> xmin ymin etamin",
// then one point per line: x y eta e[GeV/fm^3] rhoB rhoS rhoQ ux uy ueta
// Pi[GeV/fm^3] pi^xx pi^xy pi^xeta pi^yy pi^yeta pi^etaeta [GeV/fm^3], plus
// the 12 diffusion components qB0..3 qS0..3 qQ0..3 with
// input_initial_diffusion.  Blank lines and '#' lines after the header are
// skipped; points with e <= e_cutoff are dropped.
template <class Sink>
void read_ccake(Settings& s, Sink& sink)
{
  using io_detail::kHbarcGeVfm;
  if (s.input_as_entropy) check(CCAKE_ERR_UNSUPPORTED, "read_ccake (input_as_entropy)", &s);
  std::ifstream in(s.IC_file);
  if (!in.is_open()) io_detail::open_failed(s, s.IC_file);
  double x = 0, y = 0, eta = 0, col4 = 0, e = 0, rB = 0, rS = 0, rQ = 0, ux = 0, uy = 0, ueta = 0, bulk = 0;
  double pxx = 0, pxy = 0, pxe = 0, pyy = 0, pye = 0, pee = 0;
  double q[3][4] = {};
  bool header_done = false;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    if (!header_done) {
      s.headers.push_back(line);
      double skip = 0, dx = 0, dy = 0, deta = 0, x0 = 0, y0 = 0, eta0 = 0;
      iss.ignore(256, '#');
      iss >> skip >> dx >> dy >> deta >> skip >> x0 >> y0 >> eta0;
      s.stepx = dx; s.stepy = dy; s.stepEta = deta;
      s.xmin = x0; s.ymin = y0; s.etamin = eta0;
      header_done = true;
      continue;
    }
    const size_t first = line.find_first_not_of(" \t\r\n");
    if (first == std::string::npos || line[first] == '#') continue;
    iss >> x >> y >> eta >> col4 >> rB >> rS >> rQ >> ux >> uy >> ueta >> bulk
        >> pxx >> pxy >> pxe >> pyy >> pye >> pee;
    if (s.input_initial_diffusion)
      for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 4; ++k) iss >> q[c][k];
    e = col4;
    col4 = 0.0;
    e /= kHbarcGeVfm;                      // fm^-4
    pxx /= kHbarcGeVfm; pxy /= kHbarcGeVfm; pxe /= kHbarcGeVfm;
    pyy /= kHbarcGeVfm; pye /= kHbarcGeVfm; pee /= kHbarcGeVfm;
    bulk /= kHbarcGeVfm;
    Particle<2> p;
    p.r[0] = x; p.r[1] = y;
    p.u[0] = ux; p.u[1] = uy;
    p.shv[1][1] = pxx;
    p.shv[1][2] = p.shv[2][1] = pxy;
    p.shv[1][3] = p.shv[3][1] = pxe;
    p.shv[2][2] = pyy;
    p.shv[2][3] = p.shv[3][2] = pye;
    p.shv[3][3] = pee;
    p.input_e = e;
    p.input_rhoB = s.baryon_charge_enabled ? rB : 0.0;
    p.input_rhoS = s.strange_charge_enabled ? rS : 0.0;
    p.input_rhoQ = s.electric_charge_enabled ? rQ : 0.0;
    p.Bulk = bulk;
    if (s.input_initial_diffusion) {
      const bool on[3] = { s.baryon_charge_enabled, s.strange_charge_enabled, s.electric_charge_enabled };
      for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 4; ++k) p.diffusion[c][k] = on[c] ? q[c][k] : 0.0;
    }
    if (e > s.e_cutoff / kHbarcGeVfm) sink.add_particle(p);
  }
}

// ICCING format (2+1D): a header line "<ignored> dx dy <ignored> xmin ymin",
// then x y e[GeV/fm^3] rhoB rhoS rhoQ per line, fluid at rest.  Every line
// after the header is a point (no comment skipping).
template <class Sink>
void read_ICCING(Settings& s, Sink& sink)
{
  using io_detail::kHbarcGeVfm;
  std::ifstream in(s.IC_file);
  if (!in.is_open()) io_detail::open_failed(s, s.IC_file);
  double x = 0, y = 0, e = 0, rB = 0, rS = 0, rQ = 0;
  bool header_done = false;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    if (!header_done) {
      s.headers.push_back(line);
      double skip = 0, dx = 0, dy = 0, x0 = 0, y0 = 0;
      iss >> skip >> dx >> dy >> skip >> x0 >> y0;
      s.stepx = dx; s.stepy = dy; s.xmin = x0; s.ymin = y0;
      header_done = true;
      continue;
    }
    iss >> x >> y >> e >> rB >> rS >> rQ;
    e /= kHbarcGeVfm;
    Particle<2> p;
    p.r[0] = x; p.r[1] = y;
    p.input_e = e;
    p.input_rhoB = s.baryon_charge_enabled ? rB : 0.0;
    p.input_rhoS = s.strange_charge_enabled ? rS : 0.0;
    p.input_rhoQ = s.electric_charge_enabled ? rQ : 0.0;
    if (e > s.e_cutoff / kHbarcGeVfm) sink.add_particle(p);
  }
}

// BSQHydro::read_in_initial_conditions (src/BSQHydro.cpp:37-71)
template <class Sink>
void read_in_initial_conditions(Settings& s, Sink& sink)
{
  if (s.IC_type == "ccake") read_ccake(s, sink);
  else if (s.IC_type == "ICCING") read_ICCING(s, sink);
  else check(CCAKE_ERR_UNSUPPORTED, ("read_in_initial_conditions (IC type " + s.IC_type + ")").c_str(), &s);
}

}  // namespace ccake
