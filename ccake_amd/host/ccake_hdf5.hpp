// HDF5 input and output of the C++ call surface (host code; HDF5 C API).
//
//  - Hdf5SystemState: the system-state file of Output::print_system_state_to_HDF
//    (src/output.cpp:213-266) written through interface_to_HDF5
//    (include/interface_to_HDF5.h:44-176): output_directory/system_state.h5
//    holds a group /Event with the run's double attributes (h, e_cutoff, set
//    in the Output constructor, src/output.cpp:33-47) and, per call, a group
//    /Event/Frame_<k> (k zero-padded to ceil(log10(ceil(max_tau/dt)))
//    digits) with the double attributes Time, e_2_X, e_2_P (t, -10, -10) and
//    one dataset per field -- x, y (fm), T, muB, muS, muQ (MeV), e
//    (MeV/fm^3), s, B, S, Q (1/fm^3); native doubles with a variable-length
//    string attribute "Units" -- and an int dataset "labels" holding each
//    particle's EoS tag.
//  - load_flat_table_hdf5: the flat (s, nB) pre-inverted table file that
//    EoS_Interpolator::load_flat_table reads (src/eos_interpolator.cpp:291-527):
//    /dimensions = [Ns, NB] (int), /data/<name> (Ns x NB doubles, s slowest)
//    for s, nB, T, muB, muS, muQ, e, p, cs2, chiTT, chiTB, chiBB in MeV
//    units, returned as the 12 arrays of a CCAKE_EOS_PREINVERTED
//    description (include/ccake_amd.h) in that order.
// Declarations only: the definitions (ccake_hdf5.cpp) are built into
// libccake_hdf5.so against the HDF5 C library, so that a program using them
// needs neither hdf5.h nor a run-time search path to HDF5's own directory.
// Used when CCAKE_WITH_HDF5 is defined.
#pragma once
#include <string>
#include <vector>

namespace ccake {

class Hdf5SystemState {
 public:
  Hdf5SystemState() = default;
  Hdf5SystemState(const Hdf5SystemState&) = delete;
  Hdf5SystemState& operator=(const Hdf5SystemState&) = delete;
  ~Hdf5SystemState() { close(); }

  // interface_to_HDF5::initialize: create (truncate) the file and /Event
  // with the run's attributes.
  void initialize(const std::string& filename, const std::vector<double>& parameters,
                  const std::vector<std::string>& names);
  // interface_to_HDF5::output_dataset: one frame.
  void output_dataset(const std::vector<std::string>& names, const std::vector<std::string>& units,
                      const std::vector<std::vector<double>>& v, int width, int index, const std::vector<int>& labels,
                      const std::vector<double>& parameters, const std::vector<std::string>& parameter_names);
  void close();

 private:
  long long file_ = -1;     // hid_t of the open file, -1 when none
};

// The flat (s, nB) table: *Ns, *NB and the 12 arrays, MeV units, s slowest.
std::vector<double> load_flat_table_hdf5(const std::string& path, int* Ns, int* NB);

}  // namespace ccake
