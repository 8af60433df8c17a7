// Definitions of ccake_hdf5.hpp on the HDF5 C API (built into
// libccake_hdf5.so).
#include "ccake_hdf5.hpp"

#include <hdf5.h>

#include <algorithm>
#include <cstdio>
#include <stdexcept>

namespace ccake {

static void double_attribute(hid_t obj, const std::string& name, double value)
{
  const hsize_t one = 1;
  const hid_t sp = H5Screate_simple(1, &one, nullptr);
  const hid_t at = H5Acreate2(obj, name.c_str(), H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
  if (at < 0 || H5Awrite(at, H5T_NATIVE_DOUBLE, &value) < 0) throw std::runtime_error("Could not write attribute " + name);
  H5Aclose(at);
  H5Sclose(sp);
}

static void string_attribute(hid_t obj, const std::string& name, const std::string& value)
{
  const hid_t ty = H5Tcopy(H5T_C_S1);
  H5Tset_size(ty, H5T_VARIABLE);
  const hid_t sp = H5Screate(H5S_SCALAR);
  const hid_t at = H5Acreate2(obj, name.c_str(), ty, sp, H5P_DEFAULT, H5P_DEFAULT);
  const char* s = value.c_str();
  if (at < 0 || H5Awrite(at, ty, &s) < 0) throw std::runtime_error("Could not write attribute " + name);
  H5Aclose(at);
  H5Sclose(sp);
  H5Tclose(ty);
}

void Hdf5SystemState::initialize(const std::string& filename, const std::vector<double>& parameters,
                                 const std::vector<std::string>& names)
{
  close();
  const hid_t f = H5Fcreate(filename.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) throw std::runtime_error("Could not create " + filename);
  file_ = f;
  const hid_t g = H5Gcreate2(f, "/Event", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) throw std::runtime_error("Could not create /Event in " + filename);
  for (size_t k = 0; k < parameters.size(); ++k) double_attribute(g, names[k], parameters[k]);
  H5Gclose(g);
}

void Hdf5SystemState::output_dataset(const std::vector<std::string>& names, const std::vector<std::string>& units,
                                     const std::vector<std::vector<double>>& v, int width, int index,
                                     const std::vector<int>& labels, const std::vector<double>& parameters,
                                     const std::vector<std::string>& parameter_names)
{
  if (file_ < 0) throw std::runtime_error("HDF5 system state: initialize first");
  const hid_t file = (hid_t)file_;
  char pad[32];
  std::snprintf(pad, sizeof(pad), "%0*d", width > 0 ? width : 1, index);
  const std::string frame = std::string("/Event/Frame_") + pad;
  const hid_t g = H5Gcreate2(file, frame.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) throw std::runtime_error("Could not create " + frame);
  for (size_t k = 0; k < parameters.size(); ++k) double_attribute(g, parameter_names[k], parameters[k]);
  const hsize_t len = v.empty() ? 0 : (hsize_t)v[0].size();
  const hid_t space = H5Screate_simple(1, &len, nullptr);
  for (size_t k = 0; k < names.size(); ++k) {
    const hid_t ds = H5Dcreate2(g, names[k].c_str(), H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) throw std::runtime_error("Could not create " + frame + "/" + names[k]);
    string_attribute(ds, "Units", units[k]);
    if (len && H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v[k].data()) < 0)
      throw std::runtime_error("Could not write " + frame + "/" + names[k]);
    H5Dclose(ds);
  }
  const hid_t ds = H5Dcreate2(g, "labels", H5T_NATIVE_INT, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (ds < 0) throw std::runtime_error("Could not create " + frame + "/labels");
  if (len && H5Dwrite(ds, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, labels.data()) < 0)
    throw std::runtime_error("Could not write " + frame + "/labels");
  H5Dclose(ds);
  H5Sclose(space);
  H5Gclose(g);
  H5Fflush(file, H5F_SCOPE_GLOBAL);
}

void Hdf5SystemState::close()
{
  if (file_ >= 0) H5Fclose((hid_t)file_);
  file_ = -1;
}

// The flat (s, nB) table: *Ns, *NB and the 12 arrays, MeV units, s slowest.
std::vector<double> load_flat_table_hdf5(const std::string& path, int* Ns, int* NB)
{
  const hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) throw std::runtime_error("Could not open " + path);
  struct Closer {
    hid_t f;
    ~Closer() { H5Fclose(f); }
  } closer{ f };
  const hid_t dd = H5Dopen2(f, "dimensions", H5P_DEFAULT);
  if (dd < 0) throw std::runtime_error(path + ": no /dimensions");
  const hid_t dsp = H5Dget_space(dd);
  hsize_t rank_len = 0;
  const int rank = H5Sget_simple_extent_dims(dsp, &rank_len, nullptr);
  H5Sclose(dsp);
  if (rank != 1 || rank_len != 2) {
    H5Dclose(dd);
    throw std::runtime_error(path + ": /dimensions must hold (Ns, NB)");
  }
  int dims[2] = { 0, 0 };
  const herr_t rd = H5Dread(dd, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, dims);
  H5Dclose(dd);
  if (rd < 0 || dims[0] < 2 || dims[1] < 2) throw std::runtime_error(path + ": bad /dimensions");
  *Ns = dims[0];
  *NB = dims[1];
  const size_t nt = (size_t)dims[0] * dims[1];
  static const char* names[12] = { "s", "nB", "T", "muB", "muS", "muQ", "e", "p", "cs2", "chiTT", "chiTB", "chiBB" };
  std::vector<double> table(12 * nt);
  for (int k = 0; k < 12; ++k) {
    const std::string name = std::string("data/") + names[k];
    const hid_t ds = H5Dopen2(f, name.c_str(), H5P_DEFAULT);
    if (ds < 0) throw std::runtime_error(path + ": no /" + name);
    const hid_t sp = H5Dget_space(ds);
    const hssize_t npts = H5Sget_simple_extent_npoints(sp);
    H5Sclose(sp);
    if (npts != (hssize_t)nt) {
      H5Dclose(ds);
      throw std::runtime_error(path + ": /" + name + " does not hold Ns x NB values");
    }
    const herr_t r = H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, table.data() + k * nt);
    H5Dclose(ds);
    if (r < 0) throw std::runtime_error(path + ": could not read /" + name);
  }
  return table;
}

}  // namespace ccake

// C entry for callers outside C++ (the Python front-end): the flat table's
// 12 arrays into `out` (capacity `cap` doubles; with out null or cap too
// small only *Ns and *NB are set).  Returns 0, 1 for a file that cannot be
// read (message in `err`, up to err_len bytes), 7 for a short buffer.
extern "C" __attribute__((visibility("default"))) int ccake_h5_load_flat_table(const char* path, double* out,
                                                                             size_t cap, int* Ns, int* NB,
                                                                             char* err, size_t err_len)
{
  try {
    const std::vector<double> t = ccake::load_flat_table_hdf5(path, Ns, NB);
    if (!out || cap < t.size()) return 7;
    std::copy(t.begin(), t.end(), out);
    return 0;
  } catch (const std::exception& e) {
    if (err && err_len) std::snprintf(err, err_len, "%s", e.what());
    return 1;
  }
}
