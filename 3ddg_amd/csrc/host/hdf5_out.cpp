#include "hdf5_out.hpp"

#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>

namespace dg3d_host {

namespace {
typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
const unsigned kAccTrunc = 0x0002u;  // H5F_ACC_TRUNC
const hid_t kDefault = 0;            // H5P_DEFAULT, H5S_ALL

struct Api {
  herr_t (*open)();
  hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
  herr_t (*sclose)(hid_t);
  hid_t (*dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
  herr_t (*dclose)(hid_t);
  hid_t* native_double;
};
Api g_api;
}  // namespace

Hdf5Writer::Hdf5Writer() {
  const char* names[] = {getenv("DG3D_HDF5_LIB"), "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
  for (const char* n : names) {
    if (!n) continue;
    h_ = dlopen(n, RTLD_NOW | RTLD_LOCAL);
    if (h_) break;
  }
  if (!h_) {
    err_ = "HDF5 library not found (set DG3D_HDF5_LIB)";
    return;
  }
  auto sym = [&](const char* s) { return dlsym(h_, s); };
  g_api.open = (herr_t(*)())sym("H5open");
  g_api.fcreate = (hid_t(*)(const char*, unsigned, hid_t, hid_t))sym("H5Fcreate");
  g_api.fclose = (herr_t(*)(hid_t))sym("H5Fclose");
  g_api.screate_simple = (hid_t(*)(int, const hsize_t*, const hsize_t*))sym("H5Screate_simple");
  g_api.sclose = (herr_t(*)(hid_t))sym("H5Sclose");
  g_api.dcreate2 = (hid_t(*)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t))sym("H5Dcreate2");
  g_api.dwrite = (herr_t(*)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*))sym("H5Dwrite");
  g_api.dclose = (herr_t(*)(hid_t))sym("H5Dclose");
  g_api.native_double = (hid_t*)sym("H5T_NATIVE_DOUBLE_g");
  if (!g_api.open || !g_api.fcreate || !g_api.fclose || !g_api.screate_simple || !g_api.sclose || !g_api.dcreate2 ||
      !g_api.dwrite || !g_api.dclose || !g_api.native_double) {
    err_ = "HDF5 library lacks the C API symbols";
    return;
  }
  if (g_api.open() < 0) {
    err_ = "H5open failed";
    return;
  }
  ok_ = true;
}

// Error codes follow HDF5Functions.cpp:296-320: -1 file, -2 dataset, -3 dataspace.
int Hdf5Writer::write(const std::string& path, const std::string& dataset, int rank, const hsize_t* dims,
                      const double* data, double time) {
  if (!ok_) return -1;
  const hid_t dbl = *g_api.native_double;
  hid_t f = g_api.fcreate(path.c_str(), kAccTrunc, kDefault, kDefault);
  if (f < 0) return -1;
  int rc = 0;
  hid_t s = g_api.screate_simple(rank, dims, nullptr);
  if (s < 0) rc = -3;
  if (!rc) {
    hid_t d = g_api.dcreate2(f, dataset.c_str(), dbl, s, kDefault, kDefault, kDefault);
    if (d < 0 || g_api.dwrite(d, dbl, kDefault, kDefault, kDefault, data) < 0) rc = -2;
    if (d >= 0) g_api.dclose(d);
    g_api.sclose(s);
  }
  if (!rc) {
    const hsize_t one = 1;
    hid_t s2 = g_api.screate_simple(1, &one, nullptr);
    hid_t d2 = (s2 >= 0) ? g_api.dcreate2(f, "time", dbl, s2, kDefault, kDefault, kDefault) : -1;
    if (d2 < 0 || g_api.dwrite(d2, dbl, kDefault, kDefault, kDefault, &time) < 0) rc = -2;
    if (d2 >= 0) g_api.dclose(d2);
    if (s2 >= 0) g_api.sclose(s2);
  }
  g_api.fclose(f);
  return rc;
}

}  // namespace dg3d_host
