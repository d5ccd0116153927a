// HDF5 output in the reference's layout (HDF5Functions.cpp:34-397): one file per dataset and
// block ("rank"), native-double datasets [X][Y][Z] (scalars) or [3][X][Y][Z] (vectors) with
// X = elements*(N+1) in element-major node order, plus a one-value "time" dataset.
// The HDF5 C library is loaded at run time (no link-time dependency); writing is skipped with
// a message when it is not available.
#pragma once
#include <string>
#include <vector>

namespace dg3d_host {

class Hdf5Writer {
 public:
  Hdf5Writer();
  bool available() const { return ok_; }
  const std::string& error() const { return err_; }
  // data holds dims[0]*...*dims[rank-1] doubles in row-major order
  int write(const std::string& path, const std::string& dataset, int rank, const unsigned long long* dims,
            const double* data, double time);

 private:
  bool ok_ = false;
  std::string err_;
  void* h_ = nullptr;
};

}  // namespace dg3d_host
