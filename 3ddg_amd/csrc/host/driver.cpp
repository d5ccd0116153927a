// Driver with the reference's command line, parameter file, log file, progress output and
// HDF5 output layout (driver.cpp:70-291), running the block on one GPU through the C ABI.
//   dg3d_driver PARAMETER_FILENAME OUTPUT_FOLDER
// NUM_BLOCKS > 1 runs the union of the blocks as one device block and writes one file set per
// block with the block's index as the file "rank", as the reference's ranks do.
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits.h>
#include <string>
#include <vector>

#include "facade.hpp"
#include "hdf5_out.hpp"
#include "params.hpp"

using namespace dg3d_host;

namespace {

const int ERRCODE_SUCESSFUL_COMPLETION = 0;
const int ERRCODE_INVALID_ARGUMENT = 1;
const int ERRCODE_INVALID_PARAMETER = 2;
const int ERRCODE_CANNOT_READ = 3;
const int ERRCODE_CANNOT_WRITE = 4;
// Not in the reference: a HIP runtime or device failure, or an allocation failure, after the
// inputs were read (the reference has no device and maps nothing to it)
const int ERRCODE_DEVICE_FAILURE = 5;
const char* NEEDED_ARGUMENTS =
    "	PARAMETER_FILENAME : path to the file defining simulation parameters\n"
    "	OUTPUT_FILENAME : path to the file that the data will be written to";

std::string exe_dir() {
  char buf[PATH_MAX];
  const ssize_t len = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
  if (len <= 0) return ".";
  buf[len] = 0;
  std::string s(buf);
  const size_t p = s.rfind('/');
  return p == std::string::npos ? "." : s.substr(0, p);
}

struct Output {
  Hdf5Writer h5;
  std::string folder;
  const Mesh* m;
  const Params* p;

  // Scatter a field-major register copy [nf][E][n^3] into one block's [X][Y][Z] arrays
  // (HDF5Functions.cpp:64-89, 173-205).
  void block_arrays(const std::vector<double>& st, int nf, int f, int bi, int bj, int bk, std::vector<double>* out,
                    unsigned long long* dims) const {
    const int n = m->n();
    const int nb[3] = {(int)p->NUM_ELEMENTS[0], (int)p->NUM_ELEMENTS[1], (int)p->NUM_ELEMENTS[2]};
    dims[0] = (unsigned long long)nb[0] * n;
    dims[1] = (unsigned long long)nb[1] * n;
    dims[2] = (unsigned long long)nb[2] * n;
    out->assign(dims[0] * dims[1] * dims[2], 0.0);
    const int64_t np = (int64_t)n * n * n, fs = m->E() * np;
    (void)nf;
    for (int ei = 0; ei < nb[0]; ++ei)
      for (int ej = 0; ej < nb[1]; ++ej)
        for (int ek = 0; ek < nb[2]; ++ek) {
          const int64_t ge = ((int64_t)(bi * nb[0] + ei) * m->nE[1] + (bj * nb[1] + ej)) * m->nE[2] + (bk * nb[2] + ek);
          for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b)
              for (int c = 0; c < n; ++c) {
                const size_t ii = (size_t)ei * n + a, jj = (size_t)ej * n + b, kk = (size_t)ek * n + c;
                (*out)[(ii * dims[1] + jj) * dims[2] + kk] = st[f * fs + ge * np + (a * n + b) * n + c];
              }
        }
  }

  std::string name(const char* what, int rank, long step) const {
    return folder + "/raw/" + what + "_" + std::to_string(rank) + "_" + std::to_string(step) + ".h5";
  }

  void dump_state(const std::vector<double>& st, long step, double time) {
    for (int bi = 0; bi < m->nB[0]; ++bi)
      for (int bj = 0; bj < m->nB[1]; ++bj)
        for (int bk = 0; bk < m->nB[2]; ++bk) {
          const int rank = (bi * m->nB[1] + bj) * m->nB[2] + bk;
          std::vector<double> rho, vel, tmp;
          unsigned long long d[3];
          block_arrays(st, 4, 0, bi, bj, bk, &rho, d);
          double maxDRho = 0.0;
          for (double v : rho)
            if (std::fabs(v) > maxDRho || std::isnan(v)) maxDRho = std::fabs(v);
          vel.reserve(3 * rho.size());
          for (int f = 1; f <= 3; ++f) {
            block_arrays(st, 4, f, bi, bj, bk, &tmp, d);
            vel.insert(vel.end(), tmp.begin(), tmp.end());
          }
          const std::string f1 = name("dRho", rank, step), f2 = name("vel", rank, step);
          std::cout << " - Printing to " << f1 << "; Max dRho=" << maxDRho << "\n";
          if (!h5.available()) continue;
          const unsigned long long d4[4] = {3, d[0], d[1], d[2]};
          if (h5.write(f1, "rho", 3, d, rho.data(), time) || h5.write(f2, "velocity", 4, d4, vel.data(), time))
            std::cout << "HDF5 FileIException error\n";
        }
  }

  void dump_grid(const std::vector<double>& coords, const std::vector<double>& mat, double time) {
    // coords: [3][E][n^3]; mat: [4][E][n^3] = ba, diff, rho0, c0 broadcast to nodes
    for (int bi = 0; bi < m->nB[0]; ++bi)
      for (int bj = 0; bj < m->nB[1]; ++bj)
        for (int bk = 0; bk < m->nB[2]; ++bk) {
          const int rank = (bi * m->nB[1] + bj) * m->nB[2] + bk;
          std::vector<double> xyz, tmp;
          unsigned long long d[3];
          for (int f = 0; f < 3; ++f) {
            block_arrays(coords, 3, f, bi, bj, bk, &tmp, d);
            xyz.insert(xyz.end(), tmp.begin(), tmp.end());
          }
          std::cout << " - Printing to " << name("coords", rank, 0) << "\n";
          if (!h5.available()) continue;
          const unsigned long long d4[4] = {3, d[0], d[1], d[2]};
          h5.write(name("coords", rank, 0), "coords", 4, d4, xyz.data(), time);
          const char* names[4] = {"ba", "diff", "rho0", "c0"};
          for (int f = 0; f < 4; ++f) {
            block_arrays(mat, 4, f, bi, bj, bk, &tmp, d);
            h5.write(name(names[f], rank, 0), names[f], 3, d, tmp.data(), time);
          }
        }
  }
};

}  // namespace

int main(int argc, char* argv[]) {
  std::cout << " - Thread 0 created\n";
  if (argc != 3) {
    std::cerr << "Not enough command line arguments.\n" << NEEDED_ARGUMENTS;
    return ERRCODE_INVALID_ARGUMENT;
  }
  const std::string parameterFilename = argv[1], outputFolder = argv[2];
  const std::string logFilename = outputFolder + "/_log.txt";
  std::cout << " - Parameters from   : " << parameterFilename << "\n";
  std::cout << " - Log will be written to : " << logFilename << "\n";
  std::cout << " - Results in same folder as log\n";
  std::cout << "Setting up I/O\n - Checking Folder\n";
  mkdir(outputFolder.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
  mkdir((outputFolder + "/raw").c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
  std::cout << " - Preparing to write\n";
  std::ofstream logFile(logFilename);
  if (!logFile.is_open()) {
    std::cerr << "Unable to open log file " << logFilename << "\n";
    return ERRCODE_CANNOT_WRITE;
  }
  std::cout << " - Reading parameters\n";
  Params p;
  if (!parse_parameters(parameterFilename, &p, logFile, std::cerr)) {
    std::cerr << "Unable to open parameter file " << parameterFilename << "\n";
    return ERRCODE_CANNOT_READ;
  }
  logFile.close();
  if (p.ORDER < 1 || p.ORDER > 7) {
    std::cerr << "ORDER must be 1..7 on this implementation\n";
    return ERRCODE_INVALID_PARAMETER;
  }
  std::cout << " - Reading basis\n";
  Mesh mesh;
  build_mesh(p, &mesh);
  std::string err;
  if (!load_basis(p.ORDER, {".", exe_dir() + "/data"}, &mesh, &err)) {
    std::cerr << err << "\n";
    return ERRCODE_CANNOT_READ;
  }
  const unsigned NUM_TIMESTEPS = p.NUM_OUTSTEPS * p.NUM_TIME_SUBSTEPS;
  std::cout << "Creating mesh\n";
  // SCHEME = EULER runs Grid::forwardEuler, which exists only on the Block-operator path
  const bool fused = p.FUSED && p.eval_advect && p.SCHEME != 2;
  if (!fused && p.SCHEME == 1) {
    std::cerr << "SCHEME = LSRK54 runs on the fused path only (advection on)\n";
    return ERRCODE_INVALID_PARAMETER;
  }
  int periodic[3] = {p.IS_PERIODIC[0], p.IS_PERIODIC[1], p.IS_PERIODIC[2]};
  try {
    Solver solver(mesh, periodic, p.PRECISION, fused ? 1 : 4, p.eval_diff || p.SCHEME == 2);
    std::vector<double> c0, rho0, ba, sigma, diff, st;
    material_from_params(p, mesh, &c0, &rho0, &ba, &sigma, &diff);
    // D enters the stages only where the reference evaluates diffusion: Grid::RK33 under
    // EVALUATE_DIFFUSION (Grid.cpp:655), Grid::forwardEuler always; the grid dump keeps DIFF_CONST
    const std::vector<double> d_used = (p.eval_diff || p.SCHEME == 2) ? diff : std::vector<double>(diff.size(), 0.0);
    solver.check(dg3d_set_material(solver.ctx(), c0.data(), rho0.data(), ba.data(), sigma.data(), d_used.data()),
                 "set_material");
    std::vector<double> traces[6];
    gaussian_ic(p, mesh, &st, traces);
    solver.check(dg3d_set_state(solver.ctx(), 0, st.data(), 4), "set_state");
    for (int f = 0; f < 6; ++f)
      if (!traces[f].empty() && !periodic[2 - f / 2])
        solver.check(dg3d_set_halo(solver.ctx(), 0, f, traces[f].data()), "set_halo");
    if (p.SOURCE_FACE >= 0) {
      dg3d_source s{p.SOURCE_FACE, p.SOURCE[0], p.SOURCE[1], p.SOURCE[2], p.SOURCE[3], p.SOURCE[4], nullptr};
      solver.check(dg3d_set_source(solver.ctx(), &s), "set_source");
    }
    const double DT = p.CFL > 0 ? cfl_dt(mesh, p.C0, p.CFL) : (double)p.DT;

    Output out{Hdf5Writer(), outputFolder, &mesh, &p};
    if (p.BOOL_FILE && !out.h5.available()) std::cout << " - HDF5 output disabled: " << out.h5.error() << "\n";
    const int n = mesh.n();
    const int64_t np = (int64_t)n * n * n, fs = mesh.E() * np;
    if (p.BOOL_FILE) {
      std::cout << " - Logging grid\n";
      std::vector<double> coords(3 * fs), mat(4 * fs);
      std::vector<double> ec[3];
      for (int a = 0; a < 3; ++a) {
        double acc = mesh.corner[a];
        for (int i = 0; i < mesh.nE[a]; ++i) {
          ec[a].push_back(acc);
          acc += mesh.width[a][i];
        }
      }
      for (int i = 0; i < mesh.nE[0]; ++i)
        for (int j = 0; j < mesh.nE[1]; ++j)
          for (int k = 0; k < mesh.nE[2]; ++k) {
            const int64_t e = ((int64_t)i * mesh.nE[1] + j) * mesh.nE[2] + k;
            for (int nd = 0; nd < np; ++nd) {
              const int idx[3] = {nd / (n * n), (nd / n) % n, nd % n};
              const int eidx[3] = {i, j, k};
              for (int a = 0; a < 3; ++a)
                coords[a * fs + e * np + nd] = mesh.width[a][eidx[a]] * (mesh.nodes[idx[a]] / 2.0 + 0.5) + ec[a][eidx[a]];
              mat[0 * fs + e * np + nd] = ba[e];
              mat[1 * fs + e * np + nd] = diff[e];
              mat[2 * fs + e * np + nd] = rho0[e];
              mat[3 * fs + e * np + nd] = c0[e];
            }
          }
      out.dump_grid(coords, mat, 0.0);
      std::cout << " - Logging initial conditions\n";
      out.dump_state(st, 0, 0.0);
    }

    Grid g0(&solver, 0, p.eval_advect, p.eval_diff, p.eval_sponge);
    Grid g1(&solver, fused ? 0 : 1, p.eval_advect, p.eval_diff, p.eval_sponge);
    Grid k1(&solver, fused ? 0 : 2, p.eval_advect, p.eval_diff, p.eval_sponge);
    Grid k2(&solver, fused ? 0 : 3, p.eval_advect, p.eval_diff, p.eval_sponge);
    if (!fused) g1.copyFrom(g0);  // Grid myGrid2(myGrid) (driver.cpp:211)
    Grid* oldGrid = &g0;
    Grid* newGrid = &g1;
    Grid* swapGrid = &g0;
    double t = 0.0;
    // the reference's clock (driver.cpp:231,276): a float advanced by the float DT each step; it
    // is the time printed on the progress lines and written to the HDF5 'time' datasets, while
    // the solver keeps its own FP64 clock
    float t_out = 0.0F;
    const float DT_out = (float)DT;
    std::ofstream stepLog;  // RK33's log argument (unused there too)
    const auto begin = std::chrono::steady_clock::now();
    auto prev = begin;
    std::cout << "Timestepping\n";
    for (unsigned i = 0; i < NUM_TIMESTEPS; i += p.NUM_TIME_SUBSTEPS) {
      const auto now = std::chrono::steady_clock::now();
      if (i > 0) {
        std::cout << "    - Elapsed runtime:" << (int)std::chrono::duration<double>(now - begin).count() << "s\n";
        std::cout << "    - Estimated remaining:"
                  << (int)(std::chrono::duration<double>(now - prev).count() * (NUM_TIMESTEPS - i) / p.NUM_TIME_SUBSTEPS)
                  << "s\n";
      }
      prev = now;
      for (unsigned j = 0; j < p.NUM_TIME_SUBSTEPS; ++j) {
        if ((i * p.NUM_TIME_SUBSTEPS + j) % p.FREQ_TERMINAL == 0) {
          std::cout << " - Timestep ";
          std::cout.width(7);
          std::cout << i + j << "/" << NUM_TIMESTEPS << " (";
          std::cout.width(3);
          std::cout << 100 * (i + j) / NUM_TIMESTEPS << "%) : t=";
          std::cout.width(5);
          std::cout << t_out << "\n";
        }
        if (fused) {
          solver.check(dg3d_step(solver.ctx(), p.SCHEME == 1 ? DG3D_LSRK54 : DG3D_RK33, DT, 1, &t), "step");
        } else {
          swapGrid = oldGrid;
          oldGrid = newGrid;
          newGrid = swapGrid;
          if (p.SCHEME == 2) newGrid->forwardEuler(DT, oldGrid, stepLog);
          else newGrid->RK33(DT, oldGrid, &k1, &k2, stepLog);
          t += DT;
        }
        t_out += DT_out;
      }
      if (p.BOOL_FILE) {
        solver.check(dg3d_get_state(solver.ctx(), fused ? 0 : newGrid->mBlock.reg(), st.data(), 4), "get_state");
        out.dump_state(st, i + p.NUM_TIME_SUBSTEPS, (double)t_out);
      }
    }
    std::cout << "Post-processing\n";
    const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
    std::cout << "Exiting\nTotal runtime is " << total << "s\n";
  } catch (const std::bad_alloc& ex) {
    std::cerr << "host allocation failed: " << ex.what() << "\n";
    return ERRCODE_DEVICE_FAILURE;
  } catch (const std::exception& ex) {
    // everything in this block runs after the inputs were read: a failure here is the device's
    // (HIP error, device allocation) or the solver's, not an unreadable input
    std::cerr << ex.what() << "\n";
    return ERRCODE_DEVICE_FAILURE;
  }
  return ERRCODE_SUCESSFUL_COMPLETION;
}
