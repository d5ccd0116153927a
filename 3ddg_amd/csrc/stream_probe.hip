// Streaming ceiling of the device for the LSRK stage kernel's access pattern, with no arithmetic
// beyond one multiply-add per value, so the time is the memory system's. Three group-blocked
// registers of the 64^3 N=4 block (262 144 elements of 125 nodes in 128 slots; groups of 32
// elements whose 4 FP64 fields are runs of 4096 slots): each workgroup takes EPB elements of a
// group, reads their state (cached loads) and their RK register (pair-major in the group's block,
// 16-byte nontemporal loads), and writes the RK register in place and the new state
// (nontemporal), as a stage's workgroups do. Prints one JSON line: the best rate over the
// workgroup sizes, counting 128 bytes per node (4 fields read and written twice).
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr long kElems = 262144;
constexpr long kNodes = kElems * 125;
constexpr int kES = 128, kG = 32, kP = 4096;
constexpr long kSlots = kElems / kG * 4 * kP;  // per register
typedef double f64x2 __attribute__((ext_vector_type(2)));

// one lane per node slot; workgroup w owns elements [w * EPB, (w + 1) * EPB), inside one group
template <int EPB>
__global__ void __launch_bounds__(EPB * kES) k_stream(const double* __restrict__ qi, double* __restrict__ du,
                                                    double* __restrict__ qo, double a, double b) {
  const int t = threadIdx.x;
  const long e = (long)blockIdx.x * EPB + t / kES;
  const int node = t % kES;
  const long g = e / kG, eg = e % kG;
  const long s = g * 4 * kP + eg * kES + node;  // field 0; field f at + f * kP
  double q[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) q[f] = qi[s + f * kP];
  f64x2* d = reinterpret_cast<f64x2*>(du + g * 4 * kP) + eg * kES + node;  // pair 0; pair 1 at + kP
  f64x2 d0 = __builtin_nontemporal_load(d), d1 = __builtin_nontemporal_load(d + kP);
  d0[0] = a * d0[0] + 1e-9 * q[0];
  d0[1] = a * d0[1] + 1e-9 * q[1];
  d1[0] = a * d1[0] + 1e-9 * q[2];
  d1[1] = a * d1[1] + 1e-9 * q[3];
  __builtin_nontemporal_store(d0, d);
  __builtin_nontemporal_store(d1, d + kP);
  __builtin_nontemporal_store(q[0] + b * d0[0], qo + s);
  __builtin_nontemporal_store(q[1] + b * d0[1], qo + s + kP);
  __builtin_nontemporal_store(q[2] + b * d1[0], qo + s + 2 * kP);
  __builtin_nontemporal_store(q[3] + b * d1[1], qo + s + 3 * kP);
}

bool ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "stream probe: %s: %s\n", what, hipGetErrorString(e));
  return false;
}

template <int EPB>
bool run(double* r0, double* r1, double* du, hipEvent_t e0, hipEvent_t e1, double* gbs) {
  const int blocks = (int)(kElems / EPB);
  auto launch = [&](int i) {
    hipLaunchKernelGGL(k_stream<EPB>, dim3(blocks), dim3(EPB * kES), 0, 0, (i & 1) ? r1 : r0, du, (i & 1) ? r0 : r1,
                       0.5, 0.25);
  };
  for (int w = 0; w < 4; ++w) launch(w);
  if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e0, 0), "record")) return false;
  const int reps = 40;
  for (int r = 0; r < reps; ++r) launch(r);
  if (!ok(hipEventRecord(e1, 0), "record") || !ok(hipEventSynchronize(e1), "sync")) return false;
  float ms = 0.f;
  if (!ok(hipEventElapsedTime(&ms, e0, e1), "elapsed")) return false;
  *gbs = 128.0 * kNodes / (ms / reps * 1e-3) / 1e9;
  return true;
}

}  // namespace

int main() {
  const size_t bytes = sizeof(double) * kSlots;
  double *r0 = nullptr, *r1 = nullptr, *du = nullptr;
  if (!ok(hipMalloc(&r0, bytes), "hipMalloc") || !ok(hipMalloc(&r1, bytes), "hipMalloc") ||
      !ok(hipMalloc(&du, bytes), "hipMalloc"))
    return 1;
  if (!ok(hipMemset(r0, 0, bytes), "hipMemset") || !ok(hipMemset(r1, 0, bytes), "hipMemset") ||
      !ok(hipMemset(du, 0, bytes), "hipMemset"))
    return 1;
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "event") || !ok(hipEventCreate(&e1), "event")) return 1;
  double best = 0.0;
  int best_lanes = 0;
  for (int round = 0; round < 2; ++round) {
    double g = 0.0;
    if (!run<2>(r0, r1, du, e0, e1, &g)) return 1;
    if (g > best) { best = g; best_lanes = 256; }
    if (!run<4>(r0, r1, du, e0, e1, &g)) return 1;
    if (g > best) { best = g; best_lanes = 512; }
  }
  std::printf("{\"stream_GBs\": %.1f, \"lanes\": %d, \"nodes\": %ld, \"pattern\": \"LSRK stage pattern at 64^3 N=4: "
              "group-blocked FP64 state read and written (4 runs of 4096 slots per group of 32 elements), "
              "pair-major RK register read and written in place, %d-lane workgroups over whole elements, 128 B per "
              "node\"}\n",
              best, best_lanes, kNodes, best_lanes);
  (void)hipFree(r0);
  (void)hipFree(r1);
  (void)hipFree(du);
  return 0;
}
