// Streaming ceiling of the device for the stage kernel's access pattern: 8 FP64 arrays of
// 32 768 000 values read and 8 written, nontemporal, each 256-lane workgroup streaming one chunk
// of consecutive values of every array, as a stage kernel's workgroups stream their elements'
// fields (the 64^3 N=4 block has 32 768 000 nodes). No arithmetic beyond one multiply per
// value, so the time is the memory system's. Prints one JSON line: the best rate over the
// chunk sizes, counting the bytes read and written.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int kArrays = 8;
constexpr long kValues = 32768000;
// array pitch in values: the stage's field pitch rule, 256 past a multiple of 2048 values (2 KiB
// past a multiple of 16 KiB); pitches that are multiples of 1024 values stream 4-12 % slower
constexpr long kStride = (kValues + 2047) / 2048 * 2048 + 256;

__global__ void __launch_bounds__(256) k_stream(const double* __restrict__ in, double* __restrict__ out, int chunk) {
  const int t = threadIdx.x;
  if (t >= chunk) return;
  const long i = (long)blockIdx.x * chunk + t;
  if (i >= kValues) return;
  double v[kArrays];
#pragma unroll
  for (int a = 0; a < kArrays; ++a) v[a] = __builtin_nontemporal_load(in + a * kStride + i);
#pragma unroll
  for (int a = 0; a < kArrays; ++a) __builtin_nontemporal_store(v[a] * 1.0000001, out + a * kStride + i);
}

bool ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "stream probe: %s: %s\n", what, hipGetErrorString(e));
  return false;
}

}  // namespace

int main() {
  const size_t bytes = sizeof(double) * kArrays * kStride;
  double *in = nullptr, *out = nullptr;
  if (!ok(hipMalloc(&in, bytes), "hipMalloc") || !ok(hipMalloc(&out, bytes), "hipMalloc")) return 1;
  if (!ok(hipMemset(in, 0, bytes), "hipMemset") || !ok(hipMemset(out, 0, bytes), "hipMemset")) return 1;
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "event") || !ok(hipEventCreate(&e1), "event")) return 1;
  const double moved = 2.0 * kArrays * kValues * sizeof(double);
  double best = 0.0;
  int best_chunk = 0;
  for (int chunk : {256, 250, 128, 125}) {
    const int blocks = (int)((kValues + chunk - 1) / chunk);
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(k_stream, dim3(blocks), dim3(256), 0, 0, in, out, chunk);
    if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e0, 0), "record")) return 1;
    const int reps = 20;
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_stream, dim3(blocks), dim3(256), 0, 0, in, out, chunk);
    if (!ok(hipEventRecord(e1, 0), "record") || !ok(hipEventSynchronize(e1), "sync")) return 1;
    float ms = 0.f;
    if (!ok(hipEventElapsedTime(&ms, e0, e1), "elapsed")) return 1;
    const double gbs = moved / (ms / reps * 1e-3) / 1e9;
    if (gbs > best) {
      best = gbs;
      best_chunk = chunk;
    }
  }
  std::printf("{\"stream_GBs\": %.1f, \"chunk\": %d, \"arrays_read\": %d, \"arrays_written\": %d, \"values\": %ld}\n",
              best, best_chunk, kArrays, kArrays, kValues);
  (void)hipFree(in);
  (void)hipFree(out);
  return 0;
}
