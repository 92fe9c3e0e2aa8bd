// copy_probe -- a plain device-to-device copy kernel of known size.
//
//   vo_copy_probe [MiB=256] [reps=5]
//
// It serves two measurements: the stream-copy bandwidth of the device
// (printed as one JSON line), and the calibration of the FETCH_SIZE /
// WRITE_SIZE hardware counters: run under `rocprofv3 --pmc FETCH_SIZE` (then
// WRITE_SIZE), the counter reading of copy_kernel per dispatch against the
// known MiB it moves gives the scale applied to the sweeps' readings.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, long n) {
  const long stride = long(gridDim.x) * blockDim.x;
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

static void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

int main(int argc, char** argv) {
  const long mib = argc > 1 ? std::atol(argv[1]) : 256;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
  if (mib <= 0 || mib > 16384 || reps <= 0 || reps > 1000) {
    std::fprintf(stderr, "usage: vo_copy_probe [MiB in 1..16384] [reps in 1..1000]\n");
    return 2;
  }
  const size_t bytes = size_t(mib) << 20;
  const long n = long(bytes / sizeof(float4));
  void *src = nullptr, *dst = nullptr;
  check(hipMalloc(&src, bytes), "hipMalloc");
  check(hipMalloc(&dst, bytes), "hipMalloc");
  check(hipMemset(src, 1, bytes), "hipMemset");
  check(hipMemset(dst, 0, bytes), "hipMemset");
  hipEvent_t a, b;
  check(hipEventCreate(&a), "event");
  check(hipEventCreate(&b), "event");
  const dim3 grid(256 * 32), block(256);  // 32 workgroups per CU on 256 CUs, grid-stride
  hipLaunchKernelGGL(copy_kernel, grid, block, 0, nullptr, static_cast<const float4*>(src),
                     static_cast<float4*>(dst), n);  // warm-up
  check(hipGetLastError(), "launch");
  check(hipEventRecord(a, nullptr), "record");
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(copy_kernel, grid, block, 0, nullptr, static_cast<const float4*>(src),
                       static_cast<float4*>(dst), n);
  check(hipGetLastError(), "launch");
  check(hipEventRecord(b, nullptr), "record");
  check(hipEventSynchronize(b), "sync");
  float ms = 0;
  check(hipEventElapsedTime(&ms, a, b), "elapsed");
  const double per = ms / reps;
  std::printf("{\"kernel\": \"copy_kernel\", \"bytes_read\": %zu, \"bytes_written\": %zu, \"dispatches\": %d, "
              "\"avg_ms\": %.5f, \"copy_GBs\": %.1f}\n",
              bytes, bytes, reps + 1, per, 2.0 * double(bytes) / (per * 1e-3) / 1e9);
  check(hipFree(src), "hipFree");
  check(hipFree(dst), "hipFree");
  return 0;
}
