// Copy kernels of known size: the calibration of rocprofv3's FETCH_SIZE and
// WRITE_SIZE counters (KiB per dispatch) for the access widths the library's
// kernels use. Each kernel copies `n` doubles from src to dst exactly once
// (8 n bytes read, 8 n bytes written), with 8-byte (one double per lane, the
// width of the tree-solve and explicit kernels) or 16-byte (two doubles) accesses.
// Prints one line per kernel: name, bytes read, bytes written, mean ms, GB/s.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                               \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));            \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)

__global__ __launch_bounds__(256) void copy_f64(const double* __restrict__ src,
                                                double* __restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

__global__ __launch_bounds__(256) void copy_f64x2(const double2* __restrict__ src,
                                                  double2* __restrict__ dst, long long n2) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
    dst[i] = src[i];
}

int main(int argc, char** argv) {
  const long long n = argc > 1 ? std::atoll(argv[1]) : (1ll << 27);  // 1 GiB of doubles
  const int reps = 3;
  double *src = nullptr, *dst = nullptr;
  CHECK(hipMalloc(&src, n * sizeof(double)));
  CHECK(hipMalloc(&dst, n * sizeof(double)));
  CHECK(hipMemset(src, 0, n * sizeof(double)));
  int cus = 0;
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  const dim3 grid(cus * 16), block(256);
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  for (int k = 0; k < 2; k++) {
    float total = 0;
    for (int r = 0; r < reps; r++) {
      CHECK(hipEventRecord(e0));
      if (k == 0) hipLaunchKernelGGL(copy_f64, grid, block, 0, 0, src, dst, n);
      else hipLaunchKernelGGL(copy_f64x2, grid, block, 0, 0, (const double2*)src, (double2*)dst, n / 2);
      CHECK(hipGetLastError());
      CHECK(hipEventRecord(e1));
      CHECK(hipEventSynchronize(e1));
      float ms = 0;
      CHECK(hipEventElapsedTime(&ms, e0, e1));
      total += ms;
    }
    const double ms = total / reps, bytes = 8.0 * (double)n;
    std::printf("%s read_bytes %.0f write_bytes %.0f mean_ms %.4f GBps %.1f\n",
                k == 0 ? "copy_f64" : "copy_f64x2", bytes, bytes, ms, 2 * bytes / ms / 1e6);
  }
  CHECK(hipFree(src));
  CHECK(hipFree(dst));
  return 0;
}
