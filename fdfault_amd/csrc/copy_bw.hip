// HBM calibration probe: copies `mib` MiB between two device buffers with a
// plain streaming kernel, `reps` times, and prints the bytes moved per launch
// and the achieved GB/s.  Used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE
// readings (KiB per dispatch) to bytes: the copy's algorithmic traffic is
// known exactly (read n + write n bytes).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const double4* __restrict__ a, double4* __restrict__ b, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char** argv) {
  const long long mib = argc > 1 ? atoll(argv[1]) : 1024;
  const int reps = argc > 2 ? atoi(argv[2]) : 20;
  const long long bytes = mib << 20;
  const long long n = bytes / (long long)sizeof(double4);
  double4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    fprintf(stderr, "allocation failed\n");
    return 1;
  }
  (void)hipMemset(a, 0, bytes);
  (void)hipMemset(b, 0, bytes);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int blocks = 256 * 8, threads = 256;
  hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(threads), 0, 0, a, b, n);
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(threads), 0, 0, a, b, n);
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  const double per = ms / reps;
  printf("copy %lld MiB: read %lld B + write %lld B per launch, %.4f ms per launch, %.1f GB/s\n", mib, bytes, bytes, per,
         2.0 * bytes / (per * 1e-3) / 1e9);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
