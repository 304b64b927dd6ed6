// HBM calibration probe: copies `mib` MiB between two device buffers with a
// plain streaming kernel, `reps` times, and prints the bytes moved per launch
// and the achieved GB/s.  Used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE
// readings (KiB per dispatch) to bytes: the copy's algorithmic traffic is
// known exactly (read n + write n bytes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

__global__ void copy_kernel(const double4* __restrict__ a, double4* __restrict__ b, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

// the same copy with non-temporal loads and stores
__global__ void copy_kernel_nt(const double4* __restrict__ a, double4* __restrict__ b, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const double* s = reinterpret_cast<const double*>(a + i);
    double* d = reinterpret_cast<double*>(b + i);
#pragma unroll
    for (int c = 0; c < 4; c++) __builtin_nontemporal_store(__builtin_nontemporal_load(s + c), d + c);
  }
}

// SoA stream: `nin` input and `nout` output arrays of n floats (the bulk
// stage's access shape: f and df in, df and f_new out), out_q = sum of inputs
template <int NIN, int NOUT>
__global__ void soa_stream(const float* const* __restrict__ in, float* const* __restrict__ out, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < NIN; q++) s += in[q][i];
#pragma unroll
    for (int q = 0; q < NOUT; q++) out[q][i] = s + (float)q;
  }
}

template <int NIN, int NOUT>
static void soa_run(long long n, int reps) {
  float** hin = new float*[NIN];
  float** hout = new float*[NOUT];
  for (int q = 0; q < NIN; q++) { (void)hipMalloc(&hin[q], n * 4); (void)hipMemset(hin[q], 0, n * 4); }
  for (int q = 0; q < NOUT; q++) { (void)hipMalloc(&hout[q], n * 4); (void)hipMemset(hout[q], 0, n * 4); }
  float **din, **dout;
  (void)hipMalloc(&din, NIN * sizeof(float*));
  (void)hipMalloc(&dout, NOUT * sizeof(float*));
  (void)hipMemcpy(din, hin, NIN * sizeof(float*), hipMemcpyHostToDevice);
  (void)hipMemcpy(dout, hout, NOUT * sizeof(float*), hipMemcpyHostToDevice);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  // grid sizes: 8 or 32 workgroups per CU (grid-stride), or one point per thread
  const long long grids[3] = {256 * 8, 256 * 32, (n + 255) / 256};
  double per = 1e30;
  for (long long gb : grids) {
    const dim3 grid((unsigned)gb);
    hipLaunchKernelGGL((soa_stream<NIN, NOUT>), grid, dim3(256), 0, 0, (const float* const*)din, dout, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++)
      hipLaunchKernelGGL((soa_stream<NIN, NOUT>), grid, dim3(256), 0, 0, (const float* const*)din, dout, n);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    printf("  grid %lld x 256: %.4f ms\n", gb, (double)ms / reps);
    per = std::min(per, (double)ms / reps);
  }
  const double bytes = (double)(NIN + NOUT) * n * 4;
  printf("soa stream %d in + %d out arrays of %lld floats: %.4f ms per launch, %.1f GB/s\n", NIN, NOUT, n, per,
         bytes / (per * 1e-3) / 1e9);
  for (int q = 0; q < NIN; q++) (void)hipFree(hin[q]);
  for (int q = 0; q < NOUT; q++) (void)hipFree(hout[q]);
  (void)hipFree(din);
  (void)hipFree(dout);
  delete[] hin;
  delete[] hout;
}

int main(int argc, char** argv) {
  if (argc > 1 && strcmp(argv[1], "soa36") == 0) {  // "soa36 <n>": the 3-D stage's 18 in + 18 out arrays
    const long long n = argc > 2 ? atoll(argv[2]) : 16200801;
    soa_run<18, 18>(n, 10);
    return 0;
  }
  if (argc > 1 && argv[1][0] == 's') {  // "soa <npoints>": bulk-stage stream shapes
    const long long n = argc > 2 ? atoll(argv[2]) : 16200801;
    soa_run<2, 1>(n, 10);
    soa_run<9, 9>(n, 10);
    soa_run<18, 18>(n, 10);
    return 0;
  }
  const bool best = argc > 1 && strcmp(argv[1], "copybest") == 0;  // "copybest <mib>": best of a sweep
  if (best) { argv++; argc--; }
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
  double per = ms / reps;
  if (best) {  // non-temporal or not, 8 or 32 workgroups per CU or one element per thread
    const long long grids[3] = {256 * 8, 256 * 32, (n + 255) / 256};
    for (int nt = 0; nt < 2; nt++)
      for (long long gb : grids) {
        const dim3 grid((unsigned)gb);
        if (nt) hipLaunchKernelGGL(copy_kernel_nt, grid, dim3(threads), 0, 0, a, b, n);
        else hipLaunchKernelGGL(copy_kernel, grid, dim3(threads), 0, 0, a, b, n);
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < reps; r++) {
          if (nt) hipLaunchKernelGGL(copy_kernel_nt, grid, dim3(threads), 0, 0, a, b, n);
          else hipLaunchKernelGGL(copy_kernel, grid, dim3(threads), 0, 0, a, b, n);
        }
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float m2 = 0.f;
        (void)hipEventElapsedTime(&m2, e0, e1);
        printf("  %s grid %lld x %d: %.4f ms\n", nt ? "non-temporal" : "plain", gb, threads, (double)m2 / reps);
        per = std::min(per, (double)m2 / reps);
      }
  }
  printf("copy %lld MiB: read %lld B + write %lld B per launch, %.4f ms per launch, %.1f GB/s\n", mib, bytes, bytes, per,
         2.0 * bytes / (per * 1e-3) / 1e9);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
