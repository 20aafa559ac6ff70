// Elementwise vertex bookkeeping and device-side synthetic inputs.
//
//   gamma_update:  G_r = F_r - G_r after the parquet step, SRC/loop.f90:219-222
//   synth_vertex:  seeded Gamma = diag*I + X (benchmark inputs; no reference
//                  counterpart — the reference reads or initialises vertices
//                  on the host, SRC/preprocessing.f90:100-116)
//   check_finite:  NaN/Inf scan (the reference propagates them silently)
#include "internal.h"

namespace tups {

__global__ __launch_bounds__(256) void gamma_update_kernel(double2* g0, double2* g1, double2* g2,
                                                           double2* g3, const double2* f0,
                                                           const double2* f1, const double2* f2,
                                                           const double2* f3, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = blockIdx.y;
  double2* g = r == 0 ? g0 : r == 1 ? g1 : r == 2 ? g2 : g3;
  const double2* f = r == 0 ? f0 : r == 1 ? f1 : r == 2 ? f2 : f3;
  const double2 a = ld_stream(f + i), b = ld_stream(g + i);
  st_stream(g + i, make_double2(a.x - b.x, a.y - b.y));
}

int launch_gamma_update(tups_ctx* c) {
  const size_t n = (size_t)c->sz.nz * c->sz.nz * c->sz.nq;
  dim3 g((unsigned)((n + 255) / 256), 4);
  if (int rc = timer_begin(c)) return rc;
  hipLaunchKernelGGL(gamma_update_kernel, g, dim3(256), 0, c->stream, c->gamma[0], c->gamma[1],
                     c->gamma[2], c->gamma[3], c->fvert[0], c->fvert[1], c->fvert[2],
                     c->fvert[3], n);
  TUPS_HIP(hipGetLastError());
  if (int rc = timer_end(c, TUPS_K_GAMMA_UPDATE)) return rc;
  return TUPS_OK;
}

// splitmix64 finaliser (Steele, Lea, Flood 2014) used as a counter-based generator
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double unit53(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

// element i of the stream: Box-Muller on two draws, Re/Im ~ N(0, sigma^2/2)
__global__ __launch_bounds__(256) void synth_kernel(double2* dst, size_t n, int nz, uint64_t key,
                                                    double diag, double sigma) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t h1 = mix64(key ^ mix64(2 * (uint64_t)i));
  const uint64_t h2 = mix64(key ^ mix64(2 * (uint64_t)i + 1));
  const double u1 = unit53(h1), u2 = unit53(h2);
  const double rad = sqrt(-2.0 * log(1.0 - u1)) * sigma * 0.70710678118654752440;
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  double re = rad * cs, im = rad * sn;
  const size_t nz2 = (size_t)nz * nz;
  const size_t e = i % nz2;
  if (e % nz == e / nz) re += diag;
  dst[i] = make_double2(re, im);
}

int launch_synth(tups_ctx* c, double2* dst, size_t n, int nz, uint64_t seed, uint64_t stream_id,
                 double diag, double sigma) {
  const uint64_t key = seed * 0x100000001B3ull ^ (stream_id << 48);
  hipLaunchKernelGGL(synth_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, dst,
                     n, nz, key, diag, sigma);
  TUPS_HIP(hipGetLastError());
  return TUPS_OK;
}

__global__ __launch_bounds__(256) void finite_kernel(const double2* p, size_t n, int* flag) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 v = p[i];
  if (!isfinite(v.x) || !isfinite(v.y)) atomicOr(flag, 1);
}

int launch_check_finite(tups_ctx* c, const double2* p, size_t n) {
  int zero = 0, h = 0;
  TUPS_HIP(hipMemcpyAsync(c->d_flag, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(finite_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, p, n,
                     c->d_flag);
  TUPS_HIP(hipGetLastError());
  TUPS_HIP(hipMemcpyAsync(&h, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  TUPS_CHECK(h == 0, TUPS_ENONFINITE, "non-finite value in resident array");
  return TUPS_OK;
}

}  // namespace tups
