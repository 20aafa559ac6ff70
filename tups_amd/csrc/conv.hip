// Convergence check of one iteration on the device.
//
// Reference: conv(), SRC/parquet_check.f90:39-150.  e = |SigmaOld - Sigma| / |Sigma|
// elementwise, its average |sum(e) / Nt| and its maximum with the first index of
// the maximum; per channel the largest |F_r| over the resident vertex and its
// relative change against the previous call; converged when max e < f_sigma and
// every change < f_vert.
//
// Two grid-stride reductions (Sigma: Nt elements; F: four arrays of Nz^2 Nq, the
// HBM-bound part, 16 bytes read per element and nothing written) leave per-block
// partials, and one workgroup combines them in a fixed order.  The vertex
// maximum is taken on |F|^2 with the element carried along, so the reported
// value is hypot() of the element that attains it.
#include "internal.h"

#include <cmath>
#include <mutex>
#include <string>

namespace tups {

namespace {

constexpr int kConvBlocksS = 256;   // Sigma partials
constexpr int kConvBlocksF = 1024;  // vertex partials per channel

struct ConvArgs {
  const double2* old_s;
  const double2* cur_s;
  const double2* f[4];
  size_t nt, nf_elems;
  double* s_part;   // [kConvBlocksS][3]: sum, max, first index of the max
  double* f_part;   // [4][kConvBlocksF][3]: |F|^2 max, its re, im
  double* out;      // [3 + 4]: sum, max, index; |F| max per channel
};

__global__ __launch_bounds__(256) void conv_sigma(ConvArgs a) {
  __shared__ double ssum[256], smax[256], sidx[256];
  double sum = 0.0, mx = -1.0, ix = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < a.nt; i += (size_t)gridDim.x * 256) {
    const double2 o = a.old_s[i], c = a.cur_s[i];
    const double e = hypot(o.x - c.x, o.y - c.y) / hypot(c.x, c.y);
    sum += e;
    if (e > mx) {  // the first maximum (NaN never compares larger)
      mx = e;
      ix = (double)i;
    }
  }
  ssum[threadIdx.x] = sum;
  smax[threadIdx.x] = mx;
  sidx[threadIdx.x] = ix;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      const int o = threadIdx.x + h;
      ssum[threadIdx.x] += ssum[o];
      if (smax[o] > smax[threadIdx.x] || (smax[o] == smax[threadIdx.x] && sidx[o] < sidx[threadIdx.x])) {
        smax[threadIdx.x] = smax[o];
        sidx[threadIdx.x] = sidx[o];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.s_part[3 * blockIdx.x] = ssum[0];
    a.s_part[3 * blockIdx.x + 1] = smax[0];
    a.s_part[3 * blockIdx.x + 2] = sidx[0];
  }
}

// grid (kConvBlocksF, 4)
__global__ __launch_bounds__(256) void conv_fmax(ConvArgs a) {
  __shared__ double m2[256], mre[256], mim[256];
  const double2* f = a.f[blockIdx.y];
  double best = -1.0, bre = 0.0, bim = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < a.nf_elems; i += (size_t)gridDim.x * 256) {
    const double2 v = ld_stream(f + i);
    const double s = v.x * v.x + v.y * v.y;
    if (s > best) {
      best = s;
      bre = v.x;
      bim = v.y;
    }
  }
  m2[threadIdx.x] = best;
  mre[threadIdx.x] = bre;
  mim[threadIdx.x] = bim;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h && m2[threadIdx.x + h] > m2[threadIdx.x]) {
      m2[threadIdx.x] = m2[threadIdx.x + h];
      mre[threadIdx.x] = mre[threadIdx.x + h];
      mim[threadIdx.x] = mim[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* p = a.f_part + 3 * ((size_t)blockIdx.y * kConvBlocksF + blockIdx.x);
    p[0] = m2[0];
    p[1] = mre[0];
    p[2] = mim[0];
  }
}

// one workgroup of five waves: wave 0 the Sigma partials, wave 1 + r channel r's;
// lanes take the partials strided by 64, then a butterfly over the wave (a sum's
// partners add the same two values, so every lane holds the same total)
__global__ __launch_bounds__(320) void conv_final(ConvArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (wave == 0) {
    double sum = 0.0, mx = -1.0, ix = 0.0;
    for (int b = lane; b < kConvBlocksS; b += 64) {
      sum += a.s_part[3 * b];
      const double m = a.s_part[3 * b + 1], i = a.s_part[3 * b + 2];
      if (m > mx || (m == mx && i < ix)) {
        mx = m;
        ix = i;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double so = __shfl_xor(sum, off), mo = __shfl_xor(mx, off), io = __shfl_xor(ix, off);
      sum += so;
      if (mo > mx || (mo == mx && io < ix)) {
        mx = mo;
        ix = io;
      }
    }
    if (lane == 0) {
      a.out[0] = sum;
      a.out[1] = mx;
      a.out[2] = ix;
    }
  } else {
    const double* p = a.f_part + 3 * (size_t)(wave - 1) * kConvBlocksF;
    double best = -1.0, re = 0.0, im = 0.0;
    for (int b = lane; b < kConvBlocksF; b += 64)
      if (p[3 * b] > best) {
        best = p[3 * b];
        re = p[3 * b + 1];
        im = p[3 * b + 2];
      }
    for (int off = 32; off > 0; off >>= 1) {
      const double bo = __shfl_xor(best, off), ro = __shfl_xor(re, off), io = __shfl_xor(im, off);
      if (bo > best) {
        best = bo;
        re = ro;
        im = io;
      }
    }
    if (lane == 0) a.out[2 + wave] = hypot(re, im);
  }
}

}  // namespace

// ---- scratch of tups_conv_noalloc --------------------------------------------------
//
// A fixed device array replaces the context's lazily allocated workspace: the
// host Sigma pair (at most kConvNoallocMaxNt elements each), the block partials
// and the result.  Each GPU holds its own copy; its address is taken once per
// device at context creation and kept in conv_scratch_ptr.  Nothing in it must
// start at zero: every call overwrites the partials and the result before
// reading them.
constexpr size_t kConvNoallocMaxNt = 65536;
constexpr size_t kConvScratchDoubles = 4 * kConvNoallocMaxNt + 3 * kConvBlocksS + 3 * 4 * (size_t)kConvBlocksF + 8;
static __device__ __align__(256) double scratch[kConvScratchDoubles];  // 2.1 MiB

namespace {
constexpr int kMaxDevices = 64;
std::mutex conv_scratch_mu;
double* conv_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Called with `device` current (tups_create): the device address of this GPU's
// copy of the scratch array.
int conv_noalloc_setup(int device) {
  TUPS_CHECK(device >= 0 && device < kMaxDevices, TUPS_EINVAL, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lock(conv_scratch_mu);
  if (conv_scratch_ptr[device]) return TUPS_OK;
  void* p = nullptr;
  TUPS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  conv_scratch_ptr[device] = static_cast<double*>(p);
  return TUPS_OK;
}

namespace {

// The check over a workspace w: partials, result, then the host Sigma pair.
int conv_run(tups_ctx* c, double* w, const double* sigma_old, const double* sigma, double f_sigma, double f_vert,
             tups_conv_info* info) {
  const tups_sizes& s = c->sz;
  const size_t nt = (size_t)s.nx * s.ny * s.nf;
  ConvArgs a{};
  // partials and result first, so the device pair never reaches past them
  a.s_part = w;
  a.f_part = a.s_part + 3 * kConvBlocksS;
  a.out = a.f_part + 3 * 4 * (size_t)kConvBlocksF;
  double2* h_old = reinterpret_cast<double2*>(a.out + 8);
  double2* h_cur = h_old + nt;
  if (sigma) {
    TUPS_HIP(hipMemcpyAsync(h_old, sigma_old, nt * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    TUPS_HIP(hipMemcpyAsync(h_cur, sigma, nt * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    a.old_s = h_old;
    a.cur_s = h_cur;
  } else {
    TUPS_CHECK(c->sigma_old_valid, TUPS_ESTATE, "no Sigma pair on the device (tups_iteration_full first)");
    a.old_s = c->d_sigma_old;
    a.cur_s = c->d_sigma;
  }
  for (int r = 0; r < 4; ++r) a.f[r] = c->fvert[r];
  a.nt = nt;
  a.nf_elems = (size_t)s.nz * s.nz * s.nq;
  if (int rc = timer_begin(c)) return rc;
  hipLaunchKernelGGL(conv_sigma, dim3(kConvBlocksS), dim3(256), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  hipLaunchKernelGGL(conv_fmax, dim3(kConvBlocksF, 4), dim3(256), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  hipLaunchKernelGGL(conv_final, dim3(1), dim3(320), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  if (int rc = timer_end(c, TUPS_K_CONV)) return rc;
  double out[7];
  TUPS_HIP(hipMemcpyAsync(out, a.out, sizeof(out), hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  info->sigma_avg = std::fabs(out[0] / (double)nt);
  info->sigma_max = out[1];
  const long long k = (long long)out[2];  // ((kx-1) Ny + ky-1) Nf + nu-1
  info->kx = (int32_t)(k / ((long long)s.ny * s.nf)) + 1;
  info->ky = (int32_t)((k / s.nf) % s.ny) + 1;
  info->nu = (int32_t)(k % s.nf) + 1;
  bool ok = info->sigma_max < f_sigma;
  for (int r = 0; r < 4; ++r) {
    info->f_max[r] = out[3 + r];
    // abs(1 - max / max_prev); the previous maximum starts at 0 (the reference's is unset)
    info->f_rel_err[r] = std::fabs(1.0 - out[3 + r] / c->conv_fmax_prev[r]);
    ok = ok && info->f_rel_err[r] < f_vert;
    c->conv_fmax_prev[r] = out[3 + r];
  }
  info->converged = ok ? 1 : 0;
  return TUPS_OK;
}

int conv_args_ok(tups_ctx* c, const double* sigma_old, const double* sigma, tups_conv_info* info) {
  TUPS_CHECK(c && info, TUPS_EINVAL, "null argument");
  TUPS_CHECK((sigma_old == nullptr) == (sigma == nullptr), TUPS_EINVAL,
             "pass both Sigma_old and Sigma, or neither (the device pair of the last iteration)");
  return TUPS_OK;
}

}  // namespace

}  // namespace tups

using namespace tups;

extern "C" int tups_conv(tups_ctx* c, const double* sigma_old, const double* sigma, double f_sigma, double f_vert,
                         tups_conv_info* info) {
  if (int rc = conv_args_ok(c, sigma_old, sigma, info)) return rc;
  const size_t nt = (size_t)c->sz.nx * c->sz.ny * c->sz.nf;
  const size_t ws = (2 * nt * 2 + 3 * kConvBlocksS + 3 * 4 * (size_t)kConvBlocksF + 8) * sizeof(double);
  if (!c->conv_ws) TUPS_HIP(hipMalloc(&c->conv_ws, ws));
  return conv_run(c, static_cast<double*>(c->conv_ws), sigma_old, sigma, f_sigma, f_vert, info);
}

extern "C" int tups_conv_noalloc(tups_ctx* c, const double* sigma_old, const double* sigma, double f_sigma,
                                 double f_vert, tups_conv_info* info) {
  if (int rc = conv_args_ok(c, sigma_old, sigma, info)) return rc;
  const size_t nt = (size_t)c->sz.nx * c->sz.ny * c->sz.nf;
  if (sigma && nt > kConvNoallocMaxNt) {
    set_error("tups_conv_noalloc: the Sigma pair has " + std::to_string(nt) +
              " elements (Nx Ny Nf), beyond the fixed scratch of " + std::to_string(kConvNoallocMaxNt) +
              "; use tups_conv or pass the device pair (NULL, NULL)");
    return TUPS_EINVAL;
  }
  TUPS_CHECK(c->device >= 0 && c->device < kMaxDevices && conv_scratch_ptr[c->device], TUPS_ESTATE,
             "tups_conv_noalloc: no scratch address for this device (context not created by tups_create)");
  return conv_run(c, conv_scratch_ptr[c->device], sigma_old, sigma, f_sigma, f_vert, info);
}
