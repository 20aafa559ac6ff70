// Phi(q) -> Phi(R): the truncated-unity real-space projection of the
// reducible vertices.
//
// Reference: calc_PhiR / FTPhis / symop_arr, SRC/parquet_PhiR.f90:35-109,
// 196-258, 311-356, with sym_op_L (SRC/parquet_util.f90:732-890):
//   PhiR_r(z, w, R) = (1/Nx) sum_{Q in IBZ} sum_{S in star(Q)}
//                     fac_S(l1) fac_S(l2) Phi_r((l1s nu1),(l2s nu2), (Q,w)) e^{i (S Q).R}
// On one GPU the q -> omega all-to-all of communicatePhis (:114-182) is the
// identity (indexR_from_indexQ maps q onto itself when one rank owns every
// omega), so the kernel reads Phi_r(q) where the BSE left it.
//
// The reference zeroes and rewrites the whole PhiQ array once per symmetry and
// frequency; here the eight operations are folded on the host into at most
// four (source l-pair, coefficient[R]) entries per (l1, l2, Q), so every Phi
// element is read once per distinct source and PhiR is written once: the
// kernel is HBM-bound on the read of Phi.
#include "internal.h"

namespace tups {

__global__ __launch_bounds__(256) void phir_kernel(const double2* __restrict__ phi0,
                                                   const double2* __restrict__ phi1,
                                                   const double2* __restrict__ phi2,
                                                   const double2* __restrict__ phi3,
                                                   double2* out0, double2* out1, double2* out2,
                                                   double2* out3, const int* __restrict__ cnt,
                                                   const int* __restrict__ src,
                                                   const double2* __restrict__ coef, int nf,
                                                   int nl, int nw, int nibz, int nr) {
  const int nz = nf * nl;
  const size_t nz2 = (size_t)nz * nz;
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nz2 * nw) return;
  const int ch = blockIdx.y;
  const double2* phi = ch == 0 ? phi0 : ch == 1 ? phi1 : ch == 2 ? phi2 : phi3;
  double2* out = ch == 0 ? out0 : ch == 1 ? out1 : ch == 2 ? out2 : out3;
  const size_t z = gid % nz2;
  const int w = (int)(gid / nz2);
  const int row = (int)(z % nz), col = (int)(z / nz);
  const int l1 = row / nf, nu1 = row - l1 * nf;
  const int l2 = col / nf, nu2 = col - l2 * nf;

  double ar[kMaxNR], ai[kMaxNR];
#pragma unroll
  for (int r = 0; r < kMaxNR; ++r) ar[r] = ai[r] = 0.0;

  const size_t e0 = ((size_t)l1 * nl + l2) * nibz;
  for (int m = 0; m < nibz; ++m) {
    const size_t e = e0 + m;
    const double2* base = phi + nz2 * ((size_t)m * nw + w);
    const int n = cnt[e];
    for (int j = 0; j < n; ++j) {
      const int key = src[e * 4 + j];
      const int l1s = key % nl, l2s = key / nl;
      const double2 v = base[(size_t)(l1s * nf + nu1) + (size_t)nz * (l2s * nf + nu2)];
      const double2* cf = coef + (e * 4 + j) * nr;
#pragma unroll
      for (int r = 0; r < kMaxNR; ++r) {
        if (r < nr) {
          const double2 c = cf[r];
          ar[r] += c.x * v.x - c.y * v.y;
          ai[r] += c.x * v.y + c.y * v.x;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kMaxNR; ++r)
    if (r < nr) out[z + nz2 * ((size_t)w + (size_t)nw * r)] = make_double2(ar[r], ai[r]);
}

int launch_phir(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  const size_t n = (size_t)s.nz * s.nz * s.nw;
  dim3 g((unsigned)((n + 255) / 256), 4);
  if (int rc = timer_begin(c)) return rc;
  hipLaunchKernelGGL(phir_kernel, g, dim3(256), 0, c->stream, c->gamma[0], c->gamma[1],
                     c->gamma[2], c->gamma[3], c->phir[0], c->phir[1], c->phir[2], c->phir[3],
                     c->d_ph_cnt, c->d_ph_src, c->d_ph_coef, s.nf, s.nl, s.nw, s.nibz, s.nr);
  TUPS_HIP(hipGetLastError());
  if (int rc = timer_end(c, TUPS_K_PHIR)) return rc;
  return TUPS_OK;
}

}  // namespace tups
