// Transformation tensors M(l1,l2,l3,l4,Q,R) of the truncated-unity crossing.
//
// Reference: calc_M_phbar_to_ph, calc_M_pp_to_ph, calc_M_ph_to_pp,
// calc_M_phbar_to_pp, SRC/parquet_EQMatrix.f90:21-360 (run once, at ite 0,
// SRC/loop.f90:115-121).  With E(k,R,S) = expqSR (e^{-i k.SR} on the star of R):
//   phbar->ph : sum_S [sum_k E* FF(k,l1) FFi(S,k,l3) FFi(S,k+Q,l4)] [sum_k E FF(k,l2)]
//   pp->ph    : sum_S E(Q,R,S) [sum_k FF(k,l1) FFi(S,k,l3) E] [sum_k FF(k,l2) FFi(S,k,l4) E]
//   ph->pp    : as pp->ph with E* inside both k-sums
//   phbar->pp : sum_S [sum_k FF(k,l1) FFi(S,k,l3) E*] [sum_k FFi(S,Q-k,l4) FF(k,l2) E]
// The tensors depend on q only through its momentum, so they are built and
// stored once per IBZ momentum (the reference keeps one copy per local q).
// One thread per output element evaluates its k-sums directly; this is a
// set-up step of O(Ns NR NIBZ Nl^4 Nc) work.
#include "internal.h"

namespace tups {

__global__ __launch_bounds__(256) void build_m_kernel(double2* m0, double2* m1, double2* m2,
                                                      double2* m3, const double* __restrict__ ffk,
                                                      const double* __restrict__ ffik,
                                                      const double2* __restrict__ eq, int nx,
                                                      int nl, int nibz, int nr,
                                                      const int* __restrict__ ibzx,
                                                      const int* __restrict__ ibzy) {
  const int nc = nx * nx;
  const size_t nl4 = (size_t)nl * nl * nl * nl;
  const size_t total = nl4 * nibz * nr;
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const int which = blockIdx.y;
  size_t t = gid;
  const int l1 = (int)(t % nl); t /= nl;
  const int l2 = (int)(t % nl); t /= nl;
  const int l3 = (int)(t % nl); t /= nl;
  const int l4 = (int)(t % nl); t /= nl;
  const int m = (int)(t % nibz); t /= nibz;
  const int R = (int)t;
  const int qx = ibzx[m] - 1, qy = ibzy[m] - 1;
  const double* f1 = ffk + (size_t)l1 * nc;
  const double* f2 = ffk + (size_t)l2 * nc;

  double mr = 0.0, mi = 0.0;
  for (int s = 0; s < kNs; ++s) {
    const double2* e = eq + (size_t)(s * nr + R) * nc;
    const double* fi3 = ffik + (size_t)(s * nl + l3) * nc;
    const double* fi4 = ffik + (size_t)(s * nl + l4) * nc;
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
    if (which == 0) {  // phbar -> ph
      for (int ky = 0; ky < nx; ++ky)
        for (int kx = 0; kx < nx; ++kx) {
          const int k = ky * nx + kx;
          const int kq = ((ky + qy) % nx) * nx + (kx + qx) % nx;
          const double w = f1[k] * fi3[k] * fi4[kq];
          ar += e[k].x * w;
          ai -= e[k].y * w;
          br += e[k].x * f2[k];
          bi += e[k].y * f2[k];
        }
      mr += ar * br - ai * bi;
      mi += ar * bi + ai * br;
    } else if (which == 1 || which == 2) {  // pp -> ph, ph -> pp
      const double cj = which == 1 ? 1.0 : -1.0;
      for (int k = 0; k < nc; ++k) {
        const double wa = f1[k] * fi3[k];
        const double wb = f2[k] * fi4[k];
        ar += e[k].x * wa;
        ai += cj * e[k].y * wa;
        br += e[k].x * wb;
        bi += cj * e[k].y * wb;
      }
      const double2 eQ = e[qy * nx + qx];
      const double pr = ar * br - ai * bi, pi = ar * bi + ai * br;
      mr += eQ.x * pr - eQ.y * pi;
      mi += eQ.x * pi + eQ.y * pr;
    } else {  // phbar -> pp
      for (int ky = 0; ky < nx; ++ky)
        for (int kx = 0; kx < nx; ++kx) {
          const int k = ky * nx + kx;
          const int qk = ((qy - ky + nx) % nx) * nx + (qx - kx + nx) % nx;
          const double wa = f1[k] * fi3[k];
          ar += e[k].x * wa;
          ai -= e[k].y * wa;
          const double wb = fi4[qk] * f2[k];
          br += e[k].x * wb;
          bi += e[k].y * wb;
        }
      mr += ar * br - ai * bi;
      mi += ar * bi + ai * br;
    }
  }
  double2* out = which == 0 ? m0 : which == 1 ? m1 : which == 2 ? m2 : m3;
  out[gid] = make_double2(mr, mi);
}

int launch_build_m(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  const size_t total = (size_t)s.nl * s.nl * s.nl * s.nl * s.nibz * s.nr;
  const int* ibz = c->d_ibz;
  dim3 g((unsigned)((total + 255) / 256), 4);
  if (int rc = timer_begin(c)) return rc;
  hipLaunchKernelGGL(build_m_kernel, g, dim3(256), 0, c->stream, c->mten[0], c->mten[1], c->mten[2],
                     c->mten[3], c->d_ffk, c->d_ffik, c->d_eqsr, s.nx, s.nl, s.nibz, s.nr, ibz,
                     ibz + s.nibz);
  TUPS_HIP(hipGetLastError());
  if (int rc = timer_end(c, TUPS_K_BUILD_M)) return rc;
  return TUPS_OK;
}

}  // namespace tups
