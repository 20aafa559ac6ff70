// Bubbles chi_ph, chi_pp in form-factor space for every q slot:
// calc_chi (SRC/parquet_chi.f90:438-523) with get_green_coarse
// (SRC/parquet_util.f90:1303-1327), Ek_grain (SRC/parquet_util.f90:143-164)
// and the index maps index_FaddB / index_minusF (SRC/parquet_util.f90:1345-1381).
//
//   chi_k_ph(k, nu) = 1/Ng^2 sum_g G(g; k, nu) G(g; k + q, nu + w - 1)
//   chi_k_pp(k, nu) = 1/Ng^2 sum_g G(g; k, nu) G(Ng + 1 - g; q - k, Nf - nu + w)
//   chi_X(l1, l2, nu) = sum_k chi_k_X(k, nu) FF(k, l1) FF(k, l2)
// G(g; k, nu) = 1 / (i nu_n - Ek_grain(g, k) + mu - Sigma(k, nu)), with the
// Hartree value Sigma_H(k) outside the box 1..Nf.  Only the Nl x Nl blocks of
// the block-diagonal bubble are formed (the resident TUPS_CHI layout).
//
// One workgroup per q slot; the frequency range is walked in chunks: the
// k-space bubble of a chunk is formed in LDS, then projected on the l pairs
// with k summed in the reference's order (kx fastest).
#include "green.h"
#include "internal.h"

#include <cmath>

namespace tups {

struct ChiArgs {
  double2* chi_ph;
  double2* chi_pp;
  GreenParams g;
  const double* ff;       // FF[l Nc + ky Nx + kx]
  const int* qm;          // IBZ momentum slot of each q
  const int* qw;          // bosonic index of each q (1-based)
  const int* ibz;         // IBZ momenta: x[0..nibz), then y[0..nibz) (1-based)
  int nx, nf, nl, nibz, ngrain, chunk;
};

// grid: nq; dynamic LDS 2 * chunk * Nc * 16 bytes
__global__ __launch_bounds__(256) void chi_kernel(ChiArgs a) {
  extern __shared__ double2 ck[];  // [2][chunk][Nc]
  const GreenParams g = green_resolve(a.g);
  const int q = blockIdx.x;
  const int nx = a.nx, nc = nx * nx, nf = a.nf, nl = a.nl, ng = a.ngrain;
  const int m = a.qm[q];
  const int qx = a.ibz[m], qy = a.ibz[a.nibz + m], w = a.qw[q];
  double2* cph = ck;
  double2* cpp = ck + (size_t)a.chunk * nc;
  const double inv_g2 = 1.0 / ng / ng;
  for (int nu0 = 0; nu0 < nf; nu0 += a.chunk) {
    const int nch = min(a.chunk, nf - nu0);
    for (int u = threadIdx.x; u < nch * nc; u += blockDim.x) {
      const int k = u % nc, o = u / nc;
      const int kx = k % nx + 1, ky = k / nx + 1, nu = nu0 + o + 1;
      const int kqx = (kx - 1 + qx - 1) % nx + 1, kqy = (ky - 1 + qy - 1) % nx + 1;
      const int mkx = (-(kx - 1) + nx) % nx + 1, mky = (-(ky - 1) + nx) % nx + 1;
      const int qmkx = (mkx - 1 + qx - 1) % nx + 1, qmky = (mky - 1 + qy - 1) % nx + 1;
      const int nu_kq = nu + w - 1, nu_qmk = (-nu + nf + 1) + w - 1;
      double2 sph = make_double2(0.0, 0.0), spp = make_double2(0.0, 0.0);
      for (int gx = 1; gx <= ng; ++gx)
        for (int gy = 1; gy <= ng; ++gy) {
          const double2 gk = green_coarse(g, gx, gy, kx, ky, nu);
          const double2 p = cmul(gk, green_coarse(g, gx, gy, kqx, kqy, nu_kq));
          sph.x += p.x * inv_g2;
          sph.y += p.y * inv_g2;
          const double2 t = cmul(gk, green_coarse(g, ng - gx + 1, ng - gy + 1, qmkx, qmky, nu_qmk));
          spp.x += t.x * inv_g2;
          spp.y += t.y * inv_g2;
        }
      cph[(size_t)o * nc + k] = sph;
      cpp[(size_t)o * nc + k] = spp;
    }
    __syncthreads();
    for (int u = threadIdx.x; u < nch * nl * nl; u += blockDim.x) {
      const int l = u % (nl * nl), o = u / (nl * nl);
      const int l1 = l % nl, l2 = l / nl;
      const double* f1 = a.ff + (size_t)l1 * nc;
      const double* f2 = a.ff + (size_t)l2 * nc;
      double2 sph = make_double2(0.0, 0.0), spp = make_double2(0.0, 0.0);
      for (int k = 0; k < nc; ++k) {  // k = (ky - 1) Nx + kx - 1: kx fastest
        const double wgt = f1[k] * f2[k];
        const double2 xp = cph[(size_t)o * nc + k], xq = cpp[(size_t)o * nc + k];
        sph.x += xp.x * wgt;
        sph.y += xp.y * wgt;
        spp.x += xq.x * wgt;
        spp.y += xq.y * wgt;
      }
      const size_t idx = (size_t)l1 + (size_t)nl * l2 + (size_t)nl * nl * (nu0 + o) + (size_t)nl * nl * nf * q;
      a.chi_ph[idx] = sph;
      a.chi_pp[idx] = spp;
    }
    __syncthreads();
  }
}

// Sigma and Sigma_H to the device (Sigma_H defaults to Sigma at the last box
// frequency, SRC/parquet_chi.f90:103-110); synchronous w.r.t. the host buffers
int upload_sigma(tups_ctx* c, const double* sigma, const double* sigma_h, GreenParams* g, double mu, int ngrain) {
  const tups_sizes& s = c->sz;
  const int nc = s.nx * s.ny;
  const size_t nt = (size_t)nc * s.nf;
  TUPS_CHECK(sigma, TUPS_EINVAL, "null Sigma");
  TUPS_CHECK(ngrain >= 1 && ngrain <= 64, TUPS_EINVAL, "ngrain must be in 1..64");
  if (!c->d_sigma) TUPS_HIP(hipMalloc((void**)&c->d_sigma, nt * sizeof(double2)));
  if (!c->d_sigh) TUPS_HIP(hipMalloc((void**)&c->d_sigh, (size_t)nc * sizeof(double2)));
  if (!c->d_sigma_old) TUPS_HIP(hipMalloc((void**)&c->d_sigma_old, nt * sizeof(double2)));
  c->sigma_old_valid = false;
  std::vector<double> sh(2 * (size_t)nc);
  if (sigma_h) {
    std::copy(sigma_h, sigma_h + 2 * (size_t)nc, sh.begin());
  } else {
    for (int kx = 1; kx <= s.nx; ++kx)
      for (int ky = 1; ky <= s.ny; ++ky) {
        const size_t src = ((size_t)(kx - 1) * s.ny + (ky - 1)) * s.nf + (s.nf - 1);
        const size_t dst = (size_t)(kx - 1) + (size_t)s.nx * (ky - 1);
        sh[2 * dst] = sigma[2 * src];
        sh[2 * dst + 1] = sigma[2 * src + 1];
      }
  }
  TUPS_HIP(hipMemcpyAsync(c->d_sigma, sigma, nt * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  TUPS_HIP(hipMemcpyAsync(c->d_sigh, sh.data(), (size_t)nc * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  if (!c->d_gstate) TUPS_HIP(hipMalloc((void**)&c->d_gstate, sizeof(GtauState)));
  TUPS_HIP(hipMemcpyAsync(&c->d_gstate->mu, &mu, sizeof(double), hipMemcpyHostToDevice, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  g->sigma = c->d_sigma;
  g->sigh = c->d_sigh;
  g->nx = s.nx;
  g->nf = s.nf;
  g->ngrain = ngrain;
  g->mu = mu;
  g->beta = c->beta;
  g->mu_dev = nullptr;
  c->green_mu = mu;
  c->green_ngrain = ngrain;
  c->sigh_default = sigma_h == nullptr;
  c->green_set = true;
  return TUPS_OK;
}

// calc_chi from the Sigma already on the device (g)
int launch_chi_device(tups_ctx* c, const GreenParams& g) {
  const tups_sizes& s = c->sz;
  const int nc = s.nx * s.ny;
  ChiArgs a{};
  a.g = g;
  a.chi_ph = c->chi[0];
  a.chi_pp = c->chi[1];
  a.ff = c->d_ffk;
  a.qm = c->d_qm;
  a.qw = c->d_qw;
  a.ibz = c->d_ibz;
  a.nx = s.nx;
  a.nf = s.nf;
  a.nl = s.nl;
  a.nibz = s.nibz;
  a.ngrain = g.ngrain;
  a.chunk = std::max(1, std::min(s.nf, 2048 / nc));
  const size_t lds = 2 * (size_t)a.chunk * nc * sizeof(double2);
  TUPS_CHECK(lds <= 160 * 1024, TUPS_EINVAL, "calc_chi supports Nx <= 64");
  if (lds > 65536)
    TUPS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(chi_kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (int rc = timer_begin(c)) return rc;
  hipLaunchKernelGGL(chi_kernel, dim3((unsigned)s.nq), dim3(256), lds, c->stream, a);
  TUPS_HIP(hipGetLastError());
  if (int rc = timer_end(c, TUPS_K_CALC_CHI)) return rc;
  return TUPS_OK;
}

int launch_calc_chi(tups_ctx* c, const double* sigma, const double* sigma_h, double mu, int ngrain) {
  GreenParams g{};
  if (int rc = upload_sigma(c, sigma, sigma_h, &g, mu, ngrain)) return rc;
  return launch_chi_device(c, g);
}

}  // namespace tups
