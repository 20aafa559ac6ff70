// Susceptibilities of the four channels from the resident vertices, FP64.
//
// Reference: calc_sus, SRC/parquet_sus.f90:18-362, with calc_GG_ph / calc_GG_pp
// (:528-643), fillpickNu_range (:708-731) and the external form factors
// fill_FF_sdwave (SRC/parquet_formfactors.f90:792-815: 1, cos px - cos py,
// sin px, sin px sin py).  Two kernels:
//  1. sus_gg: per q slot the Green's-function pairs of the enlarged box
//       GG_ph(l, nu; ff) = 1/Ng^2 sum_{k, g} FF_sd(k, ff) FF(k, l) G(g; k, nu) G(g; k + q, nu + w - 1)
//       GG_pp(l, nu; ff) = 1/Ng^2 sum_{k, g} FF_sd(k, ff) FF(k, l) G(g; k, nu) G(Ng + 1 - g; q - k, Nf - nu + w)
//     with k split over lanes as in the SDE's auxiliary bubbles;
//  2. sus_contract: per (q, channel) sus(ff) = sum_{L1, L2} V(L1, L2) GG(L1) GG(L2) / (Nx^2 Ny^2 beta^2),
//     V = F inside the box and, for the pairs not both inside, the bare vertex
//     (U, -U, 2U, 0 for d, m, s, t) / FF11^2 on the l = 1 block plus the
//     reducible vertex Phi at the clamped frequencies.  Lanes run along L1 (so
//     the vertex columns are read coalesced), each with the four form factors'
//     partial sums; one LDS reduction per workgroup.
// The loop's G_r_LL hold Phi when the reference calls calc_sus (before G = F - G,
// SRC/loop.f90:193-222); after a device iteration they hold Gamma = F - Phi, and
// Phi is read back as F - Gamma.
#include "green.h"
#include "internal.h"

#include <cmath>

namespace tups {

namespace {

struct SusArgs {
  GreenParams g;
  const double2* fv[4];
  const double2* gam[4];  // Phi (after_update = 0) or Gamma = F - Phi (after_update = 1)
  const double* ff;       // FF[l Nc + (ky-1) Nx + kx-1]
  const int* qm;
  const int* qw;
  const int* ibz;
  double2* gg;            // [q][2 (ph, pp)][4 ff][nent]
  double2* out;           // [r][ff][q]
  int nx, nf, nl, nz, nq, nibz, ext, f_range, after_update;
  double bare[4], inv_ff11_sq, norm;
};

// grid: nq
__global__ __launch_bounds__(256) void sus_gg(SusArgs a) {
  __shared__ double2 red[8][256];
  __shared__ double fsd[4][256];  // FF_sdwave, Nc <= 256
  const GreenParams g = green_resolve(a.g);
  const int q = blockIdx.x;
  const int m = a.qm[q];
  const int qx = a.ibz[m], qy = a.ibz[a.nibz + m], w = a.qw[q];
  const int nx = a.nx, nc = nx * nx, nf = a.nf, ng = g.ngrain;
  const double pi = 3.14159265358979323846;
  for (int k = threadIdx.x; k < nc; k += 256) {
    const int kx = k % nx + 1, ky = k / nx + 1;
    const double px = 2.0 / nx * pi * (kx - 1), py = 2.0 / nx * pi * (ky - 1);
    fsd[0][k] = 1.0;
    fsd[1][k] = cos(px) - cos(py);
    fsd[2][k] = sin(px);
    fsd[3][k] = sin(px) * sin(py);
  }
  __syncthreads();
  const int nent = a.nl * a.ext;
  int ks = 1;
  while (ks * 2 * nent <= 256 && ks * 2 <= nc) ks *= 2;
  const int per = 256 / ks, part = threadIdx.x % ks, slot = threadIdx.x / ks;
  const double inv_g2 = 1.0 / ng / ng;
  double2* out = a.gg + (size_t)q * 8 * nent;
  for (int e0 = 0; e0 < nent; e0 += per) {
    const int e = e0 + slot;
    double2 acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = make_double2(0.0, 0.0);
    if (e < nent) {
      const int l = e / a.ext, nu = e - l * a.ext - a.f_range * nf + 1;
      for (int k = part; k < nc; k += ks) {
        const int kx = k % nx + 1, ky = k / nx + 1;
        const int kqx = (kx - 1 + qx - 1) % nx + 1, kqy = (ky - 1 + qy - 1) % nx + 1;
        const int mkx = (-(kx - 1) + nx) % nx + 1, mky = (-(ky - 1) + nx) % nx + 1;
        const int qmkx = (mkx - 1 + qx - 1) % nx + 1, qmky = (mky - 1 + qy - 1) % nx + 1;
        double2 sph = make_double2(0.0, 0.0), spp = make_double2(0.0, 0.0);
        for (int gy = 1; gy <= ng; ++gy)
          for (int gx = 1; gx <= ng; ++gx) {
            const double2 gk = green_coarse(g, gx, gy, kx, ky, nu);
            const double2 p = cmul(gk, green_coarse(g, gx, gy, kqx, kqy, nu + w - 1));
            const double2 t = cmul(gk, green_coarse(g, ng - gx + 1, ng - gy + 1, qmkx, qmky, -nu + nf + 1 + w - 1));
            sph.x += p.x * inv_g2;
            sph.y += p.y * inv_g2;
            spp.x += t.x * inv_g2;
            spp.y += t.y * inv_g2;
          }
        const double fl = a.ff[(size_t)l * nc + k];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          const double wgt = fsd[f][k] * fl;
          acc[f].x += wgt * sph.x;
          acc[f].y += wgt * sph.y;
          acc[4 + f].x += wgt * spp.x;
          acc[4 + f].y += wgt * spp.y;
        }
      }
    }
    if (ks == 1) {
      if (e < nent)
#pragma unroll
        for (int i = 0; i < 8; ++i) out[(size_t)i * nent + e] = acc[i];
      continue;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) red[i][threadIdx.x] = acc[i];
    __syncthreads();
    if (part == 0 && e < nent) {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        double2 s = red[i][threadIdx.x];
        for (int t = 1; t < ks; ++t) {
          s.x += red[i][threadIdx.x + t].x;
          s.y += red[i][threadIdx.x + t].y;
        }
        out[(size_t)i * nent + e] = s;
      }
    }
    __syncthreads();
  }
}

// grid: (nq, 4 channels)
__global__ __launch_bounds__(256) void sus_contract(SusArgs a) {
  __shared__ double2 red[4][256];
  const int q = blockIdx.x, r = blockIdx.y;
  const int nf = a.nf, ext = a.ext, nent = a.nl * ext;
  const size_t nz = a.nz, base = (size_t)q * nz * nz;
  const double2* gq = a.gg + ((size_t)q * 8 + (r < 2 ? 0 : 4)) * nent;  // [ff][nent]
  const double2* F = a.fv[r] + base;
  const double2* G = a.gam[r] + base;
  const double bare = a.bare[r] * a.inv_ff11_sq;
  double2 part[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) part[f] = make_double2(0.0, 0.0);
  for (int i1 = threadIdx.x; i1 < nent; i1 += 256) {
    const int l1 = i1 / ext, n1 = i1 - l1 * ext - a.f_range * nf + 1;
    const bool in1 = n1 >= 1 && n1 <= nf;
    const size_t row = (size_t)l1 * nf + (n1 < 1 ? 1 : (n1 > nf ? nf : n1)) - 1;  // pickNu
    double2 acc[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) acc[f] = make_double2(0.0, 0.0);
    for (int i2 = 0; i2 < nent; ++i2) {
      const int l2 = i2 / ext, n2 = i2 - l2 * ext - a.f_range * nf + 1;
      const bool in2 = n2 >= 1 && n2 <= nf;
      const size_t col = (size_t)l2 * nf + (n2 < 1 ? 1 : (n2 > nf ? nf : n2)) - 1;
      const size_t idx = row + nz * col;
      double2 v;
      if (in1 && in2) {
        v = F[idx];
      } else {
        v = G[idx];
        if (a.after_update) {  // Phi = F - Gamma
          const double2 fv = F[idx];
          v = make_double2(fv.x - v.x, fv.y - v.y);
        }
        if (l1 == 0 && l2 == 0) v.x += bare;
      }
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const double2 t = cmul(v, gq[(size_t)f * nent + i2]);
        acc[f].x += t.x;
        acc[f].y += t.y;
      }
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const double2 t = cmul(gq[(size_t)f * nent + i1], acc[f]);
      part[f].x += t.x;
      part[f].y += t.y;
    }
  }
#pragma unroll
  for (int f = 0; f < 4; ++f) red[f][threadIdx.x] = part[f];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h)
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        red[f][threadIdx.x].x += red[f][threadIdx.x + h].x;
        red[f][threadIdx.x].y += red[f][threadIdx.x + h].y;
      }
    __syncthreads();
  }
  if (threadIdx.x < 4) {
    const double2 s = red[threadIdx.x][0];
    a.out[((size_t)r * 4 + threadIdx.x) * a.nq + q] = make_double2(s.x * a.norm, s.y * a.norm);
  }
}

}  // namespace

}  // namespace tups

using namespace tups;

extern "C" int tups_calc_sus(tups_ctx* c, double u, int32_t f_range, int32_t after_update, double* sus_out) {
  TUPS_CHECK(c && sus_out, TUPS_EINVAL, "null argument");
  const tups_sizes& s = c->sz;
  TUPS_CHECK(s.full_ibz, TUPS_EINVAL, "the susceptibilities need every IBZ q of the box (nq = 0 at create)");
  TUPS_CHECK(f_range >= 0 && f_range <= 8, TUPS_EINVAL, "f_range must be in 0..8");
  TUPS_CHECK(s.nx * s.ny <= 256, TUPS_EINVAL, "the susceptibilities support Nx <= 16");
  int rc = ensure_tables(c);
  if (rc) return rc;
  GreenParams g{};
  if ((rc = green_params(c, &g))) return rc;
  SusArgs a{};
  a.g = g;
  a.ext = (2 * f_range + 1) * s.nf;
  const size_t nent = (size_t)s.nl * a.ext;
  const size_t need = ((size_t)s.nq * 8 * nent + 16 * (size_t)s.nq) * sizeof(double2);
  if (c->sus_ws_bytes < need) {
    if (c->sus_ws) TUPS_HIP(hipFree(c->sus_ws));
    c->sus_ws = nullptr;
    c->sus_ws_bytes = 0;
    TUPS_HIP(hipMalloc(&c->sus_ws, need));
    c->sus_ws_bytes = need;
  }
  a.gg = static_cast<double2*>(c->sus_ws);
  a.out = a.gg + (size_t)s.nq * 8 * nent;
  for (int r = 0; r < 4; ++r) {
    a.fv[r] = c->fvert[r];
    a.gam[r] = c->gamma[r];
  }
  a.ff = c->d_ffk;
  a.qm = c->d_qm;
  a.qw = c->d_qw;
  a.ibz = c->d_ibz;
  a.nx = s.nx;
  a.nf = s.nf;
  a.nl = s.nl;
  a.nz = s.nz;
  a.nq = s.nq;
  a.nibz = s.nibz;
  a.f_range = f_range;
  a.after_update = after_update ? 1 : 0;
  a.bare[0] = u;
  a.bare[1] = -u;
  a.bare[2] = 2.0 * u;
  a.bare[3] = 0.0;
  const double f11 = c->tab.ff[0];
  a.inv_ff11_sq = 1.0 / (f11 * f11);
  a.norm = 1.0 / ((double)s.nx * s.nx * s.ny * s.ny * c->beta * c->beta);
  if ((rc = timer_begin(c))) return rc;
  hipLaunchKernelGGL(sus_gg, dim3((unsigned)s.nq), dim3(256), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  hipLaunchKernelGGL(sus_contract, dim3((unsigned)s.nq, 4), dim3(256), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  if ((rc = timer_end(c, TUPS_K_SUS))) return rc;
  TUPS_HIP(hipMemcpyAsync(sus_out, a.out, 16 * (size_t)s.nq * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  return TUPS_OK;
}
