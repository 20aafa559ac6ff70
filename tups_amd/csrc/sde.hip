// Self-energy from the Schwinger-Dyson equation in form-factor space, FP64.
//
// Reference: calc_SDE, SRC/parquet_SDE.f90:27-258 (its use_U2 second-order term
// calc_order_2nd, :263-340, is added from the device Grt by gtau.hip), with
// calc_chi_aux (:346-432), calc_LChi (:436-497), calc_PhChi
// (:501-537), calc_PPChi (:540-575) and symmetrize_array
// (SRC/parquet_util.f90:1215-1272).  Three kernels:
//  1. sde_chi_aux: per IBZ q, the auxiliary bubbles
//       chi_aux_X(l4, nu) = 1/Ng^2 sum_{k, g} G(g; k, nu) G(. ; k (+/-) q, .) FF(k, l4)
//     over the enlarged box nu in (-f_range Nf, (f_range + 1) Nf];
//  2. sde_vchi: per (q, L3), PL = LChi + PhChi and PP = PPChi, i.e. the
//     resident reducible vertices contracted with the bubbles (one pass over
//     the four G_r arrays, coalesced along L3);
//  3. sde_sigma: one workgroup per IBZ (k, nu): the sum over q slots and the
//     star of each q with the Green's functions at k + Sq, -k + Sq, k - Sq,
//     reduced in LDS, scaled by -U / beta^2 / Nc^2 / 2 and written to every
//     point of k's star (the stars partition the zone, so each Sigma entry is
//     written once).
#include "green.h"
#include "internal.h"

namespace tups {

struct SdeArgs {
  GreenParams g;
  const double2* gam[4];    // reducible vertices G_r_LL (Nz, Nz, Nq)
  const double2* lam_d;     // L_d (Nf, Nf, Nw)
  const double2* lam_m;
  const double* ff;         // FF[l Nc + (ky-1) Nx + kx-1]
  const double* ffi;        // FF_inv[(s Nl + l) Nc + (ky-1) Nx + kx-1]
  const int* signs;
  const int* qm;            // IBZ momentum slot of each q
  const int* qw;            // bosonic index of each q (1-based)
  const int* ibz;           // IBZ momenta: x[0..nibz), then y[0..nibz) (1-based)
  double2* caux_ph;         // [q][l4 ext + f_range Nf + nu - 1]
  double2* caux_pp;
  double2* pl;              // [q][L3]: LChi + PhChi
  double2* pp;              // [q][L3]: PPChi
  double2* sigma_out;       // (Nx Ny Nf)
  int nx, nf, nl, nz, nq, nibz, nxi, ext, f_range, use_u2;
  double u, inv_ff11_sq, pref;
};

// list_symmetries (SRC/parquet_util.f90:1140-1207): bit s-1 set when operation s
// generates a point of the star of the IBZ point (kx, ky)
__device__ __forceinline__ unsigned star_ops(int kx, int ky, int nxi) {
  if ((kx == 1 && ky == 1) || (kx == nxi && ky == nxi)) return 0x01u;
  if (kx == nxi && ky == 1) return 0x11u;
  if (kx == nxi && ky > 1 && ky < nxi) return 0x01u | 0x02u | 0x10u | 0x80u;
  if (kx > 1 && ky < nxi) {
    if (ky == 1) return 0x01u | 0x04u | 0x10u | 0x20u;
    if (ky == kx) return 0x0fu;
    return 0xffu;
  }
  return 0u;
}

// symmetry_operation on 1-based momentum indices (SRC/parquet_util.f90:896-978)
__device__ __forceinline__ void sym_apply(int s, int ix, int iy, int nx, int& ox, int& oy) {
  const int x = ix - 1, y = iy - 1;
  int a = x, b = y;
  switch (s) {
    case 2: a = x; b = -y; break;
    case 3: a = -x; b = y; break;
    case 4: a = -x; b = -y; break;
    case 5: a = y; b = x; break;
    case 6: a = y; b = -x; break;
    case 7: a = -y; b = -x; break;
    case 8: a = -y; b = x; break;
    default: break;
  }
  ox = ((a % nx) + nx) % nx + 1;
  oy = ((b % nx) + nx) % nx + 1;
}

// grid: nq; the (l4, nu) entries of the enlarged box, each summed over k by `ks`
// lanes (k = part, part + ks, ...; ks = 256 / entries rounded down to a power of
// two, 1 once there are 256 entries or more) and the lanes' sums added in order
__global__ __launch_bounds__(256) void sde_chi_aux(SdeArgs a) {
  __shared__ double2 red[2][256];
  const GreenParams g = green_resolve(a.g);
  const int q = blockIdx.x;
  const int m = a.qm[q];
  const int qx = a.ibz[m], qy = a.ibz[a.nibz + m], w = a.qw[q];
  const int nx = a.nx, nc = nx * nx, nf = a.nf, ng = g.ngrain;
  const int nent = a.nl * a.ext;
  int ks = 1;
  while (ks * 2 * nent <= 256 && ks * 2 <= nc) ks *= 2;
  const int per = 256 / ks;  // entries per round
  const int part = threadIdx.x % ks, slot = threadIdx.x / ks;
  const double inv_g2 = 1.0 / ng / ng;
  for (int e0 = 0; e0 < nent; e0 += per) {
    const int e = e0 + slot;
    double2 sph = make_double2(0.0, 0.0), spp = make_double2(0.0, 0.0);
    if (e < nent) {
      const int l4 = e / a.ext, nu = e - l4 * a.ext - a.f_range * nf + 1;  // nu in (-f_range Nf, (f_range + 1) Nf]
      for (int k = part; k < nc; k += ks) {
        const int kx = k % nx + 1, ky = k / nx + 1;
        const double f = a.ff[(size_t)l4 * nc + k];
        const int kqx = (kx - 1 + qx - 1) % nx + 1, kqy = (ky - 1 + qy - 1) % nx + 1;
        const int mkx = (-(kx - 1) + nx) % nx + 1, mky = (-(ky - 1) + nx) % nx + 1;
        const int qmkx = (mkx - 1 + qx - 1) % nx + 1, qmky = (mky - 1 + qy - 1) % nx + 1;
        double2 tph = make_double2(0.0, 0.0), tpp = make_double2(0.0, 0.0);
        for (int gx = 1; gx <= ng; ++gx)
          for (int gy = 1; gy <= ng; ++gy) {
            const double2 gk = green_coarse(g, gx, gy, kx, ky, nu);
            const double2 p = cmul(gk, green_coarse(g, gx, gy, kqx, kqy, nu + w - 1));
            tph.x += p.x * f * inv_g2;
            tph.y += p.y * f * inv_g2;
            const double2 t = cmul(gk, green_coarse(g, ng - gx + 1, ng - gy + 1, qmkx, qmky, -nu + nf + 1 + w - 1));
            tpp.x += t.x * f * inv_g2;
            tpp.y += t.y * f * inv_g2;
          }
        sph.x += tph.x;
        sph.y += tph.y;
        spp.x += tpp.x;
        spp.y += tpp.y;
      }
    }
    if (ks == 1) {
      if (e < nent) {
        a.caux_ph[(size_t)q * nent + e] = sph;
        a.caux_pp[(size_t)q * nent + e] = spp;
      }
      continue;
    }
    red[0][threadIdx.x] = sph;
    red[1][threadIdx.x] = spp;
    __syncthreads();
    if (part == 0 && e < nent) {
      double2 tph = red[0][threadIdx.x], tpp = red[1][threadIdx.x];
      for (int i = 1; i < ks; ++i) {
        tph.x += red[0][threadIdx.x + i].x;
        tph.y += red[0][threadIdx.x + i].y;
        tpp.x += red[1][threadIdx.x + i].x;
        tpp.y += red[1][threadIdx.x + i].y;
      }
      a.caux_ph[(size_t)q * nent + e] = tph;
      a.caux_pp[(size_t)q * nent + e] = tpp;
    }
    __syncthreads();
  }
}

// grid: (ceil(Nz / 256), nq); one thread per L3 = (l3, nu)
__global__ __launch_bounds__(256) void sde_vchi(SdeArgs a) {
  const int q = blockIdx.y;
  const int i3 = blockIdx.x * blockDim.x + threadIdx.x;
  if (i3 >= a.nz) return;
  const int nf = a.nf, w = a.qw[q];
  const size_t nz = a.nz, base = (size_t)q * nz * nz;
  const double2* cph = a.caux_ph + (size_t)q * a.nl * a.ext;
  const double2* cpp = a.caux_pp + (size_t)q * a.nl * a.ext;
  double2 pl = make_double2(0.0, 0.0), pp = make_double2(0.0, 0.0);
  for (int l4 = 0; l4 < a.nl; ++l4)
    for (int o = 0; o < a.ext; ++o) {
      const int nu = o - a.f_range * nf + 1;  // l4_0
      const int col = l4 * nf + (nu < 1 ? 1 : (nu > nf ? nf : nu)) - 1;  // pickNu_range
      const size_t idx = base + (size_t)col * nz + i3;
      const double2 gd = ld_stream(a.gam[0] + idx), gm = ld_stream(a.gam[1] + idx);
      const double2 gs = ld_stream(a.gam[2] + idx), gt = ld_stream(a.gam[3] + idx);
      const double2 xp = cph[l4 * a.ext + o], xq = cpp[l4 * a.ext + o];
      const double2 vp = make_double2(gd.x - 3.0 * gm.x, gd.y - 3.0 * gm.y);
      const double2 vq = make_double2(gs.x + gt.x, gs.y + gt.y);
      double2 t = cmul(vp, xp);
      pl.x += t.x;
      pl.y += t.y;
      t = cmul(vq, xq);
      pp.x += t.x;
      pp.y += t.y;
      if (l4 == 0 && i3 < nf) {  // LChi: the l = 1 block only
        if (nu >= 1 && nu <= nf) {
          const size_t li = (size_t)i3 + (size_t)nf * (nu - 1) + (size_t)nf * nf * (w - 1);
          const double2 ld = a.lam_d[li], lm = a.lam_m[li];
          double cr = (ld.x - lm.x) * a.inv_ff11_sq, ci = (ld.y - lm.y) * a.inv_ff11_sq;
          if (a.use_u2) cr -= 2.0 * a.u * a.inv_ff11_sq;
          pl.x += cr * xp.x - ci * xp.y;
          pl.y += cr * xp.y + ci * xp.x;
        } else if (!a.use_u2) {
          pl.x += 2.0 * a.u * a.inv_ff11_sq * xp.x;
          pl.y += 2.0 * a.u * a.inv_ff11_sq * xp.y;
        }
      }
    }
  a.pl[(size_t)q * nz + i3] = pl;
  a.pp[(size_t)q * nz + i3] = pp;
}

// grid: nibz * Nf; one workgroup per IBZ (k, nu), threads over (q slot, star operation)
__global__ __launch_bounds__(256) void sde_sigma(SdeArgs a) {
  const GreenParams g = green_resolve(a.g);
  __shared__ double red[2][256];
  const int nf = a.nf, nx = a.nx, nl = a.nl, nc = nx * nx;
  const int km = blockIdx.x / nf, k0 = blockIdx.x - km * nf + 1;
  const int kx = a.ibz[km], ky = a.ibz[a.nibz + km];
  const int mkx = (-(kx - 1) + nx) % nx + 1, mky = (-(ky - 1) + nx) % nx + 1, mkw = -k0 + nf + 1;
  const int kidx = (ky - 1) * nx + (kx - 1);
  double ar = 0.0, ai = 0.0;
  for (int item = threadIdx.x; item < a.nq * 8; item += blockDim.x) {  // (q slot, operation of q's star)
    const int q = item >> 3, s = (item & 7) + 1;
    const int m = a.qm[q];
    const int qx = a.ibz[m], qy = a.ibz[a.nibz + m], w = a.qw[q];
    const unsigned ops = star_ops(qx, qy, a.nxi);
    const double2* pl = a.pl + (size_t)q * a.nz;
    const double2* pp = a.pp + (size_t)q * a.nz;
    if (!((ops >> (s - 1)) & 1u)) continue;
    int sx, sy;
    sym_apply(s, qx, qy, nx, sx, sy);
    const double2 gkq = green_plain(g, (kx - 1 + sx - 1) % nx + 1, (ky - 1 + sy - 1) % nx + 1, k0 + w - 1);
    const int qkx = (mkx - 1 + sx - 1) % nx + 1, qky = (mky - 1 + sy - 1) % nx + 1, qkw = mkw + w - 1;
    const double2 gqmk = green_plain(g, qkx, qky, qkw);
    const double2 gkmq =
        green_plain(g, (-(qkx - 1) + nx) % nx + 1, (-(qky - 1) + nx) % nx + 1, -qkw + nf + 1);
    for (int l3 = 0; l3 < nl; ++l3) {
      const double fi = a.ffi[((size_t)(s - 1) * nl + l3) * nc + kidx];
      const int i3 = l3 * nf + k0 - 1;
      const double2 x = pl[i3], y = pp[i3];
      double2 t = cmul(gkq, x);
      ar += fi * t.x;
      ai += fi * t.y;
      t = cmul(gqmk, y);
      ar += fi * t.x;
      ai += fi * t.y;
      if (w == 1) continue;
      const int i3m = l3 * nf + (-k0 + nf + 1) - 1;
      const double sg = (double)a.signs[l3];
      const double2 xm = pl[i3m], ym = pp[i3m];
      t = cmul(gkmq, make_double2(xm.x, -xm.y));
      ar += fi * sg * t.x;
      ai += fi * sg * t.y;
      t = cmul(make_double2(gkq.x, -gkq.y), make_double2(ym.x, -ym.y));
      ar += fi * sg * t.x;
      ai += fi * sg * t.y;
    }
  }
  red[0][threadIdx.x] = ar;
  red[1][threadIdx.x] = ai;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x < 8) {  // one thread per operation of k's star
    const int s = threadIdx.x + 1;
    if ((star_ops(kx, ky, a.nxi) >> (s - 1)) & 1u) {
      int sx, sy;
      sym_apply(s, kx, ky, nx, sx, sy);
      a.sigma_out[((size_t)(sx - 1) * nx + (sy - 1)) * nf + (k0 - 1)] =
          make_double2(a.pref * red[0][0], a.pref * red[1][0]);
    }
  }
}

// workspace of the SDE for f_range (allocated outside any stream capture)
int sde_reserve(tups_ctx* c, int f_range) {
  const tups_sizes& s = c->sz;
  TUPS_CHECK(f_range >= 0 && f_range <= 8, TUPS_EINVAL, "f_range must be in 0..8");
  const int nc = s.nx * s.ny;
  const size_t ext = (2 * (size_t)f_range + 1) * s.nf;
  const size_t naux = (size_t)s.nq * s.nl * ext, nv = (size_t)s.nq * s.nz, nt = (size_t)nc * s.nf;
  const size_t need = (2 * naux + 2 * nv + nt) * sizeof(double2);
  if (c->sde_ws_bytes < need) {
    drop_graphs(c);  // captured graphs hold the old workspace
    if (c->sde_ws) TUPS_HIP(hipFree(c->sde_ws));
    c->sde_ws = nullptr;
    c->sde_ws_bytes = 0;
    TUPS_HIP(hipMalloc(&c->sde_ws, need));
    c->sde_ws_bytes = need;
  }
  return TUPS_OK;
}

// Sigma output of the last SDE launch with this f_range (inside the workspace)
double2* sde_output(tups_ctx* c, int f_range) {
  const tups_sizes& s = c->sz;
  const size_t ext = (2 * (size_t)f_range + 1) * s.nf;
  const size_t naux = (size_t)s.nq * s.nl * ext, nv = (size_t)s.nq * s.nz;
  return static_cast<double2*>(c->sde_ws) + 2 * naux + 2 * nv;
}

// the SDE from the Sigma on the device (g); the workspace must be reserved
int launch_sde_device(tups_ctx* c, const GreenParams& g, int f_range, double u, int use_u2) {
  const tups_sizes& s = c->sz;
  SdeArgs a{};
  a.g = g;
  const int nc = s.nx * s.ny;
  a.ext = (2 * f_range + 1) * s.nf;
  const size_t naux = (size_t)s.nq * s.nl * a.ext, nv = (size_t)s.nq * s.nz;
  double2* ws = static_cast<double2*>(c->sde_ws);
  a.caux_ph = ws;
  a.caux_pp = ws + naux;
  a.pl = ws + 2 * naux;
  a.pp = a.pl + nv;
  a.sigma_out = a.pp + nv;
  for (int r = 0; r < 4; ++r) a.gam[r] = c->gamma[r];
  a.lam_d = c->lam[0];
  a.lam_m = c->lam[1];
  a.ff = c->d_ffk;
  a.ffi = c->d_ffik;
  a.signs = c->d_signs;
  a.qm = c->d_qm;
  a.qw = c->d_qw;
  a.ibz = c->d_ibz;
  a.nx = s.nx;
  a.nf = s.nf;
  a.nl = s.nl;
  a.nz = s.nz;
  a.nq = s.nq;
  a.nibz = s.nibz;
  a.nxi = (s.nx + 2) / 2;
  a.f_range = f_range;
  a.use_u2 = use_u2 ? 1 : 0;
  a.u = u;
  const double f11 = c->tab.ff[0];
  a.inv_ff11_sq = 1.0 / (f11 * f11);
  a.pref = -u / c->beta / c->beta / nc / nc / 2.0;
  if (int rc = timer_begin(c)) return rc;
  hipLaunchKernelGGL(sde_chi_aux, dim3((unsigned)s.nq), dim3(256), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  hipLaunchKernelGGL(sde_vchi, dim3((unsigned)((s.nz + 255) / 256), (unsigned)s.nq), dim3(256), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  hipLaunchKernelGGL(sde_sigma, dim3((unsigned)(s.nibz * s.nf)), dim3(256), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  if (int rc = timer_end(c, TUPS_K_SDE)) return rc;
  return TUPS_OK;
}

// SigmaOld <- Sigma (SRC/loop.f90:152), Sigma <- the SDE output, and with the default
// rule Sigma_H(k) = Sigma(k, Nf) (SRC/parquet_chi.f90:103-110): the device state for the
// convergence check and the next iteration
__global__ __launch_bounds__(256) void sigma_commit(const double2* __restrict__ src, double2* __restrict__ sigma,
                                                    double2* __restrict__ sigma_old, double2* __restrict__ sigh,
                                                    int nx, int nf, int sigh_default) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nt = (size_t)nx * nx * nf;
  if (i >= nt) return;
  const double2 v = src[i];
  sigma_old[i] = sigma[i];
  sigma[i] = v;
  if (sigh_default && (int)(i % nf) == nf - 1) {
    const size_t k = i / nf;  // (kx - 1) Ny + ky - 1
    const size_t kx = k / nx, ky = k % nx;
    sigh[kx + (size_t)nx * ky] = v;
  }
}

int launch_sigma_commit(tups_ctx* c, int f_range) {
  const tups_sizes& s = c->sz;
  const size_t nt = (size_t)s.nx * s.ny * s.nf;
  hipLaunchKernelGGL(sigma_commit, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream,
                     sde_output(c, f_range), c->d_sigma, c->d_sigma_old, c->d_sigh, s.nx, s.nf,
                     c->sigh_default ? 1 : 0);
  TUPS_HIP(hipGetLastError());
  return TUPS_OK;
}

int launch_sde(tups_ctx* c, const double* sigma, const double* sigma_h, double mu, int ngrain, int f_range, double u,
               int use_u2, double* sigma_out) {
  const tups_sizes& s = c->sz;
  TUPS_CHECK(sigma_out, TUPS_EINVAL, "null output");
  if (int rc = sde_reserve(c, f_range)) return rc;
  GreenParams g{};
  if (int rc = upload_sigma(c, sigma, sigma_h, &g, mu, ngrain)) return rc;
  if (int rc = launch_sde_device(c, g, f_range, u, use_u2)) return rc;
  if (use_u2) {
    TUPS_CHECK(c->grt_valid && c->gt_ngrain == ngrain, TUPS_ESTATE,
               "use_u2 needs Grt on the device for this ngrain (tups_gkw_grt first)");
    if (int rc = launch_order_2nd(c, u, sde_output(c, f_range))) return rc;
  }
  const size_t nt = (size_t)s.nx * s.ny * s.nf;
  TUPS_HIP(hipMemcpyAsync(sigma_out, sde_output(c, f_range), nt * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  return TUPS_OK;
}

}  // namespace tups
