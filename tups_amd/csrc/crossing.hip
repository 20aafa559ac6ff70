// Parquet equation in the truncated-unity basis: F_r = Phi_r + Lambda_r*Nc +
// crossed reducible vertices of the other channels.
//
// Reference: calculate_PAE, SRC/parquet_EQOuter.f90:16-118, with the eight
// contribution routines of SRC/parquet_EQContribs.f90:44-524.  The reference
// loops source channel x w2 x routine and scatters += into F along diagonals
// (nu2 a function of nu1 for each w2).  Inverting those index maps shows that
// each target entry F_t((l1,nu1),(l2,nu2),q) receives at most one term per
// routine pair (plain routine for w2 >= 1, time-reversed `_TR` for w2 >= 2):
//
//  targets d,m (ph):  phbar->ph  w2 = nu2-nu1+1         src (nu1, clamp(nu1+qw-1))
//                      _TR       w2 = nu1-nu2+1 (>=2)   conj, src (mF(nu1), clamp(mF(nu1+qw-1)))
//                     pp->ph     w2 = nu1+nu2+qw-Nf-1   src (nu1, nu2)
//                      _TR       w2 = Nf+3-qw-nu1-nu2   conj, src (mF(nu1), nu1+qw+w2-2)
//  targets s,t (pp):  ph->pp     w2 = qw-nu1-nu2+Nf+1   src (nu1, nu2)
//                      _TR       w2 = nu1+nu2-qw-Nf+1   conj, src (mF(nu1), nu1-qw-w2+2)
//                     phbar->pp  w2 = nu2-nu1+1         src (nu1, clamp(qw-nu1-w2+Nf+1))
//                      _TR       w2 = nu1-nu2+1 (>=2)   conj, src (mF(nu1), clamp(nu1-qw-w2+2))
// (1-based nu, mF(nu) = Nf+1-nu, clamp = pickNu to [1, Nf]; every routine keeps
// entries with 1 <= w2 <= Nf/2.)  Each term is
//   sum_{R2,l3,l4} M(l1,l2,l3,l4,Q,R2) * Xsrc((l3,nu_a),(l4,nu_b),w2,R2) / Nx
// where Xsrc is the linear combination of the source channels' PhiR given by
// the factor table of SRC/parquet_EQOuter.f90:85-112, and _TR terms carry
// conj() and signs(l3)*signs(l4).  One thread owns one (q, row, col) entry of a
// target pair, gathers its terms, adds Lambda*Nc on the l=1 block
// (SRC/parquet_EQOuter.f90:48-59) and writes F once: F is read (as Phi) and
// written exactly once, the minimum traffic of the crossing.
#include "internal.h"

namespace tups {

struct PaeArgs {
  const double2* phi[4];
  double2* F[4];
  const double2* phir[4];
  const double2* M[4];     // phbar_to_ph, pp_to_ph, ph_to_pp, phbar_to_pp
  const double2* lam[4];
  const int* signs;
  int nf, nw, nibz, nr;
  double inv_nx;
  double nc_aux;           // |1/FF(1,1,1)^2|
};

typedef double dv2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int clampnu(int x, int nf) { return x < 1 ? 1 : (x > nf ? nf : x); }

__device__ __forceinline__ double2 ld_stream(const double2* p) {
  const dv2 v = __builtin_nontemporal_load(reinterpret_cast<const dv2*>(p));
  return make_double2(v.x, v.y);
}

__device__ __forceinline__ void st_stream(double2* p, double2 v) {
  dv2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<dv2*>(p));
}

// acc_a += sum M * (ca0*X0 + ca1*X1); acc_b += sum M * (cb0*X0 + cb1*X1)
template <int NL>
__device__ __forceinline__ void crossing_term(const PaeArgs& a, const double2* M, int s0, int s1,
                                              double ca0, double ca1, double cb0, double cb1,
                                              int l1, int l2, int m, int w2, int nua, int nub,
                                              bool tr, double2& acc_a, double2& acc_b) {
  const int nf = a.nf, nz = nf * NL;
  const size_t nz2 = (size_t)nz * nz;
  const double2* p0 = a.phir[s0];
  const double2* p1 = a.phir[s1];
  constexpr int NL4 = NL * NL * NL * NL;
  for (int R = 0; R < a.nr; ++R) {
    const size_t pbase = nz2 * ((size_t)(w2 - 1) + (size_t)a.nw * R);
    const double2* mb = M + (size_t)NL4 * ((size_t)m + (size_t)a.nibz * R) + l1 + NL * l2;
#pragma unroll
    for (int l4 = 0; l4 < NL; ++l4) {
      const size_t colo = pbase + (size_t)nz * (l4 * nf + nub - 1) + (nua - 1);
#pragma unroll
      for (int l3 = 0; l3 < NL; ++l3) {
        const size_t pi = colo + (size_t)(l3 * nf);
        double2 x0 = p0[pi];
        double2 x1 = p1[pi];
        if (tr) {
          const double sg = (double)(a.signs[l3] * a.signs[l4]);
          x0 = make_double2(sg * x0.x, -sg * x0.y);
          x1 = make_double2(sg * x1.x, -sg * x1.y);
        }
        const double2 mv = mb[NL * NL * (l3 + NL * l4)];
        const double2 ya = make_double2(ca0 * x0.x + ca1 * x1.x, ca0 * x0.y + ca1 * x1.y);
        const double2 yb = make_double2(cb0 * x0.x + cb1 * x1.x, cb0 * x0.y + cb1 * x1.y);
        acc_a.x += mv.x * ya.x - mv.y * ya.y;
        acc_a.y += mv.x * ya.y + mv.y * ya.x;
        acc_b.x += mv.x * yb.x - mv.y * yb.y;
        acc_b.y += mv.x * yb.y + mv.y * yb.x;
      }
    }
  }
}

// grid: x -> 64 rows, y -> 4 columns, z -> (target pair, q) with the bosonic
// frequency of q slowest so that concurrently running blocks gather from the
// same PhiR(., ., w2) subset (which depends on qw only) and keep it in L2.
template <int NL>
__global__ __launch_bounds__(256) void pae_kernel(PaeArgs a) {
  const int nf = a.nf, nz = nf * NL, nw = a.nw;
  const size_t nz2 = (size_t)nz * nz;
  const int row = blockIdx.x * 64 + (threadIdx.x & 63);
  const int col = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (row >= nz || col >= nz) return;
  const int pair = blockIdx.z & 1;  // 0: targets d,m ; 1: targets s,t
  const int jq = blockIdx.z >> 1;
  const int w0 = jq / a.nibz;
  const int m = jq - w0 * a.nibz;
  const int q = m * nw + w0;        // Index_bosonic_IBZ order: momentum-major
  const int qw = w0 + 1;
  const int l1 = row / nf, nu1 = row - l1 * nf + 1;
  const int l2 = col / nf, nu2 = col - l2 * nf + 1;
  const int ta = pair == 0 ? 0 : 2, tb = ta + 1;
  const size_t idx = (size_t)row + (size_t)nz * col + nz2 * q;

  // streaming operands first: Phi (read once) and the Lambda block
  double2 fa = ld_stream(a.phi[ta] + idx);
  double2 fb = ld_stream(a.phi[tb] + idx);
  double2 la = make_double2(0.0, 0.0), lb = make_double2(0.0, 0.0);
  if (l1 == 0 && l2 == 0) {
    const size_t li = (size_t)(nu1 - 1) + (size_t)nf * ((size_t)(nu2 - 1) + (size_t)nf * (qw - 1));
    la = a.lam[ta][li];
    lb = a.lam[tb][li];
  }

  double2 acc_a = make_double2(0.0, 0.0), acc_b = make_double2(0.0, 0.0);
  if (pair == 0) {
    // phbar -> ph from d (-1/2, -1/2) and m (-3/2, +1/2)
    {
      const int dn = nu2 - nu1;
      if (dn >= 0 && dn + 1 <= nw) {
        const int w2 = dn + 1;
        crossing_term<NL>(a, a.M[0], 0, 1, -0.5, -1.5, -0.5, 0.5, l1, l2, m, w2, nu1,
                          clampnu(nu1 + qw - 1, nf), false, acc_a, acc_b);
      } else if (dn < 0 && 1 - dn <= nw) {
        const int w2 = 1 - dn;
        crossing_term<NL>(a, a.M[0], 0, 1, -0.5, -1.5, -0.5, 0.5, l1, l2, m, w2, nf + 1 - nu1,
                          clampnu(nf + 1 - (nu1 + qw - 1), nf), true, acc_a, acc_b);
      }
    }
    // pp -> ph from s (+1/2, -1/2) and t (+3/2, +1/2)
    {
      const int w2 = nu1 + nu2 + qw - nf - 1;
      const int w2t = nf + 3 - qw - nu1 - nu2;
      if (w2 >= 1 && w2 <= nw) {
        crossing_term<NL>(a, a.M[1], 2, 3, 0.5, 1.5, -0.5, 0.5, l1, l2, m, w2, nu1, nu2, false,
                          acc_a, acc_b);
      } else if (w2t >= 2 && w2t <= nw) {
        crossing_term<NL>(a, a.M[1], 2, 3, 0.5, 1.5, -0.5, 0.5, l1, l2, m, w2t, nf + 1 - nu1,
                          nu1 + qw + w2t - 2, true, acc_a, acc_b);
      }
    }
  } else {
    // ph -> pp from d (+1/2, +1/2) and m (-3/2, +1/2)
    {
      const int w2 = qw - nu1 - nu2 + nf + 1;
      const int w2t = nu1 + nu2 - qw - nf + 1;
      if (w2 >= 1 && w2 <= nw) {
        crossing_term<NL>(a, a.M[2], 0, 1, 0.5, -1.5, 0.5, 0.5, l1, l2, m, w2, nu1, nu2, false,
                          acc_a, acc_b);
      } else if (w2t >= 2 && w2t <= nw) {
        crossing_term<NL>(a, a.M[2], 0, 1, 0.5, -1.5, 0.5, 0.5, l1, l2, m, w2t, nf + 1 - nu1,
                          nu1 - qw - w2t + 2, true, acc_a, acc_b);
      }
    }
    // phbar -> pp from d (+1/2, -1/2) and m (-3/2, -1/2)
    {
      const int dn = nu2 - nu1;
      if (dn >= 0 && dn + 1 <= nw) {
        const int w2 = dn + 1;
        crossing_term<NL>(a, a.M[3], 0, 1, 0.5, -1.5, -0.5, -0.5, l1, l2, m, w2, nu1,
                          clampnu(qw - nu1 - w2 + nf + 1, nf), false, acc_a, acc_b);
      } else if (dn < 0 && 1 - dn <= nw) {
        const int w2 = 1 - dn;
        crossing_term<NL>(a, a.M[3], 0, 1, 0.5, -1.5, -0.5, -0.5, l1, l2, m, w2, nf + 1 - nu1,
                          clampnu(nu1 - qw - w2 + 2, nf), true, acc_a, acc_b);
      }
    }
  }

  fa.x += la.x * a.nc_aux + acc_a.x * a.inv_nx;
  fa.y += la.y * a.nc_aux + acc_a.y * a.inv_nx;
  fb.x += lb.x * a.nc_aux + acc_b.x * a.inv_nx;
  fb.y += lb.y * a.nc_aux + acc_b.y * a.inv_nx;
  st_stream(a.F[ta] + idx, fa);
  st_stream(a.F[tb] + idx, fb);
}

int launch_pae(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  PaeArgs a{};
  for (int r = 0; r < 4; ++r) {
    a.phi[r] = c->gamma[r];
    a.F[r] = c->fvert[r];
    a.phir[r] = c->phir[r];
    a.M[r] = c->mten[r];
    a.lam[r] = c->lam[r];
  }
  a.signs = c->d_signs;
  a.nf = s.nf;
  a.nw = s.nw;
  a.nibz = s.nibz;
  a.nr = s.nr;
  a.inv_nx = 1.0 / s.nx;
  const double f0 = c->tab.ff[0];
  a.nc_aux = std::fabs(1.0 / f0 * 1.0 / f0);
  TUPS_CHECK(2LL * s.nq <= 65535, TUPS_EINVAL, "too many q for the crossing grid");
  dim3 g((unsigned)((s.nz + 63) / 64), (unsigned)((s.nz + 3) / 4), (unsigned)(2 * s.nq));
  if (int rc = timer_begin(c)) return rc;
  if (s.nl == 1)
    hipLaunchKernelGGL(pae_kernel<1>, g, dim3(256), 0, c->stream, a);
  else if (s.nl == 5)
    hipLaunchKernelGGL(pae_kernel<5>, g, dim3(256), 0, c->stream, a);
  else
    hipLaunchKernelGGL(pae_kernel<9>, g, dim3(256), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  if (int rc = timer_end(c, TUPS_K_PAE)) return rc;
  return TUPS_OK;
}

}  // namespace tups
