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
// conj() and signs(l3)*signs(l4).  Both paths below first gather the source
// PhiR of every routine along its frequency map per bosonic frequency (the
// term depends on q only through qw), then stream Phi -> F adding the terms
// and Lambda*Nc on the l=1 block (SRC/parquet_EQOuter.f90:48-59): F is read
// (as Phi) and written exactly once, the minimum traffic of the crossing.
#include "internal.h"

namespace tups {


typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int clampnu(int x, int nf) { return x < 1 ? 1 : (x > nf ? nf : x); }


// One crossing term, branch-free: `wt` is 0 for an entry the term does not
// reach (its indices are then any in-range value), `tr` selects the
// time-reversed form conj(x) * signs(l3) signs(l4).
struct Term {
  int w2, nua, nub;
  double wt;
  bool tr;
};

__device__ __forceinline__ Term make_term(bool plain_ok, bool tr_ok, int w2p, int nap, int nbp, int w2t,
                                          int nat, int nbt) {
  Term t;
  t.tr = !plain_ok && tr_ok;
  t.wt = (plain_ok || tr_ok) ? 1.0 : 0.0;
  t.w2 = plain_ok ? w2p : (tr_ok ? w2t : 1);
  t.nua = plain_ok ? nap : (tr_ok ? nat : 1);
  t.nub = plain_ok ? nbp : (tr_ok ? nbt : 1);
  return t;
}

__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  // workgroups are dispatched round-robin over the 8 XCDs: give each XCD a
  // contiguous logical range (here: a contiguous range of bosonic frequencies)
  const int xcd = bid & 7, idx = bid >> 3;
  const int per = nblk >> 3, rem = nblk & 7;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// ---- Nl = 1: gather pre-pass + streaming crossing ---------------------------------
//
// With one form factor each routine's term at target (nu1, nu2, q) depends on
// q only through qw (the M tensor is a scalar per (Q, R)).  The pre-pass
// re-indexes the source PhiR of every routine into
//   Y[k][R][qw][nu2][nu1],  k = 2*routine + target-of-pair,
// already combined over the source channels with the factor table, with the
// time-reversed form (conj) selected and invalid entries zero.  The crossing
// then reads Y contiguously: F = Phi + Lambda Nc + (1/Nx) sum_R M(Q,R) (Y_a + Y_b).

// grid: ((per + 255) / 256, 8): blockIdx.y = k, one output entry per thread
// entry (nu1, nu2, qw, R) of Y slice k (1-based frequencies; pd..pt the PhiR of d, m, s, t)
__device__ __forceinline__ double2 y_entry(int k, int nu1, int nu2, int qw, int R, int nf, int nw,
                                           const double2* __restrict__ pd, const double2* __restrict__ pm,
                                           const double2* __restrict__ ps, const double2* __restrict__ pt) {
  const size_t nf2 = (size_t)nf * nf;
  const size_t rb = nf2 * (size_t)nw * R;  // PhiR(nu_a, nu_b, w2, R) = rb + (w2-1)*nf2 + (nub-1)*nf + nua-1
  const int dn = nu2 - nu1;
  const bool pb_plain = dn >= 0 && dn + 1 <= nw;
  const bool pb_tr = dn < 0 && 1 - dn <= nw;
  Term t;
  if (k < 2) {  // phbar->ph
    t = make_term(pb_plain, pb_tr, dn + 1, nu1, clampnu(nu1 + qw - 1, nf), 1 - dn, nf + 1 - nu1,
                  clampnu(nf + 1 - (nu1 + qw - 1), nf));
  } else if (k < 4) {  // pp->ph
    const int w2 = nu1 + nu2 + qw - nf - 1, w2t = nf + 3 - qw - nu1 - nu2;
    t = make_term(w2 >= 1 && w2 <= nw, w2t >= 2 && w2t <= nw, w2, nu1, nu2, w2t, nf + 1 - nu1, nu1 + qw + w2t - 2);
  } else if (k < 6) {  // ph->pp
    const int w2 = qw - nu1 - nu2 + nf + 1, w2t = nu1 + nu2 - qw - nf + 1;
    t = make_term(w2 >= 1 && w2 <= nw, w2t >= 2 && w2t <= nw, w2, nu1, nu2, w2t, nf + 1 - nu1, nu1 - qw - w2t + 2);
  } else {  // phbar->pp
    t = make_term(pb_plain, pb_tr, dn + 1, nu1, clampnu(qw - nu1 - dn + nf, nf), 1 - dn, nf + 1 - nu1,
                  clampnu(nu1 - qw - (1 - dn) + 2, nf));
  }
  // source channels and factors per k (SRC/parquet_EQOuter.f90:85-112)
  const double2* p0 = k == 2 || k == 3 ? ps : pd;
  const double2* p1 = k == 2 || k == 3 ? pt : pm;
  constexpr double c0s[8] = {-0.5, -0.5, 0.5, -0.5, 0.5, 0.5, 0.5, -0.5};
  constexpr double c1s[8] = {-1.5, 0.5, 1.5, 0.5, -1.5, 0.5, -1.5, -0.5};
  const double c0 = c0s[k], c1 = c1s[k];
  const size_t i = rb + (size_t)(t.w2 - 1) * nf2 + (size_t)(t.nub - 1) * nf + (t.nua - 1);
  const double2 x0 = p0[i], x1 = p1[i];
  const double im = t.tr ? -1.0 : 1.0;
  return make_double2(t.wt * (c0 * x0.x + c1 * x1.x), t.wt * im * (c0 * x0.y + c1 * x1.y));
}

__global__ __launch_bounds__(256) void pae_pregather_nl1(const double2* __restrict__ pd,
                                                         const double2* __restrict__ pm,
                                                         const double2* __restrict__ ps,
                                                         const double2* __restrict__ pt,
                                                         double2* __restrict__ Y, int nf, int nw,
                                                         int nr) {
  const size_t nf2 = (size_t)nf * nf;
  const size_t per = nf2 * nw * nr;  // one k slice
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= per) return;
  const int k = blockIdx.y;  // 2 * routine + target of the pair
  const int nu1 = (int)(gid % nf) + 1;
  const int nu2 = (int)((gid / nf) % nf) + 1;
  const int qw = (int)((gid / nf2) % nw) + 1;
  const int R = (int)(gid / (nf2 * nw));
  Y[(size_t)k * per + gid] = y_entry(k, nu1, nu2, qw, R, nf, nw, pd, pm, ps, pt);
}

struct PaeNl1Args {
  const double2* phir[4];  // PhiR of d, m, s, t (the fused variants re-index them in place)
  const double2* phi[4];
  double2* F[4];
  const double2* lam[4];
  const double2* M[4];
  const double2* Y;
  int nf, nw, nibz, nr, nq;
  double inv_nx, nc_aux;
};

// one thread per (nu1, nu2) entry of one (q, target pair), C columns per
// thread; with NR1 (one real-space shell) every load of all C columns is
// issued before the first use.
template <int C, bool NR1>
__global__ __launch_bounds__(256) void pae_stream_nl1(PaeNl1Args a) {
  const int nf = a.nf, nw = a.nw;
  const size_t nf2 = (size_t)nf * nf;
  const int tx = (nf + 63) >> 6, ty = (nf + 4 * C - 1) / (4 * C);
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = logical % (tx * ty);
  const int zq = logical / (tx * ty);
  const int nu1 = (tile % tx) * 64 + (threadIdx.x & 63);
  const int c0 = (tile / tx) * (4 * C) + (threadIdx.x >> 6);
  if (nu1 >= nf) return;
  const int pair = zq & 1;
  const int jq = zq >> 1;
  const int w0 = jq / a.nibz;
  const int m = jq - w0 * a.nibz;
  const int q = m * nw + w0;
  const int ta = pair == 0 ? 0 : 2, tb = ta + 1;
  const int ka = pair == 0 ? 0 : 4;  // routines (0,1) feed d,m ; (2,3) feed s,t
  const size_t per = nf2 * nw * a.nr;
  const double2* Ma = a.M[pair == 0 ? 0 : 2];
  const double2* Mb = a.M[pair == 0 ? 1 : 3];
  double2 fa[C], fb[C], la[C], lb[C];
  double2 y0[C], y1[C], y2[C], y3[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int nu2 = c0 + 4 * c < nf ? c0 + 4 * c : 0;
    const size_t idx = (size_t)nu1 + (size_t)nf * nu2 + nf2 * q;
    fa[c] = ld_stream(a.phi[ta] + idx);
    fb[c] = ld_stream(a.phi[tb] + idx);
    const size_t li = (size_t)nu1 + (size_t)nf * ((size_t)nu2 + (size_t)nf * w0);
    la[c] = a.lam[ta][li];
    lb[c] = a.lam[tb][li];
    if (NR1) {
      const size_t o = (size_t)nu1 + (size_t)nf * nu2 + nf2 * w0;
      y0[c] = a.Y[(ka + 0) * per + o];
      y1[c] = a.Y[(ka + 1) * per + o];
      y2[c] = a.Y[(ka + 2) * per + o];
      y3[c] = a.Y[(ka + 3) * per + o];
    }
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int nu2 = c0 + 4 * c;
    if (nu2 >= nf) continue;
    double2 sa = make_double2(0.0, 0.0), sb = make_double2(0.0, 0.0);
    for (int R = 0; R < (NR1 ? 1 : a.nr); ++R) {
      double2 ya0, yb0, ya1, yb1;
      if (NR1) {
        ya0 = y0[c];
        yb0 = y1[c];
        ya1 = y2[c];
        yb1 = y3[c];
      } else {
        const size_t o = (size_t)nu1 + (size_t)nf * nu2 + nf2 * ((size_t)w0 + (size_t)nw * R);
        ya0 = a.Y[(ka + 0) * per + o];
        yb0 = a.Y[(ka + 1) * per + o];
        ya1 = a.Y[(ka + 2) * per + o];
        yb1 = a.Y[(ka + 3) * per + o];
      }
      const double2 m0 = Ma[(size_t)m + (size_t)a.nibz * R], m1 = Mb[(size_t)m + (size_t)a.nibz * R];
      sa.x += m0.x * ya0.x - m0.y * ya0.y + m1.x * ya1.x - m1.y * ya1.y;
      sa.y += m0.x * ya0.y + m0.y * ya0.x + m1.x * ya1.y + m1.y * ya1.x;
      sb.x += m0.x * yb0.x - m0.y * yb0.y + m1.x * yb1.x - m1.y * yb1.y;
      sb.y += m0.x * yb0.y + m0.y * yb0.x + m1.x * yb1.y + m1.y * yb1.x;
    }
    double2 oa = fa[c], ob = fb[c];
    oa.x += la[c].x * a.nc_aux + sa.x * a.inv_nx;
    oa.y += la[c].y * a.nc_aux + sa.y * a.inv_nx;
    ob.x += lb[c].x * a.nc_aux + sb.x * a.inv_nx;
    ob.y += lb[c].y * a.nc_aux + sb.y * a.inv_nx;
    const size_t idx = (size_t)nu1 + (size_t)nf * nu2 + nf2 * q;
    st_stream(a.F[ta] + idx, oa);
    st_stream(a.F[tb] + idx, ob);
  }
}

// NR = 1: one thread per (nu1, nu2) of one target pair and KM momenta that
// share the bosonic frequency qw: Y and Lambda depend on q only through qw,
// so they are loaded once and reused for the KM streaming Phi -> F updates.
template <int KM>
__global__ __launch_bounds__(256) void pae_stream_nl1_mom(PaeNl1Args a) {
  const int nf = a.nf, nw = a.nw;
  const size_t nf2 = (size_t)nf * nf;
  const int tx = (nf + 63) >> 6, ty = (nf + 3) >> 2;
  const int ngm = (a.nibz + KM - 1) / KM;
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = logical % (tx * ty);
  int rest = logical / (tx * ty);
  const int mg = rest % ngm;
  rest /= ngm;
  const int pair = rest & 1;
  const int w0 = rest >> 1;
  const int nu1 = (tile % tx) * 64 + (threadIdx.x & 63);
  const int nu2 = (tile / tx) * 4 + (threadIdx.x >> 6);
  if (nu1 >= nf || nu2 >= nf) return;
  const int ta = pair == 0 ? 0 : 2, tb = ta + 1;
  const int ka = pair == 0 ? 0 : 4;
  const size_t per = nf2 * nw;  // NR = 1
  const size_t o = (size_t)nu1 + (size_t)nf * nu2 + nf2 * w0;
  const double2 ya0 = a.Y[(ka + 0) * per + o], yb0 = a.Y[(ka + 1) * per + o];
  const double2 ya1 = a.Y[(ka + 2) * per + o], yb1 = a.Y[(ka + 3) * per + o];
  const size_t li = (size_t)nu1 + (size_t)nf * ((size_t)nu2 + (size_t)nf * w0);
  const double2 la = a.lam[ta][li], lb = a.lam[tb][li];
  const double2* Ma = a.M[pair == 0 ? 0 : 2];
  const double2* Mb = a.M[pair == 0 ? 1 : 3];
  double2 fa[KM], fb[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const int m = mg * KM + k < a.nibz ? mg * KM + k : a.nibz - 1;
    const size_t idx = (size_t)nu1 + (size_t)nf * nu2 + nf2 * ((size_t)m * nw + w0);
    fa[k] = ld_stream(a.phi[ta] + idx);
    fb[k] = ld_stream(a.phi[tb] + idx);
  }
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const int m = mg * KM + k;
    if (m >= a.nibz) break;
    const double2 m0 = Ma[m], m1 = Mb[m];
    double2 oa = fa[k], ob = fb[k];
    oa.x += la.x * a.nc_aux + (m0.x * ya0.x - m0.y * ya0.y + m1.x * ya1.x - m1.y * ya1.y) * a.inv_nx;
    oa.y += la.y * a.nc_aux + (m0.x * ya0.y + m0.y * ya0.x + m1.x * ya1.y + m1.y * ya1.x) * a.inv_nx;
    ob.x += lb.x * a.nc_aux + (m0.x * yb0.x - m0.y * yb0.y + m1.x * yb1.x - m1.y * yb1.y) * a.inv_nx;
    ob.y += lb.y * a.nc_aux + (m0.x * yb0.y + m0.y * yb0.x + m1.x * yb1.y + m1.y * yb1.x) * a.inv_nx;
    const size_t idx = (size_t)nu1 + (size_t)nf * nu2 + nf2 * ((size_t)m * nw + w0);
    st_stream(a.F[ta] + idx, oa);
    st_stream(a.F[tb] + idx, ob);
  }
}

// pae_stream_nl1_mom with the re-index pass folded in: each thread forms its four
// Y entries from PhiR itself (y_entry) instead of reading the pre-pass output;
// the entries of one (nu1, nu2, qw) are recomputed by every momentum group,
// which reads PhiR through the cache in place of the pre-pass's write and read
template <int KM>
__global__ __launch_bounds__(256) void pae_stream_nl1_fused(PaeNl1Args a) {
  const int nf = a.nf, nw = a.nw;
  const size_t nf2 = (size_t)nf * nf;
  const int tx = (nf + 63) >> 6, ty = (nf + 3) >> 2;
  const int ngm = (a.nibz + KM - 1) / KM;
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = logical % (tx * ty);
  int rest = logical / (tx * ty);
  const int mg = rest % ngm;
  rest /= ngm;
  const int pair = rest & 1;
  const int w0 = rest >> 1;
  const int nu1 = (tile % tx) * 64 + (threadIdx.x & 63);
  const int nu2 = (tile / tx) * 4 + (threadIdx.x >> 6);
  if (nu1 >= nf || nu2 >= nf) return;
  const int ta = pair == 0 ? 0 : 2, tb = ta + 1;
  const int ka = pair == 0 ? 0 : 4;
  const double2 *pd = a.phir[0], *pm = a.phir[1], *ps = a.phir[2], *pt = a.phir[3];
  const double2 ya0 = y_entry(ka + 0, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
  const double2 yb0 = y_entry(ka + 1, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
  const double2 ya1 = y_entry(ka + 2, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
  const double2 yb1 = y_entry(ka + 3, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
  const size_t li = (size_t)nu1 + (size_t)nf * ((size_t)nu2 + (size_t)nf * w0);
  const double2 la = a.lam[ta][li], lb = a.lam[tb][li];
  const double2* Ma = a.M[pair == 0 ? 0 : 2];
  const double2* Mb = a.M[pair == 0 ? 1 : 3];
  double2 fa[KM], fb[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const int m = mg * KM + k < a.nibz ? mg * KM + k : a.nibz - 1;
    const size_t idx = (size_t)nu1 + (size_t)nf * nu2 + nf2 * ((size_t)m * nw + w0);
    fa[k] = ld_stream(a.phi[ta] + idx);
    fb[k] = ld_stream(a.phi[tb] + idx);
  }
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const int m = mg * KM + k;
    if (m >= a.nibz) break;
    const double2 m0 = Ma[m], m1 = Mb[m];
    double2 oa = fa[k], ob = fb[k];
    oa.x += la.x * a.nc_aux + (m0.x * ya0.x - m0.y * ya0.y + m1.x * ya1.x - m1.y * ya1.y) * a.inv_nx;
    oa.y += la.y * a.nc_aux + (m0.x * ya0.y + m0.y * ya0.x + m1.x * ya1.y + m1.y * ya1.x) * a.inv_nx;
    ob.x += lb.x * a.nc_aux + (m0.x * yb0.x - m0.y * yb0.y + m1.x * yb1.x - m1.y * yb1.y) * a.inv_nx;
    ob.y += lb.y * a.nc_aux + (m0.x * yb0.y + m0.y * yb0.x + m1.x * yb1.y + m1.y * yb1.x) * a.inv_nx;
    const size_t idx = (size_t)nu1 + (size_t)nf * nu2 + nf2 * ((size_t)m * nw + w0);
    st_stream(a.F[ta] + idx, oa);
    st_stream(a.F[tb] + idx, ob);
  }
}

// pae_stream_nl1_fused with one workgroup per (tile, pair, qw) over every
// momentum: the Y entries are formed once per (nu1, nu2, qw) and the NIBZ
// momenta stream through KC at a time, so few registers per thread carry the
// streams (eight waves per SIMD, at most 64 VGPRs: at Nx = 16, Nf = 64 the
// 2048 workgroups run as one round)
template <int KC, bool PF>
__global__ __launch_bounds__(256, 8) void pae_stream_nl1_all(PaeNl1Args a) {
  const int nf = a.nf, nw = a.nw;
  const size_t nf2 = (size_t)nf * nf;
  const int tx = (nf + 63) >> 6, ty = (nf + 3) >> 2;
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = logical % (tx * ty);
  const int rest = logical / (tx * ty);
  const int pair = rest & 1;
  const int w0 = rest >> 1;
  const int nu1 = (tile % tx) * 64 + (threadIdx.x & 63);
  const int nu2 = (tile / tx) * 4 + (threadIdx.x >> 6);
  if (nu1 >= nf || nu2 >= nf) return;
  const int ta = pair == 0 ? 0 : 2, tb = ta + 1;
  const int ka = pair == 0 ? 0 : 4;
  const size_t base = (size_t)nu1 + (size_t)nf * nu2 + nf2 * w0, mstride = nf2 * nw;
  const double2 *phia = a.phi[ta] + base, *phib = a.phi[tb] + base;
  const int nm = a.nibz;
  // PF: the next chunk's loads are issued before the current chunk is
  // combined and stored (the first chunk's before the Y entries are formed)
  double2 fa[KC], fb[KC];
  if (PF) {
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const int m = min(k, nm - 1);
      fa[k] = ld_stream(phia + mstride * m);
      fb[k] = ld_stream(phib + mstride * m);
    }
  }
  const double2 *pd = a.phir[0], *pm = a.phir[1], *ps = a.phir[2], *pt = a.phir[3];
  const double2 ya0 = y_entry(ka + 0, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
  const double2 yb0 = y_entry(ka + 1, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
  const double2 ya1 = y_entry(ka + 2, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
  const double2 yb1 = y_entry(ka + 3, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
  const size_t li = (size_t)nu1 + (size_t)nf * ((size_t)nu2 + (size_t)nf * w0);
  const double2 la = a.lam[ta][li], lb = a.lam[tb][li];
  const double2* Ma = a.M[pair == 0 ? 0 : 2];
  const double2* Mb = a.M[pair == 0 ? 1 : 3];
  for (int mb = 0; mb < nm; mb += KC) {
    double2 na[KC], nb[KC];
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const int m = min(mb + k + (PF ? KC : 0), nm - 1);
      na[k] = ld_stream(phia + mstride * m);
      nb[k] = ld_stream(phib + mstride * m);
    }
    if (!PF) {
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        fa[k] = na[k];
        fb[k] = nb[k];
      }
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const int m = mb + k;
      if (m >= nm) break;
      const double2 m0 = Ma[m], m1 = Mb[m];
      double2 oa = fa[k], ob = fb[k];
      oa.x += la.x * a.nc_aux + (m0.x * ya0.x - m0.y * ya0.y + m1.x * ya1.x - m1.y * ya1.y) * a.inv_nx;
      oa.y += la.y * a.nc_aux + (m0.x * ya0.y + m0.y * ya0.x + m1.x * ya1.y + m1.y * ya1.x) * a.inv_nx;
      ob.x += lb.x * a.nc_aux + (m0.x * yb0.x - m0.y * yb0.y + m1.x * yb1.x - m1.y * yb1.y) * a.inv_nx;
      ob.y += lb.y * a.nc_aux + (m0.x * yb0.y + m0.y * yb0.x + m1.x * yb1.y + m1.y * yb1.x) * a.inv_nx;
      st_stream(a.F[ta] + base + mstride * m, oa);
      st_stream(a.F[tb] + base + mstride * m, ob);
    }
    if (PF) {
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        fa[k] = na[k];
        fb[k] = nb[k];
      }
    }
  }
}

// pae_stream_nl1_all<1, true> with EPT entries per thread, (nu1, nu2 + 4 e)
// of a 64 x 4 EPT frequency tile: EPT times the loads in flight per thread at
// 8 / EPT waves per SIMD (at most 512 EPT / 8 VGPRs); at Nx = 16, Nf = 64 the
// grid is still one round
template <int EPT>
__global__ __launch_bounds__(256, 8 / EPT) void pae_stream_nl1_multi(PaeNl1Args a) {
  const int nf = a.nf, nw = a.nw;
  const size_t nf2 = (size_t)nf * nf;
  const int tx = (nf + 63) >> 6, ty = (nf + 4 * EPT - 1) / (4 * EPT);
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = logical % (tx * ty);
  const int rest = logical / (tx * ty);
  const int pair = rest & 1;
  const int w0 = rest >> 1;
  const int nu1 = (tile % tx) * 64 + (threadIdx.x & 63);
  const int nu2b = (tile / tx) * (4 * EPT) + (threadIdx.x >> 6);
  if (nu1 >= nf || nu2b >= nf) return;
  const int ta = pair == 0 ? 0 : 2, tb = ta + 1;
  const int ka = pair == 0 ? 0 : 4;
  const size_t mstride = nf2 * nw;
  const int nm = a.nibz;
  bool ok[EPT];
  size_t base[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int nu2 = nu2b + 4 * e;
    ok[e] = nu2 < nf;
    base[e] = (size_t)nu1 + (size_t)nf * (ok[e] ? nu2 : nu2b) + nf2 * w0;
  }
  // the first momentum's loads before the Y entries are formed
  double2 fa[EPT], fb[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    fa[e] = ld_stream(a.phi[ta] + base[e]);
    fb[e] = ld_stream(a.phi[tb] + base[e]);
  }
  const double2 *pd = a.phir[0], *pm = a.phir[1], *ps = a.phir[2], *pt = a.phir[3];
  double2 ya0[EPT], yb0[EPT], ya1[EPT], yb1[EPT], la[EPT], lb[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int nu2 = ok[e] ? nu2b + 4 * e : nu2b;
    ya0[e] = y_entry(ka + 0, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
    yb0[e] = y_entry(ka + 1, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
    ya1[e] = y_entry(ka + 2, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
    yb1[e] = y_entry(ka + 3, nu1 + 1, nu2 + 1, w0 + 1, 0, nf, nw, pd, pm, ps, pt);
    const size_t li = (size_t)nu1 + (size_t)nf * ((size_t)nu2 + (size_t)nf * w0);
    la[e] = a.lam[ta][li];
    lb[e] = a.lam[tb][li];
  }
  const double2* Ma = a.M[pair == 0 ? 0 : 2];
  const double2* Mb = a.M[pair == 0 ? 1 : 3];
  for (int m = 0; m < nm; ++m) {
    const int mn = min(m + 1, nm - 1);
    double2 na[EPT], nb[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      na[e] = ld_stream(a.phi[ta] + base[e] + mstride * mn);
      nb[e] = ld_stream(a.phi[tb] + base[e] + mstride * mn);
    }
    const double2 m0 = Ma[m], m1 = Mb[m];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      if (!ok[e]) continue;
      double2 oa = fa[e], ob = fb[e];
      oa.x += la[e].x * a.nc_aux + (m0.x * ya0[e].x - m0.y * ya0[e].y + m1.x * ya1[e].x - m1.y * ya1[e].y) * a.inv_nx;
      oa.y += la[e].y * a.nc_aux + (m0.x * ya0[e].y + m0.y * ya0[e].x + m1.x * ya1[e].y + m1.y * ya1[e].x) * a.inv_nx;
      ob.x += lb[e].x * a.nc_aux + (m0.x * yb0[e].x - m0.y * yb0[e].y + m1.x * yb1[e].x - m1.y * yb1[e].y) * a.inv_nx;
      ob.y += lb[e].y * a.nc_aux + (m0.x * yb0[e].y + m0.y * yb0[e].x + m1.x * yb1[e].y + m1.y * yb1[e].x) * a.inv_nx;
      st_stream(a.F[ta] + base[e] + mstride * m, oa);
      st_stream(a.F[tb] + base[e] + mstride * m, ob);
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      fa[e] = na[e];
      fb[e] = nb[e];
    }
  }
}

typedef void (*PaeKernel)(PaeNl1Args);
struct PaeVariant {
  int cols;   // columns per thread (column kernels)
  int km;     // momenta per thread (momentum kernels), 0 for column kernels
  bool nr1;
  PaeKernel fn;
  const char* name;
  bool fused = false;  // forms the re-indexed PhiR itself (no pre-pass)
  bool all = false;    // one workgroup covers every momentum (km at a time)
};
static const PaeVariant kPaeVariants[] = {
    {2, 0, true, pae_stream_nl1<2, true>, "nl1_c2"},
    {4, 0, true, pae_stream_nl1<4, true>, "nl1_c4"},
    {8, 0, true, pae_stream_nl1<8, true>, "nl1_c8"},
    {4, 0, false, pae_stream_nl1<4, false>, "nl1_c4_nr"},
    {1, 1, true, pae_stream_nl1_mom<1>, "nl1_m1"},
    {1, 2, true, pae_stream_nl1_mom<2>, "nl1_m2"},
    {1, 4, true, pae_stream_nl1_mom<4>, "nl1_m4"},
    {1, 8, true, pae_stream_nl1_mom<8>, "nl1_m8"},
    {1, 16, true, pae_stream_nl1_mom<16>, "nl1_m16"},
    {1, 8, true, pae_stream_nl1_fused<8>, "nl1_f8", true},
    {1, 16, true, pae_stream_nl1_fused<16>, "nl1_f16", true},
    {1, 2, true, pae_stream_nl1_all<2, false>, "nl1_a2", true, true},
    {1, 1, true, pae_stream_nl1_all<1, true>, "nl1_a1p", true, true},
    {2, 1, true, pae_stream_nl1_multi<2>, "nl1_e2p", true, true},
    {4, 1, true, pae_stream_nl1_multi<4>, "nl1_e4p", true, true},
};
constexpr int kNumPaeVariants = sizeof(kPaeVariants) / sizeof(kPaeVariants[0]);

int pae_variant_count() { return kNumPaeVariants; }
const char* pae_variant_name(int v) { return v >= 0 && v < kNumPaeVariants ? kPaeVariants[v].name : ""; }
// variant usable for this context (Nl = 1, and NR = 1 for the nr1 kernels)
bool pae_variant_ok(const tups_ctx* c, int v) {
  return v >= 0 && v < kNumPaeVariants && c->sz.nl == 1 && (!kPaeVariants[v].nr1 || c->sz.nr == 1);
}

static int pae_nl1_launch(tups_ctx* c, int v, const PaeNl1Args& b) {
  const tups_sizes& s = c->sz;
  const int cols = kPaeVariants[v].cols, km = kPaeVariants[v].km;
  const long long tiles = (long long)((s.nf + 63) / 64) * ((s.nf + 4 * cols - 1) / (4 * cols));
  const long long ngm = kPaeVariants[v].all ? 1 : (s.nibz + km - 1) / (km ? km : 1);
  const long long nb1 = km ? tiles * ngm * 2 * s.nw : tiles * 2 * s.nq;
  TUPS_CHECK(nb1 < (1LL << 31), TUPS_EINVAL, "crossing grid too large");
  hipLaunchKernelGGL(kPaeVariants[v].fn, dim3((unsigned)nb1), dim3(256), 0, c->stream, b);
  TUPS_HIP(hipGetLastError());
  return TUPS_OK;
}

static PaeNl1Args nl1_args(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  PaeNl1Args b{};
  for (int r = 0; r < 4; ++r) {
    b.phir[r] = c->phir[r];
    b.phi[r] = c->gamma[r];
    b.F[r] = c->fvert[r];
    b.lam[r] = c->lam[r];
    b.M[r] = c->mten[r];
  }
  b.Y = c->pae_y;
  b.nf = s.nf;
  b.nw = s.nw;
  b.nibz = s.nibz;
  b.nr = s.nr;
  b.nq = s.nq;
  b.inv_nx = 1.0 / s.nx;
  const double f0 = c->tab.ff[0];
  b.nc_aux = std::fabs(1.0 / f0 * 1.0 / f0);
  return b;
}

static int pae_pregather(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  const size_t per = (size_t)s.nf * s.nf * s.nw * s.nr;
  hipLaunchKernelGGL(pae_pregather_nl1, dim3((unsigned)((per + 255) / 256), 8), dim3(256), 0, c->stream,
                     c->phir[0], c->phir[1], c->phir[2], c->phir[3], c->pae_y, s.nf, s.nw, s.nr);
  TUPS_HIP(hipGetLastError());
  return TUPS_OK;
}

// Times the usable streaming variants on the resident inputs (the output F is
// rewritten identically each time), each with the re-index pre-pass it needs,
// and keeps the fastest.
int autotune_pae(tups_ctx* c, int reps, int* chosen, float* best_ms) {
  TUPS_CHECK(c->sz.nl == 1 && c->pae_y, TUPS_EINVAL, "crossing variants exist for Nl = 1 only");
  if (int rc = pae_pregather(c)) return rc;
  const PaeNl1Args b = nl1_args(c);
  float best = 0.f;
  int arg = -1;
  for (int v = 0; v < kNumPaeVariants; ++v) {
    if (!pae_variant_ok(c, v)) continue;
    const bool pre = !kPaeVariants[v].fused;  // the pre-pass counts for the variants that read it
    if (int rc = pae_nl1_launch(c, v, b)) return rc;
    TUPS_HIP(hipEventRecord(c->tev[0], c->stream));
    for (int i = 0; i < reps; ++i) {
      if (pre)
        if (int rc = pae_pregather(c)) return rc;
      if (int rc = pae_nl1_launch(c, v, b)) return rc;
    }
    TUPS_HIP(hipEventRecord(c->tev[1], c->stream));
    TUPS_HIP(hipEventSynchronize(c->tev[1]));
    float ms = 0.f;
    TUPS_HIP(hipEventElapsedTime(&ms, c->tev[0], c->tev[1]));
    ms /= reps;
    if (arg < 0 || ms < best) {
      best = ms;
      arg = v;
    }
  }
  c->pae_variant = arg;
  if (chosen) *chosen = arg;
  if (best_ms) *best_ms = best;
  return TUPS_OK;
}

// ---- Nl > 1: the crossing as complex GEMMs on FP64 MFMA --------------------------
//
// For a target pair p (0: d,m; 1: s,t) each of its two routines k contributes
//   (1/Nx) sum_{l3,l4,R} M_k(l1,l2,l3,l4,Q,R) Y_kt((l3,l4,R); nu1, nu2, qw)
// where Y_kt is the routine's PhiR gathered along its frequency map, combined
// over the source channels with target t's factors, time reversal (conj and
// signs(l3) signs(l4)) folded in and zero where the routine does not reach the
// entry: it depends on q only through qw.  With rows r = (l1, l2, Q) (Nl^2
// NIBZ), columns n = (nu1, nu2) (Nf^2) and the two routines' inner indices
// (Nl^2 NR each) concatenated, each (qw, pair, target) is one complex GEMM
//   OUT = [M_k0 | M_k1] [Y_k0t ; Y_k1t]
// whose epilogue adds Phi and Lambda Nc (l1 = l2 = 1 block) and writes F once.
// The columns are ordered in 8 x 8 blocks of (nu1, nu2), one block per 64-column
// tile: a routine reaches the entries of a band in (nu1, nu2) (|nu2 - nu1| or
// nu1 + nu2 against qw bounded), so for about a fifth of the (tile, routine)
// pairs its Y block is zero throughout, and the tile skips that half of the
// inner dimension.

struct PaeMmArgs {
  const double2* phi[4];
  double2* F[4];
  const double2* lam[4];
  const double2* M[4];
  const double2* phir[4];
  const int* signs;
  double2* Y;  // [kt = 4 pair + 2 routine + target][qw][j = l3 + Nl l4 + Nl^2 R][n]: n = 64 block + c,
               // block = bx + ceil(Nf / 8) by, nu1 = 8 bx + c % 8, nu2 = 8 by + c / 8
  const int* rows;  // GEMM row -> (l1, l2, Q) row index l1 + Nl l2 + Nl^2 m, or null for all rows in order
  int nf, nl, nw, nibz, nr, nz, kk, rr, tiles_r, tiles_n, nbf;  // nbf = ceil(Nf / 8), tiles_n = nbf^2
  double inv_nx, nc_aux;
};

// 0-based (nu1, nu2) of blocked column n; false for the padding of the last blocks
__device__ __forceinline__ bool pae_col(const PaeMmArgs& a, int n, int& nu1, int& nu2) {
  const int blk = n >> 6, c = n & 63;
  nu1 = 8 * (blk % a.nbf) + (c & 7);
  nu2 = 8 * (blk / a.nbf) + (c >> 3);
  return nu1 < a.nf && nu2 < a.nf;
}

// does routine rt of target pair p reach entry (nu1, nu2) (1-based) at qw?  The
// frequency conditions of the terms ybuild forms (weight nonzero)
__device__ __forceinline__ bool pae_reaches(int p, int rt, int nu1, int nu2, int qw, int nf, int nw) {
  if (p == rt) {  // phbar->ph (p 0, rt 0) and phbar->pp (p 1, rt 1): |nu2 - nu1| within the bosonic box
    const int dn = nu2 - nu1;
    return (dn >= 0 && dn + 1 <= nw) || (dn < 0 && 1 - dn <= nw);
  }
  int w2, w2t;
  if (p == 0) {  // pp->ph
    w2 = nu1 + nu2 + qw - nf - 1;
    w2t = nf + 3 - qw - nu1 - nu2;
  } else {  // ph->pp
    w2 = qw - nu1 - nu2 + nf + 1;
    w2t = nu1 + nu2 - qw - nf + 1;
  }
  return (w2 >= 1 && w2 <= nw) || (w2t >= 2 && w2t <= nw);
}

// Y for all (kt, qw, j, n): one thread per n for a fixed (j, qw).  The gathers of
// neighbouring n at neighbouring qw share cache lines of PhiR (the frequency
// maps shift nu1 against qw), so qw varies fastest over the workgroups of one
// XCD.  grid: ((Nf^2 + 255) / 256) * kk * nw
__global__ __launch_bounds__(256) void pae_ybuild(PaeMmArgs a) {
  const int nf = a.nf, nl = a.nl, nw = a.nw, nz = a.nz;
  const int ncol = 64 * a.tiles_n;
  const int nbn = ncol / 256 + (ncol % 256 ? 1 : 0);
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int w0 = logical % nw, rest = logical / nw;
  const int j = rest / nbn, qw = w0 + 1;
  const int n = (rest - j * nbn) * 256 + threadIdx.x;
  if (n >= ncol) return;
  const size_t nn = (size_t)ncol;
  auto at = [&](int kt) { return ((size_t)(kt * nw + w0) * a.kk + j) * nn + n; };
  int nu1, nu2;
  if (!pae_col(a, n, nu1, nu2)) {
    // padding columns: finite zeros (their products land in columns the epilogue drops)
#pragma unroll
    for (int kt = 0; kt < 8; ++kt) a.Y[at(kt)] = make_double2(0.0, 0.0);
    return;
  }
  ++nu1;
  ++nu2;
  const int l3 = j % nl, l4 = (j / nl) % nl, R = j / (nl * nl);
  const double sgn = (double)(a.signs[l3] * a.signs[l4]);
  auto comb = [&](int s0, int s1, const Term& t, double c0, double c1) {
    const size_t i = (((size_t)R * nw + (t.w2 - 1)) * nz + (size_t)(l4 * nf + t.nub - 1)) * nz +
                     (size_t)(l3 * nf + t.nua - 1);
    const double2 x0 = a.phir[s0][i], x1 = a.phir[s1][i];
    const double sg = t.wt * (t.tr ? sgn : 1.0);
    const double im = t.tr ? -1.0 : 1.0;
    return make_double2(sg * (c0 * x0.x + c1 * x1.x), sg * im * (c0 * x0.y + c1 * x1.y));
  };
  const int dn = nu2 - nu1;
  const bool pb_plain = dn >= 0 && dn + 1 <= nw;
  const bool pb_tr = dn < 0 && 1 - dn <= nw;
  const Term t0 = make_term(pb_plain, pb_tr, dn + 1, nu1, clampnu(nu1 + qw - 1, nf), 1 - dn,
                            nf + 1 - nu1, clampnu(nf + 1 - (nu1 + qw - 1), nf));
  int w2 = nu1 + nu2 + qw - nf - 1, w2t = nf + 3 - qw - nu1 - nu2;
  const Term t1 = make_term(w2 >= 1 && w2 <= nw, w2t >= 2 && w2t <= nw, w2, nu1, nu2, w2t,
                            nf + 1 - nu1, nu1 + qw + w2t - 2);
  w2 = qw - nu1 - nu2 + nf + 1;
  w2t = nu1 + nu2 - qw - nf + 1;
  const Term t2 = make_term(w2 >= 1 && w2 <= nw, w2t >= 2 && w2t <= nw, w2, nu1, nu2, w2t,
                            nf + 1 - nu1, nu1 - qw - w2t + 2);
  const Term t3 = make_term(pb_plain, pb_tr, dn + 1, nu1, clampnu(qw - nu1 - dn + nf, nf), 1 - dn,
                            nf + 1 - nu1, clampnu(nu1 - qw - (1 - dn) + 2, nf));
  a.Y[at(0)] = comb(0, 1, t0, -0.5, -1.5);  // phbar->ph -> d
  a.Y[at(1)] = comb(0, 1, t0, -0.5, 0.5);   //           -> m
  a.Y[at(2)] = comb(2, 3, t1, 0.5, 1.5);    // pp->ph    -> d
  a.Y[at(3)] = comb(2, 3, t1, -0.5, 0.5);   //           -> m
  a.Y[at(4)] = comb(0, 1, t2, 0.5, -1.5);   // ph->pp    -> s
  a.Y[at(5)] = comb(0, 1, t2, 0.5, 0.5);    //           -> t
  a.Y[at(6)] = comb(0, 1, t3, 0.5, -1.5);   // phbar->pp -> s
  a.Y[at(7)] = comb(0, 1, t3, -0.5, -0.5);  //           -> t
}

// One workgroup per 64 x 64 output tile of one (qw, pair, target); four waves of
// 2 x 2 blocks of v_mfma_f64_16x16x4f64 (four real products per complex block);
// the inner dimension walks 16 at a time with the next chunk prefetched into
// registers.  grid: tiles_r * tiles_n * nw * 4 (remapped per XCD, row tile fastest).
constexpr int MMK = 16;
// four workgroups per CU: the 128-register bound costs a few spills and gains
// occupancy (5.85 -> 5.55 ms at Nx = 8, Nl = 5)
__global__ __launch_bounds__(256, 4) void pae_mm(PaeMmArgs a) {
  __shared__ double As[2][MMK][64];
  __shared__ double Bs[2][MMK][64];
  const int tiles = a.tiles_r * a.tiles_n;
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = logical % tiles;
  const int z = logical / tiles;  // = (w0 * 2 + pair) * 2 + target
  const int t = z & 1, p = (z >> 1) & 1, w0 = z >> 2;
  // row tiles vary fastest: the tiles_r workgroups that share one column tile of
  // Y run together on one XCD, so Y comes from HBM once and the re-reads hit L2
  const int row0 = (tile % a.tiles_r) * 64, col0 = (tile / a.tiles_r) * 64;
  const int nl2 = a.nl * a.nl, kk = a.kk;
  const int kkp = (kk + 3) / 4 * 4;  // each routine's inner range padded to whole MFMA k-steps
  const size_t nn = (size_t)64 * a.tiles_n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // the routines that reach some entry of this tile's 8 x 8 frequency block
  // (lane = column; every wave forms the same ballot)
  bool reach0 = false, reach1 = false;
  {
    int nu1, nu2;
    if (pae_col(a, col0 + lane, nu1, nu2)) {
      reach0 = pae_reaches(p, 0, nu1 + 1, nu2 + 1, w0 + 1, a.nf, a.nw);
      reach1 = pae_reaches(p, 1, nu1 + 1, nu2 + 1, w0 + 1, a.nf, a.nw);
    }
  }
  const int vbeg = __any(reach0) ? 0 : kkp, vend = __any(reach1) ? 2 * kkp : kkp;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  // staging: A element (row tid & 63, k (tid >> 6) + 4 i), B element (k (tid >> 6) + 4 i, column tid & 63)
  const int sr = tid & 63, sk = tid >> 6;
  const int gr = row0 + sr;
  const bool rok = gr < a.rr;
  const int grow = rok ? (a.rows ? a.rows[gr] : gr) : 0;
  const int rl = grow % nl2, rm = grow / nl2;
  const int gn = col0 + sr;
  double2 ra[4], rb[4];
  // virtual inner index v: routine v / kkp, its index jj = v % kkp (< kk, else padding); a
  // chunk may straddle the two routines
  auto load = [&](int v0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int v = v0 + sk + 4 * i;
      const int rt = v >= kkp ? 1 : 0, jj = v - rt * kkp;
      const bool kok = jj < kk;
      const int jl = jj % nl2, jr = jj / nl2;
      const size_t mi = (size_t)rl + (size_t)nl2 * jl + (size_t)nl2 * nl2 * (rm + (size_t)a.nibz * jr);
      ra[i] = (rok && kok) ? a.M[2 * p + rt][mi] : make_double2(0.0, 0.0);
      const int kt = 4 * p + 2 * rt + t;
      rb[i] = kok ? a.Y[((size_t)(kt * a.nw + w0) * kk + jj) * nn + gn] : make_double2(0.0, 0.0);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      As[0][sk + 4 * i][sr] = ra[i].x;
      As[1][sk + 4 * i][sr] = ra[i].y;
      Bs[0][sk + 4 * i][sr] = rb[i].x;
      Bs[1][sk + 4 * i][sr] = rb[i].y;
    }
  };
  d4 acc_re[2][2], acc_im[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc_re[i][j] = acc_im[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  if (vbeg < vend) load(vbeg);
  for (int k0 = vbeg; k0 < vend; k0 += MMK) {
    store();
    __syncthreads();
    if (k0 + MMK < vend) load(k0 + MMK);
#pragma unroll
    for (int ks = 0; ks < MMK / 4; ++ks) {
      if (k0 + 4 * ks >= vend) break;  // the last chunk may end early
      const int kr = ks * 4 + lk;
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = As[0][kr][wr + i * 16 + li];
        ai[i] = As[1][kr][wr + i * 16 + li];
        br[i] = Bs[0][kr][wc + i * 16 + li];
        bi[i] = Bs[1][kr][wc + i * 16 + li];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], acc_re[i][j], 0, 0, 0);
          acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bi[j], acc_re[i][j], 0, 0, 1);
          acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], acc_im[i][j], 0, 0, 0);
          acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], acc_im[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  // epilogue: element e of lane l sits at row l / 16 + 4 e, column l % 16 of its block
  const int tgt = 2 * p + t, nf = a.nf, nz = a.nz, nl = a.nl;
  const size_t nz2 = (size_t)nz * nz;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      int nu1, nu2;
      if (!pae_col(a, col0 + wc + j * 16 + li, nu1, nu2)) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int rg = row0 + wr + i * 16 + lk + 4 * e;
        if (rg >= a.rr) continue;
        const int r = a.rows ? a.rows[rg] : rg;
        const int l = r % nl2, m = r / nl2;
        const int l1 = l % nl, l2 = l / nl;
        const size_t q = (size_t)m * a.nw + w0;
        const size_t idx = q * nz2 + (size_t)(l2 * nf + nu2) * nz + (size_t)(l1 * nf + nu1);
        double2 o = a.phi[tgt][idx];
        if (l == 0) {
          const double2 lv = a.lam[tgt][(size_t)nu1 + (size_t)nf * ((size_t)nu2 + (size_t)nf * w0)];
          o.x += lv.x * a.nc_aux;
          o.y += lv.y * a.nc_aux;
        }
        o.x += acc_re[i][j][e] * a.inv_nx;
        o.y += acc_im[i][j][e] * a.inv_nx;
        a.F[tgt][idx] = o;
      }
    }
}

static int pae_gemm_path(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  PaeMmArgs a{};
  for (int r = 0; r < 4; ++r) {
    a.phi[r] = c->gamma[r];
    a.F[r] = c->fvert[r];
    a.lam[r] = c->lam[r];
    a.M[r] = c->mten[r];
    a.phir[r] = c->phir[r];
  }
  a.signs = c->d_signs;
  a.Y = c->pae_ym;
  a.nf = s.nf;
  a.nl = s.nl;
  a.nw = s.nw;
  a.nibz = s.nibz;
  a.nr = s.nr;
  a.nz = s.nz;
  a.kk = s.nl * s.nl * s.nr;
  // the rows the caller keeps (an uneven-box iteration drops the form factors
  // beyond each momentum's Nl), else every (l1, l2, Q)
  a.rows = c->pae_rows;
  a.rr = c->pae_rows ? c->pae_nrows : s.nl * s.nl * s.nibz;
  a.tiles_r = (a.rr + 63) / 64;
  a.nbf = (s.nf + 7) / 8;
  a.tiles_n = a.nbf * a.nbf;
  a.inv_nx = 1.0 / s.nx;
  const double f0 = c->tab.ff[0];
  a.nc_aux = std::fabs(1.0 / f0 * 1.0 / f0);
  if (int rc = timer_begin(c)) return rc;
  hipLaunchKernelGGL(pae_ybuild, dim3((unsigned)((64 * a.tiles_n + 255) / 256 * a.kk * s.nw)), dim3(256), 0,
                     c->stream, a);
  TUPS_HIP(hipGetLastError());
  if (int rc = timer_end(c, TUPS_K_PAE_GATHER)) return rc;
  const long long nblk = (long long)a.tiles_r * a.tiles_n * s.nw * 4;
  TUPS_CHECK(nblk < (1LL << 31), TUPS_EINVAL, "crossing grid too large");
  if (int rc = timer_begin(c)) return rc;
  hipLaunchKernelGGL(pae_mm, dim3((unsigned)nblk), dim3(256), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  if (int rc = timer_end(c, TUPS_K_PAE)) return rc;
  return TUPS_OK;
}

int launch_pae(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  if (s.nl > 1) {
    TUPS_CHECK(c->pae_ym, TUPS_ESTATE, "crossing operand buffer missing (not a full-IBZ context)");
    return pae_gemm_path(c);
  }
  TUPS_CHECK(c->pae_y, TUPS_ESTATE, "crossing operand buffer missing (not a full-IBZ context)");
  int v = c->pae_variant;
  if (!pae_variant_ok(c, v)) v = s.nr == 1 ? 1 : 3;
  if (!kPaeVariants[v].fused) {
    if (int rc = timer_begin(c)) return rc;
    if (int rc = pae_pregather(c)) return rc;
    if (int rc = timer_end(c, TUPS_K_PAE_GATHER)) return rc;
  }
  const PaeNl1Args b = nl1_args(c);
  if (int rc = timer_begin(c)) return rc;
  if (int rc = pae_nl1_launch(c, v, b)) return rc;
  if (int rc = timer_end(c, TUPS_K_PAE)) return rc;
  return TUPS_OK;
}

}  // namespace tups
