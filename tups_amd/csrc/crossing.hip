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

// acc_a += wt sum M * (ca0*X0 + ca1*X1); acc_b += wt sum M * (cb0*X0 + cb1*X1)
template <int NL>
__device__ __forceinline__ void crossing_term(const PaeArgs& a, const double2* M, int s0, int s1,
                                              double ca0, double ca1, double cb0, double cb1,
                                              int l1, int l2, int m, const Term& t,
                                              double2& acc_a, double2& acc_b) {
  const int nf = a.nf, nz = nf * NL;
  const size_t nz2 = (size_t)nz * nz;
  const double2* p0 = a.phir[s0];
  const double2* p1 = a.phir[s1];
  constexpr int NL4 = NL * NL * NL * NL;
  const double im_sign = t.tr ? -1.0 : 1.0;
  for (int R = 0; R < a.nr; ++R) {
    const size_t pbase = nz2 * ((size_t)(t.w2 - 1) + (size_t)a.nw * R);
    const double2* mb = M + (size_t)NL4 * ((size_t)m + (size_t)a.nibz * R) + l1 + NL * l2;
    double2 ya_sum = make_double2(0.0, 0.0), yb_sum = make_double2(0.0, 0.0);
#pragma unroll 1
    for (int l4 = 0; l4 < NL; ++l4) {
      const size_t colo = pbase + (size_t)nz * (l4 * nf + t.nub - 1) + (t.nua - 1);
#pragma unroll
      for (int l3 = 0; l3 < NL; ++l3) {
        const size_t pi = colo + (size_t)(l3 * nf);
        const double2 x0 = p0[pi];
        const double2 x1 = p1[pi];
        double sg = t.wt;
        if (NL > 1) sg *= t.tr ? (double)(a.signs[l3] * a.signs[l4]) : 1.0;
        const double2 mv = mb[NL * NL * (l3 + NL * l4)];
        // y = sg * (c0 x0 + c1 x1) with conj for time reversal
        const double2 ya = make_double2(sg * (ca0 * x0.x + ca1 * x1.x), sg * im_sign * (ca0 * x0.y + ca1 * x1.y));
        const double2 yb = make_double2(sg * (cb0 * x0.x + cb1 * x1.x), sg * im_sign * (cb0 * x0.y + cb1 * x1.y));
        ya_sum.x += mv.x * ya.x - mv.y * ya.y;
        ya_sum.y += mv.x * ya.y + mv.y * ya.x;
        yb_sum.x += mv.x * yb.x - mv.y * yb.y;
        yb_sum.y += mv.x * yb.y + mv.y * yb.x;
      }
    }
    acc_a.x += ya_sum.x;
    acc_a.y += ya_sum.y;
    acc_b.x += yb_sum.x;
    acc_b.y += yb_sum.y;
  }
}

__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  // workgroups are dispatched round-robin over the 8 XCDs: give each XCD a
  // contiguous logical range (here: a contiguous range of bosonic frequencies)
  const int xcd = bid & 7, idx = bid >> 3;
  const int per = nblk >> 3, rem = nblk & 7;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// 1-D grid remapped per XCD; logical block = ((qw-major q) * 2 + pair) * tiles + tile,
// tile = 64 rows x (4 * CPT) columns of one (q, pair); a thread owns CPT columns
// of one row so that each wave keeps several streaming loads in flight.
constexpr int CPT = 4;

template <int NL>
__global__ __launch_bounds__(256) void pae_kernel(PaeArgs a) {
  const int nf = a.nf, nz = nf * NL, nw = a.nw;
  const size_t nz2 = (size_t)nz * nz;
  const int tx = (nz + 63) >> 6, ty = (nz + 4 * CPT - 1) / (4 * CPT);
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = logical % (tx * ty);
  const int zq = logical / (tx * ty);
  const int row = (tile % tx) * 64 + (threadIdx.x & 63);
  const int col0 = (tile / tx) * (4 * CPT) + (threadIdx.x >> 6);
  if (row >= nz) return;
  const int pair = zq & 1;  // 0: targets d,m ; 1: targets s,t
  const int jq = zq >> 1;
  const int w0 = jq / a.nibz;
  const int m = jq - w0 * a.nibz;
  const int q = m * nw + w0;        // Index_bosonic_IBZ order: momentum-major
  const int qw = w0 + 1;
  const int l1 = row / nf, nu1 = row - l1 * nf + 1;
  const int ta = pair == 0 ? 0 : 2, tb = ta + 1;

  // streaming operands of all CPT columns first
  double2 fa[CPT], fb[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    const int col = col0 + 4 * c;
    const size_t idx = (size_t)row + (size_t)nz * (col < nz ? col : 0) + nz2 * q;
    fa[c] = ld_stream(a.phi[ta] + idx);
    fb[c] = ld_stream(a.phi[tb] + idx);
  }
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    const int col = col0 + 4 * c;
    if (col >= nz) continue;
    const int l2 = col / nf, nu2 = col - l2 * nf + 1;
    double2 la = make_double2(0.0, 0.0), lb = make_double2(0.0, 0.0);
    if (l1 == 0 && l2 == 0) {
      const size_t li = (size_t)(nu1 - 1) + (size_t)nf * ((size_t)(nu2 - 1) + (size_t)nf * (qw - 1));
      la = a.lam[ta][li];
      lb = a.lam[tb][li];
    }
    const int dn = nu2 - nu1;
    // phbar-type term (same frequency geometry for both target pairs)
    const bool pb_plain = dn >= 0 && dn + 1 <= nw;
    const bool pb_tr = dn < 0 && 1 - dn <= nw;
    double2 acc_a = make_double2(0.0, 0.0), acc_b = make_double2(0.0, 0.0);
    if (pair == 0) {
      // phbar -> ph from d (-1/2, -1/2) and m (-3/2, +1/2)
      const Term tb1 = make_term(pb_plain, pb_tr, dn + 1, nu1, clampnu(nu1 + qw - 1, nf), 1 - dn,
                                 nf + 1 - nu1, clampnu(nf + 1 - (nu1 + qw - 1), nf));
      // pp -> ph from s (+1/2, -1/2) and t (+3/2, +1/2)
      const int w2 = nu1 + nu2 + qw - nf - 1, w2t = nf + 3 - qw - nu1 - nu2;
      const Term tp = make_term(w2 >= 1 && w2 <= nw, w2t >= 2 && w2t <= nw, w2, nu1, nu2, w2t,
                                nf + 1 - nu1, nu1 + qw + w2t - 2);
      crossing_term<NL>(a, a.M[0], 0, 1, -0.5, -1.5, -0.5, 0.5, l1, l2, m, tb1, acc_a, acc_b);
      crossing_term<NL>(a, a.M[1], 2, 3, 0.5, 1.5, -0.5, 0.5, l1, l2, m, tp, acc_a, acc_b);
    } else {
      // ph -> pp from d (+1/2, +1/2) and m (-3/2, +1/2)
      const int w2 = qw - nu1 - nu2 + nf + 1, w2t = nu1 + nu2 - qw - nf + 1;
      const Term th = make_term(w2 >= 1 && w2 <= nw, w2t >= 2 && w2t <= nw, w2, nu1, nu2, w2t,
                                nf + 1 - nu1, nu1 - qw - w2t + 2);
      // phbar -> pp from d (+1/2, -1/2) and m (-3/2, -1/2)
      const Term tb2 = make_term(pb_plain, pb_tr, dn + 1, nu1, clampnu(qw - nu1 - dn + nf, nf), 1 - dn,
                                 nf + 1 - nu1, clampnu(nu1 - qw - (1 - dn) + 2, nf));
      crossing_term<NL>(a, a.M[2], 0, 1, 0.5, -1.5, 0.5, 0.5, l1, l2, m, th, acc_a, acc_b);
      crossing_term<NL>(a, a.M[3], 0, 1, 0.5, -1.5, -0.5, -0.5, l1, l2, m, tb2, acc_a, acc_b);
    }
    double2 oa = fa[c], ob = fb[c];
    oa.x += la.x * a.nc_aux + acc_a.x * a.inv_nx;
    oa.y += la.y * a.nc_aux + acc_a.y * a.inv_nx;
    ob.x += lb.x * a.nc_aux + acc_b.x * a.inv_nx;
    ob.y += lb.y * a.nc_aux + acc_b.y * a.inv_nx;
    const size_t idx = (size_t)row + (size_t)nz * col + nz2 * q;
    st_stream(a.F[ta] + idx, oa);
    st_stream(a.F[tb] + idx, ob);
  }
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
  const int nu1 = (int)(gid % nf) + 1;
  const int nu2 = (int)((gid / nf) % nf) + 1;
  const int qw = (int)((gid / nf2) % nw) + 1;
  const int R = (int)(gid / (nf2 * nw));
  const size_t rb = nf2 * (size_t)nw * R;  // PhiR(nu_a, nu_b, w2, R) = rb + (w2-1)*nf2 + (nub-1)*nf + nua-1
  auto at = [&](const Term& t) { return rb + (size_t)(t.w2 - 1) * nf2 + (size_t)(t.nub - 1) * nf + (t.nua - 1); };
  auto comb = [&](const double2* p0, const double2* p1, const Term& t, double c0, double c1) {
    const size_t i = at(t);
    const double2 x0 = p0[i], x1 = p1[i];
    const double im = t.tr ? -1.0 : 1.0;
    return make_double2(t.wt * (c0 * x0.x + c1 * x1.x), t.wt * im * (c0 * x0.y + c1 * x1.y));
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
  const size_t o = gid;  // same (nu1, nu2, qw, R) ordering as the k slices
  Y[0 * per + o] = comb(pd, pm, t0, -0.5, -1.5);  // phbar->ph  -> d
  Y[1 * per + o] = comb(pd, pm, t0, -0.5, 0.5);   //            -> m
  Y[2 * per + o] = comb(ps, pt, t1, 0.5, 1.5);    // pp->ph     -> d
  Y[3 * per + o] = comb(ps, pt, t1, -0.5, 0.5);   //            -> m
  Y[4 * per + o] = comb(pd, pm, t2, 0.5, -1.5);   // ph->pp     -> s
  Y[5 * per + o] = comb(pd, pm, t2, 0.5, 0.5);    //            -> t
  Y[6 * per + o] = comb(pd, pm, t3, 0.5, -1.5);   // phbar->pp  -> s
  Y[7 * per + o] = comb(pd, pm, t3, -0.5, -0.5);  //            -> t
}

struct PaeNl1Args {
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

typedef void (*PaeKernel)(PaeNl1Args);
struct PaeVariant {
  int cols;   // columns per thread (column kernels)
  int km;     // momenta per thread (momentum kernels), 0 for column kernels
  bool nr1;
  PaeKernel fn;
  const char* name;
};
static const PaeVariant kPaeVariants[] = {
    {2, 0, true, pae_stream_nl1<2, true>, "nl1_c2"},
    {4, 0, true, pae_stream_nl1<4, true>, "nl1_c4"},
    {8, 0, true, pae_stream_nl1<8, true>, "nl1_c8"},
    {4, 0, false, pae_stream_nl1<4, false>, "nl1_c4_nr"},
    {1, 2, true, pae_stream_nl1_mom<2>, "nl1_m2"},
    {1, 4, true, pae_stream_nl1_mom<4>, "nl1_m4"},
    {1, 8, true, pae_stream_nl1_mom<8>, "nl1_m8"},
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
  const long long nb1 = km ? tiles * ((s.nibz + km - 1) / km) * 2 * s.nw : tiles * 2 * s.nq;
  TUPS_CHECK(nb1 < (1LL << 31), TUPS_EINVAL, "crossing grid too large");
  hipLaunchKernelGGL(kPaeVariants[v].fn, dim3((unsigned)nb1), dim3(256), 0, c->stream, b);
  TUPS_HIP(hipGetLastError());
  return TUPS_OK;
}

static PaeNl1Args nl1_args(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  PaeNl1Args b{};
  for (int r = 0; r < 4; ++r) {
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
  hipLaunchKernelGGL(pae_pregather_nl1, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, c->stream,
                     c->phir[0], c->phir[1], c->phir[2], c->phir[3], c->pae_y, s.nf, s.nw, s.nr);
  TUPS_HIP(hipGetLastError());
  return TUPS_OK;
}

// Times the usable streaming variants on the resident inputs (the output F is
// rewritten identically each time) and keeps the fastest.
int autotune_pae(tups_ctx* c, int reps, int* chosen, float* best_ms) {
  TUPS_CHECK(c->sz.nl == 1 && c->pae_y, TUPS_EINVAL, "crossing variants exist for Nl = 1 only");
  if (int rc = pae_pregather(c)) return rc;
  const PaeNl1Args b = nl1_args(c);
  float best = 0.f;
  int arg = -1;
  for (int v = 0; v < kNumPaeVariants; ++v) {
    if (!pae_variant_ok(c, v)) continue;
    if (int rc = pae_nl1_launch(c, v, b)) return rc;
    TUPS_HIP(hipEventRecord(c->tev[0], c->stream));
    for (int i = 0; i < reps; ++i)
      if (int rc = pae_nl1_launch(c, v, b)) return rc;
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
  if (s.nl == 1 && c->pae_y) {
    int v = c->pae_variant;
    if (!pae_variant_ok(c, v)) v = s.nr == 1 ? 1 : 3;
    if (int rc = timer_begin(c)) return rc;
    if (int rc = pae_pregather(c)) return rc;
    if (int rc = timer_end(c, TUPS_K_PAE_GATHER)) return rc;
    const PaeNl1Args b = nl1_args(c);
    if (int rc = timer_begin(c)) return rc;
    if (int rc = pae_nl1_launch(c, v, b)) return rc;
    if (int rc = timer_end(c, TUPS_K_PAE)) return rc;
    return TUPS_OK;
  }
  const long long nblk = (long long)((s.nz + 63) / 64) * ((s.nz + 4 * CPT - 1) / (4 * CPT)) * 2 * s.nq;
  TUPS_CHECK(nblk < (1LL << 31), TUPS_EINVAL, "crossing grid too large");
  dim3 g((unsigned)nblk);
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
