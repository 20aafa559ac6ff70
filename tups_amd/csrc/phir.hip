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
// frequency.  Here the eight operations are folded on the host into
// coefficients per symmetry block: the form factors fall into partner groups
// of one or two, and the outputs (l1, l2) of groups Ga x Gb draw only on the
// Phi entries of the same Ga x Gb.  One thread owns one (nu1, nu2, w) of one
// block and channel: per IBZ momentum it loads the block's |Ga||Gb| entries
// (lanes along nu1, so every load is a coalesced row) and accumulates all
// |Ga||Gb| outputs for eight shells R at once.  Every Phi element is read
// exactly once per eight shells and PhiR is written once, so the kernel is
// HBM-bound: bytes per launch = 16 (4 Nz^2 Nq ceil(NR / 8) + 4 Nz^2 Nw NR).
#include "internal.h"

namespace tups {

constexpr int kPhirR = 8;  // shells per thread

struct PhirArgs {
  const double2* phi[4];
  double2* out[4];
  const int* blk;         // [block][8]: la0, la1, lb0, lb1, na, nb, coefficient offset, -
  const double2* coef;    // per block [m][R][o][s]
  int nf, nl, nw, nibz, nr, first;
};

// grid: (ceil(Nf^2 Nw / 256), 4 * blocks of this shape, ceil(NR / 8))
template <int NA, int NB>
__global__ __launch_bounds__(256) void phir_block(PhirArgs a) {
  constexpr int NO = NA * NB;
  const int nf = a.nf, nz = nf * a.nl;
  const size_t nz2 = (size_t)nz * nz;
  const int ch = blockIdx.y & 3, b = a.first + (blockIdx.y >> 2);
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nf * nf * a.nw) return;
  const int nu1 = t % nf, nu2 = (t / nf) % nf, w = t / (nf * nf);
  const int* bk = a.blk + 8 * b;
  const int la[2] = {bk[0], bk[1]}, lb[2] = {bk[2], bk[3]};
  const double2* cf0 = a.coef + bk[6];
  size_t off[NO];
#pragma unroll
  for (int j = 0; j < NB; ++j)
#pragma unroll
    for (int i = 0; i < NA; ++i) off[i + NA * j] = (size_t)(la[i] * nf + nu1) + (size_t)nz * (lb[j] * nf + nu2);
  const int r0 = blockIdx.z * kPhirR, nrc = min(kPhirR, a.nr - r0);
  const double2* phi = a.phi[ch] + nz2 * w;
  const size_t mstride = nz2 * a.nw;
  double ar[NO][kPhirR], ai[NO][kPhirR];
#pragma unroll
  for (int o = 0; o < NO; ++o)
#pragma unroll
    for (int r = 0; r < kPhirR; ++r) ar[o][r] = ai[o][r] = 0.0;
#pragma unroll 2
  for (int m = 0; m < a.nibz; ++m) {
    double2 v[NO];
#pragma unroll
    for (int s = 0; s < NO; ++s) v[s] = ld_stream(phi + mstride * m + off[s]);
    const double2* cf = cf0 + ((size_t)m * a.nr + r0) * NO * NO;
#pragma unroll
    for (int r = 0; r < kPhirR; ++r) {
      if (r < nrc) {
#pragma unroll
        for (int o = 0; o < NO; ++o)
#pragma unroll
          for (int s = 0; s < NO; ++s) {
            const double2 c = cf[(r * NO + o) * NO + s];
            ar[o][r] = fma(c.x, v[s].x, fma(-c.y, v[s].y, ar[o][r]));
            ai[o][r] = fma(c.x, v[s].y, fma(c.y, v[s].x, ai[o][r]));
          }
      }
    }
  }
  double2* out = a.out[ch];
#pragma unroll
  for (int o = 0; o < NO; ++o)
#pragma unroll
    for (int r = 0; r < kPhirR; ++r)
      if (r < nrc)
        st_stream(out + off[o] + nz2 * ((size_t)w + (size_t)a.nw * (r0 + r)), make_double2(ar[o][r], ai[o][r]));
}

int launch_phir(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  PhirArgs a{};
  for (int r = 0; r < 4; ++r) {
    a.phi[r] = c->gamma[r];
    a.out[r] = c->phir[r];
  }
  a.blk = c->d_ph_blk;
  a.coef = c->d_ph_coef;
  a.nf = s.nf;
  a.nl = s.nl;
  a.nw = s.nw;
  a.nibz = s.nibz;
  a.nr = s.nr;
  const unsigned gx = (unsigned)(((size_t)s.nf * s.nf * s.nw + 255) / 256);
  const unsigned gz = (unsigned)((s.nr + kPhirR - 1) / kPhirR);
  void (*kern[4])(PhirArgs) = {phir_block<1, 1>, phir_block<2, 1>, phir_block<1, 2>, phir_block<2, 2>};
  if (int rc = timer_begin(c)) return rc;
  for (int sh = 0; sh < 4; ++sh) {
    if (!c->ph_count[sh]) continue;
    a.first = c->ph_first[sh];
    hipLaunchKernelGGL(kern[sh], dim3(gx, 4u * (unsigned)c->ph_count[sh], gz), dim3(256), 0, c->stream, a);
    TUPS_HIP(hipGetLastError());
  }
  if (int rc = timer_end(c, TUPS_K_PHIR)) return rc;
  return TUPS_OK;
}

}  // namespace tups
