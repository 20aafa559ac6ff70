// Bethe-Salpeter ladder update on FP64 MFMA (gfx950).
//
// Reference: BSE(ite), SRC/parquet_BSE.f90:62-132.  Per transfer q and channel
// r in {d, m, s, t} the reference issues four ZGEMMs and a damped mixing:
//   chiG = alpha*Gamma*chi;  T  = c_r f * chiG * F
//   chiG = alpha*chi*Gamma;  T += c_r f * F * chiG
//   Gamma <- T + (1 - f)(F - Gamma)
// with alpha = 1/(beta Nx Ny), c = 1/2, 1/2, -1/4, 1/4 and chi = chi_ph for d,m
// and chi_pp for s,t.  chi is block diagonal in the fermionic frequency
// (SRC/parquet_chi.f90:500-521), so alpha*Gamma*chi (P) and alpha*chi*Gamma (Q)
// cost O(Nl D^2); what remains is T = c f (P F + F Q): one complex GEMM with
// inner dimension 2D, operands [P | F] and [F ; Q].
//
// Kernel layout: one workgroup (4 waves) per 64x64 output tile of one (q, r)
// matrix; each wave owns a 32x32 sub-tile = 2x2 blocks of 16x16 complex.  A
// complex block uses four real MFMAs per k-step (re*re, -im*im, re*im, im*re)
// on v_mfma_f64_16x16x4f64, accumulators kept separate for Re and Im.  Operand
// tiles (16 k-columns) are staged through LDS as split Re/Im planes so each
// MFMA operand is one aligned double per lane.  The mixing term is fused into
// the epilogue.  For Nl = 1 the chi scaling of P and Q is applied while the
// tile is loaded (chi is diagonal); for Nl > 1 a prologue kernel forms P and Q.
#include "internal.h"

namespace tups {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64;
constexpr int BN = 64;
constexpr int BK = 16;

struct BseArgs {
  const double2* gam[4];   // Gamma_r, (D, D, nq)
  const double2* fv[4];    // F_r
  double2* out[4];         // Phi_r (must not alias gam)
  const double2* chi[2];   // (Nl, Nl, Nf, nq): ph, pp
  const double2* ws;       // P, Q workspace for q in [q0, q0 + nqb): [(r*nqb + qb)*2 + {0,1}] * D*D
  int D, nf, nl;
  int q0, nqb, tiles_n;
  double alpha;
  double cf[4];            // c_r * f
  double omf;              // 1 - f
};

__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  // consecutive workgroups are dispatched round-robin over the 8 XCDs; keep
  // the tiles of one matrix on one XCD so they share its L2.
  const int xcd = bid & 7, idx = bid >> 3;
  const int per = nblk >> 3, rem = nblk & 7;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <bool kScaled>
__global__ __launch_bounds__(256) void bse_ladder_kernel(BseArgs a) {
  __shared__ double As[2][BK][BM];
  __shared__ double Bs[2][BK][BN];

  const int tiles = a.tiles_n * a.tiles_n;
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int mat = logical / tiles;
  const int t = logical - mat * tiles;
  const int r = mat / a.nqb;
  const int qb = mat - r * a.nqb;
  const int q = a.q0 + qb;
  const int D = a.D;
  const size_t DD = (size_t)D * D;
  const int row0 = (t / a.tiles_n) * BM;
  const int col0 = (t % a.tiles_n) * BN;

  const double2* G = a.gam[r] + DD * q;
  const double2* F = a.fv[r] + DD * q;
  const double2* A1;
  const double2* B2;
  if (kScaled) {
    A1 = G;
    B2 = G;
  } else {
    A1 = a.ws + DD * ((size_t)(r * a.nqb + qb) * 2);
    B2 = A1 + DD;
  }
  const double2* chiq = a.chi[r >> 1] + (size_t)a.nf * a.nl * a.nl * q;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = (wave >> 1) * 32;
  const int wc = (wave & 1) * 32;
  const int li = lane & 15;
  const int lk = lane >> 4;

  d4 acc_re[2][2], acc_im[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc_re[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      acc_im[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }

  // A-tile loader: 16 columns x 64 rows, thread -> (column ka, rows ia..ia+3)
  const int ka = tid >> 4;
  const int ia = (tid & 15) * 4;
  // B-tile loader: 64 columns x 16 rows, thread -> (column jb, rows kb..kb+3)
  const int jb = tid >> 2;
  const int kb = (tid & 3) * 4;

  for (int part = 0; part < 2; ++part) {
    const double2* A = part == 0 ? A1 : F;
    const double2* B = part == 0 ? F : B2;
    for (int k0 = 0; k0 < D; k0 += BK) {
      {
        const int gk = k0 + ka;
        double2 sc = make_double2(0.0, 0.0);
        if (kScaled && part == 0 && gk < D) {
          const double2 c = chiq[gk];
          sc = make_double2(a.alpha * c.x, a.alpha * c.y);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gi = row0 + ia + j;
          double2 v = make_double2(0.0, 0.0);
          if (gi < D && gk < D) v = A[(size_t)gk * D + gi];
          if (kScaled && part == 0) v = cmul(v, sc);
          As[0][ka][ia + j] = v.x;
          As[1][ka][ia + j] = v.y;
        }
      }
      {
        const int gj = col0 + jb;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gk = k0 + kb + j;
          double2 v = make_double2(0.0, 0.0);
          if (gj < D && gk < D) {
            v = B[(size_t)gj * D + gk];
            if (kScaled && part == 1) {
              const double2 c = chiq[gk];
              v = cmul(v, make_double2(a.alpha * c.x, a.alpha * c.y));
            }
          }
          Bs[0][kb + j][jb] = v.x;
          Bs[1][kb + j][jb] = v.y;
        }
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < BK / 4; ++ks) {
        const int kr = ks * 4 + lk;
        double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          ar[i] = As[0][kr][wr + i * 16 + li];
          ai[i] = As[1][kr][wr + i * 16 + li];
          nai[i] = -ai[i];
          br[i] = Bs[0][kr][wc + i * 16 + li];
          bi[i] = Bs[1][kr][wc + i * 16 + li];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], acc_re[i][j], 0, 0, 0);
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], bi[j], acc_re[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], acc_im[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], acc_im[i][j], 0, 0, 0);
          }
      }
      __syncthreads();
    }
  }

  // epilogue: Phi = c f T + (1 - f)(F - Gamma).  v_mfma_f64_16x16x4f64 leaves
  // element e of lane l at block row (l / 16) + 4 e, column l % 16.
  const double cf = a.cf[r];
  double2* O = a.out[r] + DD * q;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gj = col0 + wc + j * 16 + li;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int gi = row0 + wr + i * 16 + lk + 4 * e;
        if (gi < D && gj < D) {
          const size_t idx = (size_t)gj * D + gi;
          const double2 f = F[idx];
          const double2 g = G[idx];
          O[idx] = make_double2(cf * acc_re[i][j][e] + a.omf * (f.x - g.x),
                                cf * acc_im[i][j][e] + a.omf * (f.y - g.y));
        }
      }
    }
}

// P = alpha*Gamma*chi, Q = alpha*chi*Gamma for Nl > 1 (block-diagonal chi).
__global__ __launch_bounds__(256) void bse_mix_kernel(BseArgs a) {
  const int D = a.D, nf = a.nf, nl = a.nl;
  const size_t DD = (size_t)D * D;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= DD) return;
  const int mat = blockIdx.y;
  const int r = mat / a.nqb;
  const int qb = mat - r * a.nqb;
  const int q = a.q0 + qb;
  const double2* G = a.gam[r] + DD * q;
  const double2* chiq = a.chi[r >> 1] + (size_t)nf * nl * nl * q;
  double2* P = const_cast<double2*>(a.ws) + DD * ((size_t)(r * a.nqb + qb) * 2);
  double2* Q = P + DD;
  const int i = (int)(e % D);
  const int c = (int)(e / D);
  // P(i, c) with c = (l2, nu): sum_l3 Gamma(i, (l3, nu)) chi_nu(l3, l2)
  {
    const int l2 = c / nf, nu = c - l2 * nf;
    const double2* cb = chiq + (size_t)nu * nl * nl;
    double sr = 0.0, si = 0.0;
    for (int l3 = 0; l3 < nl; ++l3) {
      const double2 g = G[(size_t)(l3 * nf + nu) * D + i];
      const double2 x = cb[l3 + nl * l2];
      sr += g.x * x.x - g.y * x.y;
      si += g.x * x.y + g.y * x.x;
    }
    P[e] = make_double2(a.alpha * sr, a.alpha * si);
  }
  // Q(k, c) with k = i = (l1, nu): sum_l3 chi_nu(l1, l3) Gamma((l3, nu), c)
  {
    const int l1 = i / nf, nu = i - l1 * nf;
    const double2* cb = chiq + (size_t)nu * nl * nl;
    double sr = 0.0, si = 0.0;
    for (int l3 = 0; l3 < nl; ++l3) {
      const double2 x = cb[l1 + nl * l3];
      const double2 g = G[(size_t)c * D + l3 * nf + nu];
      sr += x.x * g.x - x.y * g.y;
      si += x.x * g.y + x.y * g.x;
    }
    Q[e] = make_double2(a.alpha * sr, a.alpha * si);
  }
}

int launch_bse_ladder(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  const int D = s.nz;
  const size_t DD = (size_t)D * D;
  BseArgs a{};
  for (int r = 0; r < 4; ++r) {
    a.gam[r] = c->gamma[r];
    a.fv[r] = c->fvert[r];
    a.out[r] = c->spare[r];
  }
  a.chi[0] = c->chi[0];
  a.chi[1] = c->chi[1];
  a.D = D;
  a.nf = s.nf;
  a.nl = s.nl;
  a.tiles_n = (D + BM - 1) / BM;
  a.alpha = s.alpha;
  const double cr[4] = {0.5, 0.5, -0.25, 0.25};
  for (int r = 0; r < 4; ++r) a.cf[r] = cr[r] * c->f_damping;
  a.omf = 1.0 - c->f_damping;
  const int tiles = a.tiles_n * a.tiles_n;
  const bool scaled = s.nl == 1;
  const int qbatch = scaled ? s.nq : c->bse_qbatch;
  for (int q0 = 0; q0 < s.nq; q0 += qbatch) {
    a.q0 = q0;
    a.nqb = std::min(qbatch, s.nq - q0);
    a.ws = c->work;
    if (!scaled) {
      dim3 g((unsigned)((DD + 255) / 256), (unsigned)(4 * a.nqb));
      if (int rc = timer_begin(c)) return rc;
      hipLaunchKernelGGL(bse_mix_kernel, g, dim3(256), 0, c->stream, a);
      TUPS_HIP(hipGetLastError());
      if (int rc = timer_end(c, TUPS_K_BSE_MIX)) return rc;
    }
    const long long nblk = (long long)tiles * 4 * a.nqb;
    TUPS_CHECK(nblk < (1LL << 31), TUPS_EINVAL, "BSE grid too large");
    if (int rc = timer_begin(c)) return rc;
    if (scaled)
      hipLaunchKernelGGL(bse_ladder_kernel<true>, dim3((unsigned)nblk), dim3(256), 0, c->stream, a);
    else
      hipLaunchKernelGGL(bse_ladder_kernel<false>, dim3((unsigned)nblk), dim3(256), 0, c->stream, a);
    TUPS_HIP(hipGetLastError());
    if (int rc = timer_end(c, TUPS_K_BSE_LADDER)) return rc;
  }
  return TUPS_OK;
}

}  // namespace tups
