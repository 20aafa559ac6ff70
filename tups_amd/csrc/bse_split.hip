// Split-precision BSE ladder: the complex GEMM of the ladder on INT8 MFMA.
//
// Same update as bse.hip (reference BSE(ite), SRC/parquet_BSE.f90:62-132):
//   Phi = c_r f T + (1 - f)(F - Gamma),  T = Gamma (alpha chi F) + F (alpha chi Gamma)
// written as one GEMM T = A B with A = [Gamma | F] (D x 2D) and
// B = [alpha chi F ; alpha chi Gamma] (2D x D); chi is applied to B, whose
// columns are contiguous in memory.
//
// Method: exact integer GEMM by residues (the modular form of the Ozaki
// scheme, Ozaki, Uchino, Imamura 2025), with Gaussian moduli for the complex
// product.
//  1. Row i of A is scaled to integers A'_ik = rint(A_ik 2^(ta - e_i)), with
//     2^(e_i - 1) <= max_k |A_ik| < 2^e_i, so |A'| <= 2^ta; column j of B
//     likewise with tb and f_j.
//  2. For pairwise coprime moduli m (all of them = 1 mod 4, each <= 241) with
//     j^2 = -1 (mod m), a complex integer z = x + i y maps to the two residues
//     (x + j y) mod m and (x - j y) mod m, and a complex product becomes two
//     real products: (A'B')(+) = A'(+) B'(+), (A'B')(-) = A'(-) B'(-) (mod m).
//     Symmetric residues lie in [-120, 120], so each product is one INT8 MFMA
//     GEMM; its INT32 sums are exact for 2D * 120^2 < 2^31 (D <= 1024).
//  3. Re C' = (P+ + P-)/2 and Im C' = (P+ - P-)/(2j) (mod m); the Chinese
//     remainder theorem gives C' = sum_m u_m W_m (mod M), W_m = M/m ((M/m)^-1
//     mod m), M = prod m; the factors 1/2 and 1/(2j) are folded into the
//     weights, so only the residues of P+ + P- and P+ - P- are formed.
//     ta + tb <= log2 M - log2(8D) - 1 keeps |C'| < M/4,
//     so C' is the representative nearest zero.  The sum is held as an exact
//     high part (W rounded to 41 significant bits, a multiple of 2^L) plus a
//     low part, so C' = (S_hi - q M_hi) + (S_lo - q M_lo), q = rint(S / M),
//     carries only the final rounding.
//  4. T_ij = C'_ij 2^(e_i - ta + f_j - tb), then the damped mixing in FP64.
// The error is that of rounding A and B to ta and tb bits relative to their
// row and column maxima: 12 moduli (ta + tb = 79 at D = 320) give ~2e-11
// relative Frobenius error, 16 moduli reach FP64 rounding.
#include "internal.h"
#include "lds_dma.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <type_traits>

namespace tups {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int CRT_MAX = 16;      // most moduli
constexpr int CRT_TILE = 64;     // output tile of one workgroup

// pairwise coprime, all = 1 (mod 4): 221 = 13 * 17, 205 = 5 * 41, the rest prime
constexpr int kModuli[CRT_MAX] = {241, 233, 229, 221, 205, 197, 193, 181,
                                  173, 157, 149, 137, 113, 109, 101, 97};

struct CrtArgs {
  const double2* gam[4];
  const double2* fv[4];
  double2* out[4];
  const double2* chi[2];
  int8_t* A;                    // [mat][mod][sign][Kp / 64][Dp rows][64]: 64-byte K chunks
  int8_t* B;                    // [mat][mod][sign][Kp / 64][Dp cols][64]
  int* escale;                  // [mat][Dp] row exponents of A
  double* rmax;                 // [mat][Dp] row maxima of A
  double* rmaxp;                // [mat][Dp / 16][Dp] partial row maxima of A per 16-column block
  int* fscale;                  // [mat][Dp] column exponents of B
  int* crt;                     // large-tile GEMM: per-modulus (Re, Im) residue bytes of each tile
  // batch (null for a context): the q of this (Nf, Nl) class, and the element
  // offsets of every q's vertex and bubble blocks
  const int* qlist;
  const long long* voff;
  const long long* coff;
  int D, Dp, Kp, nf, nl;
  int q0, nqb, tiles_n;
  int nmod, ta, tb;
  double alpha;
  double cf[4];
  double omf;
  double m[CRT_MAX], inv_m[CRT_MAX], j[CRT_MAX], c26[CRT_MAX];
  // CRT weights of the Re and Im residues with the factors 1/2 and 1/(2j) of
  // Re = (P+ + P-)/2, Im = (P+ - P-)/(2j) folded in, exact high part + low part
  double wrh[CRT_MAX], wrl[CRT_MAX], wih[CRT_MAX], wil[CRT_MAX];
  unsigned nm24[CRT_MAX];       // (-m) mod 2^24
  double mhi, mlo, inv_mm;
};

// channel r of workspace matrix mat and the element offsets of its q's vertex
// block (vo) and bubble blocks (co): q after q for a context, from the offset
// tables for a batch class
__device__ __forceinline__ void crt_mat(const CrtArgs& a, int mat, int& r, size_t& vo, size_t& co) {
  r = mat / a.nqb;
  const int ql = a.q0 + (mat - r * a.nqb);
  if (a.qlist) {
    const int q = a.qlist[ql];
    vo = (size_t)a.voff[q];
    co = (size_t)a.coff[q];
  } else {
    vo = (size_t)a.D * a.D * ql;
    co = (size_t)a.nf * a.nl * a.nl * ql;
  }
}

// byte offset of inner index k of row (column) rc: a 64-row tile of one K chunk
// is 4 KiB contiguous
__device__ __forceinline__ size_t ws_off(const CrtArgs& a, int mat, int mod, int sign, int rc, int k) {
  return (((size_t)mat * a.nmod + mod) * 2 + sign) * (size_t)a.Dp * a.Kp + (size_t)(k >> 6) * a.Dp * 64 +
         (size_t)rc * 64 + (k & 63);
}

__device__ __forceinline__ int xcd_contiguous(int bid, int nblk) {
  const int xcd = bid & 7, idx = bid >> 3;
  const int per = nblk >> 3, rem = nblk & 7;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// LDS slot of the 16-byte column c of a 64-byte row: c ^ crt_swz(row).  A
// fragment read has lanes on 16 consecutive rows at one column; with this XOR
// any 8 consecutive rows touch 32 distinct 4-byte banks and any 16 consecutive
// rows 64, so the 16-byte reads are free of bank conflicts whether the LDS
// serves them 8 or 16 lanes at a time.
__device__ __forceinline__ int crt_swz(int row) { return ((row >> 1) & 3) ^ ((row >> 3) & 1); }

// symmetric residue of an integer-valued y (|y| < 2^50): exact, |r| <= (m - 1) / 2
__device__ __forceinline__ double sym_mod(double y, double m, double inv_m) {
  return fma(-m, rint(y * inv_m), y);
}

// symmetric residue of an INT32 y, as the low byte of the result: rint(y/m) is
// read from the low mantissa bits of y/m + 1.5 2^52 (exact: y/m is never within
// 2^-12 of a half integer) and y - m rint(y/m) is formed on 24-bit integers
__device__ __forceinline__ int sym_mod_lo(int y, double inv_m, unsigned nm24) {
  const double t = fma((double)y, inv_m, 0x1.8p52);
  return (int)((unsigned)y + __umul24((unsigned)__double2loint(t), nm24));
}

// four low bytes packed into one dword, byte e from x_e
__device__ __forceinline__ int pack4(int x0, int x1, int x2, int x3) {
  const unsigned lo = __builtin_amdgcn_perm((unsigned)x1, (unsigned)x0, 0x0c0c0400u);
  const unsigned hi = __builtin_amdgcn_perm((unsigned)x3, (unsigned)x2, 0x0c0c0400u);
  return (int)__builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// x (integer valued, |x| <= 2^52) = hi 2^26 + lo, 0 <= lo < 2^26
__device__ __forceinline__ void limbs(double x, double& hi, double& lo) {
  hi = floor(x * 0x1p-26);
  lo = fma(-0x1p26, hi, x);
}

// the (+, -) residues of the complex integer (xr, xi) given as limbs, as bytes
__device__ __forceinline__ void gauss_residues(const CrtArgs& a, int mod, double r1, double r0,
                                               double i1, double i0, int& rp, int& rm) {
  const double m = a.m[mod], im = a.inv_m[mod], jj = a.j[mod], c = a.c26[mod];
  const double yr = fma(r1, c, r0), yi = fma(i1, c, i0);      // |y| < 2^35
  const double yp = fma(jj, yi, yr), ym = fma(-jj, yi, yr);  // |y| < 2^43
  rp = (int)sym_mod(yp, m, im) & 0xff;
  rm = (int)sym_mod(ym, m, im) & 0xff;
}

// Row maxima of A = [Gamma | F] from the partial maxima crt_prep_b leaves per
// 16-column block (it reads every element of Gamma and F for B).  grid:
// ((Dp + 255) / 256, 4 * nqb); one thread per row.
constexpr int PREP_B_COLS = 16;  // columns per crt_prep_b workgroup
__global__ __launch_bounds__(256) void crt_rowmax_fold(CrtArgs a) {
  const int mat = blockIdx.y;
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= a.Dp) return;
  const int nb = a.Dp / PREP_B_COLS;
  const double* p = a.rmaxp + (size_t)mat * nb * a.Dp + row;
  double mx = 0.0;
  for (int b = 0; b < nb; ++b) mx = fmax(mx, p[(size_t)b * a.Dp]);
  a.rmax[(size_t)mat * a.Dp + row] = mx;
}

// A = [Gamma | F] residues: one workgroup per 64 rows x 64 inner indices; lane
// = row, wave w converts inner indices 16 w .. 16 w + 15 of the chunk.  Per
// modulus the (+, -) residue bytes of the block go through LDS (rows padded to
// 80 bytes) so the global stores are whole 64-byte rows, 1 KiB per wave and
// instruction.  grid: ((Dp / 64) * (Kp / 64), 4 * nqb)
constexpr int PREP_A_ROW = 80;
__global__ __launch_bounds__(256) void crt_prep_a(CrtArgs a) {
  __shared__ __attribute__((aligned(16))) int8_t T[2][2][64 * PREP_A_ROW];  // [buffer][sign][row][80]
  const int mat = blockIdx.y;
  int r;
  size_t vo, co;
  crt_mat(a, mat, r, vo, co);
  const int D = a.D;
  const double2* G = a.gam[r] + vo;
  const double2* F = a.fv[r] + vo;
  const int nkc = a.Kp / 64;
  const int rb = blockIdx.x / nkc, kc = blockIdx.x - rb * nkc;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row = rb * 64 + lane;
  const bool rok = row < D;
  const int k0 = kc * 64 + wave * 16;
  const double mx = a.rmax[(size_t)mat * a.Dp + row];
  int e = 0;
  if (mx > 0.0) frexp(mx, &e);
  if (kc == 0 && wave == 0) a.escale[(size_t)mat * a.Dp + row] = e;
  const double sc = ldexp(1.0, a.ta - e);
  double r1[16], r0[16], i1[16], i0[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int k = k0 + u;
    double2 v = make_double2(0.0, 0.0);
    if (rok && k < 2 * D) v = ld_stream(k < D ? G + (size_t)k * D + row : F + (size_t)(k - D) * D + row);
    limbs(rint(v.x * sc), r1[u], r0[u]);
    limbs(rint(v.y * sc), i1[u], i0[u]);
  }
  // write-out: thread tid moves 16 bytes (column tid & 3) of row tid / 4 per sign
  const int orow = tid >> 2, och = tid & 3;
  for (int mod = 0; mod < a.nmod; ++mod) {
    const int buf = mod & 1;
    int wp[4] = {0, 0, 0, 0}, wm[4] = {0, 0, 0, 0};
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      int bp, bm;
      gauss_residues(a, mod, r1[u], r0[u], i1[u], i0[u], bp, bm);
      wp[u >> 2] |= bp << (8 * (u & 3));
      wm[u >> 2] |= bm << (8 * (u & 3));
    }
    // buffer buf was last read two moduli ago, before the barrier of the previous one
    *reinterpret_cast<v4i*>(&T[buf][0][lane * PREP_A_ROW + 16 * wave]) = v4i{wp[0], wp[1], wp[2], wp[3]};
    *reinterpret_cast<v4i*>(&T[buf][1][lane * PREP_A_ROW + 16 * wave]) = v4i{wm[0], wm[1], wm[2], wm[3]};
    __syncthreads();
#pragma unroll
    for (int sg = 0; sg < 2; ++sg)
      *reinterpret_cast<v4i*>(a.A + ws_off(a, mat, mod, sg, rb * 64 + orow, kc * 64 + 16 * och)) =
          *reinterpret_cast<const v4i*>(&T[buf][sg][orow * PREP_A_ROW + 16 * och]);
  }
}

// B = [alpha chi F ; alpha chi Gamma] residues, and the row maxima of
// A = [Gamma | F] on the way (every element of Gamma and F passes here).  A
// workgroup owns 16 adjacent columns; wave w does columns 16 blockIdx.x +
// 4 w .. + 3 in turn, independently of the other waves.  With CL the q's
// bubble blocks are first staged in LDS with nu fastest ([l1][l3][nu]), so the
// mixing reads them conflict-free.  Per column the wave forms its 2D mixed
// values with lanes along nu (coalesced reads of F and Gamma) into LDS, raises
// the workgroup's running maxima of |Gamma| and |F| per row (an LDS atomic max
// on the bits of the non-negative doubles, which order as unsigned integers),
// then lane l converts inner indices 4 l + 256 n .. + 3 and stores one dword
// per (modulus, sign).  At the end the row maxima are the block's partial
// rmaxp[mat][blockIdx.x][row].  grid: (Dp / 16, 4 * nqb); dynamic LDS
// 4 Kp 16 + Dp 8 (+ nl^2 nf 16 with CL).
constexpr int PREP_B_CPW = PREP_B_COLS / 4;
template <int NL, bool CL>
__global__ __launch_bounds__(256) void crt_prep_b(CrtArgs a) {
  extern __shared__ double2 colv[];  // [4][Kp] mixed columns, [Dp] row maxima (bits), with CL [l1][l3][nu]
  const int mat = blockIdx.y;
  int r;
  size_t vo, co;
  crt_mat(a, mat, r, vo, co);
  const int D = a.D, nf = a.nf, Kp = a.Kp, Dp = a.Dp;
  constexpr int nl = NL;
  const double2* chig = a.chi[r >> 1] + co;  // [nu][l3][l1]
  unsigned long long* rmx = reinterpret_cast<unsigned long long*>(colv + 4 * (size_t)Kp);  // [Dp]
  double2* chil = reinterpret_cast<double2*>(rmx + (size_t)Dp);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int u = tid; u < Dp; u += 256) rmx[u] = 0ull;
  if (CL) {
    for (int u = tid; u < nl * nl * nf; u += 256) {
      const int nu = u % nf, l = u / nf;  // l = l1 + nl l3
      chil[u] = chig[(size_t)nu * nl * nl + l];
    }
  }
  __syncthreads();
  double2* v = colv + (size_t)wave * Kp;
  const double2* Fq = a.fv[r] + vo;
  const double2* Gq = a.gam[r] + vo;
  // loads from a clamped address, unconditionally (the caller zeroes what lies
  // outside), so the registers are written by the loads alone
  auto load_x = [&](int col, int nu, double2 (&xf)[NL], double2 (&xg)[NL]) {
    const size_t o = (size_t)(col < D ? col : 0) * D + (nu < nf ? nu : 0);
#pragma unroll
    for (int l3 = 0; l3 < NL; ++l3) {
      xf[l3] = Fq[o + l3 * nf];
      xg[l3] = Gq[o + l3 * nf];
    }
  };
  const int col0 = blockIdx.x * PREP_B_COLS + wave * PREP_B_CPW;
  // the first 64 nu of the next column are loaded before the current column's
  // conversion, so their latency hides behind it
  double2 pf[NL], pg[NL];
  load_x(col0, lane, pf, pg);
  int fkeep = 0;  // lane cc keeps column cc's exponent; stored after the loop, so no store in it
                  // holds back the prefetch (a store's data register is waited for before reuse)
  for (int cc = 0; cc < PREP_B_CPW; ++cc) {
    const int col = col0 + cc;
    // B((part, l1, nu), col) = alpha sum_l3 chi_nu(l1, l3) X((l3, nu), col), X = F (part 0), Gamma (part 1)
    double mx = 0.0;
    for (int nu0 = 0; nu0 < nf; nu0 += 64) {
      const int nu = nu0 + lane;
      double2 xf[NL], xg[NL];
      if (nu0 == 0) {
#pragma unroll
        for (int l3 = 0; l3 < NL; ++l3) {
          xf[l3] = pf[l3];
          xg[l3] = pg[l3];
        }
      } else {
        load_x(col, nu, xf, xg);
      }
      if (nu >= nf) continue;
      if (col >= D) {
#pragma unroll
        for (int l3 = 0; l3 < NL; ++l3) xf[l3] = xg[l3] = make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int l3 = 0; l3 < NL; ++l3) {
        const double am = fmax(fmax(fabs(xf[l3].x), fabs(xf[l3].y)), fmax(fabs(xg[l3].x), fabs(xg[l3].y)));
        atomicMax(&rmx[l3 * nf + nu], (unsigned long long)__double_as_longlong(am));
      }
#pragma unroll
      for (int l1 = 0; l1 < NL; ++l1) {
        double fr = 0.0, fi = 0.0, gr = 0.0, gi = 0.0;
#pragma unroll
        for (int l3 = 0; l3 < NL; ++l3) {
          const double2 c = CL ? chil[(size_t)(l1 + nl * l3) * nf + nu] : chig[(size_t)nu * nl * nl + l1 + nl * l3];
          fr = fma(c.x, xf[l3].x, fma(-c.y, xf[l3].y, fr));
          fi = fma(c.x, xf[l3].y, fma(c.y, xf[l3].x, fi));
          gr = fma(c.x, xg[l3].x, fma(-c.y, xg[l3].y, gr));
          gi = fma(c.x, xg[l3].y, fma(c.y, xg[l3].x, gi));
        }
        const double2 mf = make_double2(a.alpha * fr, a.alpha * fi);
        const double2 mg = make_double2(a.alpha * gr, a.alpha * gi);
        v[l1 * nf + nu] = mf;
        v[D + l1 * nf + nu] = mg;
        mx = fmax(mx, fmax(fmax(fabs(mf.x), fabs(mf.y)), fmax(fabs(mg.x), fabs(mg.y))));
      }
    }
    for (int k = 2 * D + lane; k < Kp; k += 64) v[k] = make_double2(0.0, 0.0);
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    int f = 0;
    if (mx > 0.0) frexp(mx, &f);
    const double sc = ldexp(1.0, a.tb - f);
    if (cc + 1 < PREP_B_CPW) load_x(col + 1, lane, pf, pg);
    // the wave's column in LDS is complete (each wave reads only its own; the
    // LDS executes one wave's accesses in order, so only the compiler needs
    // holding back)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
    for (int k0 = 4 * lane; k0 < Kp; k0 += 256) {
      double r1[4], r0[4], i1[4], i0[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double2 x = v[k0 + u];
        limbs(rint(x.x * sc), r1[u], r0[u]);
        limbs(rint(x.y * sc), i1[u], i0[u]);
      }
      for (int mod = 0; mod < a.nmod; ++mod) {
        int wp = 0, wm = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          int bp, bm;
          gauss_residues(a, mod, r1[u], r0[u], i1[u], i0[u], bp, bm);
          wp |= bp << (8 * u);
          wm |= bm << (8 * u);
        }
        *reinterpret_cast<int*>(a.B + ws_off(a, mat, mod, 0, col, k0)) = wp;
        *reinterpret_cast<int*>(a.B + ws_off(a, mat, mod, 1, col, k0)) = wm;
      }
    }
    if (lane == cc) fkeep = f;
    // the next column overwrites v only after every lane has read it
    asm volatile("" ::: "memory");
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
  }
  if (lane < PREP_B_CPW) a.fscale[(size_t)mat * Dp + col0 + lane] = fkeep;
  __syncthreads();
  for (int row = tid; row < Dp; row += 256)
    a.rmaxp[((size_t)mat * (Dp / PREP_B_COLS) + blockIdx.x) * Dp + row] = __longlong_as_double((long long)rmx[row]);
}

// The residues of one modulus (INT32 sums of the + and - products of a 32 x 32
// wave tile, 2 x 2 blocks) join the CRT sums, held per element as exact high
// and rounded low parts; the accumulators restart at zero.
__device__ __forceinline__ void crt_join(const CrtArgs& a, int mod, v4i (&accp)[2][2], v4i (&accm)[2][2],
                                         double (&shr)[16], double (&slr)[16], double (&shi)[16],
                                         double (&sli)[16]) {
  const double m = a.m[mod], im = a.inv_m[mod];
  const double wrh = a.wrh[mod], wrl = a.wrl[mod], wih = a.wih[mod], wil = a.wil[mod];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int x = 8 * i + 4 * j + e;
        const int p = accp[i][j][e], n = accm[i][j][e];
        const double ur = sym_mod((double)(p + n), m, im);
        const double ui = sym_mod((double)(p - n), m, im);
        shr[x] = fma(ur, wrh, shr[x]);
        slr[x] = fma(ur, wrl, slr[x]);
        shi[x] = fma(ui, wih, shi[x]);
        sli[x] = fma(ui, wil, sli[x]);
      }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) accp[i][j] = accm[i][j] = v4i{0, 0, 0, 0};
}

// Epilogue of a 32 x 32 wave tile at (r0, c0) of matrix mat.
// C' = (S_hi - q M_hi) + (S_lo - q M_lo); T = C' 2^(e_i - ta + f_j - tb);
// Phi = c f T + (1 - f)(F - Gamma).  Element e of lane l of a 16 x 16 INT32
// block sits at row 4 (l / 16) + e, column l % 16.
__device__ __forceinline__ void crt_tile_epilogue(const CrtArgs& a, int mat, int r0, int c0, int li, int lk,
                                                  const double (&shr)[16], const double (&slr)[16],
                                                  const double (&shi)[16], const double (&sli)[16]) {
  const int D = a.D;
  int r;
  size_t vo, co;
  crt_mat(a, mat, r, vo, co);
  const double2* Gm = a.gam[r] + vo;
  const double2* Fm = a.fv[r] + vo;
  double2* O = a.out[r] + vo;
  const double cf = a.cf[r];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gj = c0 + j * 16 + li;
      if (gj >= D) continue;
      const int fj = a.fscale[(size_t)mat * a.Dp + gj];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int gi = r0 + i * 16 + 4 * lk + e;
        if (gi >= D) continue;
        const int x = 8 * i + 4 * j + e;
        const double qr = rint((shr[x] + slr[x]) * a.inv_mm);
        const double qi = rint((shi[x] + sli[x]) * a.inv_mm);
        const double cr = fma(-qr, a.mhi, shr[x]) + fma(-qr, a.mlo, slr[x]);
        const double ci = fma(-qi, a.mhi, shi[x]) + fma(-qi, a.mlo, sli[x]);
        const double s = ldexp(cf, a.escale[(size_t)mat * a.Dp + gi] - a.ta + fj - a.tb);
        const size_t idx = (size_t)gj * D + gi;
        const double2 fv = Fm[idx], gv = Gm[idx];
        O[idx] = make_double2(fma(s, cr, a.omf * (fv.x - gv.x)), fma(s, ci, a.omf * (fv.y - gv.y)));
      }
    }
}

// One workgroup per 64 x 64 tile of one (q, r); four waves, each 32 x 32 =
// 2 x 2 blocks of v_mfma_i32_16x16x64_i8.  The loop runs over (modulus, K chunk)
// steps through a ring of four LDS stages (A+, A-, B+, B- chunks of 64 rows x
// 64 bytes, 16 KiB; the 16-byte columns of each row XOR-swizzled by crt_swz so
// the fragment reads are free of bank conflicts).  Each loop iteration computes
// two steps from two stages, then stores the two steps that arrived in
// registers during it into the two free stages and issues the loads of the two
// after them: every load has a full iteration to arrive, and all waits come
// before the next loads in program order.  After a modulus' last chunk its
// residues join the CRT sums, held per output element as exact high and
// rounded low parts.
__global__ __launch_bounds__(256, 2) void crt_gemm(CrtArgs a) {
  __shared__ __attribute__((aligned(16))) int8_t S[4][4][CRT_TILE][64];  // [stage][piece][row][byte]
  const int tiles = a.tiles_n * a.tiles_n;
  // workgroups go round-robin over the 8 XCDs: give each XCD a contiguous run
  // of logical tiles, so all tiles of one matrix share that XCD's L2
  const int bid = xcd_contiguous(blockIdx.x, gridDim.x);
  const int mat = bid / tiles;
  const int t = bid - mat * tiles;
  const int row0 = (t / a.tiles_n) * CRT_TILE, col0 = (t % a.tiles_n) * CRT_TILE;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const int nk = a.Kp / 64;
  const int nsteps = a.nmod * nk;
  // staging: thread tid moves 16 bytes (column tid & 3) of row tid >> 2 of each
  // piece v = 2 * which + sign (which 0: A rows, 1: B columns)
  const int srow = tid >> 2, sch = tid & 3;
  const int sput = 16 * (sch ^ crt_swz(srow));
  auto gptr = [&](int step, int v) -> const v4i* {
    const int md = step / nk, k0 = (step - md * nk) * 64;
    const int which = v >> 1, sign = v & 1;
    const int8_t* base = which ? a.B : a.A;
    const int rc = (which ? col0 : row0) + srow;
    return reinterpret_cast<const v4i*>(base + ws_off(a, mat, md, sign, rc, k0 + 16 * sch));
  };
  auto load = [&](int step, v4i* r) {
#pragma unroll
    for (int v = 0; v < 4; ++v) r[v] = *gptr(step, v);
  };
  auto put = [&](int st, const v4i* r) {
#pragma unroll
    for (int v = 0; v < 4; ++v) *reinterpret_cast<v4i*>(&S[st][v][srow][sput]) = r[v];
  };
  auto frag = [&](int st, int v, int row) -> v4i {
    return *reinterpret_cast<const v4i*>(&S[st][v][row][16 * (lk ^ crt_swz(row))]);
  };
  double shr[16], slr[16], shi[16], sli[16];  // CRT sums, element (i, j, e) -> 8 i + 4 j + e
#pragma unroll
  for (int x = 0; x < 16; ++x) shr[x] = slr[x] = shi[x] = sli[x] = 0.0;
  v4i accp[2][2], accm[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) accp[i][j] = accm[i][j] = v4i{0, 0, 0, 0};

  int kstep = 0, mod = 0;
  auto compute = [&](int st) {
    v4i ap[2], am[2], bp[2], bm[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      ap[i] = frag(st, 0, wr + i * 16 + li);
      am[i] = frag(st, 1, wr + i * 16 + li);
      bp[i] = frag(st, 2, wc + i * 16 + li);
      bm[i] = frag(st, 3, wc + i * 16 + li);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        accp[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ap[i], bp[j], accp[i][j], 0, 0, 0);
        accm[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(am[i], bm[j], accm[i][j], 0, 0, 0);
      }
    if (++kstep == nk) {
      crt_join(a, mod, accp, accm, shr, slr, shi, sli);
      kstep = 0;
      ++mod;
    }
  };

  // prologue: steps 0 and 1 into stages 0 and 1; steps 2 and 3 in flight
  v4i R0[4], R1[4];
  load(0, R0);
  if (nsteps > 1) load(1, R1);
  put(0, R0);
  if (nsteps > 1) put(1, R1);
  if (nsteps > 2) load(2, R0);
  if (nsteps > 3) load(3, R1);
  __syncthreads();
  for (int s = 0; s < nsteps; s += 2) {
    compute(s & 3);
    if (s + 1 < nsteps) compute((s + 1) & 3);
    // stages (s + 2) & 3 and (s + 3) & 3 were last read two iterations ago
    if (s + 2 < nsteps) put((s + 2) & 3, R0);
    if (s + 3 < nsteps) put((s + 3) & 3, R1);
    if (s + 4 < nsteps) load(s + 4, R0);
    if (s + 5 < nsteps) load(s + 5, R1);
    __syncthreads();
  }

  crt_tile_epilogue(a, mat, row0 + wr, col0 + wc, li, lk, shr, slr, shi, sli);
}

// ---- 64 x 64 tile, stages filled by direct global -> LDS loads ----------------------
//
// Same tile, waves, swizzle and CRT sums as crt_gemm; the (modulus, K chunk)
// steps go through a ring of NS LDS stages (16 KiB each: A+, A-, B+, B- of 64
// rows x 64 bytes) filled by global_load_lds_dwordx4, so no registers stage
// the data and NS - 1 steps are in flight while one is computed.  The 16
// bytes of lane l land at its wave's base (M0) + 16 l, i.e. at row tid / 4,
// slot tid % 4 of the piece, so the lane loads the 16-byte column that the
// XOR swizzle puts in that slot.  One barrier per step: each thread first
// waits for its own loads of the step (vmcnt counts them in issue order), the
// barrier then makes every wave's loads visible, and the stage refilled after
// the barrier is the one computed in the step before, which every wave has
// finished reading (its MFMAs consumed the fragments before the barrier).
// The loads are issued from asm, so the compiler, which cannot tell the
// ring's stages apart across loop iterations, adds no vmcnt(0) before the
// fragment reads; its own waits for other loads only grow more conservative
// (dma16, wait_vmcnt, stage_barrier: lds_dma.h).
typedef int8_t CrtStage[4][CRT_TILE][64];

template <int NS>
__global__ __launch_bounds__(256, 2) void crt_gemm_dma(CrtArgs a) {
  static_assert(NS >= 3 && NS <= 5, "stages");
  // one LDS object per stage, so the compiler sees which stage a read touches
  __shared__ __attribute__((aligned(16))) CrtStage S0;
  __shared__ __attribute__((aligned(16))) CrtStage S1;
  __shared__ __attribute__((aligned(16))) CrtStage S2;
  __shared__ __attribute__((aligned(16))) CrtStage S3;
  __shared__ __attribute__((aligned(16))) CrtStage S4;
  const int tiles = a.tiles_n * a.tiles_n;
  const int bid = xcd_contiguous(blockIdx.x, gridDim.x);
  const int mat = bid / tiles;
  const int t = bid - mat * tiles;
  const int row0 = (t / a.tiles_n) * CRT_TILE, col0 = (t % a.tiles_n) * CRT_TILE;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const int nk = a.Kp / 64;
  const int nsteps = a.nmod * nk;
  const int drow = tid >> 2;
  const int dch = (tid & 3) ^ crt_swz(drow);  // global 16-byte column landing in slot tid & 3
  // the byte offsets of ws_off split into a per-lane part (row drow, column
  // dch of a 64-row block) and a uniform part, so a step's four loads need no
  // vector address arithmetic
  const unsigned voff = (unsigned)(drow * 64 + 16 * dch);
  const size_t plane = (size_t)a.Dp * a.Kp;
  const int8_t* baseA = a.A + (size_t)mat * a.nmod * 2 * plane + (size_t)row0 * 64;
  const int8_t* baseB = a.B + (size_t)mat * a.nmod * 2 * plane + (size_t)col0 * 64;
  auto issue = [&](int step, CrtStage& S) {
    const int md = step / nk, kc = step - md * nk;
    const size_t off = (size_t)md * 2 * plane + (size_t)kc * a.Dp * 64;
    // pieces A+, A-, B+, B- are 4 KiB apart in the stage
    const unsigned lds0 = lds_addr(&S[0][wave * 16][0]);
    dma16(baseA + off, voff, lds0);
    dma16(baseA + off + plane, voff, lds0 + 4096);
    dma16(baseB + off, voff, lds0 + 8192);
    dma16(baseB + off + plane, voff, lds0 + 12288);
  };
  double shr[16], slr[16], shi[16], sli[16];  // CRT sums, element (i, j, e) -> 8 i + 4 j + e
#pragma unroll
  for (int x = 0; x < 16; ++x) shr[x] = slr[x] = shi[x] = sli[x] = 0.0;
  v4i accp[2][2], accm[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) accp[i][j] = accm[i][j] = v4i{0, 0, 0, 0};
  int kstep = 0, mod = 0;
  auto compute = [&](const CrtStage& S) {
    auto frag = [&](int v, int row) -> v4i {
      return *reinterpret_cast<const v4i*>(&S[v][row][16 * (lk ^ crt_swz(row))]);
    };
    v4i ap[2], am[2], bp[2], bm[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      ap[i] = frag(0, wr + i * 16 + li);
      am[i] = frag(1, wr + i * 16 + li);
      bp[i] = frag(2, wc + i * 16 + li);
      bm[i] = frag(3, wc + i * 16 + li);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        accp[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ap[i], bp[j], accp[i][j], 0, 0, 0);
        accm[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(am[i], bm[j], accm[i][j], 0, 0, 0);
      }
    if (++kstep == nk) {
      crt_join(a, mod, accp, accm, shr, slr, shi, sli);
      kstep = 0;
      ++mod;
    }
  };
  auto stage = [&](auto ic) -> CrtStage& {
    constexpr int i = decltype(ic)::value % NS;
    if constexpr (i == 0) return S0;
    else if constexpr (i == 1) return S1;
    else if constexpr (i == 2) return S2;
    else if constexpr (i == 3) return S3;
    else return S4;
  };
  // one step s held in stage I (= s mod NS)
  auto step = [&](auto ic, int s) {
    constexpr int I = decltype(ic)::value;
    if (s >= nsteps) return;
    if (s + NS - 2 < nsteps)
      wait_vmcnt<4 * (NS - 2)>();  // the 4 (NS - 2) loads after this step's may still fly
    else
      wait_vmcnt<0>();
    stage_barrier();
    if (s + NS - 1 < nsteps) issue(s + NS - 1, stage(std::integral_constant<int, I + NS - 1>{}));
    compute(stage(ic));
  };
  issue(0, S0);
  if (1 < nsteps) issue(1, S1);
  if (NS > 3 && 2 < nsteps) issue(2, S2);
  if (NS > 4 && 3 < nsteps) issue(3, S3);
  for (int s0 = 0; s0 < nsteps; s0 += NS) {
    step(std::integral_constant<int, 0>{}, s0);
    step(std::integral_constant<int, 1>{}, s0 + 1);
    step(std::integral_constant<int, 2>{}, s0 + 2);
    if constexpr (NS > 3) step(std::integral_constant<int, 3>{}, s0 + 3);
    if constexpr (NS > 4) step(std::integral_constant<int, 4>{}, s0 + 4);
  }
  crt_tile_epilogue(a, mat, row0 + wr, col0 + wc, li, lk, shr, slr, shi, sli);
}

// ---- large-tile CRT GEMM --------------------------------------------------------------
//
// Workgroup tile T = 32 TB rows x 32 TB columns; four waves 2 x 2, each a
// 16 TB x 16 TB tile = TB x TB blocks of v_mfma_i32_16x16x64_i8, so a wave
// reads 2 TB fragments from LDS per TB^2 MFMAs (the 64 x 64 kernel above reads
// 4 per 4, which leaves it bound by LDS bandwidth).  The steps run over
// (modulus, sign, 64-byte K chunk): a step stages the T rows of A_sign and the
// T columns of B_sign (2 x T x 64 bytes, 16-byte columns XOR-swizzled by
// crt_swz) into one of two LDS stages, from registers loaded three steps
// earlier.  The INT32 products of the + sign are reduced to residues and kept
// packed as bytes; after the - sign each element's (Re, Im) residue pair goes
// to a global scratch (two bytes per element and modulus) instead of CRT sums
// held in registers, and the CRT, scaling and damping run once at the end over
// the scratch, a pair of elements at a time.
template <int TB>
struct CrtBig {
  static constexpr int WT = 16 * TB;          // wave tile
  static constexpr int T = 2 * WT;            // workgroup tile
  static constexpr int VPT = T / 32;          // 16-byte pieces per thread per step (2 T rows x 4 / 256)
  static constexpr int DEPTH = 3;             // steps in flight in registers
  static constexpr int DW = 2 * TB * TB;      // scratch dwords per lane per modulus
};

template <int TB>
__global__ __launch_bounds__(256, 1) void crt_gemm_big(CrtArgs a) {
  using P = CrtBig<TB>;
  constexpr int T = P::T, WT = P::WT, VPT = P::VPT, DEPTH = P::DEPTH, DW = P::DW;
  __shared__ __attribute__((aligned(16))) int8_t S[2][2][T][64];  // [stage][A|B][row][byte]
  const int tiles = a.tiles_n * a.tiles_n;
  const int bid = xcd_contiguous(blockIdx.x, gridDim.x);
  const int mat = bid / tiles;
  const int t = bid - mat * tiles;
  const int row0 = (t / a.tiles_n) * T, col0 = (t % a.tiles_n) * T;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * WT, wc = (wave & 1) * WT;
  const int li = lane & 15, lk = lane >> 4;
  const int nk = a.Kp / 64;
  const int nsteps = a.nmod * 2 * nk;

  // piece v of a thread: u = tid + 256 v -> (A|B, row, 16-byte column).  Its
  // workspace address is a per-thread base plus a step offset common to A and
  // B ((modulus, sign, K chunk) planes have the same strides in both); both are
  // formed once, so the loop issues loads with no address arithmetic that
  // could make it wait on earlier loads
  const int8_t* pbase[VPT];
  int lofs[VPT];
#pragma unroll
  for (int v = 0; v < VPT; ++v) {
    const int u = tid + 256 * v;
    const int which = u / (4 * T), row = (u >> 2) % T, ch = u & 3;
    const size_t off = ws_off(a, mat, 0, 0, (which ? col0 : row0) + row, 16 * ch);
    pbase[v] = (which ? a.B : a.A) + off;
    lofs[v] = (which * T + row) * 64 + 16 * (ch ^ crt_swz(row));
  }
  const size_t plane = (size_t)a.Dp * a.Kp;
  auto step_off = [&](int step) __attribute__((always_inline)) -> size_t {
    const int md = step / (2 * nk), rem = step - md * 2 * nk;
    const int sign = rem / nk, kc = rem - sign * nk;
    return ((size_t)md * 2 + sign) * plane + (size_t)kc * a.Dp * 64;
  };
  auto load = [&](int step, v4i* r) __attribute__((always_inline)) {
    const size_t o = step_off(step);
#pragma unroll
    for (int v = 0; v < VPT; ++v) r[v] = *reinterpret_cast<const v4i*>(pbase[v] + o);
  };
  auto put = [&](int st, const v4i* r) __attribute__((always_inline)) {
#pragma unroll
    for (int v = 0; v < VPT; ++v) *reinterpret_cast<v4i*>(&S[st][0][0][0] + lofs[v]) = r[v];
  };
  auto frag = [&](int st, int which, int row) __attribute__((always_inline)) -> v4i {
    return *reinterpret_cast<const v4i*>(&S[st][which][row][16 * (lk ^ crt_swz(row))]);
  };

  v4i acc[TB][TB];
  int pk[TB][TB];  // residues of the + product, 4 bytes (e = 0..3) per block
#pragma unroll
  for (int i = 0; i < TB; ++i)
#pragma unroll
    for (int j = 0; j < TB; ++j) {
      acc[i][j] = v4i{0, 0, 0, 0};
      pk[i][j] = 0;
    }
  int* scr = a.crt + (size_t)bid * a.nmod * DW * 256 + tid;  // dword w of modulus md: scr[(md DW + w) 256]

  auto compute = [&](int st, int step) __attribute__((always_inline)) {
    v4i bf[TB];
#pragma unroll
    for (int j = 0; j < TB; ++j) bf[j] = frag(st, 1, wc + 16 * j + li);
#pragma unroll
    for (int i = 0; i < TB; ++i) {
      const v4i af = frag(st, 0, wr + 16 * i + li);
#pragma unroll
      for (int j = 0; j < TB; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, bf[j], acc[i][j], 0, 0, 0);
    }
    const int md = step / (2 * nk), rem = step - md * 2 * nk;
    if (rem == nk - 1) {  // + product complete: keep its residues
      const double im = a.inv_m[md];
      const unsigned nm = a.nm24[md];
#pragma unroll
      for (int i = 0; i < TB; ++i)
#pragma unroll
        for (int j = 0; j < TB; ++j) {
          pk[i][j] = pack4(sym_mod_lo(acc[i][j][0], im, nm), sym_mod_lo(acc[i][j][1], im, nm),
                           sym_mod_lo(acc[i][j][2], im, nm), sym_mod_lo(acc[i][j][3], im, nm));
          acc[i][j] = v4i{0, 0, 0, 0};
        }
    } else if (rem == 2 * nk - 1) {  // - product complete: residues of P+ + P-, P+ - P- to the scratch
      const double im = a.inv_m[md];
      const unsigned nm = a.nm24[md];
#pragma unroll
      for (int i = 0; i < TB; ++i)
#pragma unroll
        for (int j = 0; j < TB; ++j) {
          int r[4], d[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int p = (int)(signed char)(pk[i][j] >> (8 * e)), n = acc[i][j][e];
            r[e] = sym_mod_lo(p + n, im, nm);
            d[e] = sym_mod_lo(p - n, im, nm);
          }
          // dword 0: (Re, Im) bytes of elements 0, 1; dword 1: of elements 2, 3
          const int w0 = pack4(r[0], d[0], r[1], d[1]);
          const int w1 = pack4(r[2], d[2], r[3], d[3]);
          const int x = i * TB + j;  // block index; dwords 2x, 2x + 1 hold elements e = 0,1 and 2,3
          scr[((size_t)md * DW + 2 * x) * 256] = w0;
          scr[((size_t)md * DW + 2 * x + 1) * 256] = w1;
          acc[i][j] = v4i{0, 0, 0, 0};
        }
    }
  };

  // prologue: steps 0 and 1 staged in LDS; steps 2, 3, 4 in flight in registers
  // (nsteps >= 16: at least 8 moduli x 2 signs).  Loads past the last step
  // re-read the last step, so the loop issues the same loads every time.
  v4i R[DEPTH][VPT];
  load(0, R[0]);
  put(0, R[0]);
  load(1, R[1]);
  put(1, R[1]);
#pragma unroll
  for (int d = 0; d < DEPTH; ++d) load(2 + d, R[(2 + d) % DEPTH]);
  __syncthreads();
  // step s (s % DEPTH = d, a compile-time register set): compute from stage
  // s & 1; after the barrier that stage takes step s + 2 from registers, whose
  // set then loads step s + 2 + DEPTH
  auto step_body = [&](int s, auto dc) __attribute__((always_inline)) {
    constexpr int d = decltype(dc)::value;
    compute(s & 1, s);
    __syncthreads();
    put(s & 1, R[(d + 2) % DEPTH]);
    load(min(s + 2 + DEPTH, nsteps - 1), R[(d + 2) % DEPTH]);
  };
  int s0 = 0;
  for (; s0 + DEPTH <= nsteps; s0 += DEPTH) {
    step_body(s0, std::integral_constant<int, 0>{});
    step_body(s0 + 1, std::integral_constant<int, 1>{});
    step_body(s0 + 2, std::integral_constant<int, 2>{});
  }
  static_assert(DEPTH == 3, "the loop above is unrolled for three register sets");
  if (s0 < nsteps) step_body(s0, std::integral_constant<int, 0>{});
  if (s0 + 1 < nsteps) step_body(s0 + 1, std::integral_constant<int, 1>{});

  // CRT over the scratch, two elements (one dword) at a time; element e of
  // block (i, j) sits at row 4 (lane / 16) + e, column lane % 16
  const int D = a.D;
  int r;
  size_t vo, co;
  crt_mat(a, mat, r, vo, co);
  const double2* Gm = a.gam[r] + vo;
  const double2* Fm = a.fv[r] + vo;
  double2* O = a.out[r] + vo;
  const double cf = a.cf[r];
  // groups of TB dwords (2 TB elements); the scratch is read four moduli at a
  // time, all 4 TB loads of a chunk issued before any is used (a chunk past
  // nmod re-reads the last modulus with weight 0)
  for (int w0 = 0; w0 < DW; w0 += TB) {
    double shr[TB][2], slr[TB][2], shi[TB][2], sli[TB][2];
#pragma unroll
    for (int g = 0; g < TB; ++g)
#pragma unroll
      for (int h = 0; h < 2; ++h) shr[g][h] = slr[g][h] = shi[g][h] = sli[g][h] = 0.0;
    for (int m0 = 0; m0 < a.nmod; m0 += 4) {
      int dv[4][TB];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int mc = min(m0 + k, a.nmod - 1);
#pragma unroll
        for (int g = 0; g < TB; ++g) dv[k][g] = scr[((size_t)mc * DW + w0 + g) * 256];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const bool live = m0 + k < a.nmod;
        const int mk = min(m0 + k, CRT_MAX - 1);
        const double wrh = live ? a.wrh[mk] : 0.0, wrl = live ? a.wrl[mk] : 0.0;
        const double wih = live ? a.wih[mk] : 0.0, wil = live ? a.wil[mk] : 0.0;
#pragma unroll
        for (int g = 0; g < TB; ++g)
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int d = dv[k][g];
            const double ur = (double)(signed char)(d >> (16 * h)), ui = (double)(signed char)(d >> (16 * h + 8));
            shr[g][h] = fma(ur, wrh, shr[g][h]);
            slr[g][h] = fma(ur, wrl, slr[g][h]);
            shi[g][h] = fma(ui, wih, shi[g][h]);
            sli[g][h] = fma(ui, wil, sli[g][h]);
          }
      }
    }
#pragma unroll
    for (int g = 0; g < TB; ++g) {
      const int w = w0 + g;
      const int x = w >> 1, i = x / TB, j = x - i * TB, e0 = 2 * (w & 1);
      const int gj = col0 + wc + 16 * j + li;
      const int gi0 = row0 + wr + 16 * i + 4 * lk + e0;
      if (gj >= D) continue;
      const int fj = a.fscale[(size_t)mat * a.Dp + gj];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int gi = gi0 + h;
        if (gi >= D) continue;
        const double qr = rint((shr[g][h] + slr[g][h]) * a.inv_mm);
        const double qi = rint((shi[g][h] + sli[g][h]) * a.inv_mm);
        const double cr = fma(-qr, a.mhi, shr[g][h]) + fma(-qr, a.mlo, slr[g][h]);
        const double ci = fma(-qi, a.mhi, shi[g][h]) + fma(-qi, a.mlo, sli[g][h]);
        const double s = ldexp(cf, a.escale[(size_t)mat * a.Dp + gi] - a.ta + fj - a.tb);
        const size_t idx = (size_t)gj * D + gi;
        const double2 fv = Fm[idx], gv = Gm[idx];
        O[idx] = make_double2(fma(s, cr, a.omf * (fv.x - gv.x)), fma(s, ci, a.omf * (fv.y - gv.y)));
      }
    }
  }
}

// ---- host: moduli constants ---------------------------------------------------------

typedef unsigned __int128 u128;

static int bitlen(u128 x) {
  int n = 0;
  while (x) {
    ++n;
    x >>= 1;
  }
  return n;
}

// nearest double (ties away) of an integer below 2^128
static double u128_to_double(u128 x) {
  const int n = bitlen(x);
  if (n <= 53) return (double)(unsigned long long)x;
  const int sh = n - 53;
  u128 top = x >> sh;
  const u128 rem = x - (top << sh);
  if (rem >= ((u128)1 << (sh - 1))) top += 1;
  return std::ldexp((double)(unsigned long long)top, sh);
}

static long long inv_mod(long long x, long long m) {  // x^-1 mod m, gcd(x, m) = 1
  const long long v = ((x % m) + m) % m;
  for (long long y = 1; y < m; ++y)
    if (v * y % m == 1) return y;
  return 0;
}

static void crt_constants(CrtArgs& a, int nmod, int D) {
  a.nmod = nmod;
  u128 M = 1;
  for (int i = 0; i < nmod; ++i) M *= (u128)kModuli[i];
  const int L = std::max(0, bitlen(M) - 41);
  for (int i = 0; i < nmod; ++i) {
    const long long m = kModuli[i];
    long long j = 1;
    while ((j * j + 1) % m) ++j;
    const u128 Mi = M / (u128)m;
    const long long yi = inv_mod((long long)(Mi % (u128)m), m);
    a.m[i] = (double)m;
    a.inv_m[i] = 1.0 / (double)m;
    a.nm24[i] = (unsigned)((1u << 24) - (unsigned)m);
    a.j[i] = (double)j;
    a.c26[i] = (double)(((long long)1 << 26) % m);
    // residue u of Re (Im) = residue of (P+ + P-) (of P+ - P-) times 1/2 (1/(2j)) mod m;
    // its CRT term u (M/m) yi is the residue times Wr (Wi) below, all < M
    const long long h2 = (m + 1) / 2, h2j = inv_mod(2 * j, m);
    const u128 Wr = Mi * (u128)((h2 * yi) % m), Wi = Mi * (u128)((h2j * yi) % m);
    a.wrh[i] = std::ldexp((double)(unsigned long long)(Wr >> L), L);  // < 2^41 significant
    a.wrl[i] = u128_to_double(Wr - ((Wr >> L) << L));
    a.wih[i] = std::ldexp((double)(unsigned long long)(Wi >> L), L);
    a.wil[i] = u128_to_double(Wi - ((Wi >> L) << L));
  }
  a.mhi = std::ldexp((double)(unsigned long long)(M >> L), L);
  a.mlo = u128_to_double(M - ((M >> L) << L));
  a.inv_mm = 1.0 / u128_to_double(M);
  // |C'| <= 2 (2D) 2^(ta + tb) <= M / 4
  const double log2M = std::log2(u128_to_double(M));
  const int tsum = (int)std::floor(log2M - std::log2(8.0 * D) - 1.0);
  a.ta = std::min(52, tsum / 2);
  a.tb = std::min(52, tsum - tsum / 2);
}

// GEMM tile for an order D: 6 selects crt_gemm_big<6> (192 x 192 tiles) when
// padding D to 192 costs at most 7% over the 64-row padding, else 0 (the
// 64 x 64 kernel).  Measured per 256 q: D = 576 24.2 ms (big) vs 28.7 ms.  The
// autotuner (TUPS_K_SPLIT_GEMM) times every tile on the resident inputs and
// may pick another.  The padded order Dp is a multiple of the tile and of 64.
static int crt_big_tb(int D) {
  const long long d64 = (D + 63) / 64 * 64;
  const long long L = std::lcm(32 * 6, 64);
  return (D + L - 1) / L * L * 100 <= d64 * 107 ? 6 : 0;
}

// tile variants: 0 = the 64 x 64 kernel, else crt_gemm_big<tb> (32 tb x 32 tb tiles)
struct SplitVariant {
  int tb;  // 0: 64 x 64 workgroup tile; 5, 6: crt_gemm_big<tb>
  int ns;  // 64 x 64 tile: 0 register-staged (crt_gemm), else crt_gemm_dma<ns>
  const char* name;
};
static const SplitVariant kSplitVariants[] = {{0, 0, "crt_64x64"},     {5, 0, "crt_160x160"},
                                              {6, 0, "crt_192x192"},   {0, 4, "crt_64x64_dma4"},
                                              {0, 5, "crt_64x64_dma5"}};
constexpr int kNumSplitVariants = sizeof(kSplitVariants) / sizeof(kSplitVariants[0]);

int split_variant_count() { return kNumSplitVariants; }
const char* split_variant_name(int v) { return v >= 0 && v < kNumSplitVariants ? kSplitVariants[v].name : ""; }

static int split_tb(const tups_ctx* c) {
  const int v = c->split_variant;
  return v >= 0 && v < kNumSplitVariants ? kSplitVariants[v].tb : crt_big_tb(c->sz.nz);
}

// default for the 64 x 64 tile: four DMA-filled stages (measured 13% faster
// than the register-staged crt_gemm at D = 320)
static int split_ns(const tups_ctx* c) {
  const int v = c->split_variant;
  return v >= 0 && v < kNumSplitVariants ? kSplitVariants[v].ns : 4;
}

static size_t crt_padded(int D, int tb) {
  const size_t L = tb > 0 ? std::lcm(32 * tb, 64) : 64;
  return ((size_t)D + L - 1) / L * L;
}

// scratch bytes per matrix of the large-tile GEMM
static size_t crt_scratch_per_mat(int D, int nmod, int tb) {
  if (!tb) return 0;
  const size_t tn = crt_padded(D, tb) / (32 * tb);
  return tn * tn * (size_t)nmod * (2 * tb * tb) * 256 * sizeof(int);
}

// prep_b for Nl, with the bubble blocks staged in LDS when they fit beside the
// four columns and the row maxima (160 KiB per workgroup); *lds receives the
// dynamic LDS bytes
static void (*prep_b_kernel(int nl, size_t Kp, size_t Dp, int nf, size_t* lds))(CrtArgs) {
  const size_t cols = 4 * Kp * sizeof(double2) + Dp * sizeof(double), chi = (size_t)nl * nl * nf * sizeof(double2);
  const bool cl = cols + chi <= 160 * 1024;
  *lds = cols + (cl ? chi : 0);
  if (nl == 1) return cl ? crt_prep_b<1, true> : crt_prep_b<1, false>;
  if (nl == 5) return cl ? crt_prep_b<5, true> : crt_prep_b<5, false>;
  return cl ? crt_prep_b<9, true> : crt_prep_b<9, false>;
}

// workspace bytes of one matrix: residues of A and B, exponents, row maxima and
// their partials, large-tile scratch
static size_t crt_bytes_per_mat(int D, int nmod, int tb) {
  const size_t Dp = crt_padded(D, tb);
  const size_t Kp = (2 * (size_t)D + 63) / 64 * 64;
  return 4 * (size_t)nmod * Dp * Kp + 2 * Dp * sizeof(int) + (Dp / PREP_B_COLS + 1) * Dp * sizeof(double) +
         crt_scratch_per_mat(D, nmod, tb) + 16;
}

// a pass of `mats` matrices laid out in the workspace at base
static void crt_workspace(CrtArgs& a, void* base_v, size_t mats, int nmod, size_t Dp, size_t Kp) {
  int8_t* base = static_cast<int8_t*>(base_v);
  a.A = base;
  a.B = base + mats * 2 * nmod * Dp * Kp;
  a.rmax = reinterpret_cast<double*>(a.B + mats * 2 * nmod * Dp * Kp);  // 8-byte aligned: Kp % 64 == 0
  a.rmaxp = a.rmax + mats * Dp;
  a.escale = reinterpret_cast<int*>(a.rmaxp + mats * (Dp / PREP_B_COLS) * Dp);
  a.fscale = a.escale + mats * Dp;
  a.crt = reinterpret_cast<int*>((reinterpret_cast<uintptr_t>(a.fscale + mats * Dp) + 15) / 16 * 16);
}

// q points per pass and workspace of the split ladder; allocates (outside any
// stream capture) when the workspace is too small
int bse_split_reserve(tups_ctx* c, int* qbatch_out) {
  const tups_sizes& s = c->sz;
  const int nmod = c->split_moduli;
  TUPS_CHECK(nmod >= 8 && nmod <= CRT_MAX, TUPS_EINVAL, "split moduli must be in 8..16");
  TUPS_CHECK(s.nz <= 1024, TUPS_EINVAL, "split path needs Nz <= 1024 (exact INT32 sums)");
  const int tb = split_tb(c);
  const size_t Dp = crt_padded(s.nz, tb);
  const size_t Kp = (2 * (size_t)s.nz + 63) / 64 * 64;
  // per q point: 4 channels x (A and B) x 2 signs x nmod residue planes + exponents + GEMM scratch
  const size_t per_q = 4 * crt_bytes_per_mat(s.nz, nmod, tb);
  // default pass: about 4 GiB of residues, at most an eighth of the free memory
  // (fewer, larger passes measured 3-5% faster than 1 GiB ones); fixed at
  // first use, so a later call (possibly under stream capture) finds the same
  // pass size and the workspace already in place
  if (c->split_qbatch == 0 && c->split_auto_q == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
    const size_t budget =
        std::max<size_t>(per_q, std::min<size_t>((size_t)4 << 30, (free_b + c->split_ws_bytes) / 8));
    c->split_auto_q = (int)std::max<size_t>(1, std::min<size_t>(s.nq, budget / per_q));
  }
  const size_t want = c->split_qbatch > 0 ? (size_t)c->split_qbatch : (size_t)c->split_auto_q;
  const int qbatch = (int)std::max<size_t>(1, std::min<size_t>(s.nq, want));
  const size_t need = per_q * qbatch;
  if (c->split_ws_bytes < need) {
    drop_graphs(c);  // captured graphs hold the old workspace
    if (c->split_ws) TUPS_HIP(hipFree(c->split_ws));
    c->split_ws = nullptr;
    c->split_ws_bytes = 0;
    TUPS_HIP(hipMalloc(&c->split_ws, need));
    c->split_ws_bytes = need;
  }
  size_t lds_b = 0;
  void (*prep_b)(CrtArgs) = prep_b_kernel(s.nl, Kp, Dp, s.nf, &lds_b);
  if (lds_b > 65536)
    TUPS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(prep_b), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds_b));
  if (qbatch_out) *qbatch_out = qbatch;
  return TUPS_OK;
}

int launch_bse_split(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  const int nmod = c->split_moduli;
  int qbatch = 0;
  if (int rc = bse_split_reserve(c, &qbatch)) return rc;
  const int tb = split_tb(c);
  const size_t Dp = crt_padded(s.nz, tb);
  const size_t Kp = (2 * (size_t)s.nz + 63) / 64 * 64;
  CrtArgs a{};
  for (int r = 0; r < 4; ++r) {
    a.gam[r] = c->gamma[r];
    a.fv[r] = c->fvert[r];
    a.out[r] = c->spare[r];
  }
  a.chi[0] = c->chi[0];
  a.chi[1] = c->chi[1];
  a.D = s.nz;
  a.Dp = (int)Dp;
  a.Kp = (int)Kp;
  a.nf = s.nf;
  a.nl = s.nl;
  a.tiles_n = (int)(Dp / (tb > 0 ? 32 * tb : CRT_TILE));
  a.alpha = s.alpha;
  const double cr[4] = {0.5, 0.5, -0.25, 0.25};
  for (int r = 0; r < 4; ++r) a.cf[r] = cr[r] * c->f_damping;
  a.omf = 1.0 - c->f_damping;
  crt_constants(a, nmod, s.nz);
  size_t lds_b = 0;
  void (*prep_b)(CrtArgs) = prep_b_kernel(s.nl, Kp, Dp, s.nf, &lds_b);
  for (int q0 = 0; q0 < s.nq; q0 += qbatch) {
    a.q0 = q0;
    a.nqb = std::min(qbatch, s.nq - q0);
    const size_t mats = 4 * (size_t)a.nqb;
    crt_workspace(a, c->split_ws, mats, nmod, Dp, Kp);
    if (int rc = timer_begin(c)) return rc;
    hipLaunchKernelGGL(prep_b, dim3((unsigned)(Dp / PREP_B_COLS), (unsigned)mats), dim3(256), lds_b, c->stream, a);
    TUPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(crt_rowmax_fold, dim3((unsigned)((Dp + 255) / 256), (unsigned)mats), dim3(256), 0, c->stream,
                       a);
    TUPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(crt_prep_a, dim3((unsigned)((Dp / 64) * (Kp / 64)), (unsigned)mats), dim3(256), 0,
                       c->stream, a);
    TUPS_HIP(hipGetLastError());
    if (int rc = timer_end(c, TUPS_K_SPLIT_PREP)) return rc;
    const long long nblk = (long long)a.tiles_n * a.tiles_n * mats;
    TUPS_CHECK(nblk < (1LL << 31), TUPS_EINVAL, "split grid too large");
    if (int rc = timer_begin(c)) return rc;
    const int ns = split_ns(c);
    void (*gemm)(CrtArgs) = tb == 6   ? crt_gemm_big<6>
                            : tb == 5 ? crt_gemm_big<5>
                            : ns == 4 ? crt_gemm_dma<4>
                            : ns == 5 ? crt_gemm_dma<5>
                                      : crt_gemm;
    hipLaunchKernelGGL(gemm, dim3((unsigned)nblk), dim3(256), 0, c->stream, a);
    TUPS_HIP(hipGetLastError());
    if (int rc = timer_end(c, TUPS_K_SPLIT_GEMM)) return rc;
  }
  return TUPS_OK;
}

// ---- split form of a batch -------------------------------------------------------------
//
// The q of one (Nf, Nl) class share D, so each class runs as passes of the
// uniform kernels above; the kernels find a q's blocks through the batch's
// offset tables (CrtArgs::qlist, voff, coff).  Classes go largest D first,
// each with the GEMM tile its D pads to best (the 192 tile within 7%, else the
// 64 tile with four DMA stages).
static int batch_split_classes(tups_batch* b) {
  if (b->d_qlist) return TUPS_OK;
  std::vector<std::pair<int, int>> keys;  // (nf, nl)
  for (int q = 0; q < b->nq; ++q) {
    const std::pair<int, int> k{b->nf[q], b->nl[q]};
    if (std::find(keys.begin(), keys.end(), k) == keys.end()) keys.push_back(k);
  }
  std::sort(keys.begin(), keys.end(), [](const std::pair<int, int>& x, const std::pair<int, int>& y) {
    return x.first * x.second > y.first * y.second;
  });
  b->qlist.clear();
  b->class_start.clear();
  for (const auto& k : keys) {
    b->class_start.push_back((int)b->qlist.size());
    for (int q = 0; q < b->nq; ++q)
      if (b->nf[q] == k.first && b->nl[q] == k.second) b->qlist.push_back(q);
  }
  b->class_start.push_back((int)b->qlist.size());
  TUPS_HIP(hipMalloc((void**)&b->d_qlist, b->qlist.size() * sizeof(int)));
  TUPS_HIP(hipMemcpy(b->d_qlist, b->qlist.data(), b->qlist.size() * sizeof(int), hipMemcpyHostToDevice));
  return TUPS_OK;
}

int batch_split_launch(tups_batch* b, int nmod) {
  TUPS_CHECK(nmod >= 8 && nmod <= CRT_MAX, TUPS_EINVAL, "split moduli must be in 8..16");
  if (int rc = batch_split_classes(b)) return rc;
  const int ncls = (int)b->class_start.size() - 1;
  // passes of about 4 GiB of residues; the workspace holds the largest
  size_t need = 0;
  std::vector<int> qbatch(ncls);
  for (int k = 0; k < ncls; ++k) {
    const int q = b->qlist[b->class_start[k]], D = b->nf[q] * b->nl[q];
    TUPS_CHECK(D <= 1024, TUPS_EINVAL, "split path needs D <= 1024 (exact INT32 sums)");
    const size_t per_q = 4 * crt_bytes_per_mat(D, nmod, crt_big_tb(D));
    const int cnt = b->class_start[k + 1] - b->class_start[k];
    qbatch[k] = (int)std::max<size_t>(1, std::min<size_t>(cnt, ((size_t)4 << 30) / per_q));
    need = std::max(need, per_q * qbatch[k]);
  }
  if (b->split_ws_bytes < need) {
    if (b->split_ws) TUPS_HIP(hipFree(b->split_ws));
    b->split_ws = nullptr;
    b->split_ws_bytes = 0;
    TUPS_HIP(hipMalloc(&b->split_ws, need));
    b->split_ws_bytes = need;
  }
  for (int k = 0; k < ncls; ++k) {
    const int q = b->qlist[b->class_start[k]];
    const int nf = b->nf[q], nl = b->nl[q], D = nf * nl;
    const int tb = crt_big_tb(D);
    const size_t Dp = crt_padded(D, tb);
    const size_t Kp = (2 * (size_t)D + 63) / 64 * 64;
    CrtArgs a{};
    for (int r = 0; r < 4; ++r) {
      a.gam[r] = b->gamma[r];
      a.fv[r] = b->fvert[r];
      a.out[r] = b->spare[r];
    }
    a.chi[0] = b->chi[0];
    a.chi[1] = b->chi[1];
    a.qlist = b->d_qlist + b->class_start[k];
    a.voff = b->d_voff;
    a.coff = b->d_coff;
    a.D = D;
    a.Dp = (int)Dp;
    a.Kp = (int)Kp;
    a.nf = nf;
    a.nl = nl;
    a.tiles_n = (int)(Dp / (tb > 0 ? 32 * tb : CRT_TILE));
    a.alpha = b->alpha;
    const double cr[4] = {0.5, 0.5, -0.25, 0.25};
    for (int r = 0; r < 4; ++r) a.cf[r] = cr[r] * b->f_damping;
    a.omf = 1.0 - b->f_damping;
    crt_constants(a, nmod, D);
    size_t lds_b = 0;
    void (*prep_b)(CrtArgs) = prep_b_kernel(nl, Kp, Dp, nf, &lds_b);
    if (lds_b > 65536)
      TUPS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(prep_b),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_b));
    void (*gemm)(CrtArgs) = tb == 6 ? crt_gemm_big<6> : crt_gemm_dma<4>;
    const int cnt = b->class_start[k + 1] - b->class_start[k];
    for (int q0 = 0; q0 < cnt; q0 += qbatch[k]) {
      a.q0 = q0;
      a.nqb = std::min(qbatch[k], cnt - q0);
      const size_t mats = 4 * (size_t)a.nqb;
      crt_workspace(a, b->split_ws, mats, nmod, Dp, Kp);
      hipLaunchKernelGGL(prep_b, dim3((unsigned)(Dp / PREP_B_COLS), (unsigned)mats), dim3(256), lds_b, b->stream,
                         a);
      TUPS_HIP(hipGetLastError());
      hipLaunchKernelGGL(crt_rowmax_fold, dim3((unsigned)((Dp + 255) / 256), (unsigned)mats), dim3(256), 0,
                         b->stream, a);
      TUPS_HIP(hipGetLastError());
      hipLaunchKernelGGL(crt_prep_a, dim3((unsigned)((Dp / 64) * (Kp / 64)), (unsigned)mats), dim3(256), 0,
                         b->stream, a);
      TUPS_HIP(hipGetLastError());
      const long long nblk = (long long)a.tiles_n * a.tiles_n * mats;
      TUPS_CHECK(nblk < (1LL << 31), TUPS_EINVAL, "split grid too large");
      hipLaunchKernelGGL(gemm, dim3((unsigned)nblk), dim3(256), 0, b->stream, a);
      TUPS_HIP(hipGetLastError());
    }
  }
  return TUPS_OK;
}

}  // namespace tups

extern "C" int tups_batch_bse_split(tups_batch* b, int32_t moduli, float* ms) {
  using namespace tups;
  TUPS_CHECK(b, TUPS_EINVAL, "null batch");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms) {
    TUPS_HIP(hipEventCreate(&e0));
    TUPS_HIP(hipEventCreate(&e1));
    TUPS_HIP(hipEventRecord(e0, b->stream));
  }
  if (int rc = batch_split_launch(b, moduli)) return rc;
  if (ms) {
    TUPS_HIP(hipEventRecord(e1, b->stream));
    TUPS_HIP(hipEventSynchronize(e1));
    TUPS_HIP(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  for (int r = 0; r < 4; ++r) std::swap(b->gamma[r], b->spare[r]);
  return TUPS_OK;
}

namespace tups {

// Times one split step (preparation + GEMM, outputs to the ping-pong buffer) per
// tile variant on the resident inputs after a warm-up, keeps the fastest.
int autotune_split(tups_ctx* c, int* chosen, float* best_ms) {
  float best = 0.f;
  int arg = -1;
  const int saved = c->split_variant;
  for (int v = 0; v < kNumSplitVariants; ++v) {
    c->split_variant = v;
    if (int rc = launch_bse_split(c)) {
      c->split_variant = saved;
      return rc;
    }
    TUPS_HIP(hipEventRecord(c->tev[0], c->stream));
    if (int rc = launch_bse_split(c)) {
      c->split_variant = saved;
      return rc;
    }
    TUPS_HIP(hipEventRecord(c->tev[1], c->stream));
    TUPS_HIP(hipEventSynchronize(c->tev[1]));
    float ms = 0.f;
    TUPS_HIP(hipEventElapsedTime(&ms, c->tev[0], c->tev[1]));
    if (arg < 0 || ms < best) {
      best = ms;
      arg = v;
    }
  }
  c->split_variant = arg;
  if (chosen) *chosen = arg;
  if (best_ms) *best_ms = best;
  return TUPS_OK;
}

}  // namespace tups
