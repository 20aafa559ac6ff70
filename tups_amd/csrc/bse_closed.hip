// Closed form of the BSE ladder, direct (FP64) and by FP32 inverse with FP64
// iterative refinement.
//
// For fixed Gamma the damped update of BSE(ite) (SRC/parquet_BSE.f90:62-132),
//   Phi = c_r f alpha (Gamma chi F + F chi Gamma) + (1 - f)(F - Gamma),
// has the fixed point F = (I - kappa_r Gamma chi)^-1 Gamma, kappa_r = 2 c_r alpha
// (push-through identity: Gamma chi F = F chi Gamma there), where Phi = F - Gamma
// for every damping f.  Both modes solve A X = Gamma, A = I - kappa_r Gamma chi,
// per (q, r) and return Phi = X - Gamma: the converged ladder in one call.  The
// reference has no such solve; it iterates the update.
//
// Direct mode (TUPS_BSE_CLOSED_LU): blocked Gauss-Jordan elimination with
// partial pivoting on the augmented matrix [A | Gamma] (D x 2D, complex128),
// one workgroup per matrix.  Per panel of NB columns:
//   1. LU of the panel rows k0..D-1 with row pivoting (rows swapped over all
//      columns from k0 on, so the right-hand side follows the equations);
//   2. Y = L11^-1 Z on the panel rows for every trailing column Z;
//   3. rows below the panel: Z -= L21 Y                    (MFMA);
//   4. X = U11^-1 Y on the panel rows;
//   5. rows above the panel: Z -= P_above X                (MFMA);
// after the last panel the right-hand block holds the solution.  Work per
// (q, r): 12 D^3 real flops.  From D = 128 the block form (closed_solve_b64)
// runs the same elimination with 64-column blocks: the block column factored
// in a side buffer (on registers, 16 columns at a time, for D <= 512), T =
// A11^-1 from blocked triangular inverses, the block rows multiplied by T and
// every other row updated by rank-64 MFMA products whose operands stream
// into LDS by direct global-to-LDS loads.
//
// Refinement mode (TUPS_BSE_CLOSED_IR): the same elimination in FP32 on
// [fp32(A) | I] gives the inverse of the FP32-rounded A (its LU with partial
// pivoting applied to the identity, FP32 MFMA updates); then
//   X  = fp64(A32^-1 fp32(Gamma))                         (FP32 MFMA GEMM)
//   repeat: R = Gamma - A X = Gamma - X + kappa Gamma (chi X)   (FP64 MFMA GEMM)
//           stop when max over (q, r) of |R|_F / |Gamma|_F <= tol
//           X += fp64(A32^-1 fp32(R))                     (FP32 MFMA GEMM)
// FP64 is used only for A's assembly, the residual and the update of X.  Each
// step shrinks the error by about cond(A) 2^-24, so two or three steps reach
// 1e-12.  Work per (q, r): 12 D^3 FP32 (inverse) + 8 D^3 FP32 per product with
// the inverse + 8 D^3 FP64 per residual.
#include "internal.h"
#include "lds_dma.h"

#include <cmath>
#include <cstring>
#include <type_traits>

namespace tups {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int CL_NB = 16;  // panel width (closed_solve), sub-panel width (closed_solve_b64)
constexpr int CL_T = 64;   // update tile (rows and columns)
constexpr int CL_BB = 64;  // block width of closed_solve_b64

__device__ __forceinline__ double2 mkc(double x, double y) { return make_double2(x, y); }
__device__ __forceinline__ float2 mkc(float x, float y) { return make_float2(x, y); }

// Precision of an elimination or product: complex type, MFMA accumulator and
// the 16 x 16 x 4 MFMA output layout (element e of lane l sits at column l % 16
// and the row given by orow(l / 16, e)).
template <class R>
struct Prec;
template <>
struct Prec<double> {
  using C = double2;
  using V = d4;
  __device__ static V fma_add(double a, double b, V acc) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
  __device__ static V fma_sub(double a, double b, V acc) {  // A negated by the MFMA modifier
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 1);
  }
  __device__ static int orow(int lk, int e) { return lk + 4 * e; }
};
template <>
struct Prec<float> {
  using C = float2;
  using V = f4;
  __device__ static V fma_add(float a, float b, V acc) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  }
  __device__ static V fma_sub(float a, float b, V acc) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(-a, b, acc, 0, 0, 0);
  }
  __device__ static int orow(int lk, int e) { return 4 * lk + e; }
};

template <class C>
__device__ __forceinline__ C cmul(C a, C b) {
  return mkc(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class C>
__device__ __forceinline__ C csub(C a, C b) {
  return mkc(a.x - b.x, a.y - b.y);
}
template <class C>
__device__ __forceinline__ C cinv(C a) {
  // 1 / a with the larger component scaled out (Smith's method), two divisions
  if (fabs(a.x) >= fabs(a.y)) {
    const auto r = a.y / a.x, id = 1 / (a.x + a.y * r);
    return mkc(id, -r * id);
  }
  const auto r = a.x / a.y, id = 1 / (a.y + a.x * r);
  return mkc(r * id, -id);
}

// Largest 64-bit key over the wave (unsigned order), in every lane: DPP steps
// leave the row maxima in every lane of a row, row_bcast carries them to lane
// 63, which is read out.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long key) {
  unsigned hi = (unsigned)(key >> 32), lo = (unsigned)key;
  auto step = [&](auto ctrl, auto rmask) {
    constexpr int CT = decltype(ctrl)::value, RM = decltype(rmask)::value;
    const unsigned nhi = (unsigned)__builtin_amdgcn_update_dpp((int)hi, (int)hi, CT, RM, 0xf, false);
    const unsigned nlo = (unsigned)__builtin_amdgcn_update_dpp((int)lo, (int)lo, CT, RM, 0xf, false);
    if (nhi > hi || (nhi == hi && nlo > lo)) {
      hi = nhi;
      lo = nlo;
    }
  };
  step(std::integral_constant<int, 0xb1>{}, std::integral_constant<int, 0xf>{});   // quad_perm [1,0,3,2]
  step(std::integral_constant<int, 0x4e>{}, std::integral_constant<int, 0xf>{});   // quad_perm [2,3,0,1]
  step(std::integral_constant<int, 0x141>{}, std::integral_constant<int, 0xf>{});  // row_half_mirror
  step(std::integral_constant<int, 0x140>{}, std::integral_constant<int, 0xf>{});  // row_mirror
  step(std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xa>{});  // row_bcast:15 to rows 1, 3
  step(std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xc>{});  // row_bcast:31 to rows 2, 3
  hi = (unsigned)__builtin_amdgcn_readlane((int)hi, 63);
  lo = (unsigned)__builtin_amdgcn_readlane((int)lo, 63);
  return ((unsigned long long)hi << 32) | lo;
}

// Pivot key of a candidate: the bits of |.|_1 (a non-negative double, so its
// bits order as unsigned integers) with the low 10 bits replaced by 1023 -
// off (off < 1024, the row's offset in the panel), so the largest key is the
// largest magnitude (to 2^-42) and, among equal ones, the lowest row; 0 for
// no candidate and for NaN.
__device__ __forceinline__ unsigned long long pivot_key(double m, int off) {
  if (!(m >= 0.0)) return 0ull;
  return ((unsigned long long)__double_as_longlong(m) & ~0x3ffull) | (unsigned long long)(1023 - off);
}
__device__ __forceinline__ unsigned long long pivot_key(float m, int off) { return pivot_key((double)m, off); }
__device__ __forceinline__ double pivot_mag(unsigned long long key) {
  return __longlong_as_double((long long)(key & ~0x3ffull));
}

// f(std::integral_constant<int, I>{}) for I = B .. E - 1, each I a compile-time
// constant (register arrays indexed by it stay in registers)
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

struct ClosedArgs {
  const double2* gam[4];
  double2* out[4];
  const double2* chi[2];
  void* W;      // [mat][2D columns][D rows], column-major: double2 (direct) or float2 (inverse)
  void* S;      // closed_solve_b64: [mat][CL_BB columns][D rows], the block column being factored
  void* T;      // closed_solve_b64: [mat][2][CL_BB][CL_BB], L11^-1 (then T = U11^-1 L11^-1) and U11^-1
  int* info;    // first failing pivot (1-based column) over all matrices, 0 if none
  int D, nf, nl, q0, nqb;
  double kappa[4];
};

__global__ void closed_clear(int* info) { *info = 0; }

// [I - kappa Gamma chi | Gamma] (direct mode) or [fp32(I - kappa Gamma chi) | I]
// (refinement mode, A formed in FP64 and rounded).  One workgroup per (nu,
// matrix) forms the Nl columns (l2, nu) of the left block from the Nl columns
// (l1, nu) of Gamma, each read once: (Gamma chi)(row, (l2, nu)) =
// sum_l1 Gamma(row, (l1, nu)) chi_nu(l1, l2), and writes the columns (l1, nu)
// of the right block from the same reads.  grid: (Nf, 4 nqb); one thread per row.
constexpr int CL_NL_MAX = 9;
template <class T, bool IR>
__global__ __launch_bounds__(256) void closed_build_nu(ClosedArgs a) {
  const int x = blockIdx.x, mat = blockIdx.y;
  const int r = mat / a.nqb, q = a.q0 + (mat - r * a.nqb);
  const int D = a.D, nf = a.nf, nl = a.nl;
  const double2* G = a.gam[r] + (size_t)D * D * q;
  T* W = static_cast<T*>(a.W) + (size_t)mat * 2 * D * D;
  auto put = [](double re, double im) -> T {
    if constexpr (IR) return make_float2((float)re, (float)im);
    else return make_double2(re, im);
  };
  const int nu = x;
  const double2* cb = a.chi[r >> 1] + (size_t)nf * nl * nl * q + (size_t)nu * nl * nl;  // chi_nu(l1, l2) at l1 + nl l2
  const double kap = a.kappa[r];
  for (int row = threadIdx.x; row < D; row += 256) {
    double2 g[CL_NL_MAX];
#pragma unroll
    for (int l1 = 0; l1 < CL_NL_MAX; ++l1)
      if (l1 < nl) {
        const int j = l1 * nf + nu;
        g[l1] = G[(size_t)j * D + row];
        // the right-hand side: Gamma's column j (the identity's for the FP32 inverse)
        if constexpr (IR) W[(size_t)(D + j) * D + row] = put(row == j ? 1.0 : 0.0, 0.0);
        else W[(size_t)(D + j) * D + row] = g[l1];
      }
    for (int l2 = 0; l2 < nl; ++l2) {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int l1 = 0; l1 < CL_NL_MAX; ++l1)
        if (l1 < nl) {
          const double2 c = cb[l1 + nl * l2];
          sr += g[l1].x * c.x - g[l1].y * c.y;
          si += g[l1].x * c.y + g[l1].y * c.x;
        }
      const int j = l2 * nf + nu;
      W[(size_t)j * D + row] = put((row == j ? 1.0 : 0.0) - kap * sr, -kap * si);
    }
  }
}

// Z(r, c) -= sum_t W(r, k0 + t) W(k0 + t, c) for r in [r0, r1), c in [c0, c1),
// t < nb (nb <= CL_BB), by 64 x 64 tiles held in registers while the inner
// index passes through LDS CL_NB at a time; four waves of 2 x 2 blocks of
// 16 x 16 x 4 MFMAs.  A fragments: lane l holds A[l % 16][l / 16]; B: B[l / 16][l % 16].
template <class R>
__device__ void closed_update(typename Prec<R>::C* W, int D, int r0, int r1, int c0, int c1, int k0, int nb,
                              typename Prec<R>::C (*As)[CL_T + 1], typename Prec<R>::C (*Bs)[CL_T + 1]) {
  using P = Prec<R>;
  using C = typename P::C;
  using V = typename P::V;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const C zero = mkc((R)0, (R)0);
  for (int tr = r0; tr < r1; tr += CL_T)
    for (int tc = c0; tc < c1; tc += CL_T) {
      V ar[2][2], ai[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = tr + wr + i * 16 + P::orow(lk, e), col = tc + wc + j * 16 + li;
            const C z = (row < r1 && col < c1) ? W[(size_t)col * D + row] : zero;
            ar[i][j][e] = z.x;
            ai[i][j][e] = z.y;
          }
      // staging: thread tid moves A(row tid % 64, k (tid / 64) + 4 v) and
      // B(k tid % 16, column tid / 16 + 16 v), v < 4.  In FP32 the next chunk
      // is loaded into registers while the MFMAs of this one run; in FP64 that
      // costs spills in the panel kernel, and the chunk goes straight to LDS.
      constexpr bool kPrefetch = sizeof(R) == 4;
      C pa[4], pb[4];
      auto fetch = [&](int kc) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = tid & (CL_T - 1), ka = (tid >> 6) + 4 * v;
          pa[v] = (tr + row < r1 && kc + ka < nb) ? W[(size_t)(k0 + kc + ka) * D + tr + row] : zero;
          const int kb = tid % CL_NB, col = (tid / CL_NB) + 16 * v;
          pb[v] = (tc + col < c1 && kc + kb < nb) ? W[(size_t)(tc + col) * D + k0 + kc + kb] : zero;
        }
      };
      if constexpr (kPrefetch) fetch(0);
      for (int kc = 0; kc < nb; kc += CL_NB) {
        if constexpr (kPrefetch) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            As[(tid >> 6) + 4 * v][tid & (CL_T - 1)] = pa[v];
            Bs[tid % CL_NB][(tid / CL_NB) + 16 * v] = pb[v];
          }
        } else {
          for (int u = tid; u < CL_NB * CL_T; u += 256) {
            const int row = u & (CL_T - 1), ka = u >> 6;
            As[ka][row] = (tr + row < r1 && kc + ka < nb) ? W[(size_t)(k0 + kc + ka) * D + tr + row] : zero;
            const int kb = u % CL_NB, col = u / CL_NB;
            Bs[kb][col] = (tc + col < c1 && kc + kb < nb) ? W[(size_t)(tc + col) * D + k0 + kc + kb] : zero;
          }
        }
        __syncthreads();
        if (kPrefetch && kc + CL_NB < nb) fetch(kc + CL_NB);
#pragma unroll
        for (int kk = 0; kk < CL_NB; kk += 4) {
          C fa[2], fb[2];
#pragma unroll
          for (int i = 0; i < 2; ++i) fa[i] = As[kk + lk][wr + i * 16 + li];
#pragma unroll
          for (int j = 0; j < 2; ++j) fb[j] = Bs[kk + lk][wc + j * 16 + li];
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              // Z - A B: re -= ar br - ai bi, im -= ar bi + ai br
              ar[i][j] = P::fma_sub(fa[i].x, fb[j].x, ar[i][j]);
              ar[i][j] = P::fma_add(fa[i].y, fb[j].y, ar[i][j]);
              ai[i][j] = P::fma_sub(fa[i].x, fb[j].y, ai[i][j]);
              ai[i][j] = P::fma_sub(fa[i].y, fb[j].x, ai[i][j]);
            }
        }
        __syncthreads();
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = tr + wr + i * 16 + P::orow(lk, e), col = tc + wc + j * 16 + li;
            if (row < r1 && col < c1) W[(size_t)col * D + row] = mkc(ar[i][j][e], ai[i][j][e]);
          }
    }
  __syncthreads();  // the updated Z is visible to the whole workgroup
}

// The FP64 products of the block elimination with the inner index staged by
// direct global-to-LDS loads into two stages, so chunk k + 1 is in flight
// while chunk k multiplies and no registers hold it:
//   APPLY = false (closed_update): Z(r, c) -= sum_t W(r, k0 + t) W(k0 + t, c),
//     r in [r0, r1), c in [c0, c1);
//   APPLY = true (closed_apply_T): Z(k0 + i, c) <- sum_t T(i, t) Z(k0 + t, c),
//     i < nb, T column-major with leading dimension CL_BB (each tile is read
//     whole before it is written, so the product runs in place).
// A stage is the A chunk sa[CL_NB][CL_T + 1] (rows padded) and the B chunk in
// the first CL_T CL_NB slots of sb, column-major with the inner index
// XOR-swizzled by the column (a fragment read is then free of bank
// conflicts).  Rows and columns beyond the ranges load clamped addresses
// (their results are not written).  nb must be a multiple of CL_NB, so every
// chunk issues the same 8 loads per thread.
template <bool APPLY>
__device__ void closed_gemm_dma(double2* W, int D, const double2* T, int r0, int r1, int c0, int c1, int k0, int nb,
                                double2 (*s0a)[CL_T + 1], double2 (*s0b)[CL_T + 1], double2 (*s1a)[CL_T + 1],
                                double2 (*s1b)[CL_T + 1]) {
  using P = Prec<double>;
  using V = P::V;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const double2 zero = mkc(0.0, 0.0);
  const int nch = nb / CL_NB;
  // A(row, t) = Ab[(ka0 + t) lda + row]; the output row of A's row r is ro + r
  const double2* Ab = APPLY ? T : W;
  const int lda = APPLY ? CL_BB : D, ka0 = APPLY ? 0 : k0, amax = APPLY ? CL_BB : D, ro = APPLY ? k0 : 0;
  if (APPLY) {
    r0 = 0;
    r1 = nb;
  }
  for (int tr = r0; tr < r1; tr += CL_T)
    for (int tc = c0; tc < c1; tc += CL_T) {
      const int arow = min(tr + lane, amax - 1);
      auto issue = [&](int ch, double2 (*sa)[CL_T + 1], double2 (*sb)[CL_T + 1]) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {  // A: inner index ka, rows tr .. tr + 63
          const int ka = wave + 4 * v;
          dma16(Ab, (unsigned)(((size_t)(ka0 + ch * CL_NB + ka) * lda + arow) * sizeof(double2)),
                lds_addr(&sa[ka][0]));
        }
        const double2* sbf = &sb[0][0];
#pragma unroll
        for (int v = 0; v < 4; ++v) {  // B: slots 64 d .. 64 d + 63 = columns 4 d .. 4 d + 3
          const int d = wave + 4 * v, slot = 64 * d + lane, col = slot >> 4;
          const int kb = (slot & 15) ^ (col & 15);
          const int gc = min(tc + col, c1 - 1);
          dma16(W, (unsigned)(((size_t)gc * D + k0 + ch * CL_NB + kb) * sizeof(double2)), lds_addr(sbf + 64 * d));
        }
      };
      V ar[2][2], ai[2][2];
      auto compute = [&](double2 (*sa)[CL_T + 1], double2 (*sb)[CL_T + 1]) {
        const double2* sbf = &sb[0][0];
#pragma unroll
        for (int kk = 0; kk < CL_NB; kk += 4) {
          double2 fa[2], fb[2];
#pragma unroll
          for (int i = 0; i < 2; ++i) fa[i] = sa[kk + lk][wr + i * 16 + li];
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            const int col = wc + j * 16 + li;
            fb[j] = sbf[col * CL_NB + ((kk + lk) ^ (col & 15))];
          }
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              if constexpr (APPLY) {  // + A B
                ar[i][j] = P::fma_add(fa[i].x, fb[j].x, ar[i][j]);
                ar[i][j] = P::fma_sub(fa[i].y, fb[j].y, ar[i][j]);
                ai[i][j] = P::fma_add(fa[i].x, fb[j].y, ai[i][j]);
                ai[i][j] = P::fma_add(fa[i].y, fb[j].x, ai[i][j]);
              } else {  // - A B
                ar[i][j] = P::fma_sub(fa[i].x, fb[j].x, ar[i][j]);
                ar[i][j] = P::fma_add(fa[i].y, fb[j].y, ar[i][j]);
                ai[i][j] = P::fma_sub(fa[i].x, fb[j].y, ai[i][j]);
                ai[i][j] = P::fma_sub(fa[i].y, fb[j].x, ai[i][j]);
              }
            }
        }
      };
      issue(0, s0a, s0b);
      if constexpr (APPLY) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) ar[i][j] = ai[i][j] = V{};
      } else {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int row = tr + wr + i * 16 + P::orow(lk, e), col = tc + wc + j * 16 + li;
              const double2 z = (row < r1 && col < c1) ? W[(size_t)col * D + row] : zero;
              ar[i][j][e] = z.x;
              ai[i][j][e] = z.y;
            }
        // the tile's values are waited for here, with chunk 0, and not at the
        // first MFMA (where the wait would also cover chunk 1)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            asm volatile("" ::"v"(ar[i][j][0]), "v"(ar[i][j][1]), "v"(ar[i][j][2]), "v"(ar[i][j][3]),
                         "v"(ai[i][j][0]), "v"(ai[i][j][1]), "v"(ai[i][j][2]), "v"(ai[i][j][3]));
      }
      for (int ch = 0; ch < nch; ch += 2) {
        if (ch + 1 < nch) {
          issue(ch + 1, s1a, s1b);
          wait_vmcnt<8>();
        } else {
          wait_vmcnt<0>();
        }
        stage_barrier();
        compute(s0a, s0b);
        stage_barrier();
        if (ch + 1 >= nch) break;
        if (ch + 2 < nch) {
          issue(ch + 2, s0a, s0b);
          wait_vmcnt<8>();
        } else {
          wait_vmcnt<0>();
        }
        stage_barrier();
        compute(s1a, s1b);
        stage_barrier();
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = tr + wr + i * 16 + P::orow(lk, e), col = tc + wc + j * 16 + li;
            if (row < r1 && col < c1) W[(size_t)col * D + ro + row] = mkc(ar[i][j][e], ai[i][j][e]);
          }
    }
  __syncthreads();  // the result is visible to the whole workgroup
}

// the update of the block elimination: in FP64 with the DMA stages when nb
// allows, else (and in FP32) closed_update on the first stage
template <class R>
__device__ __forceinline__ void closed_update_b64(typename Prec<R>::C* W, int D, int r0, int r1, int c0, int c1,
                                                  int k0, int nb, typename Prec<R>::C (*s0a)[CL_T + 1],
                                                  typename Prec<R>::C (*s0b)[CL_T + 1],
                                                  typename Prec<R>::C (*s1a)[CL_T + 1],
                                                  typename Prec<R>::C (*s1b)[CL_T + 1]) {
  if constexpr (std::is_same<R, double>::value) {
    if (nb % CL_NB == 0) {
      closed_gemm_dma<false>(W, D, nullptr, r0, r1, c0, c1, k0, nb, s0a, s0b, s1a, s1b);
      return;
    }
  }
  closed_update<R>(W, D, r0, r1, c0, c1, k0, nb, s0a, s0b);
}

// Z(k0 + i, c) <- sum_t T(i, t) Z(k0 + t, c) for i, t < nb (nb <= CL_BB) and every
// column c in [c0, c1): T column-major with leading dimension CL_BB.  Each 64 x
// 64 tile of Z is read whole (through LDS, CL_NB rows at a time) before its
// result is written back, so the product runs in place.
template <class R>
__device__ void closed_apply_T(typename Prec<R>::C* W, const typename Prec<R>::C* T, int D, int k0, int nb, int c0,
                               int c1, typename Prec<R>::C (*As)[CL_T + 1], typename Prec<R>::C (*Bs)[CL_T + 1]) {
  using P = Prec<R>;
  using C = typename P::C;
  using V = typename P::V;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const C zero = mkc((R)0, (R)0);
  for (int tc = c0; tc < c1; tc += CL_T) {
    V ar[2][2], ai[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) ar[i][j] = ai[i][j] = V{};
    // staging as in closed_update
    constexpr bool kPrefetch = sizeof(R) == 4;
    C pa[4], pb[4];
    auto fetch = [&](int kc) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = tid & (CL_T - 1), ka = (tid >> 6) + 4 * v;
        pa[v] = (row < nb && kc + ka < nb) ? T[(size_t)(kc + ka) * CL_BB + row] : zero;
        const int kb = tid % CL_NB, col = (tid / CL_NB) + 16 * v;
        pb[v] = (tc + col < c1 && kc + kb < nb) ? W[(size_t)(tc + col) * D + k0 + kc + kb] : zero;
      }
    };
    if constexpr (kPrefetch) fetch(0);
    for (int kc = 0; kc < nb; kc += CL_NB) {
      if constexpr (kPrefetch) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          As[(tid >> 6) + 4 * v][tid & (CL_T - 1)] = pa[v];
          Bs[tid % CL_NB][(tid / CL_NB) + 16 * v] = pb[v];
        }
      } else {
        for (int u = tid; u < CL_NB * CL_T; u += 256) {
          const int row = u & (CL_T - 1), ka = u >> 6;
          As[ka][row] = (row < nb && kc + ka < nb) ? T[(size_t)(kc + ka) * CL_BB + row] : zero;
          const int kb = u % CL_NB, col = u / CL_NB;
          Bs[kb][col] = (tc + col < c1 && kc + kb < nb) ? W[(size_t)(tc + col) * D + k0 + kc + kb] : zero;
        }
      }
      __syncthreads();
      if (kPrefetch && kc + CL_NB < nb) fetch(kc + CL_NB);
#pragma unroll
      for (int kk = 0; kk < CL_NB; kk += 4) {
        C fa[2], fb[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) fa[i] = As[kk + lk][wr + i * 16 + li];
#pragma unroll
        for (int j = 0; j < 2; ++j) fb[j] = Bs[kk + lk][wc + j * 16 + li];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            // + A B: re += ar br - ai bi, im += ar bi + ai br
            ar[i][j] = P::fma_add(fa[i].x, fb[j].x, ar[i][j]);
            ar[i][j] = P::fma_sub(fa[i].y, fb[j].y, ar[i][j]);
            ai[i][j] = P::fma_add(fa[i].x, fb[j].y, ai[i][j]);
            ai[i][j] = P::fma_add(fa[i].y, fb[j].x, ai[i][j]);
          }
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = wr + i * 16 + P::orow(lk, e), col = tc + wc + j * 16 + li;
          if (row < nb && col < c1) W[(size_t)col * D + k0 + row] = mkc(ar[i][j][e], ai[i][j][e]);
        }
  }
  __syncthreads();
}

// the T products of the block elimination: in FP64 with the DMA stages when
// nb allows, else (and in FP32) closed_apply_T on the first stage
template <class R>
__device__ __forceinline__ void closed_apply_T_b64(typename Prec<R>::C* W, const typename Prec<R>::C* T, int D, int k0,
                                                   int nb, int c0, int c1, typename Prec<R>::C (*s0a)[CL_T + 1],
                                                   typename Prec<R>::C (*s0b)[CL_T + 1],
                                                   typename Prec<R>::C (*s1a)[CL_T + 1],
                                                   typename Prec<R>::C (*s1b)[CL_T + 1]) {
  if constexpr (std::is_same<R, double>::value) {
    if (nb % CL_NB == 0) {
      closed_gemm_dma<true>(W, D, T, 0, nb, c0, c1, k0, nb, s0a, s0b, s1a, s1b);
      return;
    }
  }
  closed_apply_T<R>(W, T, D, k0, nb, c0, c1, s0a, s0b);
}

// Z(k0 + i, c) <- sum_t M[i][t] Z(k0 + t, c) for i, t < nb and every column c in
// [c0, c1): one thread per column, M (NB x NB, identity-padded) in LDS.
template <class R>
__device__ void panel_rows_apply(typename Prec<R>::C* W, int D, int k0, int nb, int c0, int c1,
                                 const typename Prec<R>::C (*M)[CL_NB]) {
  using C = typename Prec<R>::C;
  for (int col = c0 + (int)threadIdx.x; col < c1; col += 256) {
    C* zc = W + (size_t)col * D + k0;
    C z[CL_NB];
#pragma unroll
    for (int i = 0; i < CL_NB; ++i) z[i] = i < nb ? zc[i] : mkc((R)0, (R)0);
#pragma unroll 1
    for (int i = 0; i < CL_NB; ++i) {
      R yr = 0, yi = 0;
#pragma unroll
      for (int t = 0; t < CL_NB; ++t) {
        const C m = M[i][t];
        yr = fma(m.x, z[t].x, fma(-m.y, z[t].y, yr));
        yi = fma(m.x, z[t].y, fma(m.y, z[t].x, yi));
      }
      if (i < nb) zc[i] = mkc(yr, yi);
    }
  }
}

// One workgroup per (q, r) matrix; grid: 4 * nqb.  PHI: the direct mode's
// final Phi = X - Gamma from the right-hand block (else the right-hand block,
// the inverse, stays in W).
template <class R, bool PHI>
__global__ __launch_bounds__(256, 2) void closed_solve(ClosedArgs a) {
  using C = typename Prec<R>::C;
  __shared__ C As[CL_NB][CL_T + 1];
  __shared__ C Bs[CL_NB][CL_T + 1];
  __shared__ C Ls[CL_NB][CL_NB];
  __shared__ C Li[CL_NB][CL_NB];
  __shared__ C Ui[CL_NB][CL_NB];
  __shared__ R red_v[256];
  __shared__ int red_i[256];
  const int mat = blockIdx.x;
  const int D = a.D, D2 = 2 * D;
  C* W = static_cast<C*>(a.W) + (size_t)mat * 2 * D * D;
  const int tid = threadIdx.x;
  const C one = mkc((R)1, (R)0), zero = mkc((R)0, (R)0);
  int info = 0;
  for (int k0 = 0; k0 < D; k0 += CL_NB) {
    const int nb = min(CL_NB, D - k0);
    // 1. panel LU, rows k0..D-1, columns k0..k0+nb-1
    for (int c = k0; c < k0 + nb; ++c) {
      R best = -1;
      int bi = c;
      for (int row = c + tid; row < D; row += 256) {
        const C v = W[(size_t)c * D + row];
        const R m = fabs(v.x) + fabs(v.y);
        if (m > best) {
          best = m;
          bi = row;
        }
      }
      red_v[tid] = best;
      red_i[tid] = bi;
      __syncthreads();
      for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
          const R o = red_v[tid + s];
          const int oi = red_i[tid + s];
          if (o > red_v[tid] || (o == red_v[tid] && oi < red_i[tid])) {
            red_v[tid] = o;
            red_i[tid] = oi;
          }
        }
        __syncthreads();
      }
      const int p = red_i[0];
      const R pmax = red_v[0];
      __syncthreads();
      if (!(pmax > 0)) {  // zero or NaN pivot column: the system is singular
        if (info == 0) info = c + 1;
        continue;
      }
      if (p != c) {
        for (int col = k0 + tid; col < D2; col += 256) {
          const C x = W[(size_t)col * D + c];
          W[(size_t)col * D + c] = W[(size_t)col * D + p];
          W[(size_t)col * D + p] = x;
        }
      }
      __syncthreads();
      const C inv = cinv(W[(size_t)c * D + c]);
      for (int row = c + 1 + tid; row < D; row += 256) {
        const C l = cmul(W[(size_t)c * D + row], inv);
        W[(size_t)c * D + row] = l;
        for (int col = c + 1; col < k0 + nb; ++col)
          W[(size_t)col * D + row] = csub(W[(size_t)col * D + row], cmul(l, W[(size_t)col * D + c]));
      }
      __syncthreads();
    }
    for (int u = tid; u < CL_NB * CL_NB; u += 256) {
      const int i = u % CL_NB, j = u / CL_NB;
      Ls[i][j] = (i < nb && j < nb) ? W[(size_t)(k0 + j) * D + k0 + i] : zero;
    }
    __syncthreads();
    // inverses of the unit lower L11 and of the upper U11, padded to NB x NB with
    // the identity: thread t < NB forms column t of both
    if (tid < CL_NB) {
      const int t = tid;
      for (int i = 0; i < CL_NB; ++i) {
        C x = i == t ? one : zero;
        if (i < nb && t < nb)
          for (int s2 = t; s2 < i; ++s2) x = csub(x, cmul(Ls[i][s2], Li[s2][t]));
        Li[i][t] = x;
      }
      for (int i = CL_NB - 1; i >= 0; --i) {
        C x = i == t ? one : zero;
        if (i < nb && t < nb) {
          for (int s2 = i + 1; s2 <= t; ++s2) x = csub(x, cmul(Ls[i][s2], Ui[s2][t]));
          x = cmul(x, cinv(Ls[i][i]));
        }
        Ui[i][t] = x;
      }
    }
    __syncthreads();
    const int j0 = k0 + nb;
    // 2. Y = L11^-1 Z on the panel rows
    panel_rows_apply<R>(W, D, k0, nb, j0, D2, Li);
    __syncthreads();
    // 3. rows below the panel
    closed_update<R>(W, D, k0 + nb, D, j0, D2, k0, nb, As, Bs);
    // 4. X = U11^-1 Y on the panel rows
    panel_rows_apply<R>(W, D, k0, nb, j0, D2, Ui);
    __syncthreads();
    // 5. rows above the panel
    closed_update<R>(W, D, 0, k0, j0, D2, k0, nb, As, Bs);
  }
  if (info && tid == 0) atomicCAS(a.info, 0, info);
  if constexpr (PHI) {
    // Phi = X - Gamma
    const int r = mat / a.nqb, q = a.q0 + (mat - r * a.nqb);
    const double2* G = a.gam[r] + (size_t)D * D * q;
    double2* O = a.out[r] + (size_t)D * D * q;
    for (size_t u = tid; u < (size_t)D * D; u += 256) O[u] = csub(W[(size_t)D * D + u], G[u]);
  }
}

// X(dst[u], col) <- X(src[u], col) for u < n, every column col in [c0, c1)
// outside [e0, e1), all with the values from before the call (the rows of a
// composed sequence of row swaps).  Whole columns go at a time, PER items per
// thread: every source of a column is read before any of its rows is written.
template <class R, int PER>
__device__ void move_rows(typename Prec<R>::C* X, int ld, int c0, int c1, int e0, int e1, const int* dst,
                          const int* src, int n) {
  using C = typename Prec<R>::C;
  const int tid = threadIdx.x;
  const int nlo = e0 - c0, ncol = (c1 - c0) - (e1 - e0);
  if (n == 0 || ncol <= 0) return;
  const int cols = max(1, 256 * PER / n);  // whole columns per round
  for (int cb = 0; cb < ncol; cb += cols) {
    const int items = n * min(cols, ncol - cb);
    C val[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int it = k * 256 + tid;
      if (it < items) {
        const int cc = cb + it / n;
        const int col = cc < nlo ? c0 + cc : e1 + (cc - nlo);
        val[k] = X[(size_t)col * ld + src[it % n]];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int it = k * 256 + tid;
      if (it < items) {
        const int cc = cb + it / n;
        const int col = cc < nlo ? c0 + cc : e1 + (cc - nlo);
        X[(size_t)col * ld + dst[it % n]] = val[k];
      }
    }
    __syncthreads();
  }
}

// As move_rows for n <= 128 and no excluded columns, without workgroup
// barriers: wave w owns the columns c0 + G (w + 4 k) .. + G - 1 and lane l the
// moved rows l and l + 64; a wave's loads of a group of columns have all
// returned (vmcnt 0) before its stores to them are issued.
template <class R, int G>
__device__ void move_rows_waves(typename Prec<R>::C* X, int ld, int c0, int c1, const int* dst, const int* src,
                                int n) {
  using C = typename Prec<R>::C;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool h0 = lane < n, h1 = lane + 64 < n;
  const int s0 = h0 ? src[lane] : 0, d0 = h0 ? dst[lane] : 0;
  const int s1 = h1 ? src[lane + 64] : 0, d1 = h1 ? dst[lane + 64] : 0;
  for (int cb = c0 + G * wave; cb < c1; cb += 4 * G) {
    C x0[G], x1[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const size_t o = (size_t)min(cb + g, c1 - 1) * ld;
      if (h0) x0[g] = X[o + s0];
      if (h1) x1[g] = X[o + s1];
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    asm volatile("" ::: "memory");
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (cb + g >= c1) break;
      const size_t o = (size_t)(cb + g) * ld;
      if (h0) X[o + d0] = x0[g];
      if (h1) X[o + d1] = x1[g];
    }
  }
}

// The rows that the swaps (r0 + i, piv[i]), i < nr, move (piv[i] >= r0 + i),
// for wave 0: map[x - lo] is the original row that row x holds after all the
// swaps, and the rows with map[x - lo] != x are listed, those of r0 .. r0 + nr
// - 1 first (in order), then those beyond, each once.  The count goes to *cnt.
__device__ __forceinline__ void list_moved_rows(const int* piv, int r0, int nr, const int* map, int lo, int* dst,
                                                int* src, int* cnt) {
  const int lane = threadIdx.x & 63;
  bool in = false, out = false;
  int p = 0, sin = 0, sout = 0;
  if (lane < nr) {
    sin = map[r0 + lane - lo];
    in = sin != r0 + lane;
    p = piv[lane];
    if (p >= r0 + nr) {
      sout = map[p - lo];
      out = sout != p;
      for (int j = 0; j < lane; ++j)
        if (piv[j] == p) out = false;
    }
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long mi = __ballot(in), mo = __ballot(out);
  const int ni = __popcll(mi);
  if (in) {
    const int u = __popcll(mi & below);
    dst[u] = r0 + lane;
    src[u] = sin;
  }
  if (out) {
    const int u = ni + __popcll(mo & below);
    dst[u] = p;
    src[u] = sout;
  }
  if (lane == 0) *cnt = ni + __popcll(mo);
}

// Block Gauss-Jordan with CL_BB-wide blocks, one workgroup per (q, r) matrix.
// For the block column k0 .. k0 + nb:
//  A. its rows k0 .. D - 1 are copied to S and factored there (LU with partial
//     pivoting, CL_NB-wide sub-panels: column-by-column within a sub-panel,
//     then the sub-panel's rank-CL_NB update of the rest of the block); each
//     row swap is applied to W from column k0 on, so W keeps the permuted
//     original block column;
//  B. T = A11^-1 = U11^-1 L11^-1 from the factors (nb x nb);
//  C. the block rows of the later columns become X = T Z1, in place;
//  D. every other row r: Z(r, .) -= W(r, k0 .. k0 + nb) X (rank-nb MFMA updates).
// Each block thus passes the later columns through memory once, where the
// CL_NB-wide form (closed_solve) passes them once per CL_NB columns.
// RP > 0 (D <= 256 RP): a sub-panel's column steps run on registers (thread t
// holds rows s0 + t + 256 i, i < RP, of the CL_NB columns).  One barrier per
// step: before it each wave leaves its best pivot candidate (by a wave
// reduction) and that candidate's whole row in LDS, and thread c - s0 its row
// c; after it every thread picks the same winner, the two rows are exchanged
// from LDS and the rank-1 update runs on registers.  The swaps are recorded
// as a row map; the rows they move (none when no row is swapped) reach the
// other columns of the block at the end of the sub-panel and W at the end of
// A.  RP = 0: every column step works on S in memory and swaps the rows of W
// at once.
template <class R, bool PHI, int RP>
__global__ __launch_bounds__(256, 2) void closed_solve_b64(ClosedArgs a) {
  using C = typename Prec<R>::C;
  constexpr int RMAX = RP > 0 ? 256 * RP : 1;  // rows of the register path
  constexpr int RQ = RP > 0 ? 1 : 0;
  // two stages of the update; the first is also the staging of the other
  // products (As, Bs), the second holds L11 and its inverse of a sub-panel
  __shared__ C stages[4][CL_NB][CL_T + 1];  // As, Bs, As1, Bs1, one object so phase B can use them whole
  C (*As)[CL_T + 1] = stages[0];
  C (*Bs)[CL_T + 1] = stages[1];
  C (*As1)[CL_T + 1] = stages[2];
  C (*Bs1)[CL_T + 1] = stages[3];
  C (*Ls)[CL_NB] = reinterpret_cast<C (*)[CL_NB]>(&As1[0][0]);
  C (*Li)[CL_NB] = reinterpret_cast<C (*)[CL_NB]>(&Bs1[0][0]);
  __shared__ R red_v[256 * (1 - RQ) + 1];
  __shared__ int red_i[256 * (1 - RQ) + 1];
  __shared__ unsigned long long pk[2][4 * RQ + 1];  // per-wave pivot keys, two buffers
  __shared__ C cand[2][4 * RQ + 1][CL_NB * RQ + 1];  // each wave's candidate pivot row, two buffers
  __shared__ C crow[2][CL_NB * RQ + 1];              // row c, which the pivot row replaces
  __shared__ int piv[CL_BB * RQ + 1];          // pivot row of each column of the block
  __shared__ int map_s[RMAX], map_b[RMAX];     // row sources after the sub-panel's / block's swaps
  __shared__ int mv_dst[2 * CL_BB * RQ + 1], mv_src[2 * CL_BB * RQ + 1];
  __shared__ int mv_n;
  const int mat = blockIdx.x;
  const int D = a.D, D2 = 2 * D;
  C* W = static_cast<C*>(a.W) + (size_t)mat * 2 * D * D;
  C* S = static_cast<C*>(a.S) + (size_t)mat * CL_BB * D;
  C* Linv = static_cast<C*>(a.T) + (size_t)mat * 2 * CL_BB * CL_BB;  // column-major, leading dimension CL_BB
  C* Uinv = Linv + CL_BB * CL_BB;
  const int tid = threadIdx.x;
  const C one = mkc((R)1, (R)0), zero = mkc((R)0, (R)0);
  int info = 0;
  for (int k0 = 0; k0 < D; k0 += CL_BB) {
    const int nb = min(CL_BB, D - k0);
    // S(row, col) for col in [k0, k0 + nb): Sk[col * D + row]
    C* Sk = S - (size_t)k0 * D;
    // A. copy, then factor
    for (int u = tid; u < nb * (D - k0); u += 256) {
      const int col = k0 + u / (D - k0), row = k0 + u % (D - k0);
      Sk[(size_t)col * D + row] = W[(size_t)col * D + row];
    }
    if constexpr (RP > 0)
      for (int x = k0 + tid; x < D; x += 256) map_b[x - k0] = x;
    __syncthreads();
    for (int s0 = k0; s0 < k0 + nb; s0 += CL_NB) {
      const int sb = min(CL_NB, k0 + nb - s0);
      if constexpr (RP > 0) {
        const int lane = tid & 63, wave = tid >> 6;
        for (int x = s0 + tid; x < D; x += 256) map_s[x - s0] = x;
        C v[RP][CL_NB];
#pragma unroll
        for (int i = 0; i < RP; ++i) {
          const int row = s0 + tid + 256 * i;
#pragma unroll
          for (int j = 0; j < CL_NB; ++j) v[i][j] = (row < D && j < sb) ? Sk[(size_t)(s0 + j) * D + row] : zero;
        }
        // the CL_NB steps unrolled: the register columns take static indices
        static_for<0, CL_NB>([&](auto jc) {
          constexpr int j = decltype(jc)::value, buf = j & 1;
          if (j >= sb) return;
          const int c = s0 + j;
          C pc[RP];  // this thread's entries of column c
#pragma unroll
          for (int i = 0; i < RP; ++i) pc[i] = v[i][j];
          unsigned long long key = 0ull;
#pragma unroll
          for (int i = 0; i < RP; ++i) {
            const int row = s0 + tid + 256 * i;
            if (row >= c && row < D) {
              const unsigned long long k2 = pivot_key(fabs(pc[i].x) + fabs(pc[i].y), row - s0);
              if (k2 > key) key = k2;
            }
          }
          key = wave_max_u64(key);  // the wave's candidate
          const int bi = key ? s0 + 1023 - (int)(key & 0x3ff) : c;
          const int tb = (bi - s0) & 255, ib = (bi - s0) >> 8;
          if (tid == tb) {
#pragma unroll
            for (int i = 0; i < RP; ++i)
              if (i == ib) {
#pragma unroll
                for (int jj = 0; jj < CL_NB; ++jj) cand[buf][wave][jj] = v[i][jj];
              }
          }
          if (tid == j) {
#pragma unroll
            for (int jj = 0; jj < CL_NB; ++jj) crow[buf][jj] = v[0][jj];
          }
          if (lane == 0) pk[buf][wave] = key;
          __syncthreads();
          unsigned long long kmax = pk[buf][0];
          int pw = 0;
#pragma unroll
          for (int w = 1; w < 4; ++w) {
            const unsigned long long k2 = pk[buf][w];
            if (k2 > kmax) {
              kmax = k2;
              pw = w;
            }
          }
          const int p = s0 + 1023 - (int)(kmax & 0x3ff);
          if (!(pivot_mag(kmax) > 0)) {
            if (info == 0) info = c + 1;
            if (tid == 0) piv[c - k0] = c;
            return;
          }
          const C* prw = cand[buf][pw];
          const int tp = (p - s0) & 255, ip = (p - s0) >> 8;
          if (p != c) {
            if (tid == j) {
#pragma unroll
              for (int jj = 0; jj < CL_NB; ++jj) v[0][jj] = prw[jj];
            }
            if (tid == tp) {
#pragma unroll
              for (int i = 0; i < RP; ++i)
                if (i == ip) {
                  pc[i] = crow[buf][j];
#pragma unroll
                  for (int jj = 0; jj < CL_NB; ++jj) v[i][jj] = crow[buf][jj];
                }
            }
          }
          if (tid == 0) {
            piv[c - k0] = p;
            if (p != c) {
              const int t1 = map_s[c - s0];
              map_s[c - s0] = map_s[p - s0];
              map_s[p - s0] = t1;
              const int t2 = map_b[c - k0];
              map_b[c - k0] = map_b[p - k0];
              map_b[p - k0] = t2;
            }
          }
          const C inv = cinv(prw[j]);
#pragma unroll
          for (int i = 0; i < RP; ++i) {
            const int row = s0 + tid + 256 * i;
            if (row > c && row < D) {
              const C l = cmul(pc[i], inv);
              v[i][j] = l;
#pragma unroll
              for (int jj = j + 1; jj < CL_NB; ++jj) v[i][jj] = csub(v[i][jj], cmul(l, prw[jj]));
            }
          }
        });
#pragma unroll
        for (int i = 0; i < RP; ++i) {
          const int row = s0 + tid + 256 * i;
          if (row < D)
#pragma unroll
            for (int j = 0; j < CL_NB; ++j)
              if (j < sb) Sk[(size_t)(s0 + j) * D + row] = v[i][j];
        }
        // the sub-panel's swaps on the other columns of the block
        if (wave == 0) list_moved_rows(piv + (s0 - k0), s0, sb, map_s, s0, mv_dst, mv_src, &mv_n);
        __syncthreads();
        move_rows<R, 8>(Sk, D, k0, k0 + nb, s0, s0 + sb, mv_dst, mv_src, mv_n);
        __syncthreads();
      } else
      for (int c = s0; c < s0 + sb; ++c) {
        R best = -1;
        int bi = c;
        for (int row = c + tid; row < D; row += 256) {
          const C v = Sk[(size_t)c * D + row];
          const R m = fabs(v.x) + fabs(v.y);
          if (m > best) {
            best = m;
            bi = row;
          }
        }
        red_v[tid] = best;
        red_i[tid] = bi;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
          if (tid < st) {
            const R o = red_v[tid + st];
            const int oi = red_i[tid + st];
            if (o > red_v[tid] || (o == red_v[tid] && oi < red_i[tid])) {
              red_v[tid] = o;
              red_i[tid] = oi;
            }
          }
          __syncthreads();
        }
        const int p = red_i[0];
        const R pmax = red_v[0];
        __syncthreads();
        if (!(pmax > 0)) {  // zero or NaN pivot column: the system is singular
          if (info == 0) info = c + 1;
          continue;
        }
        if (p != c) {
          for (int col = k0 + tid; col < D2; col += 256) {
            const C x = W[(size_t)col * D + c];
            W[(size_t)col * D + c] = W[(size_t)col * D + p];
            W[(size_t)col * D + p] = x;
          }
          for (int col = k0 + tid; col < k0 + nb; col += 256) {
            const C x = Sk[(size_t)col * D + c];
            Sk[(size_t)col * D + c] = Sk[(size_t)col * D + p];
            Sk[(size_t)col * D + p] = x;
          }
        }
        __syncthreads();
        const C inv = cinv(Sk[(size_t)c * D + c]);
        for (int row = c + 1 + tid; row < D; row += 256) {
          const C l = cmul(Sk[(size_t)c * D + row], inv);
          Sk[(size_t)c * D + row] = l;
          for (int col = c + 1; col < s0 + sb; ++col)
            Sk[(size_t)col * D + row] = csub(Sk[(size_t)col * D + row], cmul(l, Sk[(size_t)col * D + c]));
        }
        __syncthreads();
      }
      if (s0 + sb < k0 + nb) {
        // the rest of the block: U12 = L11^-1 S12 on the sub-panel rows, S22 -= L21 U12 below
        for (int u = tid; u < CL_NB * CL_NB; u += 256) {
          const int i = u % CL_NB, j = u / CL_NB;
          Ls[i][j] = (i < sb && j < sb) ? Sk[(size_t)(s0 + j) * D + s0 + i] : zero;
        }
        __syncthreads();
        if (tid < CL_NB) {
          const int t = tid;
          for (int i = 0; i < CL_NB; ++i) {
            C x = i == t ? one : zero;
            if (i < sb && t < sb)
              for (int s2 = t; s2 < i; ++s2) x = csub(x, cmul(Ls[i][s2], Li[s2][t]));
            Li[i][t] = x;
          }
        }
        __syncthreads();
        panel_rows_apply<R>(Sk, D, s0, sb, s0 + sb, k0 + nb, Li);
        __syncthreads();
        closed_update_b64<R>(Sk, D, s0 + sb, D, s0 + sb, k0 + nb, s0, sb, As, Bs, As1, Bs1);
      }
    }
    if constexpr (RP > 0) {  // the block's swaps on W from column k0 on
      if (tid < 64) list_moved_rows(piv, k0, nb, map_b, k0, mv_dst, mv_src, &mv_n);
      __syncthreads();
      move_rows_waves<R, 8>(W, D, k0, D2, mv_dst, mv_src, mv_n);
    }
    // B. T = U11^-1 L11^-1 (identity-padded to CL_BB) by 16 x 16 blocks: the
    // inverses of the four diagonal blocks of L11 and U11 from copies in LDS
    // (thread = block column), then the off-diagonal blocks by their distance d
    // from the diagonal, all blocks at one distance at once,
    //   Linv_ab = -Linv_aa sum_{c=b}^{a-1} L_ac Linv_cb      (a > b)
    //   Uinv_ab = -Uinv_aa sum_{c=a+1}^{b} U_ac Uinv_cb      (a < b),
    // then T = Uinv Linv on MFMA.
    {
      constexpr int NBB = CL_BB / CL_NB;
      // L11 (unit lower) and U11 (upper) entry (i, s), identity beyond nb
      auto Lv = [&](int i, int s2) -> C {
        if (i >= nb || s2 >= nb) return i == s2 ? one : zero;
        return i > s2 ? Sk[(size_t)(k0 + s2) * D + k0 + i] : (i == s2 ? one : zero);
      };
      auto Uv = [&](int i, int s2) -> C {
        if (i >= nb || s2 >= nb) return i == s2 ? one : zero;
        return i <= s2 ? Sk[(size_t)(k0 + s2) * D + k0 + i] : zero;
      };
      C* Ld = &As[0][0];  // the NBB diagonal blocks of L11, [blk][col][row]
      C* Ud = &Bs[0][0];  // and of U11
      for (int u = tid; u < NBB * CL_NB * CL_NB; u += 256) {
        const int blk = u / (CL_NB * CL_NB), e = u % (CL_NB * CL_NB), i = e % CL_NB, j = e / CL_NB;
        Ld[u] = Lv(blk * CL_NB + i, blk * CL_NB + j);
        Ud[u] = Uv(blk * CL_NB + i, blk * CL_NB + j);
      }
      __syncthreads();
      if (tid < 2 * NBB * CL_NB) {
        const bool up = tid >= NBB * CL_NB;
        const int blk = (tid / CL_NB) % NBB, t = tid % CL_NB, off = blk * CL_NB;
        const C* Md = (up ? Ud : Ld) + blk * CL_NB * CL_NB;
        C* out = (up ? Uinv : Linv) + (size_t)(off + t) * CL_BB + off;  // column t of the diagonal block
        C x[CL_NB];  // the column, in registers (static indices)
        if (!up) {
#pragma unroll
          for (int i = 0; i < CL_NB; ++i) {
            C v = i == t ? one : zero;
#pragma unroll
            for (int s2 = 0; s2 < i; ++s2)
              if (s2 >= t) v = csub(v, cmul(Md[s2 * CL_NB + i], x[s2]));
            x[i] = v;
          }
        } else {
#pragma unroll
          for (int i = CL_NB - 1; i >= 0; --i) {
            C v = i == t ? one : zero;
#pragma unroll
            for (int s2 = i + 1; s2 < CL_NB; ++s2)
              if (s2 <= t) v = csub(v, cmul(Md[s2 * CL_NB + i], x[s2]));
            x[i] = cmul(v, cinv(Md[i * CL_NB + i]));
          }
        }
#pragma unroll
        for (int i = 0; i < CL_NB; ++i) out[i] = x[i];
      }
      // zero blocks: above the diagonal of Linv, below it of Uinv
      for (int u = tid; u < CL_BB * CL_BB; u += 256) {
        const int i = u % CL_BB, j = u / CL_BB;
        if (i / CL_NB < j / CL_NB) Linv[u] = zero;
        if (i / CL_NB > j / CL_NB) Uinv[u] = zero;
      }
      __syncthreads();
      // the blocks of one distance through LDS (the stages are free here):
      // per pair P (L pairs, then U pairs) the block row X_P (16 x w, rows
      // padded to w + 1) and block column Y_P (w x 16) of the product, w = 16 d,
      // M_P = X_P Y_P kept in registers, then M_P and the diagonal block
      // (transposed) of the triangle's inverse for -Dinv_aa M_P
      C* stg = &As[0][0];
      for (int d = 1; d < NBB; ++d) {
        const int np = NBB - d, w = CL_NB * d, per = 2 * CL_NB * w + CL_NB;
        for (int u = tid; u < 2 * np * CL_NB * w; u += 256) {
          const int pr = u / (CL_NB * w), e = u % (CL_NB * w);
          const bool upp = pr >= np;
          const int a2 = upp ? pr - np : d + pr, b2 = upp ? a2 + d : pr;
          const int c0b = upp ? a2 + 1 : b2;  // first block of the inner range
          C* X = stg + pr * per;
          C* Y = X + CL_NB * (w + 1);
          {  // X(i, kk): row a2 * 16 + i of the triangle, column c0b * 16 + kk
            const int i = e / w, kk = e % w;
            X[i * (w + 1) + kk] = upp ? Uv(a2 * CL_NB + i, c0b * CL_NB + kk) : Lv(a2 * CL_NB + i, c0b * CL_NB + kk);
          }
          {  // Y(kk, j): the inverse's row c0b * 16 + kk, column b2 * 16 + j
            const int kk = e / CL_NB, j = e % CL_NB;
            Y[kk * CL_NB + j] = (upp ? Uinv : Linv)[(size_t)(b2 * CL_NB + j) * CL_BB + c0b * CL_NB + kk];
          }
        }
        __syncthreads();
        constexpr int MP = (2 * (NBB - 1) * CL_NB * CL_NB + 255) / 256;  // pair elements per thread, at most
        C mr[MP];
#pragma unroll
        for (int q = 0; q < MP; ++q) {
          const int u = tid + 256 * q;
          mr[q] = zero;
          if (u < 2 * np * CL_NB * CL_NB) {
            const int pr = u / (CL_NB * CL_NB), e = u % (CL_NB * CL_NB), i = e % CL_NB, j = e / CL_NB;
            const C* X = stg + pr * per;
            const C* Y = X + CL_NB * (w + 1);
            R xr = 0, xi = 0;
            for (int kk = 0; kk < w; ++kk) {
              const C x = X[i * (w + 1) + kk], y = Y[kk * CL_NB + j];
              xr = fma(x.x, y.x, fma(-x.y, y.y, xr));
              xi = fma(x.x, y.y, fma(x.y, y.x, xi));
            }
            mr[q] = mkc(xr, xi);
          }
        }
        __syncthreads();
        // M_P at stg[P 512 ..], the diagonal block a of the inverse transposed after it
        for (int u = tid; u < 2 * np * CL_NB * CL_NB; u += 256) {
          const int pr = u / (CL_NB * CL_NB), e = u % (CL_NB * CL_NB), i = e % CL_NB, k = e / CL_NB;
          const bool upp = pr >= np;
          const int a2 = upp ? pr - np : d + pr;
          const C* Dg = (upp ? Uinv : Linv) + (size_t)(a2 * CL_NB) * CL_BB + a2 * CL_NB;
          stg[pr * 2 * CL_NB * CL_NB + CL_NB * CL_NB + k * CL_NB + i] = Dg[(size_t)k * CL_BB + i];
        }
#pragma unroll
        for (int q = 0; q < MP; ++q) {
          const int u = tid + 256 * q;
          if (u < 2 * np * CL_NB * CL_NB) {
            const int pr = u / (CL_NB * CL_NB), e = u % (CL_NB * CL_NB);
            stg[pr * 2 * CL_NB * CL_NB + e] = mr[q];  // M_P(i, j) at e = j 16 + i
          }
        }
        __syncthreads();
        for (int u = tid; u < 2 * np * CL_NB * CL_NB; u += 256) {
          const int pr = u / (CL_NB * CL_NB), e = u % (CL_NB * CL_NB), i = e % CL_NB, j = e / CL_NB;
          const bool upp = pr >= np;
          const int a2 = upp ? pr - np : d + pr, b2 = upp ? a2 + d : pr;
          const C* Mp = stg + pr * 2 * CL_NB * CL_NB;
          const C* Dt = Mp + CL_NB * CL_NB;
          R xr = 0, xi = 0;
          for (int k = 0; k < CL_NB; ++k) {
            const C g = Dt[k * CL_NB + i], m = Mp[j * CL_NB + k];
            xr = fma(g.x, m.x, fma(-g.y, m.y, xr));
            xi = fma(g.x, m.y, fma(g.y, m.x, xi));
          }
          (upp ? Uinv : Linv)[(size_t)(b2 * CL_NB + j) * CL_BB + a2 * CL_NB + i] = mkc(-xr, -xi);
        }
        __syncthreads();
      }
      // T = Uinv Linv, in place in Linv
      closed_apply_T_b64<R>(Linv, Uinv, CL_BB, 0, CL_BB, 0, CL_BB, As, Bs, As1, Bs1);
    }
    const int j0 = k0 + nb;
    // C. block rows of the later columns: X = T Z1
    closed_apply_T_b64<R>(W, Linv, D, k0, nb, j0, D2, As, Bs, As1, Bs1);
    // D. rows below and above the block
    closed_update_b64<R>(W, D, k0 + nb, D, j0, D2, k0, nb, As, Bs, As1, Bs1);
    closed_update_b64<R>(W, D, 0, k0, j0, D2, k0, nb, As, Bs, As1, Bs1);
  }
  if (info && tid == 0) atomicCAS(a.info, 0, info);
  if constexpr (PHI) {
    // Phi = X - Gamma
    const int r = mat / a.nqb, q = a.q0 + (mat - r * a.nqb);
    const double2* G = a.gam[r] + (size_t)D * D * q;
    double2* O = a.out[r] + (size_t)D * D * q;
    for (size_t u = tid; u < (size_t)D * D; u += 256) O[u] = csub(W[(size_t)D * D + u], G[u]);
  }
}

// ---- refinement products ------------------------------------------------------------

struct IrArgs {
  const double2* gam[4];
  double2* x[4];             // X_r, the solution, in the output arrays
  const double2* chi[2];
  const float2* inv;         // [mat][2D][D]: the inverse is columns D..2D-1
  const double2* b;          // [mat][D][D] FP64 right-hand side of an apply (Gamma when null)
  double2* rres;             // [mat][D][D] residual R
  double2* y;                // [mat][D][D] chi X
  double* part;              // [mat][tiles][2]: sum |R|^2, sum |Gamma|^2 per tile
  double* res;               // [mat] |R|_F / |Gamma|_F
  unsigned long long* rmax;  // bits of the largest res (non-negative doubles order as integers)
  int D, nf, nl, q0, nqb, tiles_n;
  double kappa[4];
};

__global__ void ir_clear(unsigned long long* rmax) { *rmax = 0ull; }

// Y = chi X: Y((l1, nu), c) = sum_l2 chi_nu(l1, l2) X((l2, nu), c), X the
// solution in the output arrays.  One workgroup per (IR_CHI_COLS columns,
// matrix); thread t takes the pairs u = t + 256 k of (nu, c), nu fastest,
// and forms all Nl rows (l1, nu) of each from the Nl entries (l2, nu) of
// column c, each read once.  grid: (ceil(D / IR_CHI_COLS), 4 nqb)
constexpr int IR_CHI_COLS = 16;
__global__ __launch_bounds__(256) void ir_chi(IrArgs a) {
  const int mat = blockIdx.y, c0 = blockIdx.x * IR_CHI_COLS;
  const int r = mat / a.nqb, q = a.q0 + (mat - r * a.nqb);
  const int D = a.D, nf = a.nf, nl = a.nl;
  const double2* X = a.x[r] + (size_t)D * D * q;
  double2* Y = a.y + (size_t)mat * D * D;
  const double2* cq = a.chi[r >> 1] + (size_t)nf * nl * nl * q;
  const int ncols = min(IR_CHI_COLS, D - c0);
  for (int u = threadIdx.x; u < nf * ncols; u += 256) {
    const int nu = u % nf, c = c0 + u / nf;
    const double2* cb = cq + (size_t)nu * nl * nl;
    const double2* Xc = X + (size_t)c * D + nu;
    double2* Yc = Y + (size_t)c * D + nu;
    double2 x[CL_NL_MAX];
#pragma unroll
    for (int l2 = 0; l2 < CL_NL_MAX; ++l2)
      if (l2 < nl) x[l2] = Xc[(size_t)l2 * nf];
    for (int l1 = 0; l1 < nl; ++l1) {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int l2 = 0; l2 < CL_NL_MAX; ++l2)
        if (l2 < nl) {
          const double2 cc = cb[l1 + nl * l2];
          sr += cc.x * x[l2].x - cc.y * x[l2].y;
          si += cc.x * x[l2].y + cc.y * x[l2].x;
        }
      Yc[(size_t)l1 * nf] = make_double2(sr, si);
    }
  }
}

// 64 x 64 tile of a batched complex product per workgroup (four waves of 2 x 2
// blocks of 16 x 16 x 4 MFMAs), K in chunks of 16 through LDS.
//   MODE 0: X  = fp64(A32^-1 fp32(B))    MODE 1: X += fp64(A32^-1 fp32(B))
//   MODE 2: R  = Gamma - X + kappa Gamma Y (FP64), with per-tile |R|^2, |Gamma|^2
// grid: tiles_n^2 * 4 nqb
template <int MODE>
__global__ __launch_bounds__(256) void ir_gemm(IrArgs a) {
  using R = typename std::conditional<MODE == 2, double, float>::type;
  using P = Prec<R>;
  using C = typename P::C;
  using V = typename P::V;
  __shared__ C As[CL_NB][CL_T + 1];
  __shared__ C Bs[CL_NB][CL_T + 1];
  __shared__ double red[2][256];
  const int tiles = a.tiles_n * a.tiles_n;
  const int mat = blockIdx.x / tiles, t = blockIdx.x - mat * tiles;
  const int tr = (t / a.tiles_n) * CL_T, tc = (t % a.tiles_n) * CL_T;
  const int r = mat / a.nqb, q = a.q0 + (mat - r * a.nqb);
  const int D = a.D;
  const size_t DD = (size_t)D * D;
  const double2* G = a.gam[r] + DD * q;
  double2* X = a.x[r] + DD * q;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const C zero = mkc((R)0, (R)0);
  V ar[2][2], ai[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) ar[i][j][e] = ai[i][j][e] = 0;
  const double2* Bsrc = MODE == 2 ? a.y + DD * mat : (a.b ? a.b + DD * mat : G);
  for (int k0 = 0; k0 < D; k0 += CL_NB) {
    for (int u = tid; u < CL_NB * CL_T; u += 256) {
      const int row = u & (CL_T - 1), ka = u >> 6;
      const int kb = u % CL_NB, col = u / CL_NB;
      const bool aok = tr + row < D && k0 + ka < D, bok = tc + col < D && k0 + kb < D;
      if (MODE == 2) {
        As[ka][row] = aok ? reinterpret_cast<const C*>(G)[(size_t)(k0 + ka) * D + tr + row] : zero;
        Bs[kb][col] = bok ? reinterpret_cast<const C*>(Bsrc)[(size_t)(tc + col) * D + k0 + kb] : zero;
      } else {
        As[ka][row] = aok ? reinterpret_cast<const C*>(a.inv + 2 * DD * mat + DD)[(size_t)(k0 + ka) * D + tr + row]
                          : zero;
        const double2 v = bok ? Bsrc[(size_t)(tc + col) * D + k0 + kb] : make_double2(0.0, 0.0);
        Bs[kb][col] = mkc((R)v.x, (R)v.y);
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < CL_NB; kk += 4) {
      C fa[2], fb[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) fa[i] = As[kk + lk][wr + i * 16 + li];
#pragma unroll
      for (int j = 0; j < 2; ++j) fb[j] = Bs[kk + lk][wc + j * 16 + li];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          ar[i][j] = P::fma_add(fa[i].x, fb[j].x, ar[i][j]);
          ar[i][j] = P::fma_sub(fa[i].y, fb[j].y, ar[i][j]);
          ai[i][j] = P::fma_add(fa[i].x, fb[j].y, ai[i][j]);
          ai[i][j] = P::fma_add(fa[i].y, fb[j].x, ai[i][j]);
        }
    }
    __syncthreads();
  }
  double rr = 0.0, gg = 0.0;
  const double kap = a.kappa[r];
  double2* Rm = a.rres + DD * mat;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = tr + wr + i * 16 + P::orow(lk, e), col = tc + wc + j * 16 + li;
        if (row >= D || col >= D) continue;
        const size_t idx = (size_t)col * D + row;
        if (MODE == 0) {
          X[idx] = make_double2((double)ar[i][j][e], (double)ai[i][j][e]);
        } else if (MODE == 1) {
          const double2 x = X[idx];
          X[idx] = make_double2(x.x + (double)ar[i][j][e], x.y + (double)ai[i][j][e]);
        } else {
          const double2 g = G[idx], x = X[idx];
          const double2 v = make_double2(g.x - x.x + kap * ar[i][j][e], g.y - x.y + kap * ai[i][j][e]);
          Rm[idx] = v;
          rr += v.x * v.x + v.y * v.y;
          gg += g.x * g.x + g.y * g.y;
        }
      }
  if (MODE == 2) {
    red[0][tid] = rr;
    red[1][tid] = gg;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) {
        red[0][tid] += red[0][tid + s];
        red[1][tid] += red[1][tid + s];
      }
      __syncthreads();
    }
    if (tid == 0) {
      a.part[2 * ((size_t)mat * tiles + t)] = red[0][0];
      a.part[2 * ((size_t)mat * tiles + t) + 1] = red[1][0];
    }
  }
}

// per matrix |R|_F / |Gamma|_F from the tile partials (fixed order), and the
// largest over the pass; grid: ceil(4 nqb / 256)
__global__ __launch_bounds__(256) void ir_check(IrArgs a, int mats) {
  const int mat = blockIdx.x * 256 + threadIdx.x;
  if (mat >= mats) return;
  const int tiles = a.tiles_n * a.tiles_n;
  double rr = 0.0, gg = 0.0;
  for (int t = 0; t < tiles; ++t) {
    rr += a.part[2 * ((size_t)mat * tiles + t)];
    gg += a.part[2 * ((size_t)mat * tiles + t) + 1];
  }
  // a NaN residual reports as infinity, so the pass never counts as converged
  const double v = gg > 0.0 ? sqrt(rr / gg) : (rr > 0.0 ? INFINITY : 0.0);
  const double s = v == v ? v : INFINITY;
  a.res[mat] = s;
  atomicMax(a.rmax, (unsigned long long)__double_as_longlong(s));
}

// Phi = X - Gamma in place (X sits in the output array); grid (ceil(D^2 / 256), 4 nqb)
__global__ __launch_bounds__(256) void ir_final(IrArgs a) {
  const int mat = blockIdx.y;
  const int r = mat / a.nqb, q = a.q0 + (mat - r * a.nqb);
  const size_t DD = (size_t)a.D * a.D;
  const size_t u = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (u >= DD) return;
  double2* X = a.x[r] + DD * q;
  const double2 g = a.gam[r][DD * q + u], x = X[u];
  X[u] = make_double2(x.x - g.x, x.y - g.y);
}

// ---- host -------------------------------------------------------------------------

// Workspace of one pass: up to 16 GiB (1024 q at D = 320 in one launch, so
// the workgroups' tail is paid once per call), at most half of the device
// memory that is free or already held by this workspace, at least 1 GiB.
static int closed_reserve(tups_ctx* c, size_t per_q, size_t extra, int* qbatch) {
  const tups_sizes& s = c->sz;
  size_t budget = (size_t)16 << 30, fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess)
    budget = std::min(budget, std::max((size_t)1 << 30, (fr + c->closed_ws_bytes) / 2));
  const size_t want = budget / per_q;
  *qbatch = (int)std::max<size_t>(1, std::min<size_t>(s.nq, want));
  const size_t need = per_q * *qbatch + extra + 256;
  if (c->closed_ws_bytes < need) {
    if (c->closed_ws) TUPS_HIP(hipFree(c->closed_ws));
    c->closed_ws = nullptr;
    c->closed_ws_bytes = 0;
    TUPS_HIP(hipMalloc(&c->closed_ws, need));
    c->closed_ws_bytes = need;
  }
  return TUPS_OK;
}

static void closed_common(tups_ctx* c, ClosedArgs& a) {
  const tups_sizes& s = c->sz;
  for (int r = 0; r < 4; ++r) {
    a.gam[r] = c->gamma[r];
    a.out[r] = c->spare[r];
  }
  a.chi[0] = c->chi[0];
  a.chi[1] = c->chi[1];
  a.info = static_cast<int*>(c->closed_ws);
  a.W = static_cast<char*>(c->closed_ws) + 256;
  a.D = s.nz;
  a.nf = s.nf;
  a.nl = s.nl;
  const double cr[4] = {0.5, 0.5, -0.25, 0.25};
  for (int r = 0; r < 4; ++r) a.kappa[r] = 2.0 * cr[r] * s.alpha;
}

// elimination variants: 0 = CL_NB-wide panels (closed_solve), 1 = CL_BB-wide
// blocks (closed_solve_b64), 2 = the same with the sub-panels on registers
// (D <= 512; above that it runs as 1); -1 = 2 from D = 128 on
static const char* const kClosedVariants[] = {"gj_panel16", "gj_block64", "gj_block64_rp"};
constexpr int kClosedRp = 2;  // rows per thread of the register sub-panel
int closed_variant_count() { return 3; }
const char* closed_variant_name(int v) { return v >= 0 && v < 3 ? kClosedVariants[v] : ""; }
static bool closed_blocked(const tups_ctx* c) {
  return c->closed_variant >= 1 || (c->closed_variant < 0 && c->sz.nz >= 128);
}
static bool closed_regs(const tups_ctx* c) {
  return closed_blocked(c) && c->closed_variant != 1 && c->sz.nz <= 256 * kClosedRp;
}
// scratch of closed_solve_b64 per matrix, in complex elements
static size_t closed_b64_elems(size_t D) { return (size_t)CL_BB * D + 2 * (size_t)CL_BB * CL_BB; }

int launch_bse_closed(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  const size_t D = s.nz;
  TUPS_CHECK(s.nl <= CL_NL_MAX, TUPS_EINVAL, "closed form: at most 9 form factors");
  const bool blk = closed_blocked(c);
  const size_t per_mat = (2 * D * D + (blk ? closed_b64_elems(D) : 0)) * sizeof(double2);
  int qbatch = 0;
  if (int rc = closed_reserve(c, 4 * per_mat, 0, &qbatch)) return rc;
  ClosedArgs a{};
  closed_common(c, a);
  if (blk) {
    const size_t mats_max = 4 * (size_t)qbatch;
    a.S = static_cast<double2*>(a.W) + mats_max * 2 * D * D;
    a.T = static_cast<double2*>(a.S) + mats_max * CL_BB * D;
  }
  hipLaunchKernelGGL(closed_clear, dim3(1), dim3(1), 0, c->stream, a.info);
  TUPS_HIP(hipGetLastError());
  for (int q0 = 0; q0 < s.nq; q0 += qbatch) {
    a.q0 = q0;
    a.nqb = std::min(qbatch, s.nq - q0);
    const unsigned mats = 4u * (unsigned)a.nqb;
    if (int rc = timer_begin(c)) return rc;
    hipLaunchKernelGGL((closed_build_nu<double2, false>), dim3((unsigned)s.nf, mats), dim3(256), 0, c->stream, a);
    TUPS_HIP(hipGetLastError());
    if (blk && closed_regs(c))
      hipLaunchKernelGGL((closed_solve_b64<double, true, kClosedRp>), dim3(mats), dim3(256), 0, c->stream, a);
    else if (blk)
      hipLaunchKernelGGL((closed_solve_b64<double, true, 0>), dim3(mats), dim3(256), 0, c->stream, a);
    else
      hipLaunchKernelGGL((closed_solve<double, true>), dim3(mats), dim3(256), 0, c->stream, a);
    TUPS_HIP(hipGetLastError());
    if (int rc = timer_end(c, TUPS_K_BSE_CLOSED)) return rc;
  }
  int info = 0;
  TUPS_HIP(hipMemcpyAsync(&info, a.info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  TUPS_CHECK(info == 0, TUPS_ENONFINITE, "closed-form BSE: singular I - kappa Gamma chi (zero pivot column)");
  c->ir_steps = 0;
  c->ir_residual = NAN;
  return TUPS_OK;
}

int launch_bse_closed_ir(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  TUPS_CHECK(s.nl <= CL_NL_MAX, TUPS_EINVAL, "closed form: at most 9 form factors");
  const size_t D = s.nz, DD = D * D;
  const int tiles_n = (int)((D + CL_T - 1) / CL_T);
  // per matrix: [A32 | I] (2 D^2 float2), R and chi X (D^2 double2 each), tile partials, res
  const bool blk = closed_blocked(c);
  const size_t per_mat = 2 * DD * sizeof(float2) + 2 * DD * sizeof(double2) +
                         (size_t)tiles_n * tiles_n * 2 * sizeof(double) + sizeof(double) + 64 +
                         (blk ? closed_b64_elems(D) * sizeof(float2) : 0);
  int qbatch = 0;
  if (int rc = closed_reserve(c, 4 * per_mat, 256, &qbatch)) return rc;
  ClosedArgs a{};
  closed_common(c, a);
  IrArgs b{};
  for (int r = 0; r < 4; ++r) {
    b.gam[r] = c->gamma[r];
    b.x[r] = c->spare[r];
    b.kappa[r] = a.kappa[r];
  }
  b.chi[0] = c->chi[0];
  b.chi[1] = c->chi[1];
  b.D = a.D;
  b.nf = a.nf;
  b.nl = a.nl;
  b.tiles_n = tiles_n;
  const size_t mats_max = 4 * (size_t)qbatch;
  char* base = static_cast<char*>(a.W);
  b.inv = reinterpret_cast<const float2*>(base);
  base += mats_max * 2 * DD * sizeof(float2);
  b.rres = reinterpret_cast<double2*>(base);
  base += mats_max * DD * sizeof(double2);
  b.y = reinterpret_cast<double2*>(base);
  base += mats_max * DD * sizeof(double2);
  b.part = reinterpret_cast<double*>(base);
  base += mats_max * tiles_n * tiles_n * 2 * sizeof(double);
  b.res = reinterpret_cast<double*>(base);
  base += mats_max * sizeof(double);
  b.rmax = reinterpret_cast<unsigned long long*>((reinterpret_cast<uintptr_t>(base) + 63) / 64 * 64);
  if (blk) {  // FP32 scratch of the block elimination after the counter
    a.S = reinterpret_cast<float2*>(reinterpret_cast<char*>(b.rmax) + 64);
    a.T = static_cast<float2*>(a.S) + mats_max * CL_BB * D;
  }
  hipLaunchKernelGGL(closed_clear, dim3(1), dim3(1), 0, c->stream, a.info);
  TUPS_HIP(hipGetLastError());
  int steps_max = 0;
  double res_max = 0.0;
  for (int q0 = 0; q0 < s.nq; q0 += qbatch) {
    a.q0 = b.q0 = q0;
    a.nqb = b.nqb = std::min(qbatch, s.nq - q0);
    const unsigned mats = 4u * (unsigned)a.nqb;
    const unsigned gblk = (unsigned)(tiles_n * tiles_n) * mats;
    if (int rc = timer_begin(c)) return rc;
    hipLaunchKernelGGL((closed_build_nu<float2, true>), dim3((unsigned)s.nf, mats), dim3(256), 0, c->stream, a);
    TUPS_HIP(hipGetLastError());
    if (blk && closed_regs(c))
      hipLaunchKernelGGL((closed_solve_b64<float, false, kClosedRp>), dim3(mats), dim3(256), 0, c->stream, a);
    else if (blk)
      hipLaunchKernelGGL((closed_solve_b64<float, false, 0>), dim3(mats), dim3(256), 0, c->stream, a);
    else
      hipLaunchKernelGGL((closed_solve<float, false>), dim3(mats), dim3(256), 0, c->stream, a);
    TUPS_HIP(hipGetLastError());
    b.b = nullptr;  // X = A32^-1 Gamma
    hipLaunchKernelGGL(ir_gemm<0>, dim3(gblk), dim3(256), 0, c->stream, b);
    TUPS_HIP(hipGetLastError());
    int steps = 0;
    double res = INFINITY;
    for (;;) {
      hipLaunchKernelGGL(ir_chi, dim3((unsigned)((D + IR_CHI_COLS - 1) / IR_CHI_COLS), mats), dim3(256), 0,
                         c->stream, b);
      TUPS_HIP(hipGetLastError());
      hipLaunchKernelGGL(ir_clear, dim3(1), dim3(1), 0, c->stream, b.rmax);
      TUPS_HIP(hipGetLastError());
      hipLaunchKernelGGL(ir_gemm<2>, dim3(gblk), dim3(256), 0, c->stream, b);
      TUPS_HIP(hipGetLastError());
      hipLaunchKernelGGL(ir_check, dim3((mats + 255) / 256), dim3(256), 0, c->stream, b, (int)mats);
      TUPS_HIP(hipGetLastError());
      unsigned long long bits = 0;
      TUPS_HIP(hipMemcpyAsync(&bits, b.rmax, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
      TUPS_HIP(hipStreamSynchronize(c->stream));
      std::memcpy(&res, &bits, sizeof(res));
      if (res <= c->ir_tol || steps >= c->ir_max_steps) break;
      b.b = b.rres;  // X += A32^-1 R
      hipLaunchKernelGGL(ir_gemm<1>, dim3(gblk), dim3(256), 0, c->stream, b);
      TUPS_HIP(hipGetLastError());
      ++steps;
    }
    hipLaunchKernelGGL(ir_final, dim3((unsigned)((DD + 255) / 256), mats), dim3(256), 0, c->stream, b);
    TUPS_HIP(hipGetLastError());
    if (int rc = timer_end(c, TUPS_K_BSE_IR)) return rc;
    steps_max = std::max(steps_max, steps);
    res_max = std::max(res_max, res);
    if (!(res <= c->ir_tol)) break;
  }
  int info = 0;
  TUPS_HIP(hipMemcpyAsync(&info, a.info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  c->ir_steps = steps_max;
  c->ir_residual = res_max;
  TUPS_CHECK(info == 0, TUPS_ENONFINITE, "refined closed-form BSE: singular fp32(I - kappa Gamma chi) (zero pivot)");
  TUPS_CHECK(res_max <= c->ir_tol, TUPS_ENOCONV,
             "refined closed-form BSE: the residual did not reach the tolerance within the step limit");
  return TUPS_OK;
}

}  // namespace tups
