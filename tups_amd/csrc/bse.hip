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
// on v_mfma_f64_16x16x4f64, accumulators kept separate for Re and Im.
//
// The chi mixing is fused: the inner dimension is walked in chunks of G
// fermionic frequencies times all Nl form factors, so a loader thread that owns
// one (row, nu) pair of the A operand holds all Nl values that the Nl x Nl
// block chi_nu couples and applies it in registers before the tile goes to LDS:
//   part 0: A = Gamma alpha chi (:, S),  B = F(S, :)
//   part 1: A = F alpha chi (:, S),      B = Gamma(S, :)
// ((Gamma chi) F = Gamma (chi F) and F (chi Gamma) = (F chi) Gamma.)  The
// frequency of an A pair is uniform across a wave, so chi_nu comes through the
// scalar cache.  On gfx950 no VALU instruction issues while an F64 MFMA runs on
// the SIMD, so the loop keeps VALU work minimal: -Im*Im uses the MFMA negate
// modifier.  Operand tiles are stored as split Re/Im planes so each MFMA
// operand is one double per lane.  Global loads of chunk c+1 are issued before
// the MFMAs of chunk c (register prefetch), and the damped mixing is fused
// into the epilogue.
#include "internal.h"

namespace tups {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64;
constexpr int BN = 64;

struct BseArgs {
  const double2* gam[4];   // Gamma_r, (D, D, nq)
  const double2* fv[4];    // F_r
  double2* out[4];         // Phi_r (must not alias gam)
  const double2* chi[2];   // (Nl, Nl, Nf, nq): ph, pp
  int D, nf, nq, tiles_n;
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

// K3: the complex product.  0: four real MFMAs per complex block.  1 and 2:
// the three-multiplication form P1 = Re A Re B, P2 = Im A Im B,
// P3 = (Re A + Im A)(Re B + Im B), Re = P1 - P2, Im = P3 - P1 - P2, with the
// sums Re + Im held in a third LDS plane (1) or formed in registers from the
// Re and Im planes as the operands are read (2: two VALU adds per operand
// instead of LDS space, which keeps large Nl at two workgroups per CU).
template <int NL, int G, bool DB, int K3>
constexpr int ladder_lds_doubles() {
  return (DB ? 2 : 1) * (K3 == 1 ? 3 : 2) * NL * G * (BM + BN);
}

// One 64 x 64 tile of Phi for one (q, r) matrix: Gm, Fm, O are that matrix's
// Gamma, F and output (D x D, column-major), chiq its bubble blocks, lds a
// workgroup buffer of ladder_lds_doubles<...>() doubles.
template <int NL, int G, bool DB, int K3 = 0>
__device__ __forceinline__ void ladder_tile(const double2* __restrict__ Gm, const double2* __restrict__ Fm,
                                            double2* __restrict__ O, const double2* __restrict__ chiq, int D,
                                            int nf, double alpha, double cf, double omf, int row0, int col0,
                                            double* lds) {
  constexpr int KC = NL * G;   // inner-dimension chunk: G frequencies x NL form factors
  constexpr int PP = G / 4;    // (row|col, nu) pairs per thread: 64 * G / 256
  constexpr int NB = DB ? 2 : 1;
  constexpr int NPL = K3 == 1 ? 3 : 2;  // LDS planes per operand: Re, Im (, Re + Im)
  static_assert(KC % 4 == 0, "chunk must be a multiple of the MFMA k");
  auto As = reinterpret_cast<double (*)[NPL][KC][BM]>(lds);
  auto Bs = reinterpret_cast<double (*)[NPL][KC][BN]>(lds + NB * NPL * KC * BM);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = (wave >> 1) * 32;
  const int wc = (wave & 1) * 32;
  const int li = lane & 15;
  const int lk = lane >> 4;

  // K3: acc_re = P1, acc_im = P3, acc_p2 = P2
  d4 acc_re[2][2], acc_im[2][2], acc_p2[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc_re[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      acc_im[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      acc_p2[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }

  const int nch = (nf + G - 1) / G;
  const int niter = 2 * nch;
  double2 ra[PP][NL], rb[PP][NL];

  // raw global loads of chunk `it` (A pairs: (row i, nu); B pairs: (nu, column j))
  auto load = [&](int it) {
    const int part = it >= nch ? 1 : 0;
    const int nu0 = (it - part * nch) * G;
    const double2* Asrc = part ? Fm : Gm;
    const double2* Bsrc = part ? Gm : Fm;
#pragma unroll
    for (int p = 0; p < PP; ++p) {
      const int u = tid + 256 * p;
      const int i = u & 63, g = u >> 6;
      const int nu = nu0 + g, gi = row0 + i;
      const bool ok = gi < D && nu < nf;
#pragma unroll
      for (int l = 0; l < NL; ++l)
        ra[p][l] = ok ? Asrc[(size_t)(l * nf + nu) * D + gi] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int p = 0; p < PP; ++p) {
      const int u = tid + 256 * p;
      const int g = u % G, j = u / G;
      const int nu = nu0 + g, gj = col0 + j;
      const bool ok = gj < D && nu < nf;
#pragma unroll
      for (int l = 0; l < NL; ++l)
        rb[p][l] = ok ? Bsrc[(size_t)gj * D + l * nf + nu] : make_double2(0.0, 0.0);
    }
  };

  // chi mixing of the A pairs in registers, then store chunk `it` to LDS (k index = l*G + g)
  auto store = [&](int it) {
    const int part = it >= nch ? 1 : 0;
    const int sb = DB ? (it & 1) : 0;
    const int nu0 = (it - part * nch) * G;
#pragma unroll
    for (int p = 0; p < PP; ++p) {
      const int u = tid + 256 * p;
      const int i = u & 63;
      // g = u >> 6 is the wave index (+ 4p): uniform, so chi_nu is a scalar load
      const int g = __builtin_amdgcn_readfirstlane(u >> 6);
      const int nu = nu0 + g < nf ? nu0 + g : nf - 1;
      // chi is read-only for the whole launch: reading it through the constant
      // address space keeps these uniform loads on the scalar path even where
      // the compiler cannot rule out aliasing with earlier stores (a kernel
      // that loops over several tiles)
      const __attribute__((address_space(4))) double* cb =
          (const __attribute__((address_space(4))) double*)(chiq + (size_t)nu * NL * NL);
#pragma unroll
      for (int l2 = 0; l2 < NL; ++l2) {
        // (X alpha chi)(i, (l2, nu)) = alpha sum_l3 X(i, (l3, nu)) chi_nu(l3, l2)
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int l3 = 0; l3 < NL; ++l3) {
          const double xr = cb[2 * (l3 + NL * l2)], xi = cb[2 * (l3 + NL * l2) + 1];
          sr += ra[p][l3].x * xr - ra[p][l3].y * xi;
          si += ra[p][l3].x * xi + ra[p][l3].y * xr;
        }
        As[sb][0][l2 * G + g][i] = alpha * sr;
        As[sb][1][l2 * G + g][i] = alpha * si;
        if (K3 == 1) As[sb][NPL - 1][l2 * G + g][i] = alpha * sr + alpha * si;
      }
    }
#pragma unroll
    for (int p = 0; p < PP; ++p) {
      const int u = tid + 256 * p;
      const int g = u % G, j = u / G;
#pragma unroll
      for (int l1 = 0; l1 < NL; ++l1) {
        Bs[sb][0][l1 * G + g][j] = rb[p][l1].x;
        Bs[sb][1][l1 * G + g][j] = rb[p][l1].y;
        if (K3 == 1) Bs[sb][NPL - 1][l1 * G + g][j] = rb[p][l1].x + rb[p][l1].y;
      }
    }
  };

  // 16 x 16 blocks of this wave that hold some row and column below D (uniform:
  // a partial tile's blocks beyond D skip their MFMAs)
  bool blk_ok[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) blk_ok[i][j] = row0 + wr + 16 * i < D && col0 + wc + 16 * j < D;
  auto mfma_chunk = [&](int buf) {
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int kr = ks * 4 + lk;
      double ar[2], ai[2], br[2], bi[2], as[2], bs[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = As[buf][0][kr][wr + i * 16 + li];
        ai[i] = As[buf][1][kr][wr + i * 16 + li];
        br[i] = Bs[buf][0][kr][wc + i * 16 + li];
        bi[i] = Bs[buf][1][kr][wc + i * 16 + li];
        if (K3 == 1) {
          as[i] = As[buf][NPL - 1][kr][wr + i * 16 + li];
          bs[i] = Bs[buf][NPL - 1][kr][wc + i * 16 + li];
        } else if (K3 == 2) {
          as[i] = ar[i] + ai[i];
          bs[i] = br[i] + bi[i];
        }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (!blk_ok[i][j]) continue;
          if (K3) {
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], acc_re[i][j], 0, 0, 0);
            acc_p2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bi[j], acc_p2[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[i], bs[j], acc_im[i][j], 0, 0, 0);
          } else {
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], acc_re[i][j], 0, 0, 0);
            // blgp bit 0 negates the A operand of an F64 MFMA: acc -= Im(A) Im(B)
            acc_re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bi[j], acc_re[i][j], 0, 0, 1);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], acc_im[i][j], 0, 0, 0);
            acc_im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], acc_im[i][j], 0, 0, 0);
          }
        }
    }
  };

  load(0);
  if (DB) {
    // two LDS buffers, one barrier per chunk: the mixing and LDS stores of
    // chunk it+1 follow the MFMAs of chunk it in the same instruction stream
    store(0);
    __syncthreads();
    for (int it = 0; it < niter; ++it) {
      if (it + 1 < niter) load(it + 1);
      mfma_chunk(it & 1);
      if (it + 1 < niter) store(it + 1);
      __syncthreads();
    }
  } else {
    for (int it = 0; it < niter; ++it) {
      store(it);
      __syncthreads();
      if (it + 1 < niter) load(it + 1);
      mfma_chunk(0);
      __syncthreads();
    }
  }

  // epilogue: Phi = c f T + (1 - f)(F - Gamma).  v_mfma_f64_16x16x4f64 leaves
  // element e of lane l at block row (l / 16) + 4 e, column l % 16.
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
          const double2 f = Fm[idx];
          const double2 g = Gm[idx];
          double tr = acc_re[i][j][e], ti = acc_im[i][j][e];
          if (K3) {
            ti = ti - tr - acc_p2[i][j][e];
            tr = tr - acc_p2[i][j][e];
          }
          O[idx] = make_double2(cf * tr + omf * (f.x - g.x), cf * ti + omf * (f.y - g.y));
        }
      }
    }
}

template <int NL, int G, bool DB, int K3 = 0>
__global__ __launch_bounds__(256, 2) void bse_ladder_kernel(BseArgs a) {
  __shared__ double lds[ladder_lds_doubles<NL, G, DB, K3>()];
  const int tiles = a.tiles_n * a.tiles_n;
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int mat = logical / tiles;
  const int t = logical - mat * tiles;
  const int r = mat / a.nq;
  const int q = mat - r * a.nq;
  const size_t DD = (size_t)a.D * a.D;
  ladder_tile<NL, G, DB, K3>(a.gam[r] + DD * q, a.fv[r] + DD * q, a.out[r] + DD * q,
                             a.chi[r >> 1] + (size_t)a.nf * NL * NL * q, a.D, a.nf, a.alpha, a.cf[r], a.omf,
                             (t / a.tiles_n) * BM, (t % a.tiles_n) * BN, lds);
}

typedef void (*BseKernel)(BseArgs);

struct BseVariant {
  int nl;
  BseKernel fn;
  const char* name;
  bool four() const {  // the four-product complex form (the reference's ZGEMM arithmetic)
    const size_t n = std::char_traits<char>::length(name);
    return !(n > 3 && name[n - 3] == '_' && name[n - 2] == '3');
  }
};

// tile variants per form-factor count: frequency chunk G and LDS double buffering
static const BseVariant kBseVariants[] = {
    {1, bse_ladder_kernel<1, 16, false>, "nl1_g16"},
    {1, bse_ladder_kernel<1, 16, true>, "nl1_g16_db"},
    {1, bse_ladder_kernel<1, 32, false>, "nl1_g32"},
    {5, bse_ladder_kernel<5, 4, false>, "nl5_g4"},
    {5, bse_ladder_kernel<5, 4, true>, "nl5_g4_db"},
    {5, bse_ladder_kernel<5, 8, false>, "nl5_g8"},
    {9, bse_ladder_kernel<9, 4, false>, "nl9_g4"},
    {9, bse_ladder_kernel<9, 4, true>, "nl9_g4_db"},
    {1, bse_ladder_kernel<1, 16, false, true>, "nl1_g16_3m"},
    {5, bse_ladder_kernel<5, 4, false, true>, "nl5_g4_3m"},
    {5, bse_ladder_kernel<5, 8, false, true>, "nl5_g8_3m"},
    {9, bse_ladder_kernel<9, 4, false, true>, "nl9_g4_3m"},
    {1, bse_ladder_kernel<1, 16, false, 2>, "nl1_g16_3r"},
    {5, bse_ladder_kernel<5, 4, false, 2>, "nl5_g4_3r"},
    {9, bse_ladder_kernel<9, 4, false, 2>, "nl9_g4_3r"},
};
constexpr int kNumBseVariants = sizeof(kBseVariants) / sizeof(kBseVariants[0]);

static int default_variant(int nl) {
  for (int v = 0; v < kNumBseVariants; ++v)
    if (kBseVariants[v].nl == nl) return v;
  return -1;
}

static BseArgs bse_args(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  BseArgs a{};
  for (int r = 0; r < 4; ++r) {
    a.gam[r] = c->gamma[r];
    a.fv[r] = c->fvert[r];
    a.out[r] = c->spare[r];
  }
  a.chi[0] = c->chi[0];
  a.chi[1] = c->chi[1];
  a.D = s.nz;
  a.nf = s.nf;
  a.nq = s.nq;
  a.tiles_n = (s.nz + BM - 1) / BM;
  a.alpha = s.alpha;
  const double cr[4] = {0.5, 0.5, -0.25, 0.25};
  for (int r = 0; r < 4; ++r) a.cf[r] = cr[r] * c->f_damping;
  a.omf = 1.0 - c->f_damping;
  return a;
}

static int launch_variant(tups_ctx* c, int v, const BseArgs& a) {
  const long long nblk = (long long)a.tiles_n * a.tiles_n * 4 * a.nq;
  TUPS_CHECK(nblk < (1LL << 31), TUPS_EINVAL, "BSE grid too large");
  hipLaunchKernelGGL(kBseVariants[v].fn, dim3((unsigned)nblk), dim3(256), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  return TUPS_OK;
}

int launch_bse_ladder(tups_ctx* c) {
  int v = c->bse_variant;
  if (v < 0 || v >= kNumBseVariants || kBseVariants[v].nl != c->sz.nl) v = default_variant(c->sz.nl);
  TUPS_CHECK(v >= 0, TUPS_EINVAL, "no BSE kernel for this Nl");
  const BseArgs a = bse_args(c);
  if (int rc = timer_begin(c)) return rc;
  if (int rc = launch_variant(c, v, a)) return rc;
  if (int rc = timer_end(c, TUPS_K_BSE_LADDER)) return rc;
  return TUPS_OK;
}

// Times every variant for this context's shape on its resident inputs (the
// outputs go to the ping-pong buffer and are discarded) and keeps the fastest.
int autotune_bse(tups_ctx* c, int reps, int* chosen, float* best_ms) {
  const BseArgs a = bse_args(c);
  float best = 0.f;
  int arg = -1;
  for (int v = 0; v < kNumBseVariants; ++v) {
    if (kBseVariants[v].nl != c->sz.nl) continue;
    if (!c->three_mult && !kBseVariants[v].four()) continue;
    if (int rc = launch_variant(c, v, a)) return rc;  // warm-up
    TUPS_HIP(hipEventRecord(c->tev[0], c->stream));
    for (int i = 0; i < reps; ++i)
      if (int rc = launch_variant(c, v, a)) return rc;
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
  TUPS_CHECK(arg >= 0, TUPS_EINVAL, "no BSE kernel for this Nl");
  c->bse_variant = arg;
  if (chosen) *chosen = arg;
  if (best_ms) *best_ms = best;
  return TUPS_OK;
}

int bse_variant_count() { return kNumBseVariants; }

int bse_variant_nl(int v) { return v >= 0 && v < kNumBseVariants ? kBseVariants[v].nl : -1; }

const char* bse_variant_name(int v) {
  return v >= 0 && v < kNumBseVariants ? kBseVariants[v].name : "";
}

}  // namespace tups

namespace tups {

// ---- batching form: q points with their own frequency box and form-factor count ----------
//
// No reference counterpart: the reference fixes Nf and Nl for every q.  A batch
// holds nq points with Nf_q (even) and Nl_q in {1, 5, 9}, D_q = Nl_q Nf_q; the
// vertex arrays are packed q after q (D_q x D_q column-major each) and the
// bubbles as (Nl_q, Nl_q, Nf_q) blocks.  One persistent launch covers every
// (q, channel, 64 x 64 tile): the work items are sorted by D_q, largest first,
// and workgroup b takes items b, b + grid, ...; each item runs the ladder tile
// of its Nl (the same code as the uniform path).

namespace {
struct BatchArgs {
  const double2* gam[4];
  const double2* fv[4];
  double2* out[4];
  const double2* chi[2];
  const int4* items;        // 8 lists of len entries (q, r, row0, col0); q < 0 pads
  const long long* voff;    // element offset of q's D_q x D_q block
  const long long* coff;    // element offset of q's bubble blocks
  const int* nf;
  const int* nl;
  int* tickets;             // per-list counters of the dynamic order (zero at launch)
  int len;                  // entries per list
  int dyn;                  // 1: entries handed out by the per-list counters
  double alpha;
  double cf[4];
  double omf;
};
}  // namespace

template <int K1, int K5, int K9>
constexpr int batch_lds_doubles() {
  constexpr int a = ladder_lds_doubles<1, 16, false, K1>();
  constexpr int b = ladder_lds_doubles<5, 4, false, K5>();
  constexpr int c = ladder_lds_doubles<9, 4, false, K9>();
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

// workgroups per CU that the LDS footprint allows (160 KB per CU); the register
// budget follows: 256 VGPRs per lane at two, the full 512 (AGPRs included) at one
template <int K1, int K5, int K9>
constexpr int batch_occupancy() {
  return 8 * batch_lds_doubles<K1, K5, K9>() <= 80 * 1024 ? 2 : 1;
}

__device__ __forceinline__ long long uniform64(long long v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(v & 0xffffffffLL));
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// The work items come as 8 lists, one per XCD, each holding whole (q, r)
// matrices largest first, balanced by estimated work.  Workgroup b --
// dispatched to XCD x = b % 8 as that XCD's j-th workgroup, j = b / 8 -- works
// on list x, so the tiles an XCD runs at one time belong to one or two
// matrices and share that XCD's L2: entries j, j + grid/8, ... (grid/8 = len
// gives one entry per workgroup and leaves the balancing to the hardware
// dispatcher), or, with dyn set, whichever entry the list's counter hands out
// next (a resident workgroup that finishes early takes more).  The grid is a
// multiple of 8.  K1, K5, K9: the complex-product form per form-factor count.
template <int K1, int K5, int K9>
__global__ __launch_bounds__(256, (batch_occupancy<K1, K5, K9>())) void bse_batch_kernel(BatchArgs a) {
  extern __shared__ double lds[];
  __shared__ int ticket;
  const int x = blockIdx.x & 7, j = blockIdx.x >> 3;
  const int first = x * a.len + j;
  const int end = (x + 1) * a.len;
  const int step = (int)(gridDim.x >> 3);
  for (int it = first;; it += step) {
    if (a.dyn) {
      if (threadIdx.x == 0) ticket = atomicAdd(a.tickets + x, 1);
      __syncthreads();
      it = x * a.len + ticket;
      __syncthreads();  // every thread has read the ticket before the next one is drawn
      if (it >= end) break;
    } else if (it >= end) {
      break;
    }
    // the item's values are workgroup-uniform: keep them (and the addresses
    // built from them) in scalar registers, off the VALU of the MFMA loop
    int4 w = a.items[it];
    w.x = __builtin_amdgcn_readfirstlane(w.x);
    if (w.x < 0) continue;  // list padding
    w.y = __builtin_amdgcn_readfirstlane(w.y);
    w.z = __builtin_amdgcn_readfirstlane(w.z);
    w.w = __builtin_amdgcn_readfirstlane(w.w);
    const int q = w.x, r = w.y;
    const int nf = __builtin_amdgcn_readfirstlane(a.nf[q]), nl = __builtin_amdgcn_readfirstlane(a.nl[q]);
    const long long vo = uniform64(a.voff[q]), co = uniform64(a.coff[q]);
    const double2* g = a.gam[r] + vo;
    const double2* f = a.fv[r] + vo;
    double2* o = a.out[r] + vo;
    const double2* ch = a.chi[r >> 1] + co;
    if (nl == 1)
      ladder_tile<1, 16, false, K1>(g, f, o, ch, nf, nf, a.alpha, a.cf[r], a.omf, w.z, w.w, lds);
    else if (nl == 5)
      ladder_tile<5, 4, false, K5>(g, f, o, ch, 5 * nf, nf, a.alpha, a.cf[r], a.omf, w.z, w.w, lds);
    else
      ladder_tile<9, 4, false, K9>(g, f, o, ch, 9 * nf, nf, a.alpha, a.cf[r], a.omf, w.z, w.w, lds);
    __syncthreads();  // the next item reuses the LDS buffer
  }
}

__global__ void batch_tickets_clear(int* t) {
  if (threadIdx.x < 8) t[threadIdx.x] = 0;
}

struct BatchVariant {
  void (*fn)(BatchArgs);
  int lds_bytes;
  int per_cu;  // resident workgroups per CU
  const char* name;
};

// tile choice per form-factor count (Nl = 1 always runs G = 16, four products)
static const BatchVariant kBatchKernels[] = {
#define TUPS_BATCH_VARIANT(K1, K5, K9, name) \
  {bse_batch_kernel<K1, K5, K9>, 8 * batch_lds_doubles<K1, K5, K9>(), batch_occupancy<K1, K5, K9>(), name}
    TUPS_BATCH_VARIANT(0, 0, 0, "nl1_4m+nl5_4m+nl9_4m"),
    TUPS_BATCH_VARIANT(0, 1, 0, "nl1_4m+nl5_3m+nl9_4m"),
    TUPS_BATCH_VARIANT(0, 1, 2, "nl1_4m+nl5_3m+nl9_3r"),
    TUPS_BATCH_VARIANT(2, 2, 2, "nl1_3r+nl5_3r+nl9_3r"),
    TUPS_BATCH_VARIANT(0, 1, 1, "nl1_4m+nl5_3m+nl9_3m"),
#undef TUPS_BATCH_VARIANT
};
constexpr int kNumBatchKernels = sizeof(kBatchKernels) / sizeof(kBatchKernels[0]);

// variant id v = 3 * kernel + order: 0 resident grid taking its list's
// entries in static turns; 1 one workgroup per entry; 2 resident grid drawing
// entries from the per-list counters
constexpr int kBatchOrders = 3;
int batch_variant_count() { return kBatchOrders * kNumBatchKernels; }

const char* batch_variant_name(int v) {
  static const char* order[kBatchOrders] = {"_static", "_flat", "_dyn"};
  static std::string names[kBatchOrders * kNumBatchKernels];
  if (v < 0 || v >= batch_variant_count()) return "";
  if (names[v].empty()) names[v] = std::string(kBatchKernels[v / kBatchOrders].name) + order[v % kBatchOrders];
  return names[v].c_str();
}

}  // namespace tups

using namespace tups;

extern "C" int tups_batch_destroy(tups_batch* b) {
  if (!b) return TUPS_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  for (int r = 0; r < 4; ++r) {
    f(b->gamma[r]);
    f(b->fvert[r]);
    f(b->spare[r]);
  }
  f(b->chi[0]);
  f(b->chi[1]);
  f(b->d_items);
  f(b->d_voff);
  f(b->d_coff);
  f(b->d_nf);
  f(b->d_nl);
  f(b->d_tickets);
  f(b->d_qlist);
  f(b->split_ws);
  if (b->stream && b->own_stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return TUPS_OK;
}

extern "C" int tups_batch_create(int32_t nx, int32_t nq, const int32_t* nf, const int32_t* nl, double beta,
                                 double f_damping, tups_batch** out) {
  TUPS_CHECK(out && nf && nl, TUPS_EINVAL, "null argument");
  TUPS_CHECK(nx >= 2 && nq >= 1 && beta > 0.0, TUPS_EINVAL, "bad batch dimensions");
  for (int q = 0; q < nq; ++q) {
    TUPS_CHECK(nf[q] >= 2 && nf[q] % 2 == 0, TUPS_EINVAL, "Nf_q must be even and >= 2");
    TUPS_CHECK(nl[q] == 1 || nl[q] == 5 || nl[q] == 9, TUPS_EINVAL, "Nl_q must be 1, 5 or 9");
  }
  tups_batch* b = new tups_batch();
  b->nx = nx;
  b->nq = nq;
  b->alpha = 1.0 / (beta * nx * nx);
  b->f_damping = f_damping;
  b->nf.assign(nf, nf + nq);
  b->nl.assign(nl, nl + nq);
  b->voff.assign(nq + 1, 0);
  b->coff.assign(nq + 1, 0);
  for (int q = 0; q < nq; ++q) {
    const long long D = (long long)nl[q] * nf[q];
    b->voff[q + 1] = b->voff[q] + D * D;
    b->coff[q + 1] = b->coff[q] + (long long)nl[q] * nl[q] * nf[q];
  }
  // work lists: (q, r) matrices largest D first, each to the list with the
  // least work so far (work of a matrix ~ tiles x inner length = tn^2 D)
  std::vector<int> order(nq);
  for (int q = 0; q < nq; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return nl[x] * nf[x] > nl[y] * nf[y]; });
  std::vector<int4> lists[8];
  long long load[8] = {};
  long long total = 0;
  for (int q : order) {
    const int D = nl[q] * nf[q], tn = (D + BM - 1) / BM;
    for (int r = 0; r < 4; ++r) {
      const int x = (int)(std::min_element(load, load + 8) - load);
      load[x] += (long long)tn * tn * D;
      for (int t = 0; t < tn * tn; ++t) lists[x].push_back(make_int4(q, r, (t / tn) * BM, (t % tn) * BN));
      total += tn * tn;
    }
  }
  size_t len = 0;
  for (auto& l : lists) len = std::max(len, l.size());
  if (len * 8 >= (size_t)1 << 31) {
    delete b;
    set_error("batch too large for one launch");
    return TUPS_EINVAL;
  }
  std::vector<int4> items;
  items.reserve(8 * len);
  for (auto& l : lists) {
    items.insert(items.end(), l.begin(), l.end());
    items.insert(items.end(), len - l.size(), make_int4(-1, 0, 0, 0));
  }
  b->nitems = (int)total;
  b->len = (int)len;
  auto fail = [&](int rc) {
    tups_batch_destroy(b);
    return rc;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(TUPS_EHIP);
  }
  auto alloc = [&](void** p, size_t bytes) {
    if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) {
      set_error("hipMalloc failed for the batch");
      return TUPS_ENOMEM;
    }
    return (int)TUPS_OK;
  };
  const size_t vb = (size_t)b->voff[nq] * sizeof(double2), cb = (size_t)b->coff[nq] * sizeof(double2);
  int rc = 0;
  for (int r = 0; r < 4 && !rc; ++r) {
    rc = alloc((void**)&b->gamma[r], vb);
    if (!rc) rc = alloc((void**)&b->fvert[r], vb);
    if (!rc) rc = alloc((void**)&b->spare[r], vb);
  }
  for (int p = 0; p < 2 && !rc; ++p) rc = alloc((void**)&b->chi[p], cb);
  if (!rc) rc = alloc((void**)&b->d_items, items.size() * sizeof(int4));
  if (!rc) rc = alloc((void**)&b->d_voff, (size_t)(nq + 1) * sizeof(long long));
  if (!rc) rc = alloc((void**)&b->d_coff, (size_t)(nq + 1) * sizeof(long long));
  if (!rc) rc = alloc((void**)&b->d_nf, (size_t)nq * sizeof(int));
  if (!rc) rc = alloc((void**)&b->d_nl, (size_t)nq * sizeof(int));
  if (!rc) rc = alloc((void**)&b->d_tickets, 8 * sizeof(int));
  if (rc) return fail(rc);
  if (hipMemcpy(b->d_items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_voff, b->voff.data(), (nq + 1) * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_coff, b->coff.data(), (nq + 1) * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_nf, nf, nq * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_nl, nl, nq * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("batch table upload failed");
    return fail(TUPS_EHIP);
  }
  *out = b;
  return TUPS_OK;
}

extern "C" int tups_batch_query(const tups_batch* b, int64_t* vertex_elems, int64_t* chi_elems, int32_t* items) {
  TUPS_CHECK(b, TUPS_EINVAL, "null batch");
  if (vertex_elems) *vertex_elems = b->voff[b->nq];
  if (chi_elems) *chi_elems = b->coff[b->nq];
  if (items) *items = b->nitems;
  return TUPS_OK;
}

static double2* batch_array(tups_batch* b, int which, int ch, size_t* elems) {
  if (which == TUPS_GAMMA && ch >= 0 && ch < 4) {
    *elems = b->voff[b->nq];
    return b->gamma[ch];
  }
  if (which == TUPS_F && ch >= 0 && ch < 4) {
    *elems = b->voff[b->nq];
    return b->fvert[ch];
  }
  if (which == TUPS_CHI && (ch == 0 || ch == 1)) {
    *elems = b->coff[b->nq];
    return b->chi[ch];
  }
  return nullptr;
}

extern "C" int tups_batch_upload(tups_batch* b, int32_t which, int32_t ch, const double* host) {
  TUPS_CHECK(b && host, TUPS_EINVAL, "null argument");
  size_t n = 0;
  double2* p = batch_array(b, which, ch, &n);
  TUPS_CHECK(p, TUPS_EINVAL, "a batch holds GAMMA, F (4 channels) and CHI (2)");
  TUPS_HIP(hipMemcpyAsync(p, host, n * sizeof(double2), hipMemcpyHostToDevice, b->stream));
  TUPS_HIP(hipStreamSynchronize(b->stream));
  return TUPS_OK;
}

extern "C" int tups_batch_download(tups_batch* b, int32_t which, int32_t ch, double* host) {
  TUPS_CHECK(b && host, TUPS_EINVAL, "null argument");
  size_t n = 0;
  double2* p = batch_array(b, which, ch, &n);
  TUPS_CHECK(p, TUPS_EINVAL, "a batch holds GAMMA, F (4 channels) and CHI (2)");
  TUPS_HIP(hipMemcpyAsync(host, p, n * sizeof(double2), hipMemcpyDeviceToHost, b->stream));
  TUPS_HIP(hipStreamSynchronize(b->stream));
  return TUPS_OK;
}

static BatchArgs batch_args(const tups_batch* b, int dyn) {
  BatchArgs a{};
  for (int r = 0; r < 4; ++r) {
    a.gam[r] = b->gamma[r];
    a.fv[r] = b->fvert[r];
    a.out[r] = b->spare[r];
  }
  a.chi[0] = b->chi[0];
  a.chi[1] = b->chi[1];
  a.items = b->d_items;
  a.voff = b->d_voff;
  a.coff = b->d_coff;
  a.nf = b->d_nf;
  a.nl = b->d_nl;
  a.tickets = b->d_tickets;
  a.len = b->len;
  a.dyn = dyn;
  a.alpha = b->alpha;
  const double cr[4] = {0.5, 0.5, -0.25, 0.25};
  for (int r = 0; r < 4; ++r) a.cf[r] = cr[r] * b->f_damping;
  a.omf = 1.0 - b->f_damping;
  return a;
}

// one launch of variant v into the spare buffers (no swap)
int tups::batch_launch(tups_batch* b, int v) {
  const BatchVariant& k = kBatchKernels[v / kBatchOrders];
  const int order = v % kBatchOrders;
  const int flat = order == 1, dyn = order == 2;
  int dev = 0, ncu = 0;
  TUPS_HIP(hipGetDevice(&dev));
  TUPS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  // resident: as many workgroups as stay on the CUs at once (a multiple of 8);
  // flat: one workgroup per list entry
  const long long all = 8LL * b->len;
  const long long grid = flat ? all : std::min<long long>(all, std::max(8, k.per_cu * ncu / 8 * 8));
  TUPS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k.fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                               k.lds_bytes));
  if (dyn) {  // a fill kernel, not a memset node, so the launch may be captured
    hipLaunchKernelGGL(batch_tickets_clear, dim3(1), dim3(64), 0, b->stream, b->d_tickets);
    TUPS_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k.fn, dim3((unsigned)grid), dim3(256), (size_t)k.lds_bytes, b->stream,
                     batch_args(b, dyn));
  TUPS_HIP(hipGetLastError());
  return TUPS_OK;
}

extern "C" int tups_batch_bse(tups_batch* b, float* ms) {
  TUPS_CHECK(b, TUPS_EINVAL, "null batch");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms) {
    TUPS_HIP(hipEventCreate(&e0));
    TUPS_HIP(hipEventCreate(&e1));
    TUPS_HIP(hipEventRecord(e0, b->stream));
  }
  if (int rc = batch_launch(b, b->variant)) return rc;
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

extern "C" int tups_batch_set_product_forms(tups_batch* b, int32_t allow_three) {
  TUPS_CHECK(b, TUPS_EINVAL, "null batch");
  b->three_mult = allow_three != 0;
  return TUPS_OK;
}

extern "C" int tups_batch_set_variant(tups_batch* b, int32_t variant) {
  TUPS_CHECK(b, TUPS_EINVAL, "null batch");
  TUPS_CHECK(variant >= 0 && variant < batch_variant_count(), TUPS_EINVAL, "no such batch variant");
  b->variant = variant;
  return TUPS_OK;
}

// Times every variant on the resident batch (outputs go to the spare buffers
// and are discarded, the vertex state is unchanged) and keeps the fastest.
extern "C" int tups_batch_autotune(tups_batch* b, int32_t* variant, float* ms) {
  TUPS_CHECK(b, TUPS_EINVAL, "null batch");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TUPS_HIP(hipEventCreate(&e0));
  TUPS_HIP(hipEventCreate(&e1));
  float best = 0.f;
  int arg = -1, rc = TUPS_OK;
  for (int v = 0; v < batch_variant_count() && !rc; ++v) {
    if (!b->three_mult && v / kBatchOrders != 0) continue;  // kernel 0 is the four-product one
    if ((rc = batch_launch(b, v))) break;  // warm-up
    if (hipEventRecord(e0, b->stream) != hipSuccess) rc = TUPS_EHIP;
    for (int i = 0; i < 2 && !rc; ++i) rc = batch_launch(b, v);
    if (rc) break;
    float t = 0.f;
    if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
      set_error("batch autotune timing failed");
      rc = TUPS_EHIP;
      break;
    }
    t *= 0.5f;
    if (arg < 0 || t < best) {
      best = t;
      arg = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  b->variant = arg;
  if (variant) *variant = arg;
  if (ms) *ms = best;
  return TUPS_OK;
}

extern "C" int tups_batch_sync(tups_batch* b) {
  TUPS_CHECK(b, TUPS_EINVAL, "null batch");
  TUPS_HIP(hipStreamSynchronize(b->stream));
  return TUPS_OK;
}

