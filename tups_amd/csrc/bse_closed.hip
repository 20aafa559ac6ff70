// Closed form of the BSE ladder.
//
// For fixed Gamma the damped update of BSE(ite) (SRC/parquet_BSE.f90:62-132),
//   Phi = c_r f alpha (Gamma chi F + F chi Gamma) + (1 - f)(F - Gamma),
// has the fixed point F = (I - kappa_r Gamma chi)^-1 Gamma, kappa_r = 2 c_r alpha
// (push-through identity: Gamma chi F = F chi Gamma there), where Phi = F - Gamma
// for every damping f.  This mode solves (I - kappa_r Gamma chi) X = Gamma per
// (q, r) and returns Phi = X - Gamma: the converged ladder in one call.  The
// reference has no such solve; it iterates the update.
//
// Method: blocked Gauss-Jordan elimination with partial pivoting on the
// augmented matrix [I - kappa Gamma chi | Gamma] (D x 2D, complex128), one
// workgroup per matrix.  Per panel of NB columns:
//   1. LU of the panel rows k0..D-1 with row pivoting (rows swapped over all
//      columns from k0 on, so the right-hand side follows the equations);
//   2. Y = L11^-1 Z on the panel rows for every trailing column Z;
//   3. rows below the panel: Z -= L21 Y                    (FP64 MFMA);
//   4. X = U11^-1 Y on the panel rows;
//   5. rows above the panel: Z -= P_above X                (FP64 MFMA);
// after the last panel the right-hand block holds the solution.  Work per
// (q, r): 12 D^3 real flops (Gauss-Jordan on D right-hand sides).
#include "internal.h"

namespace tups {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int CL_NB = 16;  // panel width
constexpr int CL_T = 64;   // update tile (rows and columns)

struct ClosedArgs {
  const double2* gam[4];
  double2* out[4];
  const double2* chi[2];
  double2* W;   // [mat][2D columns][D rows], column-major
  int* info;    // first failing pivot (1-based column) over all matrices, 0 if none
  int D, nf, nl, q0, nqb;
  double kappa[4];
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cinv(double2 a) {
  // 1 / a with the larger component scaled out (Smith's method)
  if (fabs(a.x) >= fabs(a.y)) {
    const double r = a.y / a.x, d = a.x + a.y * r;
    return make_double2(1.0 / d, -r / d);
  }
  const double r = a.x / a.y, d = a.y + a.x * r;
  return make_double2(r / d, -1.0 / d);
}

__global__ void closed_clear(int* info) { *info = 0; }

// [I - kappa Gamma chi | Gamma]; column j of the left block is (l2, nu) = (j / Nf, j % Nf),
// (Gamma chi)(:, (l2, nu)) = sum_l1 Gamma(:, (l1, nu)) chi_nu(l1, l2).
// grid: (2D, 4 * nqb); one thread per row.
__global__ __launch_bounds__(256) void closed_build(ClosedArgs a) {
  const int j = blockIdx.x, mat = blockIdx.y;
  const int r = mat / a.nqb, q = a.q0 + (mat - r * a.nqb);
  const int D = a.D, nf = a.nf, nl = a.nl;
  const double2* G = a.gam[r] + (size_t)D * D * q;
  double2* Wc = a.W + (size_t)mat * 2 * D * D + (size_t)j * D;
  if (j >= D) {
    for (int row = threadIdx.x; row < D; row += 256) Wc[row] = G[(size_t)(j - D) * D + row];
    return;
  }
  const int l2 = j / nf, nu = j - l2 * nf;
  const double2* cb = a.chi[r >> 1] + (size_t)nf * nl * nl * q + (size_t)nu * nl * nl;
  const double kap = a.kappa[r];
  for (int row = threadIdx.x; row < D; row += 256) {
    double sr = 0.0, si = 0.0;
    for (int l1 = 0; l1 < nl; ++l1) {
      const double2 g = G[(size_t)(l1 * nf + nu) * D + row];
      const double2 c = cb[l1 + nl * l2];
      sr += g.x * c.x - g.y * c.y;
      si += g.x * c.y + g.y * c.x;
    }
    Wc[row] = make_double2((row == j ? 1.0 : 0.0) - kap * sr, -kap * si);
  }
}

// Z(r, c) -= sum_t W(r, k0 + t) W(k0 + t, c) for r in [r0, r1), c in [c0, c1),
// t < nb, by 64 x 64 tiles; four waves of 2 x 2 blocks of v_mfma_f64_16x16x4f64.
// A fragments: lane l holds A[l % 16][l / 16]; B: B[l / 16][l % 16]; output
// element e of lane l: row l / 16 + 4 e, column l % 16.
__device__ void closed_update(double2* W, int D, int r0, int r1, int c0, int c1, int k0, int nb,
                              double2 (*As)[CL_T + 1], double2 (*Bs)[CL_T + 1]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  for (int tr = r0; tr < r1; tr += CL_T)
    for (int tc = c0; tc < c1; tc += CL_T) {
      for (int u = tid; u < CL_NB * CL_T; u += 256) {
        const int row = u & (CL_T - 1), ka = u >> 6;   // A: rows contiguous
        As[ka][row] = (tr + row < r1 && ka < nb) ? W[(size_t)(k0 + ka) * D + tr + row] : make_double2(0.0, 0.0);
        const int kb = u % CL_NB, col = u / CL_NB;     // B: k contiguous in a column
        Bs[kb][col] = (tc + col < c1 && kb < nb) ? W[(size_t)(tc + col) * D + k0 + kb] : make_double2(0.0, 0.0);
      }
      __syncthreads();
      d4 ar[2][2], ai[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = tr + wr + i * 16 + lk + 4 * e, col = tc + wc + j * 16 + li;
            const double2 z = (row < r1 && col < c1) ? W[(size_t)col * D + row] : make_double2(0.0, 0.0);
            ar[i][j][e] = z.x;
            ai[i][j][e] = z.y;
          }
#pragma unroll
      for (int kk = 0; kk < CL_NB; kk += 4) {
        double2 fa[2], fb[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) fa[i] = As[kk + lk][wr + i * 16 + li];
#pragma unroll
        for (int j = 0; j < 2; ++j) fb[j] = Bs[kk + lk][wc + j * 16 + li];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            // Z - A B: re -= ar br - ai bi, im -= ar bi + ai br (blgp 1 negates A)
            ar[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].x, fb[j].x, ar[i][j], 0, 0, 1);
            ar[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].y, fb[j].y, ar[i][j], 0, 0, 0);
            ai[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].x, fb[j].y, ai[i][j], 0, 0, 1);
            ai[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].y, fb[j].x, ai[i][j], 0, 0, 1);
          }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = tr + wr + i * 16 + lk + 4 * e, col = tc + wc + j * 16 + li;
            if (row < r1 && col < c1) W[(size_t)col * D + row] = make_double2(ar[i][j][e], ai[i][j][e]);
          }
      __syncthreads();
    }
}

// Z(k0 + i, c) <- sum_t M[i][t] Z(k0 + t, c) for i, t < nb and every column c in
// [c0, c1): one thread per column, M (NB x NB, identity-padded) in LDS.
__device__ void panel_rows_apply(double2* W, int D, int k0, int nb, int c0, int c1,
                                 const double2 (*M)[CL_NB]) {
  for (int col = c0 + (int)threadIdx.x; col < c1; col += 256) {
    double2* zc = W + (size_t)col * D + k0;
    double2 z[CL_NB];
#pragma unroll
    for (int i = 0; i < CL_NB; ++i) z[i] = i < nb ? zc[i] : make_double2(0.0, 0.0);
#pragma unroll 1
    for (int i = 0; i < CL_NB; ++i) {
      double yr = 0.0, yi = 0.0;
#pragma unroll
      for (int t = 0; t < CL_NB; ++t) {
        const double2 m = M[i][t];
        yr = fma(m.x, z[t].x, fma(-m.y, z[t].y, yr));
        yi = fma(m.x, z[t].y, fma(m.y, z[t].x, yi));
      }
      if (i < nb) zc[i] = make_double2(yr, yi);
    }
  }
}

// one workgroup per (q, r) matrix; grid: 4 * nqb
__global__ __launch_bounds__(256, 2) void closed_solve(ClosedArgs a) {
  __shared__ double2 As[CL_NB][CL_T + 1];
  __shared__ double2 Bs[CL_NB][CL_T + 1];
  __shared__ double2 Ls[CL_NB][CL_NB];
  __shared__ double2 Li[CL_NB][CL_NB];
  __shared__ double2 Ui[CL_NB][CL_NB];
  __shared__ double red_v[256];
  __shared__ int red_i[256];
  const int mat = blockIdx.x;
  const int D = a.D, D2 = 2 * D;
  double2* W = a.W + (size_t)mat * 2 * D * D;
  const int tid = threadIdx.x;
  int info = 0;
  for (int k0 = 0; k0 < D; k0 += CL_NB) {
    const int nb = min(CL_NB, D - k0);
    // 1. panel LU, rows k0..D-1, columns k0..k0+nb-1
    for (int c = k0; c < k0 + nb; ++c) {
      double best = -1.0;
      int bi = c;
      for (int row = c + tid; row < D; row += 256) {
        const double2 v = W[(size_t)c * D + row];
        const double m = fabs(v.x) + fabs(v.y);
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
          const double o = red_v[tid + s];
          const int oi = red_i[tid + s];
          if (o > red_v[tid] || (o == red_v[tid] && oi < red_i[tid])) {
            red_v[tid] = o;
            red_i[tid] = oi;
          }
        }
        __syncthreads();
      }
      const int p = red_i[0];
      const double pmax = red_v[0];
      __syncthreads();
      if (!(pmax > 0.0)) {  // zero or NaN pivot column: the system is singular
        if (info == 0) info = c + 1;
        continue;
      }
      if (p != c) {
        for (int col = k0 + tid; col < D2; col += 256) {
          const double2 x = W[(size_t)col * D + c];
          W[(size_t)col * D + c] = W[(size_t)col * D + p];
          W[(size_t)col * D + p] = x;
        }
      }
      __syncthreads();
      const double2 inv = cinv(W[(size_t)c * D + c]);
      for (int row = c + 1 + tid; row < D; row += 256) {
        const double2 l = cmul(W[(size_t)c * D + row], inv);
        W[(size_t)c * D + row] = l;
        for (int col = c + 1; col < k0 + nb; ++col)
          W[(size_t)col * D + row] = csub(W[(size_t)col * D + row], cmul(l, W[(size_t)col * D + c]));
      }
      __syncthreads();
    }
    for (int u = tid; u < CL_NB * CL_NB; u += 256) {
      const int i = u % CL_NB, j = u / CL_NB;
      Ls[i][j] = (i < nb && j < nb) ? W[(size_t)(k0 + j) * D + k0 + i] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    // inverses of the unit lower L11 and of the upper U11, padded to NB x NB with
    // the identity: thread t < NB forms column t of both
    if (tid < CL_NB) {
      const int t = tid;
      for (int i = 0; i < CL_NB; ++i) {
        double2 x = make_double2(i == t ? 1.0 : 0.0, 0.0);
        if (i < nb && t < nb)
          for (int s2 = t; s2 < i; ++s2) x = csub(x, cmul(Ls[i][s2], Li[s2][t]));
        Li[i][t] = x;
      }
      for (int i = CL_NB - 1; i >= 0; --i) {
        double2 x = make_double2(i == t ? 1.0 : 0.0, 0.0);
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
    panel_rows_apply(W, D, k0, nb, j0, D2, Li);
    __syncthreads();
    // 3. rows below the panel
    closed_update(W, D, k0 + nb, D, j0, D2, k0, nb, As, Bs);
    // 4. X = U11^-1 Y on the panel rows
    panel_rows_apply(W, D, k0, nb, j0, D2, Ui);
    __syncthreads();
    // 5. rows above the panel
    closed_update(W, D, 0, k0, j0, D2, k0, nb, As, Bs);
  }
  if (info && tid == 0) atomicCAS(a.info, 0, info);
  // Phi = X - Gamma
  const int r = mat / a.nqb, q = a.q0 + (mat - r * a.nqb);
  const double2* G = a.gam[r] + (size_t)D * D * q;
  double2* O = a.out[r] + (size_t)D * D * q;
  for (size_t u = tid; u < (size_t)D * D; u += 256) O[u] = csub(W[(size_t)D * D + u], G[u]);
}

int launch_bse_closed(tups_ctx* c) {
  const tups_sizes& s = c->sz;
  const size_t D = s.nz;
  const size_t per_q = 4 * 2 * D * D * sizeof(double2);
  const size_t want = ((size_t)4 << 30) / per_q;  // about 4 GiB of augmented matrices per pass
  const int qbatch = (int)std::max<size_t>(1, std::min<size_t>(s.nq, want));
  const size_t need = per_q * qbatch + 256;
  if (c->closed_ws_bytes < need) {
    if (c->closed_ws) TUPS_HIP(hipFree(c->closed_ws));
    c->closed_ws = nullptr;
    c->closed_ws_bytes = 0;
    TUPS_HIP(hipMalloc(&c->closed_ws, need));
    c->closed_ws_bytes = need;
  }
  ClosedArgs a{};
  for (int r = 0; r < 4; ++r) {
    a.gam[r] = c->gamma[r];
    a.out[r] = c->spare[r];
  }
  a.chi[0] = c->chi[0];
  a.chi[1] = c->chi[1];
  a.info = static_cast<int*>(c->closed_ws);
  a.W = reinterpret_cast<double2*>(static_cast<char*>(c->closed_ws) + 256);
  a.D = s.nz;
  a.nf = s.nf;
  a.nl = s.nl;
  const double cr[4] = {0.5, 0.5, -0.25, 0.25};
  for (int r = 0; r < 4; ++r) a.kappa[r] = 2.0 * cr[r] * s.alpha;
  hipLaunchKernelGGL(closed_clear, dim3(1), dim3(1), 0, c->stream, a.info);
  TUPS_HIP(hipGetLastError());
  for (int q0 = 0; q0 < s.nq; q0 += qbatch) {
    a.q0 = q0;
    a.nqb = std::min(qbatch, s.nq - q0);
    const unsigned mats = 4u * (unsigned)a.nqb;
    if (int rc = timer_begin(c)) return rc;
    hipLaunchKernelGGL(closed_build, dim3((unsigned)(2 * D), mats), dim3(256), 0, c->stream, a);
    TUPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(closed_solve, dim3(mats), dim3(256), 0, c->stream, a);
    TUPS_HIP(hipGetLastError());
    if (int rc = timer_end(c, TUPS_K_BSE_CLOSED)) return rc;
  }
  int info = 0;
  TUPS_HIP(hipMemcpyAsync(&info, a.info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  TUPS_CHECK(info == 0, TUPS_ENONFINITE, "closed-form BSE: singular I - kappa Gamma chi (zero pivot column)");
  return TUPS_OK;
}

}  // namespace tups
