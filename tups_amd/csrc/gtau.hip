// G in imaginary time, the chemical-potential bisection and the SDE's
// second-order term, on the device, FP64.
//
// Reference: calc_Gkw_Grt, SRC/parquet_chi.f90:18-210: Sigma_H(k) <- Sigma(k, Nf);
// mu bisected between mu -/+ U until |n - nParticle| <= 1e-6, n = -2 Grt(r = 0,
// tau = beta); G(k, tau) on the fine (coarse-graining) grid from the Matsubara sum
// with the bare G subtracted and added back in closed form; fftf2d per tau slice
// (SRC/MATH_mod.f90:215-236: forward transform divided by n1 n2).
// calc_order_2nd, SRC/parquet_SDE.f90:263-340: GGG(r, tau) = Grt(r, tau)^2
// Grt(-r, beta - tau) U^2, fftb2d (unscaled backward transform) read at the
// coarse momenta, then FDfit + nfourier (SRC/MATH_mod.f90:95-212: a cubic spline
// through the Nf tau points with 5-point end slopes, integrated exactly against
// e^{i w tau}) and added to Sigma at +/- nu.
//
// The bisection only needs n(mu) = -2/N sum_k G(k, beta), and at tau = beta the
// Matsubara phase is e^{-i nu_n beta} = -1, so
//   G(k, beta) = -1/beta sum_nu Re[G - G0](k, nu) - 1 / (1 + e^{beta (e_k - mu)}).
// One pass evaluates the next six levels of the bisection tree at once (63
// candidate mu, one row of workgroups each); one workgroup then walks the tree
// as the sequential loop would, stops at the first mu within the tolerance, and
// lays out the next subtree.  Small grids run every pass inside one workgroup
// of 1024 lanes (16 per candidate) and stop as soon as the walk ends.  The mu
// found is the one the reference's loop stops at; the number of launches is fixed (passes after convergence exit at once),
// so the whole search sits in a HIP graph.
//
// The 2D transforms are direct DFTs (n1 n2 (n1 + n2) Nf complex MACs, small next
// to the ladder) with twiddles from exact integer phase reduction.
#include "green.h"
#include "internal.h"

#include <cmath>

namespace tups {

namespace {

constexpr int kDepth = 6;
constexpr int kCand = (1 << kDepth) - 1;  // candidates per pass
constexpr int kMaxSteps = 64;             // evaluations before the search gives up
constexpr int kPasses = (kMaxSteps + kDepth - 1) / kDepth;
constexpr double kTol = 1e-6;             // |n - nParticle|, SRC/parquet_chi.f90:117
constexpr double kSmallWork = 1 << 19;    // 63 Nfine Nf up to this: the one-workgroup search

struct GtauArgs {
  const double2* sigma;  // (Nx Ny Nf)
  double2* sigh;         // Sigma_H(Nx, Ny)
  GtauState* st;
  double* etab;          // fine-grid energies [j2 n1 + j1]
  double* partial;       // [cand][block]
  double* greal;         // G(k_fine, tau) [(it n2 + j2) n1 + j1]
  double2* xbuf;         // first-axis transform, same layout
  double2* grt;          // Grt(j1, j2, it), same layout
  double* u2data;        // [kidx Nf + it]: Re fftb2d(GGG) at the coarse momenta
  double2* sigma_out;    // Sigma receiving the second-order term
  const double* wfine;   // weight of each fine point in the forward transform (null: 1)
  int nx, nf, ng, n1, n2, nfine, nbk;
  double beta, u;
};

__device__ __forceinline__ double dev_mu(const GtauState* st) { return st->mu; }

// 1 / (1 + e^{x}) written so that neither branch overflows
__device__ __forceinline__ double fermi(double x) {
  if (x > 0.0) {
    const double t = exp(-x);
    return t / (1.0 + t);
  }
  return 1.0 / (1.0 + exp(x));
}

// coarse momentum index ((kx-1) Ny + ky-1) of fine point (j1, j2): the inverse of
// idx = (k-1) Ng + (g - (Ng+1)/2) + 1 wrapped into 1..Nx Ng (SRC/parquet_chi.f90:135-139)
__device__ __forceinline__ int coarse_of(const GtauArgs& a, int j1, int j2) {
  const int half = (a.ng + 1) / 2;
  const int t1 = (j1 + half - 1) % a.n1, t2 = (j2 + half - 1) % a.n2;
  return (t1 / a.ng) * a.nx + t2 / a.ng;
}

// Matsubara frequency of box index nu (1-based), get_freq (SRC/parquet_util.f90:1332-1339)
__device__ __forceinline__ double wn(int nu, int nf, double beta) {
  return 3.14159265358979323846 / beta * (2.0 * (nu - nf / 2 - 1) + 1.0);
}

// Re[1 / (i w + mu - e - s) - 1 / (i w + mu - e)]
__device__ __forceinline__ double2 dress(double w, double mu, double e, double2 s) {
  const double a = mu - e - s.x, b = w - s.y;
  const double d = a * a + b * b;
  const double a0 = mu - e;
  const double d0 = a0 * a0 + w * w;
  return make_double2(a / d - a0 / d0, -b / d + w / d0);
}

// the bisection tree below (hi, lo): node i (1-based, heap order; left child =
// "n too large", hi <- mu) holds the midpoint of its interval
__device__ void fill_tree(double* cand, double hi, double lo, int i) {
  if (i >= kCand) return;
  const int node = i + 1;
  const int depth = 31 - __clz(node);
  double h = hi, l = lo;
  for (int b = depth - 1; b >= 0; --b) {
    const double m = 0.5 * (h + l);
    if ((node >> b) & 1) l = m;
    else h = m;
  }
  cand[i] = 0.5 * (h + l);
}

// Sigma_H, the fine-grid energies and the start of the search; one workgroup.
// mode 0: bisection, 1: one evaluation at the midpoint (the reference's cal_chi
// exit), 2: no search (G(tau) at the current mu)
__global__ __launch_bounds__(256) void gt_setup(GtauArgs a, int mu_from_host, double mu_host, double n_target,
                                                int mode, int set_hartree) {
  GtauState* st = a.st;
  const int tid = threadIdx.x;
  if (set_hartree) {
    for (int k = tid; k < a.nx * a.nx; k += blockDim.x) {
      const int kx = k / a.nx, ky = k % a.nx;
      a.sigh[kx + a.nx * ky] = a.sigma[(size_t)k * a.nf + a.nf - 1];
    }
  }
  const double pi = 3.14159265358979323846;
  const double dk = 2.0 * pi / a.nx;
  const int half = (a.ng + 1) / 2;
  for (int j = tid; j < a.nfine; j += blockDim.x) {
    const int j1 = j % a.n1, j2 = j / a.n1;
    const int t1 = (j1 + half - 1) % a.n1, t2 = (j2 + half - 1) % a.n2;
    const int kx = t1 / a.ng, gx = t1 % a.ng + 1, ky = t2 / a.ng, gy = t2 % a.ng + 1;
    const double px = dk * kx + (gx - half) * dk / (double)a.ng;
    const double py = dk * ky + (gy - half) * dk / (double)a.ng;
    a.etab[j] = -2.0 * cos(px) - 2.0 * cos(py);
  }
  __shared__ double bounds[2];
  if (tid == 0) {
    const double mu0 = mu_from_host ? mu_host : st->mu;
    bounds[0] = mu0 + a.u;
    bounds[1] = mu0 - a.u;
    st->hi = mu0 + a.u;
    st->lo = mu0 - a.u;
    st->n_target = n_target;
    st->steps = 0;
    st->status = 0;
    st->single = mode == 1;
    st->done = mode == 2;
    st->t1 = 0.0;
    st->mu = mu0;
    st->arrived = 0;
  }
  __syncthreads();
  fill_tree(st->cand, bounds[0], bounds[1], tid);
}

// G(k, beta) of fine point j at mu: the summand of the particle number
__device__ __forceinline__ double g_beta(const GtauArgs& a, double mu, int j) {
  const double e = a.etab[j];
  const int kc = coarse_of(a, j % a.n1, j / a.n1);
  const double2* s = a.sigma + (size_t)kc * a.nf;
  double acc = 0.0;
  for (int nu = 1; nu <= a.nf; ++nu) acc += dress(wn(nu, a.nf, a.beta), mu, e, s[nu - 1]).x;
  return -acc / a.beta - fermi(a.beta * (e - mu));
}

__device__ int walk_tree(GtauState* st, const double* cand, const double* t1c, double* bounds);

// grid (nbk, kCand): partial sums of G(k, beta) over the fine grid per candidate.
// The last workgroup to finish (counted in st->arrived, which it resets) adds
// the partials in a fixed order, walks the subtree and lays out the next one.
// Every workgroup reads st->done before any of them can count itself in, so a
// pass runs whole or not at all.
__global__ __launch_bounds__(256) void gt_density(GtauArgs a) {
  __shared__ double red[256];
  __shared__ double t1c[kCand];
  __shared__ double bounds[2];
  __shared__ int last, finished;
  GtauState* st = a.st;
  if (st->done) return;
  const int tid = threadIdx.x;
  const int c = blockIdx.y;
  const double mu = st->cand[c];
  const int j = blockIdx.x * 256 + tid;
  red[tid] = j < a.nfine ? g_beta(a, mu, j) : 0.0;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) {
    a.partial[(size_t)c * a.nbk + blockIdx.x] = red[0];
    __threadfence();
    last = atomicAdd(&st->arrived, 1u) == gridDim.x * gridDim.y - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (tid < kCand) {
    double s = 0.0;
    for (int b = 0; b < a.nbk; ++b) s += a.partial[(size_t)tid * a.nbk + b];
    t1c[tid] = -2.0 * s / (double)a.nfine;
    st->t1c[tid] = t1c[tid];
  }
  __syncthreads();
  if (tid == 0) {
    st->arrived = 0;
    finished = walk_tree(st, st->cand, t1c, bounds);
  }
  __syncthreads();
  if (!finished) fill_tree(st->cand, bounds[0], bounds[1], tid);
}

// one lane: walks the evaluated subtree (candidates cand, particle numbers t1c) as the
// sequential loop of SRC/parquet_chi.f90:117-198 would; returns whether the search ended
// and leaves the new bracket in bounds
__device__ int walk_tree(GtauState* st, const double* cand, const double* t1c, double* bounds) {
  double h = st->hi, l = st->lo;
  const double target = st->n_target;
  int node = 1, steps = st->steps, done = 0, status = 0;
  double mu = st->mu, t1 = st->t1;
  for (int level = 0; level < kDepth; ++level) {
    mu = cand[node - 1];
    t1 = t1c[node - 1];
    steps += 1;
    if (st->single) {
      done = 1;
      break;
    }
    // the loop condition abs(t1 - nParticle) > 1e-6 is false for a NaN t1 as well
    if (!(fabs(t1 - target) > kTol)) {
      done = 1;
      if (!isfinite(t1)) status = 2;
      break;
    }
    if (t1 > target) {
      h = mu;
      node = 2 * node;
    } else {
      l = mu;
      node = 2 * node + 1;
    }
    if (steps >= kMaxSteps) {
      done = 1;
      status = 1;
      break;
    }
  }
  st->mu = mu;
  st->t1 = t1;
  st->steps = steps;
  st->hi = h;
  st->lo = l;
  st->status = status;
  st->done = done;
  bounds[0] = h;
  bounds[1] = l;
  return done;
}

// the whole search in one workgroup of 1024 for small grids (63 Nfine Nf within
// kSmallWork): 16 lanes per candidate, passes until the walk ends, one launch
__global__ __launch_bounds__(1024) void gt_search_small(GtauArgs a) {
  __shared__ double cand[kCand + 1];
  __shared__ double t1c[kCand + 1];
  __shared__ double part[1024];
  __shared__ double bounds[2];
  __shared__ int finished;
  GtauState* st = a.st;
  if (st->done) return;
  const int tid = threadIdx.x, grp = tid >> 4, lane = tid & 15;
  if (tid < kCand) cand[tid] = st->cand[tid];
  for (int pass = 0; pass < kPasses; ++pass) {
    __syncthreads();
    double v = 0.0;
    if (grp < kCand)
      for (int j = lane; j < a.nfine; j += 16) v += g_beta(a, cand[grp], j);
    part[tid] = v;
    __syncthreads();
    if (lane == 0 && grp < kCand) {
      double s = 0.0;
      for (int i = 0; i < 16; ++i) s += part[tid + i];
      t1c[grp] = -2.0 * s / (double)a.nfine;
    }
    __syncthreads();
    if (tid == 0) finished = walk_tree(st, cand, t1c, bounds);
    __syncthreads();
    if (finished) break;
    fill_tree(cand, bounds[0], bounds[1], tid);
  }
}

// grid: nfine; G(k_fine, tau) for every tau at the final mu (SRC/parquet_chi.f90:124-170)
__global__ __launch_bounds__(64) void gt_fine(GtauArgs a) {
  extern __shared__ double2 dnu[];  // [Nf] dressed-minus-bare G, then [2 (Nf - 1)] phases
  const int j = blockIdx.x;
  const int nf = a.nf;
  const int per = 2 * (nf - 1);
  double2* ph = dnu + nf;  // e^{-i pi m / (Nf - 1)} as (cos, sin) of pi m / (Nf - 1)
  const double mu = dev_mu(a.st), beta = a.beta, e = a.etab[j];
  const int kc = coarse_of(a, j % a.n1, j / a.n1);
  const double2* s = a.sigma + (size_t)kc * nf;
  for (int nu = threadIdx.x + 1; nu <= nf; nu += blockDim.x) dnu[nu - 1] = dress(wn(nu, nf, beta), mu, e, s[nu - 1]);
  for (int m = threadIdx.x; m < per; m += blockDim.x) {
    double sn, cs;
    sincospi((double)m / (double)(nf - 1), &sn, &cs);
    ph[m] = make_double2(cs, sn);
  }
  __syncthreads();
  for (int it = threadIdx.x; it < nf; it += blockDim.x) {
    double acc = 0.0;
    // e^{-i nu_n tau_it}: phase pi (2 nu - Nf - 1) it / (Nf - 1), reduced exactly; the
    // index steps by 2 it (mod 2 (Nf - 1)) from nu to nu + 1
    int m = ((1 - nf) * it) % per;
    if (m < 0) m += per;
    const int step = (2 * it) % per;
    for (int nu = 1; nu <= nf; ++nu) {
      const double2 t = ph[m], d = dnu[nu - 1];
      acc += (t.x * d.x + t.y * d.y) / beta;
      m += step;
      if (m >= per) m -= per;
    }
    const double tau = beta / (nf - 1) * it;
    const double x = e - mu;
    double g0;
    if (x > 0.0) g0 = exp(-tau * x) / (exp(-beta * x) + 1.0);
    else g0 = exp((beta - tau) * x) / (1.0 + exp(beta * x));
    a.greal[((size_t)it * a.n2 + j / a.n1) * a.n1 + j % a.n1] = acc - g0;
  }
}

// grid (n2, Nf): transform along the first axis, X(k1, j2, it) = sum_j1 G e^{-2 pi i k1 j1 / n1}
__global__ __launch_bounds__(64) void gt_dft1(GtauArgs a) {
  extern __shared__ double2 lds2[];
  double2* tw = lds2;                                    // [n1]
  double* row = reinterpret_cast<double*>(lds2 + a.n1);  // [n1]
  const int n1 = a.n1;
  const size_t base = ((size_t)blockIdx.y * a.n2 + blockIdx.x) * n1;
  for (int i = threadIdx.x; i < n1; i += blockDim.x) {
    row[i] = a.wfine ? a.greal[base + i] * a.wfine[(size_t)blockIdx.x * n1 + i] : a.greal[base + i];
    double sn, cs;
    sincospi(2.0 * i / n1, &sn, &cs);
    tw[i] = make_double2(cs, -sn);
  }
  __syncthreads();
  for (int k1 = threadIdx.x; k1 < n1; k1 += blockDim.x) {
    double re = 0.0, im = 0.0;
    int m = 0;
    for (int j1 = 0; j1 < n1; ++j1) {
      const double2 t = tw[m];
      re += row[j1] * t.x;
      im += row[j1] * t.y;
      m += k1;
      if (m >= n1) m -= n1;
    }
    a.xbuf[base + k1] = make_double2(re, im);
  }
}

// grid (n2, Nf): second axis and the 1 / (n1 n2) of fftf2d; block (k2, it), threads over k1
__global__ __launch_bounds__(64) void gt_dft2(GtauArgs a) {
  extern __shared__ double2 tw2[];  // [n2]
  const int n1 = a.n1, n2 = a.n2, k2 = blockIdx.x;
  for (int i = threadIdx.x; i < n2; i += blockDim.x) {
    double sn, cs;
    sincospi(2.0 * i / n2, &sn, &cs);
    tw2[i] = make_double2(cs, -sn);
  }
  __syncthreads();
  const double inv = 1.0 / ((double)n1 * n2);
  const size_t slab = (size_t)blockIdx.y * n2 * n1;
  for (int k1 = threadIdx.x; k1 < n1; k1 += blockDim.x) {
    double re = 0.0, im = 0.0;
    int m = 0;
    for (int j2 = 0; j2 < n2; ++j2) {
      const double2 x = a.xbuf[slab + (size_t)j2 * n1 + k1];
      const double2 t = tw2[m];
      re += x.x * t.x - x.y * t.y;
      im += x.x * t.y + x.y * t.x;
      m += k2;
      if (m >= n2) m -= n2;
    }
    a.grt[slab + (size_t)k2 * n1 + k1] = make_double2(re * inv, im * inv);
  }
}

// grid (Nc, Nf): Re fftb2d(GGG) at coarse momentum kidx = (kx-1) Ny + ky-1, slice it
// (SRC/parquet_SDE.f90:290-315; the fine index of kx is (kx-1) Ng + 1)
__global__ __launch_bounds__(256) void u2_ggg(GtauArgs a) {
  __shared__ double red[256];
  const int kidx = blockIdx.x, it = blockIdx.y;
  const int nx = a.nx, n1 = a.n1, n2 = a.n2, nf = a.nf;
  const int kx = kidx / nx, ky = kidx % nx;  // 0-based
  const size_t slab = (size_t)it * n1 * n2, mslab = (size_t)(nf - 1 - it) * n1 * n2;
  const double u2 = a.u * a.u;
  double acc = 0.0;
  for (int j = threadIdx.x; j < a.nfine; j += blockDim.x) {
    const int j1 = j % n1, j2 = j / n1;
    const int m1 = j1 == 0 ? 0 : n1 - j1, m2 = j2 == 0 ? 0 : n2 - j2;
    const double2 g = a.grt[slab + j];
    const double2 gm = a.grt[mslab + (size_t)m2 * n1 + m1];
    const double2 gg = cmul(cmul(g, g), gm);
    // e^{+2 pi i (kx Ng j1 / n1 + ky Ng j2 / n2)} = e^{2 pi i (kx j1 + ky j2) / Nx}
    const int m = (kx * j1 + ky * j2) % nx;
    double sn, cs;
    sincospi(2.0 * m / nx, &sn, &cs);
    acc += (gg.x * cs - gg.y * sn) * u2;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.u2data[(size_t)kidx * nf + it] = red[0];
}

// Clamped cubic spline through r[0..L] on the uniform grid t_k = k*h (the method of
// Burden & Faires, Numerical Analysis, sec. 3.5): the second derivatives m_k solve
//   2 m_0 + m_1               = (6/h) ((r_1 - r_0)/h - s_0)
//   m_{k-1} + 4 m_k + m_{k+1} = (6/h^2) (r_{k+1} - 2 r_k + r_{k-1}),   0 < k < L
//   m_{L-1} + 2 m_L           = (6/h) (s_L - (r_L - r_{L-1})/h)
// with the end slopes s_0, s_L from the five-point one-sided difference
// f'(t_0) = (-25 f_0 + 48 f_1 - 36 f_2 + 16 f_3 - 3 f_4)/(12 h) + O(h^4) (and its mirror
// at t_L) — the end condition of the reference's tau quadrature (FDfit and nfourier,
// SRC/MATH_mod.f90:95-160).  The tridiagonal system is solved by the Thomas algorithm on
// one lane: cp and dp hold the eliminated super-diagonal and right-hand side.
__device__ void spline_fit(const double* r, double* m, double* cp, double* dp, int L, double h) {
  const double s0 = (-25.0 * r[0] + 48.0 * r[1] - 36.0 * r[2] + 16.0 * r[3] - 3.0 * r[4]) / (12.0 * h);
  const double sL = (25.0 * r[L] - 48.0 * r[L - 1] + 36.0 * r[L - 2] - 16.0 * r[L - 3] + 3.0 * r[L - 4]) / (12.0 * h);
  const double g = 6.0 / (h * h);
  // row 0: diagonal 2, super-diagonal 1
  cp[0] = 0.5;
  dp[0] = 0.5 * (6.0 / h) * ((r[1] - r[0]) / h - s0);
  // rows 1..L-1: sub-diagonal 1, diagonal 4, super-diagonal 1
  for (int k = 1; k < L; ++k) {
    const double piv = 4.0 - cp[k - 1];
    cp[k] = 1.0 / piv;
    dp[k] = (g * (r[k + 1] - 2.0 * r[k] + r[k - 1]) - dp[k - 1]) / piv;
  }
  // row L: sub-diagonal 1, diagonal 2
  const double dL = (6.0 / h) * (sL - (r[L] - r[L - 1]) / h);
  m[L] = (dL - dp[L - 1]) / (2.0 - cp[L - 1]);
  for (int k = L - 1; k >= 0; --k) m[k] = dp[k] - cp[k] * m[k + 1];
}

// Exact integral of the spline against e^{i om t} over the intervals j = j0, j0 + js, ...
// <= L (interval j spans [t_{j-1}, t_j]), added to (re, im).  On interval j the spline is
// p(u) = a + b u + c u^2 + d u^3, u = t - t_{j-1}, with
//   a = r_{j-1}, b = (r_j - r_{j-1})/h - h (2 m_{j-1} + m_j)/6, c = m_{j-1}/2, d = (m_j - m_{j-1})/(6h),
// and repeated integration by parts gives the antiderivative
//   int p e^{i om t} dt = e^{i om t} [ (p1/om^2 - p3/om^4) + i (p2/om^3 - p/om) ]
// (pk the k-th derivative of p), evaluated at u = h and u = 0; for om -> 0 the plain
// polynomial integral.
__device__ void spline_integral(const double* r, const double* m, int L, double h, double om, int j0, int js,
                                double& re, double& im) {
  const bool zero = fabs(om) < 1e-6;
  const double i1 = zero ? 0.0 : 1.0 / om, i2 = i1 * i1, i3 = i2 * i1, i4 = i2 * i2;
  for (int j = j0; j <= L; j += js) {
    const double a = r[j - 1], c = 0.5 * m[j - 1];
    const double b = (r[j] - r[j - 1]) / h - h * (2.0 * m[j - 1] + m[j]) / 6.0;
    const double d = (m[j] - m[j - 1]) / (6.0 * h);
    if (zero) {
      re += h * (a + h * (b / 2.0 + h * (c / 3.0 + h * d / 4.0)));
      continue;
    }
    // p and its derivatives at the interval's end (u = h) and start (u = 0)
    const double p_h = a + h * (b + h * (c + h * d)), p1_h = b + h * (2.0 * c + 3.0 * h * d);
    const double p2_h = 2.0 * c + 6.0 * h * d, p3 = 6.0 * d;
    const double f1r = p1_h * i2 - p3 * i4, f1i = p2_h * i3 - p_h * i1;
    const double f0r = b * i2 - p3 * i4, f0i = 2.0 * c * i3 - a * i1;
    double s1, c1, s0, c0;
    sincos(om * h * j, &s1, &c1);
    sincos(om * h * (j - 1), &s0, &c0);
    re += (c1 * f1r - s1 * f1i) - (c0 * f0r - s0 * f0i);
    im += (c1 * f1i + s1 * f1r) - (c0 * f0i + s0 * f0r);
  }
}

// lanes per frequency for nw frequencies of L intervals in a workgroup of 256
__device__ __forceinline__ int lanes_per_freq(int nw, int L) {
  int ks = 1;
  while (ks * 2 * nw <= 256 && ks * 2 <= L) ks *= 2;
  return ks;
}

// grid Nc: FDfit + nfourier of one coarse momentum, added to Sigma at +/- nu
// (SRC/parquet_SDE.f90:317-333)
__global__ __launch_bounds__(256) void u2_spline(GtauArgs a) {
  extern __shared__ double sp[];
  __shared__ double2 red[256];
  const int nf = a.nf, L = nf - 1, kidx = blockIdx.x;
  double* r = sp;             // [L + 1]
  double* xm = sp + nf;       // [L + 1]
  const double beta = a.beta, delta = beta / L;
  for (int i = threadIdx.x; i < nf; i += blockDim.x) r[i] = a.u2data[(size_t)kidx * nf + i];
  __syncthreads();
  if (threadIdx.x == 0) spline_fit(r, xm, sp + 2 * nf, sp + 3 * nf, L, delta);
  __syncthreads();
  // the Nf/2 fermionic frequencies, each summed over the L intervals by `ks` lanes
  // whose sums are then added in order
  const double pi = 3.14159265358979323846;
  double2* out = a.sigma_out + (size_t)kidx * nf;
  const int nw = nf / 2, ks = lanes_per_freq(nw, L);
  const int per = 256 / ks, part = threadIdx.x % ks, slot = threadIdx.x / ks;
  for (int i0 = 0; i0 < nw; i0 += per) {
    const int i = i0 + slot;
    double re = 0.0, im = 0.0;
    if (i < nw) spline_integral(r, xm, L, delta, (2.0 * (i + 1) - 1.0) * pi / beta, 1 + part, ks, re, im);
    red[threadIdx.x] = make_double2(re, im);
    __syncthreads();
    if (part == 0 && i < nw) {
      for (int t = 1; t < ks; ++t) {
        re += red[threadIdx.x + t].x;
        im += red[threadIdx.x + t].y;
      }
      // ret(i) lands at nu = Nf/2 - i (conjugated) and at Nf + 1 - nu
      const int nu = nw - i;
      double2 v = out[nu - 1];
      out[nu - 1] = make_double2(v.x + re, v.y - im);
      v = out[nf - nu];
      out[nf - nu] = make_double2(v.x + re, v.y + im);
    }
    __syncthreads();
  }
}

// ---- the bare bubbles of calc_Gkw_Grt with cal_chi (SRC/parquet_chi.f90:216-430) ----

constexpr int kC0 = 10;  // ph, pp, ph_ff[4], pp_ff[4]

struct C0Args {
  const double2* grt;     // Grt(j1, j2, it)
  const double2* ffg;     // [ff][nfs]: FFG of the four external form factors
  const int* ibz;         // IBZ momenta x[0..nibz), y[0..nibz) (1-based)
  double* data;           // [m][kC0][Nf]: Re fftb2d of the products at IBZ momentum m
  double2* out;           // [kC0][nibz Nw]
  int nx, nf, n1, n2, nfine, nibz;
  double beta;
};

// FF_sdwave(kx, ky, ff)^2 of each fine point's coarse momentum (SRC/parquet_formfactors.f90:792-815)
__global__ __launch_bounds__(256) void c0_weights(GtauArgs a, double* w) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= a.nfine) return;
  const int kc = coarse_of(a, j % a.n1, j / a.n1);
  const int kx = kc / a.nx, ky = kc % a.nx;  // 0-based
  const double pi = 3.14159265358979323846;
  const double px = 2.0 / a.nx * pi * kx, py = 2.0 / a.nx * pi * ky;
  const double f[4] = {1.0, cos(px) - cos(py), sin(px), sin(px) * sin(py)};
#pragma unroll
  for (int ff = 0; ff < 4; ++ff) w[(size_t)ff * a.nfine + j] = f[ff] * f[ff];
}

// grid (nibz, Nf): the ten products at slice it, transformed back at IBZ momentum m
__global__ __launch_bounds__(256) void c0_pair(C0Args a) {
  __shared__ double red[kC0][256];
  const int m = blockIdx.x, it = blockIdx.y;
  const int nx = a.nx, n1 = a.n1, n2 = a.n2, nf = a.nf;
  const int qx = a.ibz[m] - 1, qy = a.ibz[a.nibz + m] - 1;
  const size_t nfs = (size_t)a.nfine * nf;
  const size_t slab = (size_t)it * n1 * n2, mslab = (size_t)(nf - 1 - it) * n1 * n2;
  double acc[kC0];
#pragma unroll
  for (int c = 0; c < kC0; ++c) acc[c] = 0.0;
  for (int j = threadIdx.x; j < a.nfine; j += blockDim.x) {
    const int j1 = j % n1, j2 = j / n1;
    const int m1 = j1 == 0 ? 0 : n1 - j1, m2 = j2 == 0 ? 0 : n2 - j2;
    const size_t jm = (size_t)m2 * n1 + m1;
    const double2 g = a.grt[slab + j], gm = a.grt[mslab + jm];
    // e^{+2 pi i (qx Ng j1 / n1 + qy Ng j2 / n2)} = e^{2 pi i (qx j1 + qy j2) / Nx}
    double sn, cs;
    sincospi(2.0 * ((qx * j1 + qy * j2) % nx) / nx, &sn, &cs);
    auto re_of = [&](double2 p, double c) __attribute__((always_inline)) { return c * (p.x * cs - p.y * sn); };
    acc[0] += re_of(cmul(g, gm), -1.0);
    acc[1] += re_of(cmul(g, g), -0.5);
#pragma unroll
    for (int ff = 0; ff < 4; ++ff) {
      const double2* F = a.ffg + (size_t)ff * nfs;
      acc[2 + ff] += re_of(cmul(g, F[mslab + jm]), -1.0);
      acc[6 + ff] += re_of(cmul(g, F[slab + j]), -0.5);
    }
  }
#pragma unroll
  for (int c = 0; c < kC0; ++c) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h)
#pragma unroll
      for (int c = 0; c < kC0; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < kC0) a.data[((size_t)m * kC0 + threadIdx.x) * nf + it] = red[threadIdx.x][0];
}

// grid (nibz, kC0): FDfit + bosonic nfourier of one momentum and quantity
__global__ __launch_bounds__(256) void c0_spline(C0Args a) {
  extern __shared__ double sp[];
  __shared__ double2 red[256];
  const int nf = a.nf, L = nf - 1, m = blockIdx.x, c = blockIdx.y;
  double* r = sp;
  double* xm = sp + nf;
  const double beta = a.beta, delta = beta / L;
  for (int i = threadIdx.x; i < nf; i += blockDim.x) r[i] = a.data[((size_t)m * kC0 + c) * nf + i];
  __syncthreads();
  if (threadIdx.x == 0) spline_fit(r, xm, sp + 2 * nf, sp + 3 * nf, L, delta);
  __syncthreads();
  const double pi = 3.14159265358979323846;
  const int nw = nf / 2, ks = lanes_per_freq(nw, L);
  const int per = 256 / ks, part = threadIdx.x % ks, slot = threadIdx.x / ks;
  double2* out = a.out + (size_t)c * a.nibz * nw + (size_t)m * nw;
  for (int i0 = 0; i0 < nw; i0 += per) {
    const int i = i0 + slot;
    double re = 0.0, im = 0.0;
    if (i < nw) spline_integral(r, xm, L, delta, 2.0 * i * pi / beta, 1 + part, ks, re, im);
    red[threadIdx.x] = make_double2(re, im);
    __syncthreads();
    if (part == 0 && i < nw) {
      for (int t = 1; t < ks; ++t) {
        re += red[threadIdx.x + t].x;
        im += red[threadIdx.x + t].y;
      }
      out[i] = make_double2(re, im);
    }
    __syncthreads();
  }
}

void fill_args(tups_ctx* c, GtauArgs* a, double u) {
  const tups_sizes& s = c->sz;
  a->sigma = c->d_sigma;
  a->sigh = c->d_sigh;
  a->st = c->d_gstate;
  a->nx = s.nx;
  a->nf = s.nf;
  a->ng = c->gt_ngrain;
  a->n1 = s.nx * c->gt_ngrain;
  a->n2 = s.ny * c->gt_ngrain;
  a->nfine = a->n1 * a->n2;
  a->nbk = (a->nfine + 255) / 256;
  a->beta = c->beta;
  a->u = u;
  // every piece starts on a 256-byte boundary
  char* w = static_cast<char*>(c->gt_ws);
  const size_t nfine = (size_t)a->nfine, nfs = nfine * s.nf;
  auto take = [&w](size_t bytes) {
    char* p = w;
    w += (bytes + 255) & ~(size_t)255;
    return p;
  };
  a->etab = reinterpret_cast<double*>(take(nfine * sizeof(double)));
  a->partial = reinterpret_cast<double*>(take((size_t)kCand * a->nbk * sizeof(double)));
  a->greal = reinterpret_cast<double*>(take(nfs * sizeof(double)));
  a->xbuf = reinterpret_cast<double2*>(take(nfs * sizeof(double2)));
  a->grt = reinterpret_cast<double2*>(take(nfs * sizeof(double2)));
  a->u2data = reinterpret_cast<double*>(take((size_t)s.nx * s.ny * s.nf * sizeof(double)));
}

}  // namespace

// the G(tau) workspace for Ngrain (allocated outside any stream capture)
int gtau_reserve(tups_ctx* c, int ngrain) {
  const tups_sizes& s = c->sz;
  TUPS_CHECK(ngrain >= 1 && ngrain <= 64, TUPS_EINVAL, "ngrain must be in 1..64");
  TUPS_CHECK(s.nx == s.ny, TUPS_EINVAL, "G(tau) needs Nx = Ny");
  TUPS_CHECK(s.nf >= 6 && s.nf <= 1024, TUPS_EINVAL, "G(tau) needs 6 <= Nf <= 1024 (five-point end slopes)");
  TUPS_CHECK((size_t)s.nx * ngrain <= 1024, TUPS_EINVAL, "G(tau) needs Nx Ngrain <= 1024");
  if (!c->d_gstate) TUPS_HIP(hipMalloc((void**)&c->d_gstate, sizeof(GtauState)));
  const size_t nfine = (size_t)s.nx * ngrain * s.ny * ngrain, nfs = nfine * s.nf;
  const size_t nbk = (nfine + 255) / 256;
  const size_t need = (nfine + kCand * nbk + nfs) * sizeof(double) + 2 * nfs * sizeof(double2) +
                      (size_t)s.nx * s.ny * s.nf * sizeof(double) + 6 * 256;  // six pieces, 256-byte aligned
  if (c->gt_ws_bytes < need) {
    drop_graphs(c);  // captured graphs hold the old workspace
    if (c->gt_ws) TUPS_HIP(hipFree(c->gt_ws));
    c->gt_ws = nullptr;
    c->gt_ws_bytes = 0;
    if (hipMalloc(&c->gt_ws, need) != hipSuccess) {
      (void)hipGetLastError();
      set_error("G(tau) workspace: device allocation failed");
      return TUPS_ENOMEM;
    }
    c->gt_ws_bytes = need;
  }
  if (c->gt_ngrain != ngrain) c->grt_valid = false;
  c->gt_ngrain = ngrain;
  return TUPS_OK;
}

// calc_Gkw_Grt on the device: mode 0 bisection, 1 one midpoint evaluation, 2 the
// current mu; mu from the host value when mu_from_host, else the device mu.
// The workspace must be reserved for the context's Ngrain.
int launch_gtau(tups_ctx* c, int mode, int mu_from_host, double mu_host, double u, double n_target,
                int set_hartree) {
  GtauArgs a{};
  fill_args(c, &a, u);
  const int nf = c->sz.nf;
  if (int rc = timer_begin(c)) return rc;
  hipLaunchKernelGGL(gt_setup, dim3(1), dim3(256), 0, c->stream, a, mu_from_host, mu_host, n_target, mode,
                     set_hartree);
  TUPS_HIP(hipGetLastError());
  const int passes = mode == 2 ? 0 : (mode == 1 ? 1 : kPasses);
  const bool small = (double)kCand * a.nfine * nf <= kSmallWork;
  if (passes > 0 && small) {
    hipLaunchKernelGGL(gt_search_small, dim3(1), dim3(1024), 0, c->stream, a);
    TUPS_HIP(hipGetLastError());
  }
  for (int p = 0; p < (small ? 0 : passes); ++p) {
    hipLaunchKernelGGL(gt_density, dim3((unsigned)a.nbk, kCand), dim3(256), 0, c->stream, a);
    TUPS_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(gt_fine, dim3((unsigned)a.nfine), dim3(64), (size_t)(3 * nf) * sizeof(double2), c->stream, a);
  TUPS_HIP(hipGetLastError());
  hipLaunchKernelGGL(gt_dft1, dim3((unsigned)a.n2, (unsigned)nf), dim3(64),
                     (size_t)a.n1 * (sizeof(double) + sizeof(double2)), c->stream, a);
  TUPS_HIP(hipGetLastError());
  hipLaunchKernelGGL(gt_dft2, dim3((unsigned)a.n2, (unsigned)nf), dim3(64), (size_t)a.n2 * sizeof(double2),
                     c->stream, a);
  TUPS_HIP(hipGetLastError());
  if (int rc = timer_end(c, TUPS_K_GTAU)) return rc;
  c->grt_valid = true;
  return TUPS_OK;
}

// calc_order_2nd added to sigma_out (the SDE's Sigma); needs the device Grt
int launch_order_2nd(tups_ctx* c, double u, double2* sigma_out) {
  TUPS_CHECK(c->grt_valid, TUPS_ESTATE, "use_u2 needs Grt on the device (tups_gkw_grt first)");
  GtauArgs a{};
  fill_args(c, &a, u);
  a.sigma_out = sigma_out;
  const int nf = c->sz.nf, nc = c->sz.nx * c->sz.ny;
  hipLaunchKernelGGL(u2_ggg, dim3((unsigned)nc, (unsigned)nf), dim3(256), 0, c->stream, a);
  TUPS_HIP(hipGetLastError());
  hipLaunchKernelGGL(u2_spline, dim3((unsigned)nc), dim3(256), 4 * (size_t)nf * sizeof(double), c->stream, a);
  TUPS_HIP(hipGetLastError());
  return TUPS_OK;
}

}  // namespace tups

using namespace tups;

extern "C" int tups_chi0(tups_ctx* c, double* chi0_ph, double* chi0_pp, double* chi0_ph_ff, double* chi0_pp_ff) {
  TUPS_CHECK(c && chi0_ph && chi0_pp && chi0_ph_ff && chi0_pp_ff, TUPS_EINVAL, "null argument");
  TUPS_CHECK(c->grt_valid, TUPS_ESTATE, "the bare bubbles need Grt on the device (tups_gkw_grt first)");
  const tups_sizes& s = c->sz;
  GtauArgs a{};
  fill_args(c, &a, 0.0);
  const size_t nfine = (size_t)a.nfine, nfs = nfine * s.nf, nw = s.nf / 2, nred = (size_t)s.nibz * nw;
  const size_t need = (4 * nfine + (size_t)s.nibz * kC0 * s.nf) * sizeof(double) +
                      (4 * nfs + kC0 * nred) * sizeof(double2) + 4 * 256;
  if (c->c0_ws_bytes < need) {
    if (c->c0_ws) TUPS_HIP(hipFree(c->c0_ws));
    c->c0_ws = nullptr;
    c->c0_ws_bytes = 0;
    TUPS_HIP(hipMalloc(&c->c0_ws, need));
    c->c0_ws_bytes = need;
  }
  char* w = static_cast<char*>(c->c0_ws);
  auto take = [&w](size_t bytes) {
    char* p = w;
    w += (bytes + 255) & ~(size_t)255;
    return p;
  };
  double* wts = reinterpret_cast<double*>(take(4 * nfine * sizeof(double)));
  C0Args b{};
  b.ffg = reinterpret_cast<double2*>(take(4 * nfs * sizeof(double2)));
  b.data = reinterpret_cast<double*>(take((size_t)s.nibz * kC0 * s.nf * sizeof(double)));
  b.out = reinterpret_cast<double2*>(take(kC0 * nred * sizeof(double2)));
  b.grt = a.grt;
  b.ibz = c->d_ibz;
  b.nx = s.nx;
  b.nf = s.nf;
  b.n1 = a.n1;
  b.n2 = a.n2;
  b.nfine = a.nfine;
  b.nibz = s.nibz;
  b.beta = c->beta;
  hipLaunchKernelGGL(c0_weights, dim3((unsigned)((nfine + 255) / 256)), dim3(256), 0, c->stream, a, wts);
  TUPS_HIP(hipGetLastError());
  for (int ff = 0; ff < 4; ++ff) {  // FFG: fftf2d of G(k, tau) FF_sdwave(k, ff)^2
    GtauArgs f = a;
    f.wfine = wts + ff * nfine;
    f.grt = const_cast<double2*>(b.ffg) + ff * nfs;
    hipLaunchKernelGGL(gt_dft1, dim3((unsigned)f.n2, (unsigned)s.nf), dim3(64),
                       (size_t)f.n1 * (sizeof(double) + sizeof(double2)), c->stream, f);
    TUPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(gt_dft2, dim3((unsigned)f.n2, (unsigned)s.nf), dim3(64), (size_t)f.n2 * sizeof(double2),
                       c->stream, f);
    TUPS_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(c0_pair, dim3((unsigned)s.nibz, (unsigned)s.nf), dim3(256), 0, c->stream, b);
  TUPS_HIP(hipGetLastError());
  hipLaunchKernelGGL(c0_spline, dim3((unsigned)s.nibz, kC0), dim3(256), 4 * (size_t)s.nf * sizeof(double), c->stream,
                     b);
  TUPS_HIP(hipGetLastError());
  const size_t z = nred * sizeof(double2);
  TUPS_HIP(hipMemcpyAsync(chi0_ph, b.out, z, hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipMemcpyAsync(chi0_pp, b.out + nred, z, hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipMemcpyAsync(chi0_ph_ff, b.out + 2 * nred, 4 * z, hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipMemcpyAsync(chi0_pp_ff, b.out + 6 * nred, 4 * z, hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  return TUPS_OK;
}

namespace tups {

const double2* gtau_grt(const tups_ctx* c) {
  GtauArgs a{};
  fill_args(const_cast<tups_ctx*>(c), &a, 0.0);
  return a.grt;
}

}  // namespace tups
