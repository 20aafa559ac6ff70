// Parquet iteration on uneven boxes: every IBZ transfer q = (m, w) of an Nx
// lattice with its own fermionic box Nf_w (per bosonic frequency) and its own
// form-factor count Nl_m (per IBZ momentum).
//
// The reference fixes one Nf and one Nl for every q (SRC/parquet_ini.f90:
// 289-348), so the BSE of each q is the reference's (SRC/parquet_BSE.f90:
// 62-132) on its own box, but the crossing has no reference semantics here.
// It is defined as the uniform crossing on the largest box (Nf = max Nf_w,
// Nl = max Nl_m) of the vertices embedded in that box, restricted to each q's
// box afterwards:
//   embed:   Phi_max((l1, n1), (l2, n2); q) = Phi_q((l1, c_q(n1)), (l2, c_q(n2)))
//            for l1, l2 < Nl_q, else 0; c_q clamps the fermionic frequency index
//            n (box-centred, n = nu - Nf/2) to q's box, the reference's
//            treatment of frequencies outside the box (pickNu,
//            SRC/parquet_EQContribs.f90:17-39);
//   cross:   calc_PhiR -> calculate_PAE on the largest box (SRC/parquet_PhiR.f90,
//            SRC/parquet_EQOuter.f90:16-118), F_max = Phi_max + crossing + Lambda Nc;
//   extract: F_q((l1, n1), (l2, n2)) = F_max((l1, n1), (l2, n2); q) for q's box.
// On uniform boxes embed and extract are the identity, so the iteration is the
// reference's.  One iteration: batched BSE (one persistent launch for every
// (q, channel), tups_batch_*) -> embed -> PhiR -> PAE -> extract -> G = F - G
// (SRC/loop.f90:219-222) on the uneven boxes, all on one stream, so that it can
// be captured once per buffer parity as a HIP graph and replayed.
#include "internal.h"

#include <algorithm>
#include <string>
#include <vector>

struct tups_uneven {
  tups_ctx* ctx = nullptr;     // the largest box, full IBZ
  tups_batch* batch = nullptr; // q = m Nw + w, the reference's q order
  int nw = 0, nibz = 0, nf_max = 0, nl_max = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int parity = 0;
  int cap_batch_variant = -1, cap_pae_variant = -1;  // the variants the graphs were captured with
  const void* cap_tables = nullptr;                   // and the tables (their device pointers)
};

namespace tups {

namespace {

struct UnevenArgs {
  const double2* bsrc[4];   // batch arrays (packed boxes)
  double2* bdst[4];
  const double2* usrc[4];   // largest-box arrays (Nz^2 Nq)
  double2* udst[4];
  const long long* voff;
  const int* nf;
  const int* nl;
  int nf_max, nz, nq;
};

// batch Phi -> largest box; grid (ceil(Nz^2 / 256), Nq, 4)
__global__ __launch_bounds__(256) void uneven_embed(UnevenArgs a) {
  const int q = blockIdx.y, ch = blockIdx.z;
  const size_t nz2 = (size_t)a.nz * a.nz;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= nz2) return;
  const int row = (int)(e % a.nz), col = (int)(e / a.nz);
  const int l1 = row / a.nf_max, l2 = col / a.nf_max;
  const int nfq = a.nf[q], nlq = a.nl[q], h = a.nf_max / 2, hq = nfq / 2;
  double2 v = make_double2(0.0, 0.0);
  if (l1 < nlq && l2 < nlq) {
    const int n1 = min(max(row - l1 * a.nf_max - h, -hq), hq - 1) + hq;
    const int n2 = min(max(col - l2 * a.nf_max - h, -hq), hq - 1) + hq;
    const int D = nlq * nfq;
    v = a.bsrc[ch][a.voff[q] + (size_t)(l1 * nfq + n1) + (size_t)D * (l2 * nfq + n2)];
  }
  a.udst[ch][nz2 * q + e] = v;
}

// largest-box F -> batch boxes; grid (ceil(Dmax^2 / 256), Nq, 4)
__global__ __launch_bounds__(256) void uneven_extract(UnevenArgs a) {
  const int q = blockIdx.y, ch = blockIdx.z;
  const int nfq = a.nf[q], D = a.nl[q] * nfq;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)D * D) return;
  const int rq = (int)(e % D), cq = (int)(e / D);
  const int l1 = rq / nfq, l2 = cq / nfq;
  const int sh = a.nf_max / 2 - nfq / 2;
  const int row = l1 * a.nf_max + rq - l1 * nfq + sh, col = l2 * a.nf_max + cq - l2 * nfq + sh;
  a.bdst[ch][a.voff[q] + e] = a.usrc[ch][(size_t)a.nz * a.nz * q + (size_t)row + (size_t)a.nz * col];
}

// G = F - G on the batch (Gamma = F - Phi); grid-stride over the packed boxes
__global__ __launch_bounds__(256) void uneven_mix(const double2* f0, const double2* f1, const double2* f2,
                                                  const double2* f3, double2* g0, double2* g1, double2* g2,
                                                  double2* g3, size_t n) {
  const int ch = blockIdx.y;
  const double2* f = ch == 0 ? f0 : ch == 1 ? f1 : ch == 2 ? f2 : f3;
  double2* g = ch == 0 ? g0 : ch == 1 ? g1 : ch == 2 ? g2 : g3;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const double2 x = f[i], y = g[i];
    g[i] = make_double2(x.x - y.x, x.y - y.y);
  }
}

UnevenArgs uneven_args(tups_uneven* u) {
  UnevenArgs a{};
  tups_batch* b = u->batch;
  tups_ctx* c = u->ctx;
  for (int r = 0; r < 4; ++r) {
    a.bsrc[r] = b->gamma[r];  // Phi after the batched BSE
    a.bdst[r] = b->fvert[r];
    a.usrc[r] = c->fvert[r];
    a.udst[r] = c->gamma[r];
  }
  a.voff = b->d_voff;
  a.nf = b->d_nf;
  a.nl = b->d_nl;
  a.nf_max = u->nf_max;
  a.nz = c->sz.nz;
  a.nq = c->sz.nq;
  return a;
}

int uneven_sequence(tups_uneven* u, bool* swapped) {
  tups_batch* b = u->batch;
  tups_ctx* c = u->ctx;
  if (int rc = batch_launch(b, b->variant)) return rc;
  for (int r = 0; r < 4; ++r) std::swap(b->gamma[r], b->spare[r]);
  u->parity ^= 1;
  *swapped = true;
  const UnevenArgs a = uneven_args(u);
  const size_t nz2 = (size_t)c->sz.nz * c->sz.nz;
  hipLaunchKernelGGL(uneven_embed, dim3((unsigned)((nz2 + 255) / 256), (unsigned)c->sz.nq, 4), dim3(256), 0,
                     c->stream, a);
  TUPS_HIP(hipGetLastError());
  if (int rc = launch_phir(c)) return rc;
  if (int rc = launch_pae(c)) return rc;
  long long dmax = 0;
  for (int q = 0; q < b->nq; ++q) dmax = std::max<long long>(dmax, (long long)b->nl[q] * b->nf[q]);
  hipLaunchKernelGGL(uneven_extract, dim3((unsigned)((dmax * dmax + 255) / 256), (unsigned)b->nq, 4), dim3(256), 0,
                     c->stream, a);
  TUPS_HIP(hipGetLastError());
  const size_t n = (size_t)b->voff[b->nq];
  const unsigned gx = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(uneven_mix, dim3(gx, 4), dim3(256), 0, c->stream, b->fvert[0], b->fvert[1], b->fvert[2],
                     b->fvert[3], b->gamma[0], b->gamma[1], b->gamma[2], b->gamma[3], n);
  TUPS_HIP(hipGetLastError());
  return TUPS_OK;
}

void uneven_unswap(tups_uneven* u) {
  for (int r = 0; r < 4; ++r) std::swap(u->batch->gamma[r], u->batch->spare[r]);
  u->parity ^= 1;
}

void uneven_drop_graphs(tups_uneven* u) {
  for (auto& g : u->graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

}  // namespace

}  // namespace tups

using namespace tups;

extern "C" int tups_uneven_destroy(tups_uneven* u) {
  if (!u) return TUPS_OK;
  uneven_drop_graphs(u);
  if (u->batch) tups_batch_destroy(u->batch);
  if (u->ctx) tups_destroy(u->ctx);
  delete u;
  return TUPS_OK;
}

extern "C" int tups_uneven_create(int32_t nx, int32_t nf_max, const int32_t* nf_w, const int32_t* nl_m, double beta,
                                  double f_damping, tups_uneven** out) {
  TUPS_CHECK(out && nf_w && nl_m, TUPS_EINVAL, "null argument");
  *out = nullptr;
  TUPS_CHECK(nx >= 2 && nx % 2 == 0, TUPS_EINVAL, "Nx must be even and >= 2");
  TUPS_CHECK(nf_max >= 2 && nf_max % 2 == 0, TUPS_EINVAL, "the largest box Nf must be even and >= 2");
  const int nw = nf_max / 2, nxi = (nx + 2) / 2, nibz = nxi * (nxi + 1) / 2;
  int nl_max = 0;
  bool top = false;
  for (int w = 0; w < nw; ++w) {
    TUPS_CHECK(nf_w[w] >= 2 && nf_w[w] % 2 == 0 && nf_w[w] <= nf_max, TUPS_EINVAL,
               "Nf_w must be even, >= 2 and at most the largest box");
    top = top || nf_w[w] == nf_max;
  }
  TUPS_CHECK(top, TUPS_EINVAL, "some bosonic frequency must have the largest box");
  for (int m = 0; m < nibz; ++m) {
    TUPS_CHECK(nl_m[m] == 1 || nl_m[m] == 5 || nl_m[m] == 9, TUPS_EINVAL, "Nl_m must be 1, 5 or 9");
    nl_max = std::max(nl_max, nl_m[m]);
  }
  tups_uneven* u = new tups_uneven();
  u->nw = nw;
  u->nibz = nibz;
  u->nf_max = nf_max;
  u->nl_max = nl_max;
  tups_dims d{};
  d.nx = nx;
  d.nf = nf_max;
  d.nl = nl_max;
  d.beta = beta;
  d.f_damping = f_damping;
  int rc = tups_create(&d, &u->ctx);
  std::vector<int32_t> nf(nibz * nw), nl(nibz * nw);
  for (int m = 0; m < nibz; ++m)
    for (int w = 0; w < nw; ++w) {
      nf[m * nw + w] = nf_w[w];
      nl[m * nw + w] = nl_m[m];
    }
  if (!rc) rc = tups_batch_create(nx, nibz * nw, nf.data(), nl.data(), beta, f_damping, &u->batch);
  if (!rc) {
    // one stream for the whole iteration, so that it can be captured
    (void)hipStreamSynchronize(u->batch->stream);
    (void)hipStreamDestroy(u->batch->stream);
    u->batch->stream = u->ctx->stream;
    u->batch->own_stream = false;
    rc = tups_build_m(u->ctx);
  }
  if (!rc && nl_max > 1) {
    // the crossing GEMM forms only the rows (l1, l2, m) that the extraction keeps
    std::vector<int> rows;
    for (int m = 0; m < nibz; ++m)
      for (int l2 = 0; l2 < nl_m[m]; ++l2)
        for (int l1 = 0; l1 < nl_m[m]; ++l1) rows.push_back(l1 + nl_max * l2 + nl_max * nl_max * m);
    if (hipMalloc(&u->ctx->pae_rows, rows.size() * sizeof(int)) != hipSuccess ||
        hipMemcpy(u->ctx->pae_rows, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("uneven-box iteration: crossing row table upload failed");
      rc = TUPS_EHIP;
    } else {
      u->ctx->pae_nrows = (int)rows.size();
    }
  }
  if (rc) {
    tups_uneven_destroy(u);
    return rc;
  }
  *out = u;
  return TUPS_OK;
}

extern "C" int tups_uneven_parts(tups_uneven* u, tups_batch** batch, tups_ctx** ctx) {
  TUPS_CHECK(u, TUPS_EINVAL, "null uneven-box iteration");
  if (batch) *batch = u->batch;
  if (ctx) *ctx = u->ctx;
  return TUPS_OK;
}

extern "C" int tups_uneven_iteration(tups_uneven* u, int32_t use_graph) {
  TUPS_CHECK(u, TUPS_EINVAL, "null uneven-box iteration");
  tups_ctx* c = u->ctx;
  int rc = ensure_tables(c);
  if (rc) return rc;
  bool swapped = false;
  if (!use_graph) {
    rc = uneven_sequence(u, &swapped);
    if (!rc) c->phir_valid = true;
    return rc;
  }
  // tile variants and tables are part of the captured launches
  if (u->cap_batch_variant != u->batch->variant || u->cap_pae_variant != c->pae_variant ||
      u->cap_tables != c->d_ph_coef) {
    uneven_drop_graphs(u);
    u->cap_batch_variant = u->batch->variant;
    u->cap_pae_variant = c->pae_variant;
    u->cap_tables = c->d_ph_coef;
  }
  if (!u->graph[u->parity]) {
    const int slot = u->parity;
    const bool timing = c->timing;
    c->timing = false;  // no event nodes in the graph
    TUPS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = uneven_sequence(u, &swapped);
    hipGraph_t gr = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &gr);
    c->timing = timing;
    if (swapped) uneven_unswap(u);  // capture launched nothing: restore the buffers
    if (rc || ec != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      if (!rc) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        rc = TUPS_EHIP;
      }
      return rc;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ei != hipSuccess) {
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      return TUPS_EHIP;
    }
    u->graph[slot] = ex;
  }
  TUPS_HIP(hipGraphLaunch(u->graph[u->parity], c->stream));
  uneven_unswap(u);  // the graph's BSE wrote Phi into the partner buffers
  c->phir_valid = true;
  return TUPS_OK;
}

extern "C" int tups_uneven_sync(tups_uneven* u) {
  TUPS_CHECK(u, TUPS_EINVAL, "null uneven-box iteration");
  TUPS_HIP(hipStreamSynchronize(u->ctx->stream));
  return TUPS_OK;
}
