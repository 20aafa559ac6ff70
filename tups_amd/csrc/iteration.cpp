// One parquet iteration of the hot path, eager or as a replayed HIP graph.
//
// The reference's loop (SRC/loop.f90:78-139,219-222) calls, per iteration,
// BSE(ite), calc_PhiR(ite), calculate_PAE and G_r_LL = F_r_LL - G_r_LL; here
// the same four steps run back to back on the context's stream.  With a graph
// the first call of each (ladder mode, buffer parity) captures the launch
// sequence into a hipGraph and later calls replay it: the BSE writes Phi into
// the ping-pong partner of Gamma, so the buffers alternate and two graphs per
// mode cover every iteration.  Nothing is allocated while capturing (the split
// workspace is reserved beforehand) and no memset node is recorded.
#include "green.h"
#include "internal.h"

using namespace tups;

namespace tups {

void swap_vertex_buffers(tups_ctx* c) {
  for (int r = 0; r < 4; ++r) std::swap(c->gamma[r], c->spare[r]);
  c->parity ^= 1;
}

void drop_graphs(tups_ctx* c) {
  for (auto& g : c->graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  for (auto& g : c->graph_full) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

int green_params(const tups_ctx* c, GreenParams* g) {
  TUPS_CHECK(c->green_set, TUPS_ESTATE, "no Sigma on the device (tups_set_sigma, tups_calc_chi or tups_sde first)");
  g->sigma = c->d_sigma;
  g->sigh = c->d_sigh;
  g->nx = c->sz.nx;
  g->nf = c->sz.nf;
  g->ngrain = c->green_ngrain;
  g->mu = c->green_mu;
  g->beta = c->beta;
  return TUPS_OK;
}

static int iteration_sequence(tups_ctx* c, int mode, bool* swapped) {
  int rc = mode == TUPS_BSE_LADDER_SPLIT ? launch_bse_split(c) : launch_bse_ladder(c);
  if (rc) return rc;
  swap_vertex_buffers(c);
  *swapped = true;
  if ((rc = launch_phir(c))) return rc;
  if ((rc = launch_pae(c))) return rc;
  return launch_gamma_update(c);
}

// calc_chi -> BSE -> calc_PhiR -> calculate_PAE -> SDE -> Sigma handed over ->
// G = F - G, with Sigma and mu on the device (SRC/loop.f90:67-222 without the
// mu bisection of calc_Gkw_Grt, the convergence check and the output)
static int full_sequence(tups_ctx* c, int mode, int f_range, double u, int use_u2, bool* swapped) {
  GreenParams g{};
  int rc = green_params(c, &g);
  if (rc) return rc;
  if ((rc = launch_chi_device(c, g))) return rc;
  if ((rc = mode == TUPS_BSE_LADDER_SPLIT ? launch_bse_split(c) : launch_bse_ladder(c))) return rc;
  swap_vertex_buffers(c);
  *swapped = true;
  if ((rc = launch_phir(c))) return rc;
  if ((rc = launch_pae(c))) return rc;
  if ((rc = launch_sde_device(c, g, f_range, u, use_u2))) return rc;
  if ((rc = launch_sigma_commit(c, f_range))) return rc;
  return launch_gamma_update(c);
}

}  // namespace tups

extern "C" int tups_set_sigma(tups_ctx* c, const double* sigma, const double* sigma_h, double mu, int32_t ngrain) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  GreenParams g{};
  return upload_sigma(c, sigma, sigma_h, &g, mu, ngrain);
}

extern "C" int tups_get_sigma(tups_ctx* c, double* sigma) {
  TUPS_CHECK(c && sigma, TUPS_EINVAL, "null argument");
  TUPS_CHECK(c->green_set, TUPS_ESTATE, "no Sigma on the device");
  const size_t nt = (size_t)c->sz.nx * c->sz.ny * c->sz.nf;
  TUPS_HIP(hipMemcpyAsync(sigma, c->d_sigma, nt * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  return TUPS_OK;
}

extern "C" int tups_iteration_full(tups_ctx* c, int32_t mode, int32_t use_graph, int32_t f_range, double u,
                                   int32_t use_u2) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_CHECK(mode == TUPS_BSE_LADDER_FP64 || mode == TUPS_BSE_LADDER_SPLIT, TUPS_EINVAL,
             "an iteration runs one of the ladder modes");
  TUPS_CHECK(c->sz.full_ibz, TUPS_EINVAL, "an iteration needs every IBZ q of the box (nq = 0 at create)");
  TUPS_CHECK(c->m_built, TUPS_ESTATE, "iteration before the M tensors were built");
  TUPS_CHECK(c->green_set, TUPS_ESTATE, "no Sigma on the device (tups_set_sigma first)");
  int rc = ensure_tables(c);
  if (rc) return rc;
  if (mode == TUPS_BSE_LADDER_SPLIT && (rc = bse_split_reserve(c, nullptr))) return rc;
  if ((rc = sde_reserve(c, f_range))) return rc;
  bool swapped = false;
  if (!use_graph) {
    rc = full_sequence(c, mode, f_range, u, use_u2 ? 1 : 0, &swapped);
    if (!rc) c->phir_valid = true;
    return rc;
  }
  if (c->full_f_range != f_range || c->full_u != u || c->full_use_u2 != (use_u2 ? 1 : 0)) {
    drop_graphs(c);  // the settings are part of the captured launches
    c->full_f_range = f_range;
    c->full_u = u;
    c->full_use_u2 = use_u2 ? 1 : 0;
  }
  const int slot = 2 * mode + c->parity;
  if (!c->graph_full[slot]) {
    const bool timing = c->timing;
    c->timing = false;  // no event nodes in the graph
    TUPS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = full_sequence(c, mode, f_range, u, use_u2 ? 1 : 0, &swapped);
    hipGraph_t gr = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &gr);
    c->timing = timing;
    if (swapped) swap_vertex_buffers(c);  // capture launched nothing: restore the buffers
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
    c->graph_full[slot] = ex;
  }
  TUPS_HIP(hipGraphLaunch(c->graph_full[slot], c->stream));
  swap_vertex_buffers(c);  // the graph's BSE wrote Phi into the partner buffers
  c->phir_valid = true;
  return TUPS_OK;
}

extern "C" int tups_iteration(tups_ctx* c, int32_t mode, int32_t use_graph) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_CHECK(mode == TUPS_BSE_LADDER_FP64 || mode == TUPS_BSE_LADDER_SPLIT, TUPS_EINVAL,
             "an iteration runs one of the ladder modes");
  TUPS_CHECK(c->sz.full_ibz, TUPS_EINVAL, "an iteration needs every IBZ q of the box (nq = 0 at create)");
  TUPS_CHECK(c->m_built, TUPS_ESTATE, "iteration before the M tensors were built");
  int rc = ensure_tables(c);
  if (rc) return rc;
  if (mode == TUPS_BSE_LADDER_SPLIT && (rc = bse_split_reserve(c, nullptr))) return rc;
  bool swapped = false;
  if (!use_graph) {
    rc = iteration_sequence(c, mode, &swapped);
    if (!rc) c->phir_valid = true;
    return rc;
  }
  const int slot = 2 * mode + c->parity;
  if (!c->graph[slot]) {
    const bool timing = c->timing;
    c->timing = false;  // no event nodes in the graph
    TUPS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = iteration_sequence(c, mode, &swapped);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    c->timing = timing;
    if (swapped) swap_vertex_buffers(c);  // capture launched nothing: restore the buffers
    if (rc || ec != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (!rc) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        rc = TUPS_EHIP;
      }
      return rc;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      return TUPS_EHIP;
    }
    c->graph[slot] = ex;
  }
  TUPS_HIP(hipGraphLaunch(c->graph[slot], c->stream));
  swap_vertex_buffers(c);  // the graph's BSE wrote Phi into the partner buffers
  c->phir_valid = true;
  return TUPS_OK;
}
