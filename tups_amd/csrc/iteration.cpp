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

#include <cmath>

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
  g->mu_dev = &c->d_gstate->mu;
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

// [calc_Gkw_Grt ->] calc_chi -> BSE -> calc_PhiR -> calculate_PAE -> SDE [+ its
// second-order term] -> Sigma handed over -> G = F - G, with Sigma and mu on the
// device (SRC/loop.f90:67-222 without the convergence check and the output).
// n_particle >= 0 bisects mu as calc_Gkw_Grt does; otherwise mu stays and G(tau)
// is formed only when use_u2 needs it.
static int full_sequence(tups_ctx* c, int mode, int f_range, double u, int use_u2, double n_particle,
                         bool* swapped) {
  GreenParams g{};
  int rc = green_params(c, &g);
  if (rc) return rc;
  if (n_particle >= 0.0) {
    if ((rc = launch_gtau(c, 0, 0, 0.0, u, n_particle, 1))) return rc;
  } else if (use_u2) {
    if ((rc = launch_gtau(c, 2, 0, 0.0, u, 0.0, 1))) return rc;
  }
  if ((rc = launch_chi_device(c, g))) return rc;
  if ((rc = mode == TUPS_BSE_LADDER_SPLIT ? launch_bse_split(c) : launch_bse_ladder(c))) return rc;
  swap_vertex_buffers(c);
  *swapped = true;
  if ((rc = launch_phir(c))) return rc;
  if ((rc = launch_pae(c))) return rc;
  if ((rc = launch_sde_device(c, g, f_range, u, use_u2))) return rc;
  if (use_u2 && (rc = launch_order_2nd(c, u, sde_output(c, f_range)))) return rc;
  if ((rc = launch_sigma_commit(c, f_range))) return rc;
  return launch_gamma_update(c);
}

// the device search state, with its failure turned into a status code
static int read_gstate(tups_ctx* c, GtauState* st) {
  TUPS_HIP(hipMemcpyAsync(st, c->d_gstate, sizeof(GtauState), hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  c->green_mu = st->mu;
  if (st->status == 1) {
    set_error("mu bisection: |n - nParticle| > 1e-6 after 64 steps (mu between mu -/+ U does not bracket it)");
    return TUPS_ENOCONV;
  }
  TUPS_CHECK(st->status != 2, TUPS_ENONFINITE, "mu bisection: the particle number is not finite");
  return TUPS_OK;
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

extern "C" int tups_gkw_grt(tups_ctx* c, double u, double n_particle, int32_t cal_chi, double* mu_inout,
                            int32_t* steps_out, double* grt_out) {
  TUPS_CHECK(c && mu_inout, TUPS_EINVAL, "null argument");
  TUPS_CHECK(c->green_set, TUPS_ESTATE, "no Sigma on the device (tups_set_sigma, tups_calc_chi or tups_sde first)");
  TUPS_CHECK(std::isfinite(u) && std::isfinite(*mu_inout) && std::isfinite(n_particle), TUPS_EINVAL,
             "u, mu and n_particle must be finite");
  int rc = gtau_reserve(c, c->green_ngrain);
  if (rc) return rc;
  if ((rc = launch_gtau(c, cal_chi ? 1 : 0, 1, *mu_inout, u, n_particle, 1))) return rc;
  c->sigh_default = true;  // Sigma_H(k) = Sigma(k, Nf) from here on, as in the reference
  if (grt_out) {
    const size_t n = (size_t)c->sz.nx * c->gt_ngrain * c->sz.ny * c->gt_ngrain * c->sz.nf;
    TUPS_HIP(hipMemcpyAsync(grt_out, gtau_grt(c), n * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  }
  GtauState st{};
  rc = read_gstate(c, &st);
  *mu_inout = st.mu;
  if (steps_out) *steps_out = st.steps;
  return rc;
}

extern "C" int tups_get_mu(tups_ctx* c, double* mu, int32_t* steps) {
  TUPS_CHECK(c && mu, TUPS_EINVAL, "null argument");
  TUPS_CHECK(c->green_set && c->d_gstate, TUPS_ESTATE, "no mu on the device (tups_set_sigma first)");
  GtauState st{};
  const int rc = read_gstate(c, &st);
  *mu = st.mu;
  if (steps) *steps = st.steps;
  return rc;
}

extern "C" int tups_iteration_full(tups_ctx* c, int32_t mode, int32_t use_graph, int32_t f_range, double u,
                                   int32_t use_u2, double n_particle) {
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
  if (!(n_particle >= 0.0)) n_particle = -1.0;  // negative or NaN: mu stays
  const bool gtau = n_particle >= 0.0 || use_u2;
  if (gtau && (rc = gtau_reserve(c, c->green_ngrain))) return rc;
  bool swapped = false;
  if (!use_graph) {
    rc = full_sequence(c, mode, f_range, u, use_u2 ? 1 : 0, n_particle, &swapped);
    if (!rc) c->phir_valid = c->sigma_old_valid = true;
    return rc;
  }
  const int sigh = c->sigh_default ? 1 : 0;
  if (c->full_f_range != f_range || c->full_u != u || c->full_use_u2 != (use_u2 ? 1 : 0) ||
      c->full_n_particle != n_particle || c->full_ngrain != c->green_ngrain || c->full_sigh_default != sigh) {
    drop_graphs(c);  // the settings are part of the captured launches
    c->full_f_range = f_range;
    c->full_u = u;
    c->full_use_u2 = use_u2 ? 1 : 0;
    c->full_n_particle = n_particle;
    c->full_ngrain = c->green_ngrain;
    c->full_sigh_default = sigh;
  }
  const int slot = 2 * mode + c->parity;
  if (!c->graph_full[slot]) {
    const bool timing = c->timing;
    c->timing = false;  // no event nodes in the graph
    TUPS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = full_sequence(c, mode, f_range, u, use_u2 ? 1 : 0, n_particle, &swapped);
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
  c->phir_valid = c->sigma_old_valid = true;
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
