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

}  // namespace tups

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
