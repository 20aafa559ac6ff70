// Internal declarations shared by the TUPS MI355X library sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tups_gpu.h"

namespace tups {

constexpr int kNs = 8;       // C4v operations, SRC/parquet_ini.f90:26
constexpr int kMaxNR = 16;   // real-space shells (the PhiR kernel takes them eight at a time)
constexpr int kEvents = 16;

void set_error(const std::string& msg);

#define TUPS_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::tups::set_error(std::string(#call) + ": " + hipGetErrorString(e_));        \
      return TUPS_EHIP;                                                             \
    }                                                                               \
  } while (0)

#define TUPS_CHECK(cond, code, msg)  \
  do {                               \
    if (!(cond)) {                   \
      ::tups::set_error(msg);        \
      return (code);                 \
    }                                \
  } while (0)

// One-time host tables (the reference computes these on the host as well).
struct Tables {
  int nx = 0, nf = 0, nl = 0, nr = 0, nw = 0, nibz = 0, nc = 0, nx_ibz = 0;
  std::vector<int> ibz_x, ibz_y;        // IBZ momenta (1-based), index m
  std::vector<uint8_t> sym_list;        // [m*8 + s] star operations of IBZ point m
  std::vector<double> ff;               // FF(kx,ky,l) Fortran order (Nx,Ny,Nl)
  std::vector<double> ff_inv;           // FF_inv(s,kx,ky,l) Fortran order (Ns,Nx,Ny,Nl)
  std::vector<int> signs, reps;         // per l
  std::vector<int> lsym, lfac;          // [s*nl + l] (0-based l)
  std::vector<double> exp_sqr;          // expSqR(NIBZ,NR,Ns) complex, interleaved
  std::vector<double> exp_qsr;          // expqSR(Nc,NR,Ns) complex, interleaved
  std::vector<int> qx, qy, qw;          // per q slot, 1-based
};

int nl_supported(int nl);
int nr_default(int nl);
int build_tables(int nx, int nf, int nl, int nr, Tables& t);
int rebuild_sym_l(Tables& t);
void build_q_map_full(Tables& t);

// PhiR source table: for every (l1,l2,m) at most four distinct permuted sources.
struct PhirTable {
  std::vector<int> cnt;       // [(l1*nl+l2)*nibz + m]
  std::vector<int> src;       // [... *4 + slot] = l1s + nl*l2s
  std::vector<double> coef;   // [((... *4 + slot)*nr + R)] complex interleaved, includes 1/Nx
};
void build_phir_table(const Tables& t, PhirTable& out);

// The same table by symmetry blocks: the form factors fall into partner groups
// of one or two (the C4v operations map l only inside its group), so for
// groups Ga, Gb the outputs (l1, l2) in Ga x Gb draw only on sources in
// Ga x Gb.  Per block: la0, la1, lb0, lb1 (la1 = la0 for a group of one),
// na, nb, and the offset of its coefficients, laid out [m][R][o][s] with
// o = i + na j for (l1, l2) = (Ga[i], Gb[j]) and s likewise for the source.
// Blocks are ordered by shape (na, nb) = (1,1), (2,1), (1,2), (2,2).
struct PhirBlocks {
  std::vector<int> blk;        // [block][8]
  std::vector<double> coef;    // complex interleaved
  int first[4] = {0, 0, 0, 0}, count[4] = {0, 0, 0, 0};
};
int build_phir_blocks(const Tables& t, const PhirTable& pt, PhirBlocks& out);

// Loads and stores of data touched once per kernel, marked non-temporal: a
// plain G = F - G stream reaches 81% of HBM peak with them against 73% without.
typedef double stream2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_stream(const double2* p) {
  const stream2 v = __builtin_nontemporal_load(reinterpret_cast<const stream2*>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(double2* p, double2 v) {
  stream2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<stream2*>(p));
}

// device state of the chemical-potential search of calc_Gkw_Grt: the current mu,
// the bracket, the last particle number and the candidates of the current pass
struct GtauState {
  double mu, hi, lo, t1, n_target;
  int steps, done, status, single;  // status 0 ok, 1 no convergence, 2 non-finite n
  unsigned arrived;                  // workgroups of the running density pass that are done
  double cand[63];
  double t1c[63];
};

}  // namespace tups

struct tups_ctx {
  tups_sizes sz{};
  double beta = 0, f_damping = 0;
  tups::Tables tab;
  hipStream_t stream = nullptr;
  hipEvent_t ev[tups::kEvents] = {};
  int device = 0;

  // resident arrays (interleaved complex as double2)
  double2* gamma[4] = {};
  double2* fvert[4] = {};
  double2* spare[4] = {};   // ping-pong partner of gamma (Phi output)
  double2* chi[2] = {};
  double2* phir[4] = {};
  double2* mten[4] = {};
  double2* lam[4] = {};

  // device copies of the tables
  double* d_ffk = nullptr;      // [l*nc + k], k = ky*nx + kx
  double* d_ffik = nullptr;     // [(s*nl + l)*nc + k]
  double2* d_eqsr = nullptr;    // [(s*nr + R)*nc + k]
  int* d_qm = nullptr;          // IBZ momentum slot of each q
  int* d_qw = nullptr;          // bosonic frequency of each q (1-based)
  int* d_ibz = nullptr;         // IBZ momenta: x[0..nibz), then y[0..nibz) (1-based)
  int* d_signs = nullptr;
  int* d_ph_blk = nullptr;        // PhiR symmetry blocks (PhirBlocks::blk)
  double2* d_ph_coef = nullptr;   // their coefficients
  int ph_first[4] = {0, 0, 0, 0}, ph_count[4] = {0, 0, 0, 0};
  int* d_flag = nullptr;
  double2* d_sigma = nullptr;   // calc_chi: Sigma(Nx Ny Nf)
  double2* d_sigh = nullptr;    // calc_chi: Sigma_H(Nx, Ny)
  double2* pae_ym = nullptr;    // Nl > 1 crossing: gathered PhiR, 8 x (Nl^2 NR, 64 ceil(Nf / 8)^2) per qw
  double2* pae_y = nullptr;     // Nl = 1 crossing: re-indexed PhiR, 8 x (Nf,Nf,Nw,NR)


  int bse_variant = -1;          // kernel variant chosen by the autotuner (-1: default)
  bool three_mult = false;       // the autotuner may pick the three-multiplication product forms
  int pae_variant = -1;
  int closed_variant = -1;       // closed-form elimination: 0 panels of 16, 1 blocks of 64, 2 blocks with register sub-panels (-1: by D)
  int* pae_rows = nullptr;       // Nl > 1 crossing: the (l1, l2, Q) rows to form (null: all); F's other rows
                                 // are not written (set only by the uneven-box iteration, which reads no other)
  int pae_nrows = 0;
  int split_moduli = 12;         // split ladder: CRT moduli (8..16)
  int split_qbatch = 0;          // split ladder: q points per pass (0: default)
  int split_auto_q = 0;          // split ladder: the default, fixed at first use until set_split
  int split_variant = -1;        // split ladder: GEMM tile variant (-1: chosen by the matrix order)
  void* split_ws = nullptr;      // split ladder: residues and exponents of one pass
  void* sde_ws = nullptr;        // SDE: auxiliary bubbles, contracted vertices, Sigma
  size_t sde_ws_bytes = 0;
  bool green_set = false;        // d_sigma / d_sigh hold a Sigma (calc_chi, sde or set_sigma)
  bool sigh_default = true;      // Sigma_H follows Sigma(k, Nf) when Sigma is replaced on the device
  double green_mu = 0.0;
  int green_ngrain = 1;
  tups::GtauState* d_gstate = nullptr;  // device mu and its search state
  void* gt_ws = nullptr;         // G(tau): energies, partial sums, G(k, tau), transforms, Grt
  size_t gt_ws_bytes = 0;
  int gt_ngrain = 0;             // Ngrain the G(tau) workspace is laid out for
  bool grt_valid = false;        // Grt of the last calc_Gkw_Grt is in the workspace
  void* c0_ws = nullptr;         // bare bubbles: weights, FFG, transformed products, result
  size_t c0_ws_bytes = 0;
  double2* d_sigma_old = nullptr;  // Sigma before the last device SDE update (SigmaOld of conv)
  bool sigma_old_valid = false;
  void* sus_ws = nullptr;        // susceptibilities: Green's-function pairs, result
  size_t sus_ws_bytes = 0;
  void* conv_ws = nullptr;       // convergence check: host Sigma pair, partials, result
  double conv_fmax_prev[4] = {};  // max |F_r| of the previous convergence check
  double full_n_particle = -1.0; // full-iteration graphs: nParticle and Ngrain they were captured with
  int full_ngrain = 0;
  int full_sigh_default = -1;    // and the Sigma_H rule of the Sigma commit (sigh_default)
  hipGraphExec_t graph_full[4] = {};  // full iteration, per (ladder mode, buffer parity)
  int full_f_range = -1, full_use_u2 = 0;  // settings the full-iteration graphs were captured with
  double full_u = 0.0;
  size_t split_ws_bytes = 0;
  void* closed_ws = nullptr;     // closed form: augmented matrices of one pass
  size_t closed_ws_bytes = 0;
  double ir_tol = 1e-12;         // refined closed form: residual bound |R|_F / |Gamma|_F
  int ir_max_steps = 10;         // and the most refinement steps
  int ir_steps = 0;              // last closed-form call: refinement steps taken
  double ir_residual = 0.0;      // and the largest |R|_F / |Gamma|_F reached (NaN for the direct mode)
  bool timing = false;
  hipEvent_t tev[2] = {};
  double t_ms[TUPS_K_COUNT] = {};
  long long t_n[TUPS_K_COUNT] = {};

  int parity = 0;                // vertex ping-pong: number of swaps mod 2
  hipGraphExec_t graph[4] = {};  // one iteration per (ladder mode, parity)

  bool tables_dirty = true;
  bool m_built = false;
  bool phir_valid = false;
};

namespace tups {
// kernels launchers (defined in the .hip sources)
int launch_bse_ladder(tups_ctx* c);
int launch_bse_split(tups_ctx* c);
int bse_split_reserve(tups_ctx* c, int* qbatch);
void swap_vertex_buffers(tups_ctx* c);
int ensure_tables(tups_ctx* c);  // rebuild and upload the tables when stale
void drop_graphs(tups_ctx* c);
int launch_bse_closed(tups_ctx* c);
int launch_bse_closed_ir(tups_ctx* c);
int launch_calc_chi(tups_ctx* c, const double* sigma, const double* sigma_h, double mu, int ngrain);
struct GreenParams;
int upload_sigma(tups_ctx* c, const double* sigma, const double* sigma_h, GreenParams* g, double mu, int ngrain);
int launch_sde(tups_ctx* c, const double* sigma, const double* sigma_h, double mu, int ngrain, int f_range, double u,
               int use_u2, double* sigma_out);
int launch_chi_device(tups_ctx* c, const GreenParams& g);
int sde_reserve(tups_ctx* c, int f_range);
int launch_sde_device(tups_ctx* c, const GreenParams& g, int f_range, double u, int use_u2);
int launch_sigma_commit(tups_ctx* c, int f_range);
int green_params(const tups_ctx* c, GreenParams* g);
int gtau_reserve(tups_ctx* c, int ngrain);
int launch_gtau(tups_ctx* c, int mode, int mu_from_host, double mu_host, double u, double n_target, int set_hartree);
int launch_order_2nd(tups_ctx* c, double u, double2* sigma_out);
const double2* gtau_grt(const tups_ctx* c);
double2* sde_output(tups_ctx* c, int f_range);
int autotune_bse(tups_ctx* c, int reps, int* chosen, float* best_ms);
const char* bse_variant_name(int v);
int bse_variant_count();
int bse_variant_nl(int v);
const char* batch_variant_name(int v);
int batch_variant_count();
int autotune_pae(tups_ctx* c, int reps, int* chosen, float* best_ms);
int autotune_split(tups_ctx* c, int* chosen, float* best_ms);
int split_variant_count();
const char* split_variant_name(int v);
int pae_variant_count();
const char* pae_variant_name(int v);
bool pae_variant_ok(const tups_ctx* c, int v);
int launch_phir(tups_ctx* c);
int launch_build_m(tups_ctx* c);
int launch_pae(tups_ctx* c);
int launch_gamma_update(tups_ctx* c);
int launch_synth(tups_ctx* c, double2* dst, size_t n, int nz, uint64_t seed, uint64_t stream_id,
                 double diag, double sigma);
int launch_check_finite(tups_ctx* c, const double2* p, size_t n);
int upload_tables(tups_ctx* c);
// bracket one kernel launch when per-kernel timing is on
int timer_begin(tups_ctx* c);
// device address of the scratch array of tups_conv_noalloc, once per device (conv.hip)
int conv_noalloc_setup(int device);
int closed_variant_count();
const char* closed_variant_name(int v);
int timer_end(tups_ctx* c, int kernel);
// one launch of the batching form's ladder, variant v, into the spare buffers (bse.hip)
int batch_launch(tups_batch* b, int v);
}  // namespace tups

// the batching form (tups_batch_*): q points with their own Nf_q, Nl_q, packed
// q after q; defined here so the uneven-box iteration can drive it
struct tups_batch {
  int nx = 0, nq = 0;
  double alpha = 0, f_damping = 0;
  std::vector<int> nf, nl;
  std::vector<long long> voff, coff;  // offsets, with totals at [nq]
  hipStream_t stream = nullptr;
  bool own_stream = true;             // false when the stream belongs to an uneven-box iteration
  double2* gamma[4] = {};
  double2* fvert[4] = {};
  double2* spare[4] = {};
  double2* chi[2] = {};
  int4* d_items = nullptr;
  long long* d_voff = nullptr;
  long long* d_coff = nullptr;
  int* d_nf = nullptr;
  int* d_nl = nullptr;
  int* d_tickets = nullptr;
  int nitems = 0;  // work items (tiles)
  int len = 0;     // entries per XCD list, padding included
  int variant = 0;
  bool three_mult = false;            // the autotuner may pick the three-multiplication product forms
  // split form (tups_batch_bse_split): q grouped by (Nf, Nl) class, largest D first
  std::vector<int> qlist;             // q of class 0, then class 1, ...
  std::vector<int> class_start;       // class k holds qlist[class_start[k] .. class_start[k + 1])
  int* d_qlist = nullptr;
  void* split_ws = nullptr;
  size_t split_ws_bytes = 0;
};
