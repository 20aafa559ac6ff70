/*
 * tups_gpu.h — C ABI of the MI355X vertex hot path of the TUPS parquet solver.
 *
 * The reference (annakauch/TUPS, Fortran) runs one parquet iteration from
 * SRC/loop.f90:29-249.  Its hot path is five routines that take no array
 * arguments and operate on module variables of parquet_ini / parquet_EQMatrix:
 *
 *   BSE(ite)                 SRC/parquet_BSE.f90:19          -> tups_bse
 *   calculate_expSqR/expqSR  SRC/parquet_PhiR.f90:360,400    -> tups_create (tables)
 *   calc_PhiR(ite)           SRC/parquet_PhiR.f90:35         -> tups_phir
 *   calc_M_*()               SRC/parquet_EQMatrix.f90:21,126,198,274 -> tups_build_m
 *   calculate_PAE()          SRC/parquet_EQOuter.f90:16      -> tups_pae
 *   G_r = F_r - G_r          SRC/loop.f90:219-222            -> tups_gamma_update
 *
 * This library keeps those arrays resident in HBM behind an opaque context.
 * The Fortran side (tups_amd/fortran/tups_gpu_iface.f90) binds every entry
 * point below with ISO_C_BINDING; the drop-in module replacing the reference
 * routines is tups_amd/fortran/tups_dropin.f90.
 *
 * Conventions
 *  - All arrays are the reference's Fortran column-major layouts:
 *      vertex  X(Nz, Nz, Nq)           complex(8)   (G_r_LL, F_r_LL)
 *      chi     chi(Nl, Nl, Nf, Nq)     complex(8)   compact l-blocks of the
 *              block-diagonal bubble chi_LL(l1 nu, l2 nu) (SRC/parquet_chi.f90:500-521)
 *      PhiR    PhiR(Nz*Nz, Nw, NR)     complex(8)   (SRC/parquet_ini.f90:263)
 *      Lambda  L(Nf, Nf, Nw)           complex(8)   (SRC/parquet_util.f90:302-305)
 *      M       M(Nl,Nl,Nl,Nl,NIBZ,NR)  complex(8)   stored per IBZ momentum; the
 *              reference's per-q copy M(...,q,R) equals M(...,momentum(q),R)
 *    complex(8) is passed as double* pointing at interleaved (re, im).
 *  - Integers are 32-bit (the reference uses default integer / MKL lp64).
 *  - Every entry point returns 0 on success and a negative TUPS_E* code on
 *    failure; tups_last_error() returns a message for the calling thread.
 *  - One host thread drives one context; calls are stream-ordered on the
 *    context's HIP stream and return without synchronising unless stated.
 */
#ifndef TUPS_GPU_H
#define TUPS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TUPS_OK 0
#define TUPS_EINVAL -1    /* bad argument / shape (the reference stops, SRC/preprocessing.f90:95-98) */
#define TUPS_EHIP -2      /* HIP runtime error */
#define TUPS_ENOMEM -3    /* device allocation failed */
#define TUPS_ESTATE -4    /* call order violated (e.g. PAE before M) */
#define TUPS_ENONFINITE -5/* non-finite value found by tups_check_finite */
#define TUPS_ENOCONV -6   /* the mu bisection did not reach |n - nParticle| <= 1e-6 in 64 steps */

/* channel numbering of the reference: d, m, s, t (SRC/parquet_BSE.f90:69-128) */
#define TUPS_CH_D 0
#define TUPS_CH_M 1
#define TUPS_CH_S 2
#define TUPS_CH_T 3

/* resident arrays addressed by tups_upload / tups_download */
#define TUPS_GAMMA 0   /* G_r_LL (Nz,Nz,Nq) per channel; holds Phi_r after tups_bse */
#define TUPS_F 1       /* F_r_LL (Nz,Nz,Nq) per channel */
#define TUPS_CHI 2     /* chi(Nl,Nl,Nf,Nq); channel 0 = ph bubble, 1 = pp bubble */
#define TUPS_PHIR 3    /* PhiR_r (Nz*Nz,Nw,NR) per channel */
#define TUPS_M 4       /* M(Nl,Nl,Nl,Nl,NIBZ,NR); channel 0..3 = phbar_to_ph,
                          pp_to_ph, ph_to_pp, phbar_to_pp */
#define TUPS_LAMBDA 5  /* L_r (Nf,Nf,Nw) per channel */

/* BSE modes */
#define TUPS_BSE_LADDER_FP64 0  /* the reference's damped ladder, FP64 MFMA */
#define TUPS_BSE_LADDER_SPLIT 1 /* same update; the GEMM runs as exact INT8 MFMA products of
                                   residues modulo Gaussian moduli, recombined in FP64 */
#define TUPS_BSE_CLOSED_LU 2    /* closed form: the ladder's fixed point for fixed Gamma,
                                   Phi = (I - kappa_r Gamma chi)^-1 Gamma - Gamma, kappa_r =
                                   2 c_r alpha, by pivoted Gauss-Jordan (FP64 MFMA updates);
                                   returns TUPS_ENONFINITE on a zero pivot column */
#define TUPS_BSE_CLOSED_IR 3    /* the same closed form by FP32 LU (pivoted Gauss-Jordan on
                                   [fp32(A) | I], FP32 MFMA) and FP64 iterative refinement
                                   X += A32^-1 (Gamma - A X) until |R|_F / |Gamma|_F <= the
                                   tolerance of tups_set_refinement (default 1e-12, at most
                                   10 steps; TUPS_ENOCONV otherwise) */

typedef struct tups_dims {
  int32_t nx;        /* Nx = Ny (even), SRC/parquet_ini.f90:304-305 */
  int32_t nf;        /* Nf fermionic box (even) */
  int32_t nl;        /* Nl form factors: 1, 5 or 9 (sym_op_L, SRC/parquet_util.f90:756) */
  int32_t nr;        /* NR real-space shells; 0 = reference choice {1,4,6} */
  int32_t nq;        /* q points held; 0 = all of IBZ x Nf/2 (one rank holds everything) */
  int32_t reserved;
  double beta;       /* inverse temperature */
  double f_damping;  /* BSE damping f, prms line 11 */
} tups_dims;

typedef struct tups_ctx tups_ctx;

/* The run-parameter file (set_parameters, SRC/parquet_ini.f90:289-329): 21
 * records, one list-directed value each, in this order.  Logicals are 0/1;
 * character values are cut to the reference's 100 characters. */
typedef struct tups_prms {
  double beta, u, n_particle;             /* beta, xU, nParticle */
  int32_t nx, nf, neighbours, readinpick; /* Nx (= Ny), Nf, neighbours, readinpick */
  int32_t ngrain, f_range, ite_max, dga;  /* Ngrain, f_range, ite_max, DGA */
  int32_t use_sig_dmft, use_u2, save_f, oneshot, old, sde_old;
  double f_damping;
  char pick_file[104], fname_l[104], fname_sig_dmft[104];
} tups_prms;
/* Reads the file at `path` with the reference's list-directed rules (first
 * value of each record, D or E exponents, .true./T/F logicals, quoted or bare
 * strings); TUPS_EINVAL names the line and field of the first bad value. */
int tups_read_prms(const char* path, tups_prms* prms);
/* The context sizes the parameters imply: Nx, Nf, Nl from neighbours (1, 2, 3
 * -> 1, 5, 9 form factors; SRC/parquet_formfactors.f90:495-515), beta,
 * f_damping; NR and Nq left to the library's defaults. */
int tups_prms_dims(const tups_prms* prms, tups_dims* dims);

/* derived sizes returned by tups_query: */
typedef struct tups_sizes {
  int32_t nx, ny, nf, nl, nz, nr, nw, nibz, nq, nc, full_ibz;
  int32_t reserved;
  double alpha; /* 1/(beta*Nx*Ny), SRC/parquet_BSE.f90:72 */
} tups_sizes;

const char* tups_last_error(void);

/* Host-only (no device): the one-time tables for `dims`, in the reference's
 * layouts FF(Nx,Ny,Nl), FF_inv(8,Nx,Ny,Nl), signs(Nl), reps(Nl),
 * expSqR(NIBZ,NR,8) and expqSR(Nx*Ny,NR,8) complex (interleaved); any output
 * pointer may be NULL.  Replaces FF_array/init_FF_inv
 * (SRC/parquet_formfactors.f90:11-90,545-569) and calculate_expSqR/expqSR
 * (SRC/parquet_PhiR.f90:360-435). */
int tups_build_tables(const tups_dims* dims, double* ff, double* ff_inv, int32_t* signs,
                      int32_t* reps, double* exp_sqr, double* exp_qsr);

/* --- context: replaces set_parameters/Initialization's allocations and the
 *     one-time tables of SRC/preprocessing.f90:19-133, SRC/parquet_PhiR.f90:360-435 */
int tups_create(const tups_dims* dims, tups_ctx** out);
int tups_destroy(tups_ctx* ctx);
int tups_query(const tups_ctx* ctx, tups_sizes* out);
/* host tables computed by the library (FF, FF_inv, signs, reps, expSqR, expqSR,
 * Index_bosonic_IBZ) — a caller that computed its own may overwrite them: */
int tups_get_tables(const tups_ctx* ctx, double* ff, double* ff_inv, int32_t* signs,
                    int32_t* reps, double* exp_sqr, double* exp_qsr);
int tups_set_tables(tups_ctx* ctx, const double* ff, const double* ff_inv,
                    const int32_t* signs, const int32_t* reps, const double* exp_sqr,
                    const double* exp_qsr);
int tups_get_q_map(const tups_ctx* ctx, int32_t* qx, int32_t* qy, int32_t* qw);
/* synthetic q batches (BSE only): explicit (qx,qy,qw) per slot, 1-based */
int tups_set_q_map(tups_ctx* ctx, const int32_t* qx, const int32_t* qy, const int32_t* qw);

/* --- host <-> HBM transfers (synchronous w.r.t. the host buffer) */
int tups_upload(tups_ctx* ctx, int32_t which, int32_t channel, const double* host);
int tups_download(tups_ctx* ctx, int32_t which, int32_t channel, double* host);
/* pack a dense (Nz,Nz,nqs) bubble block for slots [q0, q0+nqs) into TUPS_CHI */
int tups_upload_chi_dense(tups_ctx* ctx, int32_t pp, int32_t q0, int32_t nqs,
                          const double* dense);

/* --- the hot path (stream-ordered) */
/* BSE(ite): Gamma_r <- Phi_r for r = d,m,s,t; SRC/parquet_BSE.f90:62-132 */
int tups_bse(tups_ctx* ctx, int32_t mode);
/* TUPS_BSE_CLOSED_IR: refine until |Gamma - A X|_F / |Gamma|_F <= tol for every
 * (q, r), with at most max_steps refinement steps (defaults 1e-12, 10) */
int tups_set_refinement(tups_ctx* ctx, double tol, int32_t max_steps);
/* after a closed-form tups_bse: the refinement steps the last call took (the
 * most over its passes) and the largest relative residual it reached; 0 and
 * NaN after the direct mode, which computes no residual */
int tups_bse_solve_info(const tups_ctx* ctx, int32_t* steps, double* residual);
/* calc_chi for every q slot (SRC/parquet_chi.f90:438-523 with get_green_coarse,
 * SRC/parquet_util.f90:1303-1327): fills the resident TUPS_CHI bubbles (ph:
 * channel 0, pp: channel 1) from the self-energy.  sigma: Sigma(Nt) complex,
 * Nt = Nx Ny Nf, index ((kx-1) Ny + ky-1) Nf + nu (the reference's layout);
 * sigma_h: Sigma_H(Nx, Ny) complex column-major, used outside the frequency box,
 * or NULL for the reference's rule Sigma_H(kx, ky) = Sigma(k, Nf)
 * (SRC/parquet_chi.f90:103-110); mu the chemical potential; ngrain the
 * coarse-graining Ngrain.  Synchronous w.r.t. the host buffers. */
int tups_calc_chi(tups_ctx* ctx, const double* sigma, const double* sigma_h, double mu, int32_t ngrain);

/* calc_SDE (SRC/parquet_SDE.f90:27-258): the self-energy over the whole zone
 * from the resident reducible vertices G_r_LL (the GAMMA arrays, holding Phi
 * after the crossing as at SRC/loop.f90:154) and Lambda (L_d, L_m), in FP64.
 * sigma, sigma_h, mu, ngrain as for tups_calc_chi (they build G); f_range the
 * enlarged frequency box of the auxiliary bubbles (0..8); u the Hubbard U;
 * use_u2 selects the reference's use_U2 form of LChi and adds the second-order
 * term calc_order_2nd (:263-340) from the Grt of the last tups_gkw_grt (the
 * reference passes calc_Gkw_Grt's Grt to calc_SDE; TUPS_ESTATE without one for
 * this ngrain).  sigma_out receives Sigma (Nx Ny Nf complex, the reference's
 * Sigma layout).  Needs a full-IBZ context. */
int tups_sde(tups_ctx* ctx, const double* sigma, const double* sigma_h, double mu, int32_t ngrain, int32_t f_range,
             double u, int32_t use_u2, double* sigma_out);

/* Split ladder settings (no counterpart in the reference): `moduli` CRT moduli
 * (8..16, default 12: relative error ~2e-11 at Nz = 320; 13 ~1e-12; 16 FP64
 * rounding), `qbatch` q points per residue+GEMM pass (0 = library default:
 * about 4 GiB of residues, at most an eighth of the free device memory, fixed
 * at the first split call).  The workspace holds about 8 N P^2 + 72 P bytes per
 * (q, channel) of one pass, P = Nz rounded up to a multiple of 64 (of 192 when
 * the 192 x 192 GEMM tile is used), plus 2 N P^2 bytes of GEMM scratch in that
 * case; the split mode needs Nz <= 1024. */
int tups_set_split(tups_ctx* ctx, int32_t moduli, int32_t qbatch);
/* One iteration of the path (no counterpart as one call; SRC/loop.f90:78-139,
 * 219-222 calls the steps one by one): BSE(mode) -> calc_PhiR -> calculate_PAE
 * -> G = F - G, for the ladder modes.  use_graph != 0 captures the sequence
 * into a HIP graph on first use (per mode and vertex-buffer parity) and replays
 * it afterwards.  Needs a full-IBZ context with the M tensors built. */
int tups_iteration(tups_ctx* ctx, int32_t mode, int32_t use_graph);
/* Sigma on the device (no counterpart: the reference keeps Sigma in host
 * module arrays): tups_set_sigma uploads Sigma (Nx Ny Nf complex), Sigma_H
 * (NULL: Sigma(k, Nf), also after every device update), mu and ngrain for
 * tups_iteration_full; tups_get_sigma downloads the current Sigma. */
int tups_set_sigma(tups_ctx* ctx, const double* sigma, const double* sigma_h, double mu, int32_t ngrain);
int tups_get_sigma(tups_ctx* ctx, double* sigma);
/* calc_Gkw_Grt (SRC/parquet_chi.f90:18-210) from the Sigma on the device
 * (tups_set_sigma, tups_calc_chi or tups_sde): Sigma_H(k) <- Sigma(k, Nf) on the
 * device; mu bisected from *mu_inout between mu - u and mu + u until the
 * particle number n = -2 Grt(r = 0, tau = beta) is within 1e-6 of n_particle
 * (cal_chi != 0: one evaluation at the midpoint and no search, as the
 * reference's early exit); G(k, tau) on the Nx Ngrain x Ny Ngrain grid and its
 * 2D transform Grt.  *mu_inout receives the new mu, which also becomes the
 * device mu; steps_out (may be NULL) the number of evaluations; grt_out (may
 * be NULL) Grt(Nx Ngrain, Ny Ngrain, Nf) complex, column-major.  Grt stays on
 * the device for tups_sde's use_u2 term.  TUPS_ENOCONV when 64 evaluations do
 * not reach the tolerance (the reference loops on), TUPS_ENONFINITE when n is
 * not finite.  Needs Nx = Ny, 6 <= Nf <= 1024, Nx Ngrain <= 1024.
 * Synchronous. */
int tups_gkw_grt(tups_ctx* ctx, double u, double n_particle, int32_t cal_chi, double* mu_inout, int32_t* steps_out,
                 double* grt_out);
/* The bare bubbles of calc_Gkw_Grt with cal_chi (SRC/parquet_chi.f90:216-430)
 * from the Grt (and G(k, tau)) of the last tups_gkw_grt: Chi0_ph, Chi0_pp
 * (Nred = NIBZ Nf/2 complex each, the reference's IBZ order) and Chi0_ph_FF,
 * Chi0_pp_FF (Nred, 4) complex, column-major, for the external form factors
 * 1, cos kx - cos ky, sin kx, sin kx sin ky.  Synchronous. */
int tups_chi0(tups_ctx* ctx, double* chi0_ph, double* chi0_pp, double* chi0_ph_ff, double* chi0_pp_ff);
/* The device mu (no counterpart: the reference keeps mu in a module variable)
 * and the number of evaluations of the last search; returns the search's
 * failure, if any, as tups_gkw_grt would.  Synchronous. */
int tups_get_mu(tups_ctx* ctx, double* mu, int32_t* steps);
/* conv() (SRC/parquet_check.f90:39-150) on the device.  e = |SigmaOld - Sigma| /
 * |Sigma| elementwise over Nx Ny Nf: sigma_avg = |sum(e)| / Nt, sigma_max =
 * max(e) at 1-based (kx, ky, nu) (the first maximum, as maxloc); per channel
 * f_max = max |F_r| over the resident F and f_rel_err = |1 - f_max / f_max of
 * the previous call on this context| (infinite on the first call, where the
 * reference's previous value is unset); converged = sigma_max < f_sigma and
 * every f_rel_err < f_vert (the reference uses 1e-8 for both,
 * SRC/parquet_ini.f90:31-32).  sigma_old and sigma: host arrays (Nx Ny Nf
 * complex), or both NULL for the pair the last tups_iteration_full left on the
 * device.  Synchronous. */
typedef struct tups_conv_info {
  double sigma_avg, sigma_max;
  int32_t kx, ky, nu, converged;
  double f_max[4], f_rel_err[4]; /* channels d, m, s, t */
} tups_conv_info;
int tups_conv(tups_ctx* ctx, const double* sigma_old, const double* sigma, double f_sigma, double f_vert,
              tups_conv_info* info);
/* tups_conv without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight loops, harnesses that forbid hipMalloc).  Same arguments, results and
 * errors as tups_conv; the block partials, the result and the staged host Sigma
 * pair live in a fixed device array of this library (2.1 MiB, one copy per
 * GPU, held until the process ends) instead of a workspace allocated at the
 * first call.  Host Sigma arrays are limited to Nx Ny Nf <= 65536 elements
 * (TUPS_EINVAL beyond; the device pair, NULL and NULL, has no limit).  Every
 * call on one GPU shares that array: calls that overlap on one GPU (from two
 * host threads, or on two contexts' streams) race and must be serialised by
 * the caller.  Synchronous. */
int tups_conv_noalloc(tups_ctx* ctx, const double* sigma_old, const double* sigma, double f_sigma, double f_vert,
                      tups_conv_info* info);
/* calc_sus (SRC/parquet_sus.f90:18-362) from the resident vertices and the
 * Sigma and mu on the device: sus_out receives sus_X_IBZ(q, ff) for X = d, m,
 * s, t, as [X][ff][q] complex (4 x 4 x Nq; ff the external form factors 1,
 * cos kx - cos ky, sin kx, sin kx sin ky of fill_FF_sdwave,
 * SRC/parquet_formfactors.f90:792-815).  u the Hubbard U of the bare vertex
 * outside the box; f_range the enlarged box; after_update = 0 when the GAMMA
 * arrays hold Phi (the reference calls calc_sus before G = F - G,
 * SRC/loop.f90:193-222), 1 when they hold Gamma = F - Phi (after
 * tups_iteration_full).  The reference precedes it with calc_Gkw_Grt(cal_chi):
 * tups_gkw_grt with cal_chi != 0 sets that mu and Sigma_H.  Needs a full-IBZ
 * context, Nx <= 16.  Synchronous. */
int tups_calc_sus(tups_ctx* ctx, double u, int32_t f_range, int32_t after_update, double* sus_out);
/* The loop body of SRC/loop.f90:67-222 on the device (no counterpart as one
 * call): calc_Gkw_Grt's mu search when n_particle >= 0 (else mu stays as set)
 * -> calc_chi from the device Sigma -> BSE(mode) -> calc_PhiR -> calculate_PAE
 * -> SDE (tups_sde's arguments f_range, u, use_u2; with use_u2 the
 * second-order term from this iteration's Grt) -> the new Sigma replaces the
 * old -> G = F - G.  use_graph != 0 captures the sequence once per (mode,
 * buffer parity, SDE and search settings, Ngrain) as a HIP graph; the search
 * then runs a fixed number of launches, so a failed search is reported by
 * tups_get_mu. */
int tups_iteration_full(tups_ctx* ctx, int32_t mode, int32_t use_graph, int32_t f_range, double u, int32_t use_u2,
                        double n_particle);

/* --- batching form (no counterpart in the reference, which fixes Nf and Nl for
 * every q): a batch of nq transfer points, each with its own even Nf_q and
 * Nl_q in {1, 5, 9}, D_q = Nl_q Nf_q.  GAMMA and F (4 channels) are packed q
 * after q as D_q x D_q column-major blocks, CHI (ph: 0, pp: 1) as (Nl_q, Nl_q,
 * Nf_q) blocks.  tups_batch_bse runs the damped ladder of every (q, channel)
 * in one persistent launch, largest D_q first; Phi replaces GAMMA as in
 * tups_bse.  `ms` (may be NULL) receives the launch time. */
typedef struct tups_batch tups_batch;
int tups_batch_create(int32_t nx, int32_t nq, const int32_t* nf, const int32_t* nl, double beta,
                      double f_damping, tups_batch** out);
int tups_batch_destroy(tups_batch* batch);
int tups_batch_query(const tups_batch* batch, int64_t* vertex_elems, int64_t* chi_elems, int32_t* items);
int tups_batch_upload(tups_batch* batch, int32_t which, int32_t channel, const double* host);
int tups_batch_download(tups_batch* batch, int32_t which, int32_t channel, double* host);
int tups_batch_bse(tups_batch* batch, float* ms);
/* Split (INT8 CRT, `moduli` in 8..16) form of tups_batch_bse: the same update
 * of every (q, channel) (SRC/parquet_BSE.f90:62-132), the q of one (Nf_q, Nl_q)
 * class run together as passes of the context's split kernels
 * (tups_bse(ctx, TUPS_BSE_LADDER_SPLIT)); 12 moduli keep it within about 1e-10
 * of the FP64 form.  Allocates its residue workspace on first use. */
int tups_batch_bse_split(tups_batch* batch, int32_t moduli, float* ms);
int tups_batch_sync(tups_batch* batch);
/* Project autotuner for the batch: times every tile variant (tile choice per
 * Nl, XCD-striped or plain item order) on the resident batch without changing
 * it and keeps the fastest; tups_variant_name(TUPS_K_BSE_BATCH, v) names v. */
int tups_batch_autotune(tups_batch* batch, int32_t* variant, float* ms);
int tups_batch_set_variant(tups_batch* batch, int32_t variant);
int tups_batch_set_product_forms(tups_batch* batch, int32_t allow_three);

/* --- uneven-box iteration (no counterpart in the reference, which fixes one Nf
 * and Nl for every q): the IBZ q = (m, w) of an Nx lattice, q = m Nw + w as in
 * the reference (Nw = nf_max / 2), with the fermionic box nf_w[w] (even, at
 * most nf_max, nf_max for some w) per bosonic frequency and nl_m[m] in {1, 5, 9}
 * per IBZ momentum.  It owns a batch (the per-q boxes; upload GAMMA, F and CHI
 * there through tups_batch_upload) and a context for the largest box (Nf =
 * nf_max, Nl = max nl_m, full IBZ; upload LAMBDA there; its M tensors are
 * built at creation), both on one stream.  tups_uneven_parts hands them out
 * (owned by the iteration: do not destroy them; the context's crossing forms
 * only the rows (l1, l2, Q) with l1, l2 < Nl of Q, the ones the boxes keep).
 * tups_uneven_iteration runs BSE (one persistent launch,
 * SRC/parquet_BSE.f90:62-132 per q) -> crossing -> G = F - G
 * (SRC/loop.f90:219-222) with the crossing defined as the reference's
 * calc_PhiR + calculate_PAE on the largest box of the vertices embedded in it
 * (form factors beyond Nl_q zero, frequencies outside q's box clamped to its
 * edge, as pickNu does) and restricted to each q's box; on uniform boxes this
 * is the reference's iteration.  use_graph != 0 captures it once per buffer
 * parity as a HIP graph and replays it.  Stream-ordered: tups_uneven_sync
 * waits. */
typedef struct tups_uneven tups_uneven;
int tups_uneven_create(int32_t nx, int32_t nf_max, const int32_t* nf_w, const int32_t* nl_m, double beta,
                       double f_damping, tups_uneven** out);
int tups_uneven_destroy(tups_uneven* u);
int tups_uneven_parts(tups_uneven* u, tups_batch** batch, tups_ctx** ctx);
int tups_uneven_iteration(tups_uneven* u, int32_t use_graph);
int tups_uneven_sync(tups_uneven* u);

/* calc_PhiR(ite): PhiR_r from Gamma_r (holding Phi_r); SRC/parquet_PhiR.f90:35-109 */
int tups_phir(tups_ctx* ctx);
/* calc_M_*: the four transformation tensors; SRC/parquet_EQMatrix.f90:21-360 */
int tups_build_m(tups_ctx* ctx);
/* calculate_PAE: F_r <- Phi_r + Lambda_r*Nc + crossing; SRC/parquet_EQOuter.f90:16-118 */
int tups_pae(tups_ctx* ctx);
/* G_r = F_r - G_r; SRC/loop.f90:219-222 */
int tups_gamma_update(tups_ctx* ctx);

/* Project autotuner: times the tile variants of `kernel` (TUPS_K_BSE_LADDER,
 * TUPS_K_SPLIT_GEMM — the GEMM tile of the split ladder, timed as whole split
 * steps — or TUPS_K_PAE for Nl = 1 after PhiR and M exist) on the context's
 * resident inputs and keeps the fastest for later calls.
 * Synchronous.  `variant` receives the chosen variant id, `ms` its time per
 * launch; tups_variant_name() names a variant id. */
int tups_autotune(tups_ctx* ctx, int32_t kernel, int32_t* variant, float* ms);
/* Whether the autotuner of the FP64 ladder (and of a batch,
 * tups_batch_set_product_forms) may pick the three-multiplication complex
 * products (Re = P1 - P2, Im = P3 - P1 - P2).  They meet the relative Frobenius
 * bar of the complex result, but a small imaginary part next to a large real
 * one loses relative accuracy (7e-9 where the four-product form, the
 * reference's ZGEMM arithmetic, keeps 1e-15), so the default is 0: the four-
 * product variants only.  tups_set_variant can still force any variant. */
int tups_set_product_forms(tups_ctx* ctx, int32_t allow_three);
const char* tups_variant_name(int32_t kernel, int32_t variant);
/* forces a variant id (must match the context's Nl); -1 restores the default.
 * TUPS_K_BSE_CLOSED and TUPS_K_BSE_IR choose the elimination of the closed
 * forms: 0 "gj_panel16" (16-column panels), 1 "gj_block64" (64-column blocks,
 * each passing the later columns through memory once), 2 "gj_block64_rp" (the
 * same with each 16-column sub-panel factored on registers, D <= 512, above
 * that as 1; the default from D = 128 on).  The closed forms work in passes
 * whose workspace is at most 16 GiB and half of the free device memory. */
int tups_set_variant(tups_ctx* ctx, int32_t kernel, int32_t variant);
/* number of variant ids of `kernel` (ids 0 .. n-1; a ladder or crossing
 * variant serves one Nl, a split-GEMM or elimination variant every Nl) */
int tups_variant_count(int32_t kernel);

/* --- stream control and timing */
int tups_sync(tups_ctx* ctx);
/* records event `slot` (0..15) on the context stream */
int tups_event_record(tups_ctx* ctx, int32_t slot);
/* milliseconds between two recorded events (synchronises on `b`) */
int tups_event_elapsed(tups_ctx* ctx, int32_t a, int32_t b, float* ms);
/* per-kernel timing: when on, every kernel launch is bracketed by HIP events
 * on the context stream and synchronised (serialises launches; for profiling
 * runs only).  tups_set_timing resets the counters. */
#define TUPS_K_BSE_LADDER 0   /* FP64 MFMA ladder GEMM, chi mixing and damping fused */
#define TUPS_K_PHIR 1
#define TUPS_K_BUILD_M 2
#define TUPS_K_PAE 3
#define TUPS_K_GAMMA_UPDATE 4
#define TUPS_K_PAE_GATHER 5   /* Nl = 1: PhiR re-indexed for contiguous crossing reads */
#define TUPS_K_SPLIT_PREP 6   /* split ladder: operands scaled to integers and reduced to residues */
#define TUPS_K_SPLIT_GEMM 7   /* split ladder: INT8 MFMA residue products, CRT, damping */
#define TUPS_K_BSE_CLOSED 8   /* closed form: build of [I - kappa Gamma chi | Gamma] and its solve */
#define TUPS_K_CALC_CHI 9     /* bubbles chi_ph, chi_pp for every q slot */
#define TUPS_K_BSE_BATCH 10   /* batching form: variant names only (tups_variant_name/count) */
#define TUPS_K_SDE 11         /* self-energy: auxiliary bubbles, vertex contraction, star sums */
#define TUPS_K_GTAU 12        /* G(tau): mu bisection, fine-grid G(k, tau), 2D transform */
#define TUPS_K_CONV 13        /* convergence check: Sigma change, vertex maxima */
#define TUPS_K_SUS 14         /* susceptibilities: Green's-function pairs and their contraction */
#define TUPS_K_BSE_IR 15      /* refined closed form: FP32 inverse, products with it, FP64 residuals */
#define TUPS_K_COUNT 16
int tups_set_timing(tups_ctx* ctx, int32_t on);
int tups_get_timing(const tups_ctx* ctx, int32_t kernel, double* total_ms, int64_t* launches);
/* scans the resident array for NaN/Inf; returns TUPS_ENONFINITE if any */
int tups_check_finite(tups_ctx* ctx, int32_t which, int32_t channel);

/* --- synthetic inputs on the device (benchmarks): Gamma_r = U*I + X_r with
 * X_r complex Gaussian, Re/Im iid N(0, sigma^2/2), from a counter-based
 * splitmix64 stream (seed, array, channel, element).  which = TUPS_GAMMA or TUPS_F. */
int tups_synth_vertex(tups_ctx* ctx, int32_t which, int32_t channel, uint64_t seed,
                      double diag, double sigma);
/* copies resident array (which_src, ch) onto (which_dst, ch) */
int tups_copy(tups_ctx* ctx, int32_t which_dst, int32_t which_src, int32_t channel);

#ifdef __cplusplus
}
#endif
#endif /* TUPS_GPU_H */
