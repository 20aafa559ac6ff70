// C ABI of the MI355X vertex hot path: context, resident arrays, transfers.
//
// The reference keeps every array as a module variable (SRC/parquet_ini.f90:128-137,
// 263-270; SRC/parquet_EQMatrix.f90:12-16) allocated in Initialization
// (SRC/parquet_util.f90:349-368).  Here they live in HBM behind an opaque
// context so that an iteration never leaves the device; the host copies in and
// out only at explicit tups_upload / tups_download points.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace tups {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

namespace {

size_t array_elems(const tups_ctx* c, int which) {
  const tups_sizes& s = c->sz;
  switch (which) {
    case TUPS_GAMMA:
    case TUPS_F:
      return (size_t)s.nz * s.nz * s.nq;
    case TUPS_CHI:
      return (size_t)s.nl * s.nl * s.nf * s.nq;
    case TUPS_PHIR:
      return (size_t)s.nz * s.nz * s.nw * s.nr;
    case TUPS_M:
      return (size_t)s.nl * s.nl * s.nl * s.nl * s.nibz * s.nr;
    case TUPS_LAMBDA:
      return (size_t)s.nf * s.nf * s.nw;
    default:
      return 0;
  }
}

double2* array_ptr(tups_ctx* c, int which, int ch) {
  switch (which) {
    case TUPS_GAMMA: return ch >= 0 && ch < 4 ? c->gamma[ch] : nullptr;
    case TUPS_F: return ch >= 0 && ch < 4 ? c->fvert[ch] : nullptr;
    case TUPS_CHI: return ch >= 0 && ch < 2 ? c->chi[ch] : nullptr;
    case TUPS_PHIR: return ch >= 0 && ch < 4 ? c->phir[ch] : nullptr;
    case TUPS_M: return ch >= 0 && ch < 4 ? c->mten[ch] : nullptr;
    case TUPS_LAMBDA: return ch >= 0 && ch < 4 ? c->lam[ch] : nullptr;
    default: return nullptr;
  }
}

int dalloc(double2** p, size_t elems) {
  if (elems == 0) {
    *p = nullptr;
    return TUPS_OK;
  }
  hipError_t e = hipMalloc((void**)p, elems * sizeof(double2));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc of ") + std::to_string(elems * sizeof(double2)) +
              " bytes failed: " + hipGetErrorString(e));
    *p = nullptr;
    return TUPS_ENOMEM;
  }
  return TUPS_OK;
}

template <class T>
int upload_vec(T** dst, const std::vector<T>& src, hipStream_t st) {
  if (*dst) {
    hipFree(*dst);
    *dst = nullptr;
  }
  TUPS_HIP(hipMalloc((void**)dst, std::max<size_t>(1, src.size()) * sizeof(T)));
  if (!src.empty())
    TUPS_HIP(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return TUPS_OK;
}

}  // namespace

int upload_tables(tups_ctx* c) {
  drop_graphs(c);  // graphs hold the table pointers
  Tables& t = c->tab;
  const int nl = t.nl, nc = t.nc;
  int rc;
  std::vector<double> ffik((size_t)kNs * nl * nc);
  for (int s = 0; s < kNs; ++s)
    for (int l = 0; l < nl; ++l)
      for (int k = 0; k < nc; ++k) {
        const int kx = k % t.nx, ky = k / t.nx;
        ffik[((size_t)s * nl + l) * nc + k] = t.ff_inv[s + kNs * (kx + (size_t)t.nx * (ky + (size_t)t.nx * l))];
      }
  if ((rc = upload_vec(&c->d_ffk, t.ff, c->stream))) return rc;
  if ((rc = upload_vec(&c->d_ffik, ffik, c->stream))) return rc;
  std::vector<double> eq(t.exp_qsr);
  if (c->d_eqsr) hipFree(c->d_eqsr);
  c->d_eqsr = nullptr;
  TUPS_HIP(hipMalloc((void**)&c->d_eqsr, std::max<size_t>(1, eq.size()) * sizeof(double)));
  TUPS_HIP(hipMemcpyAsync(c->d_eqsr, eq.data(), eq.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if ((rc = upload_vec(&c->d_qw, t.qw, c->stream))) return rc;
  // IBZ momentum slot of each q (PAE indexes M by momentum)
  std::vector<int> qm(t.qx.size());
  for (size_t q = 0; q < qm.size(); ++q) qm[q] = ((t.qx[q] - 1) * t.qx[q]) / 2 + t.qy[q] - 1;
  if ((rc = upload_vec(&c->d_qm, qm, c->stream))) return rc;
  std::vector<int> ibz(2 * t.nibz);
  for (int m = 0; m < t.nibz; ++m) {
    ibz[m] = t.ibz_x[m];
    ibz[t.nibz + m] = t.ibz_y[m];
  }
  if ((rc = upload_vec(&c->d_ibz, ibz, c->stream))) return rc;
  if ((rc = upload_vec(&c->d_signs, t.signs, c->stream))) return rc;
  PhirTable pt;
  build_phir_table(t, pt);
  PhirBlocks pb;
  if ((rc = build_phir_blocks(t, pt, pb))) return rc;
  if ((rc = upload_vec(&c->d_ph_blk, pb.blk, c->stream))) return rc;
  for (int k = 0; k < 4; ++k) {
    c->ph_first[k] = pb.first[k];
    c->ph_count[k] = pb.count[k];
  }
  if (c->d_ph_coef) hipFree(c->d_ph_coef);
  c->d_ph_coef = nullptr;
  TUPS_HIP(hipMalloc((void**)&c->d_ph_coef, std::max<size_t>(1, pb.coef.size()) * sizeof(double)));
  TUPS_HIP(hipMemcpyAsync(c->d_ph_coef, pb.coef.data(), pb.coef.size() * sizeof(double),
                          hipMemcpyHostToDevice, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  c->tables_dirty = false;
  return TUPS_OK;
}

int timer_begin(tups_ctx* c) {
  if (c->timing) TUPS_HIP(hipEventRecord(c->tev[0], c->stream));
  return TUPS_OK;
}

int timer_end(tups_ctx* c, int kernel) {
  if (!c->timing) return TUPS_OK;
  TUPS_HIP(hipEventRecord(c->tev[1], c->stream));
  TUPS_HIP(hipEventSynchronize(c->tev[1]));
  float ms = 0.f;
  TUPS_HIP(hipEventElapsedTime(&ms, c->tev[0], c->tev[1]));
  c->t_ms[kernel] += ms;
  c->t_n[kernel] += 1;
  return TUPS_OK;
}

int ensure_tables(tups_ctx* c) {
  if (!c->tables_dirty) return TUPS_OK;
  return upload_tables(c);
}

static bool full_ibz(const tups_ctx* c) { return c->sz.full_ibz != 0; }

}  // namespace tups

using namespace tups;

extern "C" {

const char* tups_last_error(void) { return g_err.c_str(); }

int tups_build_tables(const tups_dims* d, double* ff, double* ff_inv, int32_t* signs, int32_t* reps,
                      double* exp_sqr, double* exp_qsr) {
  TUPS_CHECK(d, TUPS_EINVAL, "null argument");
  TUPS_CHECK(d->nx >= 2 && d->nx % 2 == 0, TUPS_EINVAL, "Nx must be even and >= 2 (list_symmetries)");
  TUPS_CHECK(d->nf >= 2 && d->nf % 2 == 0, TUPS_EINVAL, "Nf must be even and >= 2");
  TUPS_CHECK(nl_supported(d->nl), TUPS_EINVAL, "Nl must be 1, 5 or 9 (sym_op_L tables)");
  const int nx_ibz = (d->nx + 2) / 2;
  const int nr = d->nr > 0 ? d->nr : nr_default(d->nl);
  TUPS_CHECK(nr <= nx_ibz * (nx_ibz + 1) / 2, TUPS_EINVAL, "NR exceeds the IBZ size");
  Tables t;
  int rc = build_tables(d->nx, d->nf, d->nl, nr, t);
  if (rc) return rc;
  if (ff) std::memcpy(ff, t.ff.data(), t.ff.size() * sizeof(double));
  if (ff_inv) std::memcpy(ff_inv, t.ff_inv.data(), t.ff_inv.size() * sizeof(double));
  if (signs) std::copy(t.signs.begin(), t.signs.end(), signs);
  if (reps) std::copy(t.reps.begin(), t.reps.end(), reps);
  if (exp_sqr) std::memcpy(exp_sqr, t.exp_sqr.data(), t.exp_sqr.size() * sizeof(double));
  if (exp_qsr) std::memcpy(exp_qsr, t.exp_qsr.data(), t.exp_qsr.size() * sizeof(double));
  return TUPS_OK;
}

int tups_create(const tups_dims* d, tups_ctx** out) {
  TUPS_CHECK(d && out, TUPS_EINVAL, "null argument");
  *out = nullptr;
  TUPS_CHECK(d->nx >= 2 && d->nx % 2 == 0, TUPS_EINVAL, "Nx must be even and >= 2 (list_symmetries)");
  TUPS_CHECK(d->nf >= 2 && d->nf % 2 == 0, TUPS_EINVAL, "Nf must be even and >= 2");
  TUPS_CHECK(nl_supported(d->nl), TUPS_EINVAL, "Nl must be 1, 5 or 9 (sym_op_L tables)");
  TUPS_CHECK(d->beta > 0.0, TUPS_EINVAL, "beta must be positive");
  const int nx_ibz = (d->nx + 2) / 2;
  const int nibz = nx_ibz * (nx_ibz + 1) / 2;
  const int nr = d->nr > 0 ? d->nr : nr_default(d->nl);
  TUPS_CHECK(nr <= nibz && nr <= kMaxNR, TUPS_EINVAL, "NR exceeds the IBZ size or the kernel limit (16)");
  const int nw = d->nf / 2;
  const int nq_full = nibz * nw;
  const int nq = d->nq > 0 ? d->nq : nq_full;

  tups_ctx* c = new tups_ctx();
  c->beta = d->beta;
  c->f_damping = d->f_damping;
  tups_sizes& s = c->sz;
  s.nx = d->nx;
  s.ny = d->nx;
  s.nf = d->nf;
  s.nl = d->nl;
  s.nz = d->nl * d->nf;
  s.nr = nr;
  s.nw = nw;
  s.nibz = nibz;
  s.nq = nq;
  s.nc = d->nx * d->nx;
  s.full_ibz = nq == nq_full ? 1 : 0;
  s.alpha = 1.0 / (d->beta * d->nx * d->nx);

  int rc = build_tables(s.nx, s.nf, s.nl, s.nr, c->tab);
  if (rc) {
    delete c;
    return rc;
  }
  if (nq != nq_full) {
    // synthetic batch: slot q cycles through the IBZ map until set_q_map
    std::vector<int> qx(nq), qy(nq), qw(nq);
    for (int q = 0; q < nq; ++q) {
      qx[q] = c->tab.qx[q % nq_full];
      qy[q] = c->tab.qy[q % nq_full];
      qw[q] = c->tab.qw[q % nq_full];
    }
    c->tab.qx = qx;
    c->tab.qy = qy;
    c->tab.qw = qw;
  }

  hipError_t he = hipGetDevice(&c->device);
  if (he == hipSuccess && (rc = conv_noalloc_setup(c->device)) != TUPS_OK) {
    tups_destroy(c);
    return rc;
  }
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  for (int i = 0; i < kEvents && he == hipSuccess; ++i) he = hipEventCreate(&c->ev[i]);
  for (int i = 0; i < 2 && he == hipSuccess; ++i) he = hipEventCreate(&c->tev[i]);
  if (he != hipSuccess) {
    set_error(std::string("HIP initialisation failed: ") + hipGetErrorString(he));
    tups_destroy(c);
    return TUPS_EHIP;
  }
  const size_t nv = array_elems(c, TUPS_GAMMA);
  for (int r = 0; r < 4 && rc == 0; ++r) {
    rc = dalloc(&c->gamma[r], nv);
    if (!rc) rc = dalloc(&c->fvert[r], nv);
    if (!rc) rc = dalloc(&c->spare[r], nv);
    if (!rc) rc = dalloc(&c->lam[r], array_elems(c, TUPS_LAMBDA));
    if (!rc && s.full_ibz) rc = dalloc(&c->phir[r], array_elems(c, TUPS_PHIR));
    if (!rc && s.full_ibz) rc = dalloc(&c->mten[r], array_elems(c, TUPS_M));
  }
  for (int p = 0; p < 2 && rc == 0; ++p) rc = dalloc(&c->chi[p], array_elems(c, TUPS_CHI));
  if (!rc && s.full_ibz && s.nl == 1) rc = dalloc(&c->pae_y, (size_t)8 * s.nf * s.nf * s.nw * s.nr);
  if (!rc && s.full_ibz && s.nl > 1) {
    // columns in 8 x 8 blocks of (nu1, nu2), padded to whole blocks
    const size_t nblk = (size_t)(s.nf + 7) / 8;
    rc = dalloc(&c->pae_ym, (size_t)8 * 64 * nblk * nblk * s.nw * s.nl * s.nl * s.nr);
  }
  if (!rc) {
    he = hipMalloc((void**)&c->d_flag, sizeof(int));
    if (he != hipSuccess) rc = TUPS_ENOMEM;
  }
  if (!rc) {
    for (int r = 0; r < 4 && !rc; ++r) {
      if (hipMemsetAsync(c->gamma[r], 0, nv * sizeof(double2), c->stream) != hipSuccess ||
          hipMemsetAsync(c->fvert[r], 0, nv * sizeof(double2), c->stream) != hipSuccess ||
          hipMemsetAsync(c->lam[r], 0, array_elems(c, TUPS_LAMBDA) * sizeof(double2), c->stream) != hipSuccess)
        rc = TUPS_EHIP;
    }
    for (int p = 0; p < 2 && !rc; ++p)
      if (hipMemsetAsync(c->chi[p], 0, array_elems(c, TUPS_CHI) * sizeof(double2), c->stream) != hipSuccess)
        rc = TUPS_EHIP;
  }
  if (!rc) rc = upload_tables(c);
  if (rc) {
    tups_destroy(c);
    return rc;
  }
  *out = c;
  return TUPS_OK;
}

int tups_destroy(tups_ctx* c) {
  if (!c) return TUPS_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graphs(c);
  auto f = [](void* p) {
    if (p) hipFree(p);
  };
  for (int r = 0; r < 4; ++r) {
    f(c->gamma[r]);
    f(c->fvert[r]);
    f(c->spare[r]);
    f(c->phir[r]);
    f(c->mten[r]);
    f(c->lam[r]);
  }
  f(c->chi[0]);
  f(c->chi[1]);
  f(c->split_ws);
  f(c->sde_ws);
  f(c->gt_ws);
  f(c->conv_ws);
  f(c->sus_ws);
  f(c->c0_ws);
  f(c->d_sigma_old);
  f(c->d_gstate);
  f(c->closed_ws);
  f(c->d_sigma);
  f(c->d_sigh);
  f(c->d_ffk);
  f(c->d_ffik);
  f(c->d_eqsr);
  f(c->d_qm);
  f(c->d_qw);
  f(c->d_ibz);
  f(c->d_signs);
  f(c->d_ph_blk);
  f(c->d_ph_coef);
  f(c->d_flag);
  f(c->pae_y);
  f(c->pae_ym);
  f(c->pae_rows);
  for (int i = 0; i < kEvents; ++i)
    if (c->ev[i]) hipEventDestroy(c->ev[i]);
  for (int i = 0; i < 2; ++i)
    if (c->tev[i]) hipEventDestroy(c->tev[i]);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return TUPS_OK;
}

int tups_query(const tups_ctx* c, tups_sizes* out) {
  TUPS_CHECK(c && out, TUPS_EINVAL, "null argument");
  *out = c->sz;
  return TUPS_OK;
}

int tups_get_tables(const tups_ctx* c, double* ff, double* ff_inv, int32_t* signs, int32_t* reps,
                    double* exp_sqr, double* exp_qsr) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  const Tables& t = c->tab;
  if (ff) std::memcpy(ff, t.ff.data(), t.ff.size() * sizeof(double));
  if (ff_inv) std::memcpy(ff_inv, t.ff_inv.data(), t.ff_inv.size() * sizeof(double));
  if (signs) std::copy(t.signs.begin(), t.signs.end(), signs);
  if (reps) std::copy(t.reps.begin(), t.reps.end(), reps);
  if (exp_sqr) std::memcpy(exp_sqr, t.exp_sqr.data(), t.exp_sqr.size() * sizeof(double));
  if (exp_qsr) std::memcpy(exp_qsr, t.exp_qsr.data(), t.exp_qsr.size() * sizeof(double));
  return TUPS_OK;
}

int tups_set_tables(tups_ctx* c, const double* ff, const double* ff_inv, const int32_t* signs,
                    const int32_t* reps, const double* exp_sqr, const double* exp_qsr) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  Tables& t = c->tab;
  if (reps) {  // validate before any table changes, so a rejected call leaves the context as it was
    for (int l = 0; l < t.nl; ++l) TUPS_CHECK(reps[l] >= 1 && reps[l] <= 6, TUPS_EINVAL, "reps out of range");
    Tables trial = t;
    std::copy(reps, reps + t.nl, trial.reps.begin());
    if (int rc = rebuild_sym_l(trial)) return rc;
  }
  if (ff) std::memcpy(t.ff.data(), ff, t.ff.size() * sizeof(double));
  if (ff_inv) std::memcpy(t.ff_inv.data(), ff_inv, t.ff_inv.size() * sizeof(double));
  if (signs) std::copy(signs, signs + t.nl, t.signs.begin());
  if (reps) {
    std::copy(reps, reps + t.nl, t.reps.begin());
    (void)rebuild_sym_l(t);  // validated above
  }
  if (exp_sqr) std::memcpy(t.exp_sqr.data(), exp_sqr, t.exp_sqr.size() * sizeof(double));
  if (exp_qsr) std::memcpy(t.exp_qsr.data(), exp_qsr, t.exp_qsr.size() * sizeof(double));
  c->tables_dirty = true;
  c->m_built = false;
  c->phir_valid = false;  // PhiR was formed with the old symmetry/Fourier tables
  return ensure_tables(c);
}

int tups_get_q_map(const tups_ctx* c, int32_t* qx, int32_t* qy, int32_t* qw) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  if (qx) std::copy(c->tab.qx.begin(), c->tab.qx.end(), qx);
  if (qy) std::copy(c->tab.qy.begin(), c->tab.qy.end(), qy);
  if (qw) std::copy(c->tab.qw.begin(), c->tab.qw.end(), qw);
  return TUPS_OK;
}

int tups_set_q_map(tups_ctx* c, const int32_t* qx, const int32_t* qy, const int32_t* qw) {
  TUPS_CHECK(c && qx && qy && qw, TUPS_EINVAL, "null argument");
  TUPS_CHECK(!c->sz.full_ibz, TUPS_EINVAL, "q map of a full-IBZ context is fixed by Index_bosonic_IBZ");
  const int nxi = (c->sz.nx + 2) / 2;
  for (int q = 0; q < c->sz.nq; ++q) {
    TUPS_CHECK(qx[q] >= 1 && qx[q] <= nxi && qy[q] >= 1 && qy[q] <= qx[q] && qw[q] >= 1 &&
                   qw[q] <= c->sz.nw,
               TUPS_EINVAL, "q map entry outside the IBZ x bosonic box");
  }
  c->tab.qx.assign(qx, qx + c->sz.nq);
  c->tab.qy.assign(qy, qy + c->sz.nq);
  c->tab.qw.assign(qw, qw + c->sz.nq);
  c->tables_dirty = true;
  return ensure_tables(c);
}

int tups_upload(tups_ctx* c, int32_t which, int32_t ch, const double* host) {
  TUPS_CHECK(c && host, TUPS_EINVAL, "null argument");
  double2* p = array_ptr(c, which, ch);
  TUPS_CHECK(p, TUPS_EINVAL, "array/channel not resident in this context");
  TUPS_HIP(hipMemcpyAsync(p, host, array_elems(c, which) * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  if (which == TUPS_M) c->m_built = true;
  if (which == TUPS_PHIR) c->phir_valid = true;
  return TUPS_OK;
}

int tups_download(tups_ctx* c, int32_t which, int32_t ch, double* host) {
  TUPS_CHECK(c && host, TUPS_EINVAL, "null argument");
  double2* p = array_ptr(c, which, ch);
  TUPS_CHECK(p, TUPS_EINVAL, "array/channel not resident in this context");
  TUPS_HIP(hipMemcpyAsync(host, p, array_elems(c, which) * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  return TUPS_OK;
}

int tups_upload_chi_dense(tups_ctx* c, int32_t pp, int32_t q0, int32_t nqs, const double* dense) {
  TUPS_CHECK(c && dense && (pp == 0 || pp == 1), TUPS_EINVAL, "bad argument");
  const tups_sizes& s = c->sz;
  TUPS_CHECK(q0 >= 0 && nqs >= 0 && q0 + nqs <= s.nq, TUPS_EINVAL, "q range outside the context");
  const size_t blk = (size_t)s.nl * s.nl * s.nf;
  std::vector<double> packed(2 * blk * nqs);
  for (int q = 0; q < nqs; ++q)
    for (int nu = 0; nu < s.nf; ++nu)
      for (int l2 = 0; l2 < s.nl; ++l2)
        for (int l1 = 0; l1 < s.nl; ++l1) {
          const size_t src = (size_t)(l1 * s.nf + nu) + (size_t)s.nz * ((l2 * s.nf + nu) + (size_t)s.nz * q);
          const size_t dst = l1 + (size_t)s.nl * (l2 + (size_t)s.nl * (nu + (size_t)s.nf * q));
          packed[2 * dst] = dense[2 * src];
          packed[2 * dst + 1] = dense[2 * src + 1];
        }
  TUPS_HIP(hipMemcpyAsync(c->chi[pp] + blk * q0, packed.data(), packed.size() * sizeof(double),
                          hipMemcpyHostToDevice, c->stream));
  TUPS_HIP(hipStreamSynchronize(c->stream));
  return TUPS_OK;
}

int tups_bse(tups_ctx* c, int32_t mode) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_CHECK(mode == TUPS_BSE_LADDER_FP64 || mode == TUPS_BSE_LADDER_SPLIT || mode == TUPS_BSE_CLOSED_LU ||
                 mode == TUPS_BSE_CLOSED_IR,
             TUPS_EINVAL, "unknown BSE mode");
  int rc = ensure_tables(c);
  if (rc) return rc;
  rc = mode == TUPS_BSE_LADDER_SPLIT ? launch_bse_split(c)
       : mode == TUPS_BSE_CLOSED_LU  ? launch_bse_closed(c)
       : mode == TUPS_BSE_CLOSED_IR  ? launch_bse_closed_ir(c)
                                     : launch_bse_ladder(c);
  if (rc) return rc;
  // Phi overwrites G_r_LL (SRC/parquet_BSE.f90:81): swap the ping-pong pair
  swap_vertex_buffers(c);
  c->phir_valid = false;
  return TUPS_OK;
}

int tups_set_product_forms(tups_ctx* c, int32_t allow_three) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  c->three_mult = allow_three != 0;
  return TUPS_OK;
}

int tups_set_refinement(tups_ctx* c, double tol, int32_t max_steps) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_CHECK(tol > 0.0 && max_steps >= 0 && max_steps <= 100, TUPS_EINVAL,
             "refinement tolerance must be positive and the step limit in 0..100");
  c->ir_tol = tol;
  c->ir_max_steps = max_steps;
  return TUPS_OK;
}

int tups_bse_solve_info(const tups_ctx* c, int32_t* steps, double* residual) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  if (steps) *steps = c->ir_steps;
  if (residual) *residual = c->ir_residual;
  return TUPS_OK;
}

int tups_calc_chi(tups_ctx* c, const double* sigma, const double* sigma_h, double mu, int32_t ngrain) {
  TUPS_CHECK(c && sigma, TUPS_EINVAL, "null argument");
  int rc = ensure_tables(c);
  if (rc) return rc;
  return launch_calc_chi(c, sigma, sigma_h, mu, ngrain);
}

int tups_sde(tups_ctx* c, const double* sigma, const double* sigma_h, double mu, int32_t ngrain, int32_t f_range,
             double u, int32_t use_u2, double* sigma_out) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_CHECK(full_ibz(c), TUPS_EINVAL, "the SDE sums over every IBZ q of the box (nq = 0 at create)");
  int rc = ensure_tables(c);
  if (rc) return rc;
  return launch_sde(c, sigma, sigma_h, mu, ngrain, f_range, u, use_u2, sigma_out);
}

int tups_set_split(tups_ctx* c, int32_t moduli, int32_t qbatch) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_CHECK(moduli >= 8 && moduli <= 16, TUPS_EINVAL, "split moduli must be in 8..16");
  TUPS_CHECK(qbatch >= 0, TUPS_EINVAL, "qbatch must be >= 0");
  drop_graphs(c);
  c->split_moduli = moduli;
  c->split_qbatch = qbatch;
  c->split_auto_q = 0;
  return TUPS_OK;
}

int tups_phir(tups_ctx* c) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_CHECK(full_ibz(c), TUPS_EINVAL, "calc_PhiR needs every IBZ q of the box (nq = 0 at create)");
  int rc = ensure_tables(c);
  if (rc) return rc;
  rc = launch_phir(c);
  if (!rc) c->phir_valid = true;
  return rc;
}

int tups_build_m(tups_ctx* c) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_CHECK(full_ibz(c), TUPS_EINVAL, "M tensors need a full-IBZ context");
  int rc = ensure_tables(c);
  if (rc) return rc;
  rc = launch_build_m(c);
  if (!rc) c->m_built = true;
  return rc;
}

int tups_pae(tups_ctx* c) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_CHECK(full_ibz(c), TUPS_EINVAL, "calculate_PAE needs a full-IBZ context");
  TUPS_CHECK(c->m_built, TUPS_ESTATE, "calculate_PAE before the M tensors were built");
  TUPS_CHECK(c->phir_valid, TUPS_ESTATE, "calculate_PAE before calc_PhiR");
  int rc = ensure_tables(c);
  if (rc) return rc;
  return launch_pae(c);
}

int tups_gamma_update(tups_ctx* c) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  return launch_gamma_update(c);
}

int tups_autotune(tups_ctx* c, int32_t kernel, int32_t* variant, float* ms) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_CHECK(kernel == TUPS_K_BSE_LADDER || kernel == TUPS_K_PAE || kernel == TUPS_K_SPLIT_GEMM, TUPS_EINVAL,
             "only the BSE ladder, its split GEMM and the crossing have tile variants");
  int rc = ensure_tables(c);
  if (rc) return rc;
  int v = -1;
  float t = 0.f;
  drop_graphs(c);  // a new variant changes the captured launches
  if (kernel == TUPS_K_PAE) {
    TUPS_CHECK(c->m_built && c->phir_valid, TUPS_ESTATE, "crossing autotune needs M and PhiR");
    rc = autotune_pae(c, 3, &v, &t);
  } else if (kernel == TUPS_K_SPLIT_GEMM) {
    rc = autotune_split(c, &v, &t);
  } else {
    rc = autotune_bse(c, 2, &v, &t);
  }
  if (rc) return rc;
  if (variant) *variant = v;
  if (ms) *ms = t;
  return TUPS_OK;
}

const char* tups_variant_name(int32_t kernel, int32_t variant) {
  if (kernel == TUPS_K_BSE_BATCH) return batch_variant_name(variant);
  if (kernel == TUPS_K_SPLIT_GEMM) return split_variant_name(variant);
  if (kernel == TUPS_K_BSE_CLOSED || kernel == TUPS_K_BSE_IR) return closed_variant_name(variant);
  return kernel == TUPS_K_BSE_LADDER ? bse_variant_name(variant)
                                     : (kernel == TUPS_K_PAE ? pae_variant_name(variant) : "");
}

int tups_set_variant(tups_ctx* c, int32_t kernel, int32_t variant) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  drop_graphs(c);
  if (kernel == TUPS_K_PAE) {
    TUPS_CHECK(variant == -1 || pae_variant_ok(c, variant), TUPS_EINVAL, "crossing variant not usable here");
    c->pae_variant = variant;
    return TUPS_OK;
  }
  if (kernel == TUPS_K_SPLIT_GEMM) {
    TUPS_CHECK(variant >= -1 && variant < split_variant_count(), TUPS_EINVAL, "no such split GEMM variant");
    c->split_variant = variant;
    return TUPS_OK;
  }
  if (kernel == TUPS_K_BSE_CLOSED || kernel == TUPS_K_BSE_IR) {
    TUPS_CHECK(variant >= -1 && variant < closed_variant_count(), TUPS_EINVAL, "no such elimination variant");
    c->closed_variant = variant;
    return TUPS_OK;
  }
  TUPS_CHECK(kernel == TUPS_K_BSE_LADDER, TUPS_EINVAL, "kernel has no tile variants");
  TUPS_CHECK(variant == -1 || bse_variant_nl(variant) == c->sz.nl, TUPS_EINVAL,
             "variant does not exist for this Nl");
  c->bse_variant = variant;
  return TUPS_OK;
}

int tups_variant_count(int32_t kernel) {
  if (kernel == TUPS_K_BSE_BATCH) return batch_variant_count();
  if (kernel == TUPS_K_SPLIT_GEMM) return split_variant_count();
  if (kernel == TUPS_K_BSE_CLOSED || kernel == TUPS_K_BSE_IR) return closed_variant_count();
  return kernel == TUPS_K_BSE_LADDER ? bse_variant_count() : (kernel == TUPS_K_PAE ? pae_variant_count() : 0);
}

int tups_sync(tups_ctx* c) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_HIP(hipStreamSynchronize(c->stream));
  return TUPS_OK;
}

int tups_event_record(tups_ctx* c, int32_t slot) {
  TUPS_CHECK(c && slot >= 0 && slot < kEvents, TUPS_EINVAL, "bad event slot");
  TUPS_HIP(hipEventRecord(c->ev[slot], c->stream));
  return TUPS_OK;
}

int tups_event_elapsed(tups_ctx* c, int32_t a, int32_t b, float* ms) {
  TUPS_CHECK(c && ms && a >= 0 && a < kEvents && b >= 0 && b < kEvents, TUPS_EINVAL, "bad event slot");
  TUPS_HIP(hipEventSynchronize(c->ev[b]));
  TUPS_HIP(hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
  return TUPS_OK;
}

int tups_set_timing(tups_ctx* c, int32_t on) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  TUPS_HIP(hipStreamSynchronize(c->stream));
  c->timing = on != 0;
  for (int k = 0; k < TUPS_K_COUNT; ++k) {
    c->t_ms[k] = 0.0;
    c->t_n[k] = 0;
  }
  return TUPS_OK;
}

int tups_get_timing(const tups_ctx* c, int32_t kernel, double* total_ms, int64_t* launches) {
  TUPS_CHECK(c && kernel >= 0 && kernel < TUPS_K_COUNT, TUPS_EINVAL, "bad kernel id");
  if (total_ms) *total_ms = c->t_ms[kernel];
  if (launches) *launches = c->t_n[kernel];
  return TUPS_OK;
}

int tups_check_finite(tups_ctx* c, int32_t which, int32_t ch) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  double2* p = array_ptr(c, which, ch);
  TUPS_CHECK(p, TUPS_EINVAL, "array/channel not resident in this context");
  return launch_check_finite(c, p, array_elems(c, which));
}

int tups_synth_vertex(tups_ctx* c, int32_t which, int32_t ch, uint64_t seed, double diag, double sigma) {
  TUPS_CHECK(c && (which == TUPS_GAMMA || which == TUPS_F) && ch >= 0 && ch < 4, TUPS_EINVAL, "bad argument");
  double2* p = array_ptr(c, which, ch);
  return launch_synth(c, p, array_elems(c, which), c->sz.nz, seed, (uint64_t)(which * 4 + ch), diag, sigma);
}

int tups_copy(tups_ctx* c, int32_t wd, int32_t ws, int32_t ch) {
  TUPS_CHECK(c, TUPS_EINVAL, "null context");
  double2* d = array_ptr(c, wd, ch);
  double2* s = array_ptr(c, ws, ch);
  TUPS_CHECK(d && s && array_elems(c, wd) == array_elems(c, ws), TUPS_EINVAL, "incompatible arrays");
  TUPS_HIP(hipMemcpyAsync(d, s, array_elems(c, wd) * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
  return TUPS_OK;
}

}  // extern "C"
