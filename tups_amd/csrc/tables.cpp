// One-time host tables of the truncated-unity basis.
//
// These are set-up quantities the reference also computes on the host before
// the first iteration; the GPU path consumes device copies of them.
//   IBZ momentum list / Index_bosonic_IBZ   SRC/parquet_util.f90:110-117
//   symmetry_operation                      SRC/parquet_util.f90:896-978
//   list_symmetries                         SRC/parquet_util.f90:1140-1207
//   create_list + sort                      SRC/parquet_formfactors.f90:594-646
//   initialize_bonds                        SRC/parquet_formfactors.f90:95-360
//   FF_array                                SRC/parquet_formfactors.f90:11-90
//   init_FF_inv                             SRC/parquet_formfactors.f90:545-569
//   char_table                              SRC/parquet_formfactors.f90:572-588
//   sym_op_L                                SRC/parquet_util.f90:732-890
//   calculate_expSqR / calculate_expqSR     SRC/parquet_PhiR.f90:360-435
#include <algorithm>
#include <cmath>
#include <complex>

#include "internal.h"

namespace tups {

namespace {

// char_table(s, rep_column), rows = operations 1..8, columns = A1, A2, B1, B2, E
const int kChar[8][5] = {
    {1, 1, 1, 1, 2},  {1, -1, 1, -1, 0}, {1, -1, 1, -1, 0}, {1, 1, 1, 1, -2},
    {1, -1, -1, 1, 0}, {1, 1, -1, -1, 0}, {1, -1, -1, 1, 0}, {1, 1, -1, -1, 0},
};

// operation s (0-based) acting on an integer vector
void op_vec(int s, int x, int y, int& ox, int& oy) {
  switch (s) {
    case 0: ox = x;  oy = y;  break;   // identity
    case 1: ox = x;  oy = -y; break;   // (kx,ky) -> (kx,-ky)
    case 2: ox = -x; oy = y;  break;   // (kx,ky) -> (-kx,ky)
    case 3: ox = -x; oy = -y; break;   // C2
    case 4: ox = y;  oy = x;  break;   // (kx,ky) -> (ky,kx)
    case 5: ox = y;  oy = -x; break;   // C4, inverse of op 7 (0-based)
    case 6: ox = -y; oy = -x; break;   // (kx,ky) -> (-ky,-kx)
    default: ox = -y; oy = x; break;   // C4, inverse of op 5 (0-based)
  }
}

const int kInv[8] = {0, 1, 2, 3, 4, 7, 6, 5};

int wrap(int v, int n) { return ((v % n) + n) % n; }

void star_ops(int kx, int ky, int nx_ibz, uint8_t* out) {
  for (int s = 0; s < kNs; ++s) out[s] = 0;
  if ((kx == 1 && ky == 1) || (kx == nx_ibz && ky == nx_ibz)) {
    out[0] = 1;
  } else if (kx == nx_ibz && ky == 1) {
    out[0] = out[4] = 1;
  } else if (kx == nx_ibz && ky > 1 && ky < nx_ibz) {
    out[0] = out[1] = out[4] = out[7] = 1;
  } else if (kx > 1 && ky < nx_ibz) {
    if (ky == 1) {
      out[0] = out[2] = out[4] = out[5] = 1;
    } else if (ky == kx) {
      out[0] = out[1] = out[2] = out[3] = 1;
    } else {
      for (int s = 0; s < kNs; ++s) out[s] = 1;
    }
  }
}

struct Bond {
  int x, y, col, rep, sign;
};

// First bond of each form factor in the order the reference generates them
// (lattice sites by distance, then the irreps each site supports).
bool make_bonds(int nx, int nl, std::vector<Bond>& bonds) {
  struct Site { int x, y, d; };
  std::vector<Site> sites;
  for (int i = 0; i <= nx / 2; ++i)
    for (int j = 0; j <= i; ++j) sites.push_back({i, j, i * i + j * j});
  std::stable_sort(sites.begin(), sites.end(), [](const Site& a, const Site& b) { return a.d < b.d; });
  const double half = nx / 2.0;
  bonds.clear();
  for (const Site& st : sites) {
    const int x = st.x, y = st.y;
    std::vector<Bond> c;
    const bool at_edge = std::fabs(x - half) < 0.01;
    if (x == 0 && y == 0) {
      c = {{x, y, 1, 1, 1}};
    } else if (at_edge && std::fabs(y - half) < 0.01) {
      c = {{x, y, 1, 1, 1}};
    } else if (at_edge && y == 0) {
      c = {{x, y, 1, 1, 1}, {x, y, 3, 3, 1}};
    } else if (x == y) {
      c = {{x, y, 1, 1, 1}, {x, y, 4, 4, 1}, {x, y, 5, 5, -1}, {-y, x, 5, 6, -1}};
    } else if (y == 0 || at_edge) {
      c = {{x, y, 1, 1, 1}, {x, y, 3, 3, 1}, {x, y, 5, 5, -1}, {-y, x, 5, 6, -1}};
    } else {
      c = {{x, y, 1, 1, 1},  {x, y, 2, 2, 1},   {x, y, 3, 3, 1}, {x, y, 4, 4, 1},
           {x, y, 5, 5, -1}, {x, -y, 5, 6, -1}, {y, x, 5, 5, -1}, {-y, x, 5, 6, -1}};
    }
    for (const Bond& b : c) {
      bonds.push_back(b);
      if ((int)bonds.size() == nl) return true;
    }
  }
  return false;
}

// (l_sym, fac) of l under operation s; both 0-based.  Two-dimensional irreps
// map l to its partner for the diagonal-type operations.
void sym_l(int s, int l, const std::vector<int>& reps, int& lsym, int& fac) {
  static const int l5_low[8] = {1, 1, -1, -1, 1, 1, -1, -1};
  static const int l5_hi_f[8] = {1, -1, 1, -1, 1, 1, -1, -1};
  static const int l5_hi_d[8] = {0, 1, 1, 0, 0, 1, 0, 1};
  static const int l6_low[8] = {1, -1, 1, -1, 1, -1, -1, 1};
  static const int l6_hi_f[8] = {1, -1, 1, -1, -1, -1, 1, 1};
  static const int l6_hi_d[8] = {0, -1, -1, 0, 0, -1, 0, -1};
  const int rep = reps[l];
  const bool low = (l + 1) <= 5;
  if (rep >= 1 && rep <= 4) {
    lsym = l;
    fac = kChar[s][rep - 1];
  } else if (rep == 5) {
    if (low) {
      lsym = s < 4 ? l : l + 1;
      fac = l5_low[s];
    } else {
      lsym = l + l5_hi_d[s];
      fac = l5_hi_f[s];
    }
  } else {
    if (low) {
      lsym = s < 4 ? l : l - 1;
      fac = l6_low[s];
    } else {
      lsym = l + l6_hi_d[s];
      fac = l6_hi_f[s];
    }
  }
}

}  // namespace

int nl_supported(int nl) { return nl == 1 || nl == 5 || nl == 9; }

int nr_default(int nl) { return nl == 1 ? 1 : (nl == 5 ? 4 : 6); }

int rebuild_sym_l(Tables& t) {
  t.lsym.assign(kNs * t.nl, 0);
  t.lfac.assign(kNs * t.nl, 0);
  for (int s = 0; s < kNs; ++s)
    for (int l = 0; l < t.nl; ++l) {
      sym_l(s, l, t.reps, t.lsym[s * t.nl + l], t.lfac[s * t.nl + l]);
      // a two-dimensional irrep must have its partner inside the basis
      if (t.lsym[s * t.nl + l] < 0 || t.lsym[s * t.nl + l] >= t.nl) {
        set_error("reps: two-dimensional irrep without a partner form factor");
        return TUPS_EINVAL;
      }
    }
  return TUPS_OK;
}

void build_q_map_full(Tables& t) {
  const int nq = t.nibz * t.nw;
  t.qx.assign(nq, 0);
  t.qy.assign(nq, 0);
  t.qw.assign(nq, 0);
  for (int m = 0; m < t.nibz; ++m)
    for (int w = 0; w < t.nw; ++w) {
      const int q = m * t.nw + w;
      t.qx[q] = t.ibz_x[m];
      t.qy[q] = t.ibz_y[m];
      t.qw[q] = w + 1;
    }
}

int build_tables(int nx, int nf, int nl, int nr, Tables& t) {
  const double pi = std::acos(-1.0);
  t.nx = nx;
  t.nf = nf;
  t.nl = nl;
  t.nr = nr;
  t.nw = nf / 2;
  t.nx_ibz = (nx + 2) / 2;
  t.nibz = t.nx_ibz * (t.nx_ibz + 1) / 2;
  t.nc = nx * nx;
  const int ny = nx, nc = t.nc;

  // IBZ points in the order of Index_bosonic_IBZ
  t.ibz_x.clear();
  t.ibz_y.clear();
  for (int kx = 1; kx <= t.nx_ibz; ++kx)
    for (int ky = 1; ky <= kx; ++ky) {
      t.ibz_x.push_back(kx);
      t.ibz_y.push_back(ky);
    }
  t.sym_list.assign(t.nibz * kNs, 0);
  for (int m = 0; m < t.nibz; ++m) star_ops(t.ibz_x[m], t.ibz_y[m], t.nx_ibz, &t.sym_list[m * kNs]);

  // form factors
  std::vector<Bond> bonds;
  if (!make_bonds(nx, nl, bonds)) {
    set_error("lattice too small for the requested number of form factors");
    return TUPS_EINVAL;
  }
  t.ff.assign((size_t)nx * ny * nl, 0.0);
  t.signs.assign(nl, 1);
  t.reps.assign(nl, 1);
  for (int l = 0; l < nl; ++l) {
    const Bond& b = bonds[l];
    t.signs[l] = b.sign;
    t.reps[l] = b.rep;
    std::vector<std::complex<double>> comp((size_t)nx * ny);
    double norm = 0.0;
    for (int kx = 0; kx < nx; ++kx)
      for (int ky = 0; ky < ny; ++ky) {
        const double px = 2.0 / nx * pi * kx, py = 2.0 / ny * pi * ky;
        std::complex<double> acc = 0.0;
        for (int s = 0; s < kNs; ++s) {
          int sx, sy;
          op_vec(s, b.x, b.y, sx, sy);
          const double ph = sx * px + sy * py;
          acc += (double)kChar[s][b.col - 1] * std::complex<double>(std::cos(ph), std::sin(ph));
        }
        comp[kx + (size_t)nx * ky] = acc;
        norm += std::abs(acc) * std::abs(acc);
      }
    const double inv = 1.0 / std::sqrt(norm);
    for (size_t i = 0; i < comp.size(); ++i)
      t.ff[i + (size_t)nx * ny * l] = inv * comp[i].real() + inv * comp[i].imag();
  }

  // FF_inv(s,kx,ky,l) = FF(S_s^-1 k, l)
  t.ff_inv.assign((size_t)kNs * nx * ny * nl, 0.0);
  for (int l = 0; l < nl; ++l)
    for (int ky = 0; ky < ny; ++ky)
      for (int kx = 0; kx < nx; ++kx)
        for (int s = 0; s < kNs; ++s) {
          int sx, sy;
          op_vec(kInv[s], kx, ky, sx, sy);
          sx = wrap(sx, nx);
          sy = wrap(sy, ny);
          t.ff_inv[s + kNs * (kx + (size_t)nx * (ky + (size_t)ny * l))] = t.ff[sx + (size_t)nx * (sy + (size_t)ny * l)];
        }
  if (int rc = rebuild_sym_l(t)) return rc;

  // expSqR(Q, R, s) = e^{i (S Q).R} on the star of Q
  t.exp_sqr.assign((size_t)2 * t.nibz * nr * kNs, 0.0);
  for (int m = 0; m < t.nibz; ++m)
    for (int s = 0; s < kNs; ++s) {
      if (!t.sym_list[m * kNs + s]) continue;
      int sx, sy;
      op_vec(s, t.ibz_x[m] - 1, t.ibz_y[m] - 1, sx, sy);
      sx = wrap(sx, nx);
      sy = wrap(sy, ny);
      for (int r = 0; r < nr; ++r) {
        const double ph = (2.0 * pi / nx) * sx * (t.ibz_x[r] - 1) + (2.0 * pi / ny) * sy * (t.ibz_y[r] - 1);
        const size_t idx = m + (size_t)t.nibz * (r + (size_t)nr * s);
        t.exp_sqr[2 * idx] = std::cos(ph);
        t.exp_sqr[2 * idx + 1] = std::sin(ph);
      }
    }

  // expqSR(k, R, s) = e^{-i k.(S R)} on the star of R, k = (ky-1)*Nx + kx
  t.exp_qsr.assign((size_t)2 * nc * nr * kNs, 0.0);
  for (int r = 0; r < nr; ++r)
    for (int s = 0; s < kNs; ++s) {
      if (!t.sym_list[r * kNs + s]) continue;
      int sx, sy;
      op_vec(s, t.ibz_x[r] - 1, t.ibz_y[r] - 1, sx, sy);
      sx = wrap(sx, nx);
      sy = wrap(sy, ny);
      for (int ky = 0; ky < ny; ++ky)
        for (int kx = 0; kx < nx; ++kx) {
          const double ph = (2.0 * pi / nx) * kx * sx + (2.0 * pi / ny) * ky * sy;
          const size_t idx = (ky * nx + kx) + (size_t)nc * (r + (size_t)nr * s);
          t.exp_qsr[2 * idx] = std::cos(ph);
          t.exp_qsr[2 * idx + 1] = -std::sin(ph);
        }
    }
  build_q_map_full(t);
  return TUPS_OK;
}

void build_phir_table(const Tables& t, PhirTable& out) {
  const int nl = t.nl, nm = t.nibz, nr = t.nr;
  const size_t ne = (size_t)nl * nl * nm;
  out.cnt.assign(ne, 0);
  out.src.assign(ne * 4, 0);
  out.coef.assign(ne * 4 * nr * 2, 0.0);
  for (int l1 = 0; l1 < nl; ++l1)
    for (int l2 = 0; l2 < nl; ++l2)
      for (int m = 0; m < nm; ++m) {
        const size_t e = ((size_t)l1 * nl + l2) * nm + m;
        for (int s = 0; s < kNs; ++s) {
          if (!t.sym_list[m * kNs + s]) continue;
          const int l1s = t.lsym[s * nl + l1], l2s = t.lsym[s * nl + l2];
          const double f = (double)(t.lfac[s * nl + l1] * t.lfac[s * nl + l2]);
          const int key = l1s + nl * l2s;
          int slot = -1;
          for (int j = 0; j < out.cnt[e]; ++j)
            if (out.src[e * 4 + j] == key) slot = j;
          if (slot < 0) {
            slot = out.cnt[e]++;
            out.src[e * 4 + slot] = key;
          }
          for (int r = 0; r < nr; ++r) {
            const size_t xi = m + (size_t)nm * (r + (size_t)nr * s);
            const size_t ci = ((e * 4 + slot) * nr + r) * 2;
            out.coef[ci] += f * t.exp_sqr[2 * xi] / t.nx;
            out.coef[ci + 1] += f * t.exp_sqr[2 * xi + 1] / t.nx;
          }
        }
      }
}

int build_phir_blocks(const Tables& t, const PhirTable& pt, PhirBlocks& out) {
  const int nl = t.nl, nm = t.nibz, nr = t.nr;
  // partner groups: the images of l under the eight operations
  std::vector<int> gid(nl, -1);
  std::vector<std::vector<int>> groups;
  for (int l = 0; l < nl; ++l) {
    if (gid[l] >= 0) continue;
    std::vector<int> g;
    for (int s = 0; s < kNs; ++s) {
      const int ls = t.lsym[s * nl + l];
      if (std::find(g.begin(), g.end(), ls) == g.end()) g.push_back(ls);
    }
    std::sort(g.begin(), g.end());
    TUPS_CHECK(g.size() <= 2, TUPS_EINVAL, "PhiR: a form-factor partner group larger than two");
    for (int x : g) {
      TUPS_CHECK(gid[x] < 0, TUPS_EINVAL, "PhiR: form-factor partner groups overlap");
      gid[x] = (int)groups.size();
    }
    groups.push_back(g);
  }
  out.blk.clear();
  out.coef.clear();
  for (int sh = 0; sh < 4; ++sh) {
    const int na = 1 + (sh & 1), nb = 1 + (sh >> 1), no = na * nb;
    out.first[sh] = (int)(out.blk.size() / 8);
    out.count[sh] = 0;
    for (const auto& ga : groups)
      for (const auto& gb : groups) {
        if ((int)ga.size() != na || (int)gb.size() != nb) continue;
        const size_t off = out.coef.size() / 2;
        out.blk.insert(out.blk.end(), {ga[0], ga[na - 1], gb[0], gb[nb - 1], na, nb, (int)off, 0});
        out.coef.resize(out.coef.size() + (size_t)2 * nm * nr * no * no, 0.0);
        for (int i = 0; i < na; ++i)
          for (int j = 0; j < nb; ++j)
            for (int m = 0; m < nm; ++m) {
              const size_t e = ((size_t)ga[i] * nl + gb[j]) * nm + m;
              const int o = i + na * j;
              for (int slot = 0; slot < pt.cnt[e]; ++slot) {
                const int key = pt.src[e * 4 + slot], l1s = key % nl, l2s = key / nl;
                const int is = (int)(std::find(ga.begin(), ga.end(), l1s) - ga.begin());
                const int js = (int)(std::find(gb.begin(), gb.end(), l2s) - gb.begin());
                TUPS_CHECK(is < na && js < nb, TUPS_EINVAL, "PhiR: a source outside its symmetry block");
                const int sidx = is + na * js;
                for (int r = 0; r < nr; ++r) {
                  const size_t ci = 2 * (off + (((size_t)m * nr + r) * no + o) * no + sidx);
                  const size_t pi = ((e * 4 + slot) * nr + r) * 2;
                  out.coef[ci] += pt.coef[pi];
                  out.coef[ci + 1] += pt.coef[pi + 1];
                }
              }
            }
        ++out.count[sh];
      }
  }
  return TUPS_OK;
}

}  // namespace tups
