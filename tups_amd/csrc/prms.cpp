// The run-parameter file of the reference.
//
// Reference: set_parameters, SRC/parquet_ini.f90:289-329.  The file holds 21
// records, each read by one list-directed READ of one item: beta, xU, nParticle,
// Nx (Ny = Nx), Nf, neighbours, readinpick, pick_file, Ngrain, f_range,
// f_damping, ite_max, DGA, fname_L, use_sig_dmft, fname_sig_dmft, use_U2,
// save_F, oneshot, old, SDE_old.
//
// List-directed input of one item per READ, as the reading here implements it:
// each READ starts at the next record; blank records are skipped; the value is
// the first token (ended by a blank, comma, slash or the end of the record),
// and the rest of the record is ignored (so the trailing "!comment"s of the
// reference's file are never read); "r*value" repeats a value; a null value (a
// leading comma or slash) leaves the variable undefined in Fortran and is an
// error here.  Reals take D or E exponents; integers are optionally signed
// digits; logicals are an optional period then T or F, case-insensitive, with
// any further characters; character values are either delimited by ' or "
// (a doubled delimiter stands for one) and may then span blanks, or
// undelimited up to the first blank, comma or slash; values longer than the
// reference's character(len=100) are cut to 100 characters.
#include "internal.h"

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

namespace {

constexpr size_t kFortranLen = 100;  // character(len=100), SRC/parquet_ini.f90:58-64,281

struct Reader {
  std::ifstream in;
  int line = 0;
  std::string field;

  // the next non-blank record, from its first non-blank character
  bool next_record(std::string* rec) {
    std::string s;
    while (std::getline(in, s)) {
      ++line;
      if (!s.empty() && s.back() == '\r') s.pop_back();
      const size_t p = s.find_first_not_of(" \t");
      if (p == std::string::npos) continue;
      *rec = s.substr(p);
      return true;
    }
    return false;
  }

  int fail(const std::string& what) {
    tups::set_error("prms line " + std::to_string(line) + " (" + field + "): " + what);
    return TUPS_EINVAL;
  }

  // first value of the next record; character values keep their delimiters off
  int token(std::string* tok, bool character) {
    std::string rec;
    if (!next_record(&rec)) return fail("the file ends before this value");
    if (rec[0] == ',' || rec[0] == '/') return fail("null value (the reference leaves the variable undefined)");
    size_t i = 0;
    // r*value
    size_t j = 0;
    while (j < rec.size() && std::isdigit((unsigned char)rec[j])) ++j;
    if (j > 0 && j < rec.size() && rec[j] == '*') {
      i = j + 1;
      if (i >= rec.size() || rec[i] == ' ' || rec[i] == ',' || rec[i] == '/')
        return fail("null value after a repeat count");
    }
    if (character && (rec[i] == '\'' || rec[i] == '"')) {
      const char d = rec[i];
      std::string v;
      size_t k = i + 1;
      for (;;) {
        if (k >= rec.size()) return fail("unterminated character value");
        if (rec[k] == d) {
          if (k + 1 < rec.size() && rec[k + 1] == d) {
            v.push_back(d);
            k += 2;
            continue;
          }
          break;
        }
        v.push_back(rec[k++]);
      }
      *tok = v;
      return TUPS_OK;
    }
    size_t k = i;
    while (k < rec.size() && rec[k] != ' ' && rec[k] != '\t' && rec[k] != ',' && rec[k] != '/') ++k;
    *tok = rec.substr(i, k - i);
    return TUPS_OK;
  }

  int real(const char* name, double* out) {
    field = name;
    std::string t;
    if (int rc = token(&t, false)) return rc;
    for (char& ch : t)
      if (ch == 'd' || ch == 'D' || ch == 'q' || ch == 'Q') ch = 'e';
    if (t.find_first_not_of("+-.0123456789eE") != std::string::npos ||
        t.find_first_of("0123456789") == std::string::npos)
      return fail("not a real value: '" + t + "'");
    errno = 0;
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    if (end != t.c_str() + t.size() || errno == ERANGE) return fail("not a real value: '" + t + "'");
    *out = v;
    return TUPS_OK;
  }

  int integer(const char* name, int32_t* out) {
    field = name;
    std::string t;
    if (int rc = token(&t, false)) return rc;
    size_t p = (t[0] == '+' || t[0] == '-') ? 1 : 0;
    if (p >= t.size() || t.find_first_not_of("0123456789", p) != std::string::npos)
      return fail("not an integer value: '" + t + "'");
    errno = 0;
    const long v = std::strtol(t.c_str(), nullptr, 10);
    if (errno == ERANGE || v < INT32_MIN || v > INT32_MAX) return fail("integer out of range: '" + t + "'");
    *out = (int32_t)v;
    return TUPS_OK;
  }

  int logical(const char* name, int32_t* out) {
    field = name;
    std::string t;
    if (int rc = token(&t, false)) return rc;
    size_t p = t[0] == '.' ? 1 : 0;
    const char c = p < t.size() ? (char)std::toupper((unsigned char)t[p]) : '\0';
    if (c != 'T' && c != 'F') return fail("not a logical value: '" + t + "'");
    *out = c == 'T' ? 1 : 0;
    return TUPS_OK;
  }

  int character(const char* name, char* out, size_t cap) {
    field = name;
    std::string t;
    if (int rc = token(&t, true)) return rc;
    if (t.size() > kFortranLen) t.resize(kFortranLen);
    std::snprintf(out, cap, "%s", t.c_str());
    return TUPS_OK;
  }
};

}  // namespace

extern "C" int tups_read_prms(const char* path, tups_prms* p) {
  TUPS_CHECK(path && p, TUPS_EINVAL, "null argument");
  Reader r;
  r.in.open(path);
  TUPS_CHECK(r.in.good(), TUPS_EINVAL, std::string("cannot open the parameter file ") + path);
  tups_prms v{};
  int rc = 0;
  if ((rc = r.real("beta", &v.beta)) || (rc = r.real("xU", &v.u)) || (rc = r.real("nParticle", &v.n_particle)) ||
      (rc = r.integer("Nx", &v.nx)) || (rc = r.integer("Nf", &v.nf)) ||
      (rc = r.integer("neighbours", &v.neighbours)) || (rc = r.logical("readinpick", &v.readinpick)) ||
      (rc = r.character("pick_file", v.pick_file, sizeof(v.pick_file))) ||
      (rc = r.integer("Ngrain", &v.ngrain)) || (rc = r.integer("f_range", &v.f_range)) ||
      (rc = r.real("f_damping", &v.f_damping)) || (rc = r.integer("ite_max", &v.ite_max)) ||
      (rc = r.logical("DGA", &v.dga)) || (rc = r.character("fname_L", v.fname_l, sizeof(v.fname_l))) ||
      (rc = r.logical("use_sig_dmft", &v.use_sig_dmft)) ||
      (rc = r.character("fname_sig_dmft", v.fname_sig_dmft, sizeof(v.fname_sig_dmft))) ||
      (rc = r.logical("use_U2", &v.use_u2)) || (rc = r.logical("save_F", &v.save_f)) ||
      (rc = r.logical("oneshot", &v.oneshot)) || (rc = r.logical("old", &v.old)) ||
      (rc = r.logical("SDE_old", &v.sde_old)))
    return rc;
  *p = v;
  return TUPS_OK;
}

extern "C" int tups_prms_dims(const tups_prms* p, tups_dims* d) {
  TUPS_CHECK(p && d, TUPS_EINVAL, "null argument");
  // neighbours -> number of form factors (SRC/parquet_formfactors.f90:495-515: the origin
  // counts 1, the shells (1,0) and (1,1) four each while Nx >= 4), i.e. 1 -> 1, 2 -> 5, 3 -> 9
  static const int nl_of[4] = {0, 1, 5, 9};
  TUPS_CHECK(p->neighbours >= 1 && p->neighbours <= 3, TUPS_EINVAL,
             "neighbours must be 1, 2 or 3 (1, 5 or 9 form factors)");
  TUPS_CHECK(p->neighbours == 1 || p->nx >= 4, TUPS_EINVAL,
             "neighbours > 1 needs Nx >= 4 (smaller lattices fold the shells onto the zone edge)");
  TUPS_CHECK(!p->readinpick, TUPS_EINVAL, "a form-factor pick list (readinpick = .true.) is not supported");
  tups_dims v{};
  v.nx = p->nx;
  v.nf = p->nf;
  v.nl = nl_of[p->neighbours];
  v.nr = 0;
  v.nq = 0;
  v.beta = p->beta;
  v.f_damping = p->f_damping;
  *d = v;
  return TUPS_OK;
}
