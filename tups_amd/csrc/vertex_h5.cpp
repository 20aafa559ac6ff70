// The reference's HDF5 vertex dump (write_F, SRC/parquet_plot.f90:237-276) for the
// vertices resident on the device.  Host code over the HDF5 C library; a
// library of its own so that the compute library does not depend on HDF5.
#include "../../include/tups_h5.h"

#include <hdf5.h>

#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
  g_err = m;
  return TUPS_EINVAL;
}

// the compound {"r", "i"} of hdf5_create_complex_datatype (SRC/hdf5_wrapper.f90:285-287)
hid_t complex_type() {
  const hid_t t = H5Tcreate(H5T_COMPOUND, 2 * sizeof(double));
  if (t < 0) return t;
  if (H5Tinsert(t, "r", 0, H5T_NATIVE_DOUBLE) < 0 || H5Tinsert(t, "i", sizeof(double), H5T_NATIVE_DOUBLE) < 0) {
    H5Tclose(t);
    return -1;
  }
  return t;
}

// the group of "a/b/name" (created as needed); *name receives the last part
hid_t open_groups(hid_t file, const std::string& full, std::string* name) {
  const size_t p = full.rfind('/');
  *name = p == std::string::npos ? full : full.substr(p + 1);
  hid_t loc = H5Gopen2(file, "/", H5P_DEFAULT);
  if (p == std::string::npos || loc < 0) return loc;
  size_t b = 0;
  while (b < p) {
    size_t e = full.find('/', b);
    if (e == std::string::npos || e > p) e = p;
    const std::string g = full.substr(b, e - b);
    b = e + 1;
    if (g.empty()) continue;
    hid_t next = H5Lexists(loc, g.c_str(), H5P_DEFAULT) > 0 ? H5Gopen2(loc, g.c_str(), H5P_DEFAULT)
                                                           : H5Gcreate2(loc, g.c_str(), H5P_DEFAULT, H5P_DEFAULT,
                                                                        H5P_DEFAULT);
    H5Gclose(loc);
    if (next < 0) return next;
    loc = next;
  }
  return loc;
}

}  // namespace

extern "C" const char* tups_h5_last_error(void) { return g_err.c_str(); }

extern "C" int tups_h5_write_vertex(const char* path, const char* dset, const double* data, int32_t nf, int32_t nl,
                                    int32_t nb, int32_t create) {
  if (!path || !dset || !data) return fail("null argument");
  if (nf <= 0 || nl <= 0 || nb <= 0) return fail("sizes must be positive");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors come back as codes
  const hid_t file = create ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                            : H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
  if (file < 0) return fail(std::string("cannot ") + (create ? "create " : "open ") + path);
  std::string name;
  const hid_t grp = open_groups(file, dset, &name);
  const hid_t type = complex_type();
  // Fortran (nf, nl, nf, nl, nb) is the C dataspace (nb, nl, nf, nl, nf)
  const hsize_t dims[5] = {(hsize_t)nb, (hsize_t)nl, (hsize_t)nf, (hsize_t)nl, (hsize_t)nf};
  const hid_t space = H5Screate_simple(5, dims, dims);
  int rc = 0;
  if (grp < 0 || type < 0 || space < 0) {
    rc = fail(std::string("cannot set up dataset ") + dset);
  } else {
    const hid_t d = H5Dcreate2(grp, name.c_str(), type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (d < 0) {
      rc = fail(std::string("cannot create dataset ") + dset + " (does it exist already?)");
    } else {
      if (H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) rc = fail(std::string("cannot write ") + dset);
      H5Dclose(d);
    }
  }
  if (space >= 0) H5Sclose(space);
  if (type >= 0) H5Tclose(type);
  if (grp >= 0) H5Gclose(grp);
  if (H5Fclose(file) < 0 && rc == 0) rc = fail(std::string("cannot close ") + path);
  return rc;
}

extern "C" int tups_h5_write_f(tups_ctx* ctx, const char* path) {
  if (!ctx || !path) return fail("null argument");
  tups_sizes s{};
  int rc = tups_query(ctx, &s);
  if (rc) {
    g_err = std::string("tups_query: ") + tups_last_error();
    return rc;
  }
  const size_t n = (size_t)s.nz * s.nz * s.nq;
  std::vector<double> buf(2 * n);
  static const char* names[4] = {"F/F_d", "F/F_m", "F/F_s", "F/F_t"};
  for (int r = 0; r < 4; ++r) {
    if ((rc = tups_download(ctx, TUPS_F, r, buf.data())) || (rc = tups_sync(ctx))) {
      g_err = std::string("download of F: ") + tups_last_error();
      return rc;
    }
    if ((rc = tups_h5_write_vertex(path, names[r], buf.data(), s.nf, s.nl, s.nq, r == 0))) return rc;
  }
  return TUPS_OK;
}
