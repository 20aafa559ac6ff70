/*
 * tups_h5.h — the reference's HDF5 vertex dump, from the device vertices.
 *
 * write_F (SRC/parquet_plot.f90:237-276) writes each channel's full vertex as
 * reshape(F_r_LL, (/ Nf, Nl, Nf, Nl, Nb /)) into datasets F/F_d, F/F_m, F/F_s,
 * F/F_t of one file, through hdf5_write_data (SRC/hdf5_wrapper.f90:2410-2445):
 * a compound type {"r": double at 0, "i": double at 8}
 * (SRC/hdf5_wrapper.f90:285-287) and, as the Fortran HDF5 API stores the
 * dimensions in reverse, the dataspace (Nb, Nl, Nf, Nl, Nf) in C order.  The
 * reshape is the resident (Nz, Nz, Nq) layout itself, so the device array is
 * written unchanged.  Built as libtups_h5.so against the HDF5 C library (1.10).
 */
#ifndef TUPS_H5_H
#define TUPS_H5_H

#include <stdint.h>

#include "tups_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* hdf5_write_data for one complex(8) array of shape (nf, nl, nf, nl, nb)
 * (interleaved re, im, Fortran order) as dataset `dset` ("group/name", groups
 * created as needed) of the file at `path`; create != 0 makes a new file
 * (truncating, as hdf5_create_file), else the file is opened for writing.
 * Returns 0 or TUPS_EINVAL; tups_h5_last_error() describes a failure. */
int tups_h5_write_vertex(const char* path, const char* dset, const double* data, int32_t nf, int32_t nl,
                         int32_t nb, int32_t create);
/* write_F: the four resident F_r of the context into F/F_d .. F/F_t of a new
 * file at `path` (downloads them through tups_download). */
int tups_h5_write_f(tups_ctx* ctx, const char* path);
const char* tups_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
