// Green's functions of the bubble and self-energy kernels, on the device.
//
// get_green (SRC/parquet_util.f90:1278-1299, Gkw of SRC/parquet_chi.f90:208)
// and get_green_coarse (SRC/parquet_util.f90:1303-1327, Ek_grain of
// SRC/parquet_util.f90:143-164): G = 1 / (i nu_n - e_k + mu - Sigma(k, nu)),
// with the Hartree value Sigma_H(k) outside the box 1..Nf; the coarse form
// shifts k by the grain offset.
#pragma once
#include <hip/hip_runtime.h>

namespace tups {

struct GreenParams {
  const double2* sigma;  // (Nx Ny Nf): ((kx-1) Ny + ky-1) Nf + nu-1
  const double2* sigh;   // Sigma_H(Nx, Ny), column-major
  int nx, nf, ngrain;
  double mu, beta;
  const double* mu_dev;  // non-null: mu lives on the device (the full iteration's search)
};

// the GreenParams a kernel works with: mu read once from the device when it lives there
__device__ __forceinline__ GreenParams green_resolve(const GreenParams& a) {
  GreenParams g = a;
  if (g.mu_dev) g.mu = *g.mu_dev;
  return g;
}

__device__ __forceinline__ double2 green_at(const GreenParams& a, double ek, int kx, int ky, int nu) {
  const double pi = 3.14159265358979323846;
  const double wn = pi / a.beta * (2.0 * (nu - a.nf / 2 - 1) + 1.0);
  double2 s;
  if (nu < 1 || nu > a.nf)
    s = a.sigh[(kx - 1) + (size_t)a.nx * (ky - 1)];
  else
    s = a.sigma[((size_t)(kx - 1) * a.nx + (ky - 1)) * a.nf + (nu - 1)];
  const double re = -ek + a.mu - s.x, im = wn - s.y;  // z = re + i im
  const double d = re * re + im * im;
  return make_double2(re / d, -im / d);
}

// get_green_coarse: 1-based grain (gx, gy), momentum (kx, ky), frequency nu (may lie outside 1..Nf)
__device__ __forceinline__ double2 green_coarse(const GreenParams& a, int gx, int gy, int kx, int ky, int nu) {
  const double pi = 3.14159265358979323846;
  const int nx = a.nx, ng = a.ngrain;
  const double dk = 2.0 * pi / nx;
  const int half = (ng + 1) / 2;
  const double px = dk * (kx - 1) + (gx - half) * dk / (double)ng;
  const double py = dk * (ky - 1) + (gy - half) * dk / (double)ng;
  return green_at(a, -2.0 * cos(px) - 2.0 * cos(py), kx, ky, nu);
}

// get_green: the dispersion at the lattice point itself
__device__ __forceinline__ double2 green_plain(const GreenParams& a, int kx, int ky, int nu) {
  const double pi = 3.14159265358979323846;
  const double dk = 2.0 * pi / a.nx;
  return green_at(a, -2.0 * cos(dk * (kx - 1)) - 2.0 * cos(dk * (ky - 1)), kx, ky, nu);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

}  // namespace tups
