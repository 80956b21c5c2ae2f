// Pointwise kernels of the pseudo-spectral Navier-Stokes solver
// (examples/navier-stokes-MPI/main.cpp, examples/utils/math_utils.hpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace dfft {

// One virtual rank's block of a spectral (x, y, z/2+1) or physical (x, y, z)
// field set: `nf` fields of extents ext[3] starting at global index start[3].
struct NsBlock {
  int32_t nf;
  int32_t ext[3];
  int32_t start[3];
  // divisions by ext[1] and ext[2] (set by set_divisors; block-local indices
  // stay below 2^31)
  FastDiv div1, div2;
  void set_divisors() {
    div1 = FastDiv(static_cast<uint32_t>(ext[1]));
    div2 = FastDiv(static_cast<uint32_t>(ext[2]));
  }
};

struct NsGrid {
  int32_t n;       // grid points per direction
  int32_t cutoff;  // 2/3-rule cutoff on |integer wavenumber index|
  double kscale;   // wavenumber = kscale * signed index (2 / lx)
  double nu;
  double dt;
  double coef;     // (lx pi)^3 / n^6, kinetic-energy normalisation
};

// u (3 fields) -> the 6 products u_i u_j (i <= j: uu, uv, uw, vv, vw, ww)
hipError_t ns_products(const double* u3, double* p6, int64_t npts, hipStream_t s);
// u (3) + 9 gradients (du/dx, du/dy, du/dz, dv/dx, ...) -> advection u.grad(u) (3)
hipError_t ns_advection(const double* f12, double* adv3, int64_t npts, hipStream_t s);
// spectral state y (3) -> [y, i kx y_0, i ky y_0, i kz y_0, i kx y_1, ...] (12)
hipError_t ns_gradients(const NsGrid& g, const NsBlock& b, const double2* y3, double2* f12, hipStream_t s);

// Right-hand side of the RK stage and the RK4 update (main.cpp:572-666 and
// math_utils.hpp:89-141).  form 0: `nl` holds the 6 product spectra
// (divergence form, N_i = i k_j (u_i u_j)^); form 1: `nl` holds the 3
// advection spectra.  y is the stage value (updated in place), yc/k1/k2/k3
// the RK registers.
struct NsStage {
  int32_t form;
  int32_t stage;
  const double2* nl;
  double2* y;
  double2* yc;
  double2* k1;
  double2* k2;
  double2* k3;
};
hipError_t ns_rhs_rk(const NsGrid& g, const NsBlock& b, const NsStage& st, hipStream_t s);

// Projection u - k (k.u)/k^2 with the dealias mask (main.cpp:235-254, 214-225).
hipError_t ns_dealias_project(const NsGrid& g, const NsBlock& b, double2* y3, hipStream_t s);

// Taylor-Green field on a physical block (main.cpp:349-353): grid spacing h.
hipError_t ns_taylor_green(const NsBlock& b, double h, double v0, double* u3, hipStream_t s);

// Kinetic energy partial sums of a spectral block (main.cpp:855-910, global kz
// index for the Hermitian weight); partials has `nblocks` doubles.
hipError_t ns_energy_partials(const NsGrid& g, const NsBlock& b, const double2* y3, double* partials,
                              int nblocks, hipStream_t s);
hipError_t ns_sum(const double* partials, int n, double* out, hipStream_t s);

// ---- tangent / adjoint pieces (divergence form) ---------------------------
// p_ij = du_i u_j + u_i du_j (6 fields) from u (3) and du (3)
hipError_t ns_products_tangent(const double* u3, const double* du3, double* p6, int64_t npts, hipStream_t s);
// R(y) only (no RK update): out = P(-D N - nu k^2 D y), N from 6 product spectra
hipError_t ns_rhs_only(const NsGrid& g, const NsBlock& b, const double2* nl6, const double2* y3, double2* out3,
                       hipStream_t s);
// mu = D P lambda (written to mu3) and the 6 symmetric fields
// sigma_mj = -i (k_j mu_m + k_m mu_j), m <= j (written to sig6)
hipError_t ns_adjoint_prepare(const NsGrid& g, const NsBlock& b, const double2* lam3, double2* mu3,
                              double2* sig6, hipStream_t s);
// c_m = sum_j u_j S_mj (physical)
hipError_t ns_adjoint_contract(const double* u3, const double* S6, double* c3, int64_t npts, hipStream_t s);
// out = -D chat - nu k^2 mu
hipError_t ns_adjoint_finish(const NsGrid& g, const NsBlock& b, const double2* chat3, const double2* mu3,
                             double2* out3, hipStream_t s);
// x = a*x + b*y over n complex values; dealias / dealias+project helpers on 3 fields
hipError_t ns_axpby(int64_t n, double a, double2* x, double b, const double2* y, hipStream_t s);
// dst = D src (src == dst allowed)
hipError_t ns_dealias(const NsGrid& g, const NsBlock& b, const double2* src3, double2* dst3, hipStream_t s);
// x *= a over n doubles
hipError_t ns_scale(int64_t n, double a, double* x, hipStream_t s);

// reverse RK4 step pieces: t = -D chat - nu k^2 mu; mode 1: acc += t,
// fb = dt (a lam + c t); mode 2: lam = D (acc + t)
hipError_t ns_adjoint_finish_rk(const NsGrid& g, const NsBlock& b, const double2* chat3, const double2* mu3,
                                double2* acc3, double2* lam3, double2* fb3, double a, double c, int mode,
                                hipStream_t s);
// fb = c lam, acc = lam over n complex values
hipError_t ns_rk_adjoint_init(int64_t n, const double2* lam, double2* fb, double2* acc, double c, hipStream_t s);

// diagnostics (main.cpp:781-910)
constexpr int kMaxSpectrumBins = 256;
// w = curl of y: i (ky w - kz v), i (kz u - kx w), i (kx v - ky u)
hipError_t ns_vorticity(const NsGrid& g, const NsBlock& b, const double2* y3, double2* w3, hipStream_t s);
// per-block partial sums of 0.5 f (|u|^2 + |v|^2 + |w|^2) over the half-open
// radial bins [bins[i], bins[i+1]), f = 1 on the global kz = 0 plane, 2 elsewhere
hipError_t ns_energy_spectrum_partials(const NsGrid& g, const NsBlock& b, const double2* y3, const double* bins,
                                       int nb, double* partials, int nblocks, hipStream_t s);
hipError_t ns_bin_sum(const double* partials, int rows, int nb, double* out, hipStream_t s);

}  // namespace dfft
