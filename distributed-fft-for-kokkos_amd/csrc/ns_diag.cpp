// Diagnostics of the Navier-Stokes solver (examples/navier-stokes-MPI/
// main.cpp:781-935): physical velocity and vorticity, the radial kinetic
// energy spectrum, and the reference's raw per-rank field files.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <vector>

#include "navier_stokes.hpp"

namespace dfft {

namespace {
void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(DFFT_ERROR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr int kSpecBlocks = 128;  // blocks (rows of partial sums) per rank

// IO::zfill + to_binary_file naming (main.cpp:918-934)
std::string dat_name(const std::string& dir, const std::string& label, int rx, int ry, int iter) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "_rx%03d_ry%03d_%010d.dat", rx, ry, iter);
  return dir + "/" + label + buf;
}

void write_raw(const std::string& name, const void* data, std::size_t bytes) {
  std::ofstream ofs(name, std::ios::binary);
  if (!ofs) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "Failed to open file: " + name);
  ofs.write(static_cast<const char*>(data), static_cast<std::streamsize>(bytes));
  if (!ofs) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "Failed to write file: " + name);
}
}  // namespace

void NavierStokes::velocity(void* const* u3, hipStream_t s) {
  Plan* pb3 = plan_b3_ ? plan_b3_.get() : plan_b_.get();
  pb3->execute(y_.data(), u3, DFFT_BACKWARD, s);
}

void NavierStokes::vorticity(void* const* w3, hipStream_t s) {
  ensure_aux();
  for (int r = 0; r < P_; ++r)
    hip_check(ns_vorticity(g_, sblk_[r], static_cast<const double2*>(y_[r]), static_cast<double2*>(aux_s3a_[r]), s),
              "ns_vorticity");
  Plan* pb3 = plan_b3_ ? plan_b3_.get() : plan_b_.get();
  pb3->execute(aux_s3a_.data(), w3, DFFT_BACKWARD, s);
}

void NavierStokes::energy_spectrum(int nb, double* e_spec, double* k_vals, hipStream_t s) {
  if (nb < 1 || nb > kMaxSpectrumBins)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "number of bins must be in [1, " + std::to_string(kMaxSpectrumBins) + "]");
  // k_max over the whole spectrum: |kx| = |ky| = n/2 (fftfreq's -n/2 entry), kz = n/2
  const double a = g_.kscale * (n_ / 2);
  const double k_max = std::sqrt(a * a + a * a + a * a);
  // Math::linspace(0, k_max, nb + 1, endpoint = true) (math_utils.hpp:14-33)
  std::vector<double> bins(nb + 1);
  const double delta = k_max / nb;
  for (int i = 0; i < nb; ++i) bins[i] = delta * i;
  bins[nb] = k_max;
  if (!spec_partials_) {
    hip_check(hipMalloc(&spec_partials_, sizeof(double) * kSpecBlocks * kMaxSpectrumBins * P_), "hipMalloc");
    hip_check(hipMalloc(&spec_bins_, sizeof(double) * (kMaxSpectrumBins + 1 + kMaxSpectrumBins)), "hipMalloc");
  }
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  hip_check(hipMemcpy(spec_bins_, bins.data(), sizeof(double) * (nb + 1), hipMemcpyHostToDevice), "hipMemcpy");
  double* sums = spec_bins_ + kMaxSpectrumBins + 1;
  for (int r = 0; r < P_; ++r)
    hip_check(ns_energy_spectrum_partials(g_, sblk_[r], static_cast<const double2*>(y_[r]), spec_bins_, nb,
                                          spec_partials_ + static_cast<std::size_t>(r) * kSpecBlocks * nb,
                                          kSpecBlocks, s),
              "ns_energy_spectrum");
  hip_check(ns_bin_sum(spec_partials_, kSpecBlocks * P_, nb, sums, s), "ns_bin_sum");
  hip_check(hipMemcpyAsync(e_spec, sums, sizeof(double) * nb, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (k_vals)
    for (int i = 0; i < nb; ++i) k_vals[i] = 0.5 * (bins[i] + bins[i + 1]);
}

double NavierStokes::diag(const std::string& out_dir, int diag_it, hipStream_t s) {
  namespace fs = std::filesystem;
  const std::string dir = out_dir + "_px" + std::to_string(px_) + "_py" + std::to_string(py_);
  std::error_code ec;
  fs::create_directories(dir, ec);
  if (ec) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "cannot create " + dir + ": " + ec.message());
  ensure_aux();
  const char* vel[3] = {"u", "v", "w"};
  const char* vor[3] = {"vor_x", "vor_y", "vor_z"};
  std::vector<double> host;
  auto dump = [&](const char* const* labels) {
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    for (int r = 0; r < P_; ++r) {
      const std::size_t n = static_cast<std::size_t>(phys_pts(r));
      host.resize(3 * n);
      if (n > 0)
        hip_check(hipMemcpy(host.data(), aux_p3_[r], sizeof(double) * 3 * n, hipMemcpyDeviceToHost), "hipMemcpy");
      const int rx = r / py_, ry = r % py_;  // rank_to_coord({px, py}, r)
      for (int c = 0; c < 3; ++c) write_raw(dat_name(dir, labels[c], rx, ry, diag_it), host.data() + c * n,
                                            sizeof(double) * n);
    }
  };
  velocity(aux_p3_.data(), s);
  dump(vel);
  vorticity(aux_p3_.data(), s);
  dump(vor);
  const int nb = 50;  // m_num_bins (main.cpp:437)
  std::vector<double> e_spec(nb), k_vals(nb);
  energy_spectrum(nb, e_spec.data(), k_vals.data(), s);
  write_raw(dat_name(dir, "E_spec", 0, 0, diag_it), e_spec.data(), sizeof(double) * nb);
  write_raw(dat_name(dir, "k_vals", 0, 0, diag_it), k_vals.data(), sizeof(double) * nb);
  return energy(s);
}

}  // namespace dfft
