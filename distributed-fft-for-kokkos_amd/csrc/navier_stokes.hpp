// Pseudo-spectral incompressible Navier-Stokes solver on P virtual ranks
// (the caller of the distributed plan in examples/navier-stokes-MPI and
// examples/navier-stokes-MPI-batched).
//
// Every field set is held as in the batched example: a leading field axis,
// physical blocks in the (1, px, py, 1) topology, spectral blocks in
// (1, 1, px, py).  Each RK stage runs two batched plan executions: the
// backward transform of all fields it needs and the forward transform of the
// nonlinear terms -- each line pass is one kernel launch over every field.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "ns_kernels.hpp"
#include "plan.hpp"

namespace dfft {

class NavierStokes {
 public:
  // form 0: divergence form, 3 C2R + 6 R2C per stage (u_i u_j products);
  // form 1: the reference's advective form, 12 C2R + 3 R2C per stage.
  NavierStokes(int n, int px, int py, double lx, double nu, double dt, int form);
  ~NavierStokes();
  NavierStokes(const NavierStokes&) = delete;
  NavierStokes& operator=(const NavierStokes&) = delete;

  void init_taylor_green(double v0, hipStream_t s);
  void autotune(hipStream_t s);  // tunes every plan on scratch buffers; keeps the state
  // "fuse_products" (1, default): the divergence form's 6 products are formed
  // in the forward transform's first pass from the 3 velocity fields where
  // the plan allows it; 0: a separate products kernel writes the 6 fields
  void set_option(const std::string& key, long value);
  long get_option(const std::string& key) const;
  void step(int nsteps, bool use_graph, hipStream_t s);
  double energy(hipStream_t s);
  void get_state(void* const* dst, hipStream_t s) const;
  void set_state(void* const* src, hipStream_t s);
  void set_physical(void* const* u3, hipStream_t s);  // physical velocity -> dealiased, projected spectrum
  const NsBlock& spectral_block(int r) const { return sblk_.at(r); }
  const NsBlock& physical_block(int r) const { return pblk_.at(r); }
  int nranks() const { return P_; }
  // throws DFFT_ERROR_MISMATCH unless every rank's allocation holds 3 fields
  // of its spectral (complex) or physical (real) block
  void check_fields(void* const* f3, bool spectral, const char* what) const;

  // diagnostics (ns_diag.cpp; main.cpp:812-935)
  void velocity(void* const* u3, hipStream_t s);   // physical velocity of the state
  void vorticity(void* const* w3, hipStream_t s);  // physical vorticity
  // radial spectrum of 0.5 f |u_k|^2 over num_bins bins of [0, k_max]; k_vals = bin midpoints
  void energy_spectrum(int num_bins, double* e_spec, double* k_vals, hipStream_t s);
  // the reference's diag(): per-rank u, v, w, vor_x, vor_y, vor_z and the
  // E_spec / k_vals files under <out_dir>_px<px>_py<py>; returns the energy
  double diag(const std::string& out_dir, int diag_it, hipStream_t s);

  // reverse mode (ns_adjoint.cpp)
  double gradient(int nsteps, int checkpoint_every, void* const* u0, void* const* grad, hipStream_t s);
  void apply_rhs(void* const* y3, void* const* out3, hipStream_t s);          // R(y)
  void apply_rhs_tangent(void* const* y3, void* const* d3, void* const* out3, hipStream_t s);  // R'(y) d
  void apply_rhs_adjoint(void* const* y3, void* const* l3, void* const* out3, hipStream_t s);  // R'(y)^* l

 private:
  // one RK4 stage; u_keep (optional, 3 physical fields per rank) receives the
  // stage's physical velocity IFFT(y) instead of the scratch buffer
  void rk_stage(int stage, hipStream_t s, void* const* u_keep = nullptr);
  // R'(y)^* l (divergence form); u3 = IFFT(D y) if already known (else from
  // y3); finish: 0 writes t to out3, 1/2 fuse the reverse RK4 combination
  struct AdjFinish {
    int mode = 0;
    void* const* out3 = nullptr;  // mode 0
    void* const* acc = nullptr;   // modes 1, 2
    void* const* lam = nullptr;
    void* const* fb = nullptr;  // mode 1: written (may alias l3)
    double a = 0.0, c = 0.0;
  };
  void adjoint_core(const void* const* y3, const void* const* l3, const void* const* u3, const AdjFinish& f,
                    hipStream_t s);
  void one_step(hipStream_t s);
  int64_t spec_pts(int r) const;
  int64_t phys_pts(int r) const;
  std::vector<void*> alloc_set(bool spectral, int nf);
  void ensure_aux();

  // forward transform of the 6 products of the velocity blocks u (fused into
  // the R2C pass when enabled and supported, else products kernel + transform)
  void forward_products(void* const* u, hipStream_t s);
  bool fuse_products_ = true;
  int n_, px_, py_, P_, form_;
  NsGrid g_{};
  double lx_;
  std::unique_ptr<Plan> plan_b_, plan_f_;  // backward (nf_b fields), forward (nf_f fields)
  std::unique_ptr<Plan> plan_b3_, plan_f3_;  // 3-field plans (init, adjoint)
  std::unique_ptr<Plan> plan_b6_;            // 6-field backward (adjoint)
  int nf_b_, nf_f_;
  std::vector<NsBlock> sblk_, pblk_;  // per rank (3-field blocks; nf is set per use)
  // device buffers, one pointer per rank
  std::vector<void*> y_, yc_, k1_, k2_, k3_, nl_, bin_, pb_, pf_;
  std::vector<void*> aux_s3a_, aux_s3b_, aux_p3_;  // tangent/adjoint work fields (allocated on first use)
  std::vector<void*> all_;
  double* partials_ = nullptr;
  double* spec_partials_ = nullptr;  // diagnostics: per-block spectrum rows (allocated on first use)
  double* spec_bins_ = nullptr;
  double* esum_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipStream_t graph_stream_ = nullptr;
};

}  // namespace dfft
