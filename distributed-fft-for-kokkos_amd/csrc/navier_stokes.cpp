// Pseudo-spectral Navier-Stokes driver (examples/navier-stokes-MPI/main.cpp,
// examples/navier-stokes-MPI-batched/main.cpp) over the fused plans.
#include "navier_stokes.hpp"

#include "buffer_check.hpp"
#include "trace.hpp"

#include <cmath>
#include <cstring>

namespace dfft {

namespace {
void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(DFFT_ERROR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

dfft_desc ns_desc(int nf, int n, int px, int py) {
  dfft_desc d{};
  d.rank = 4;
  d.fft_rank = 3;
  d.axes[0] = 1;
  d.axes[1] = 2;
  d.axes[2] = 3;
  d.transform = DFFT_R2C;
  d.precision = DFFT_FP64;
  d.layout = DFFT_LAYOUT_RIGHT;
  const size_t ein[4] = {size_t(nf), size_t(n), size_t(n), size_t(n)};
  const size_t eout[4] = {size_t(nf), size_t(n), size_t(n), size_t(n / 2 + 1)};
  // main-batched.cpp:102-107: (1, px, py, 1) -> (1, 1, px, py)
  const size_t tin[4] = {1, size_t(px), size_t(py), 1};
  const size_t tout[4] = {1, 1, size_t(px), size_t(py)};
  for (int a = 0; a < 4; ++a) {
    d.in_extents[a] = ein[a];
    d.out_extents[a] = eout[a];
    d.in_topology[a] = tin[a];
    d.out_topology[a] = tout[a];
  }
  d.in_topology_layout = DFFT_LAYOUT_RIGHT;
  d.out_topology_layout = DFFT_LAYOUT_RIGHT;
  d.norm = DFFT_NORM_BACKWARD;
  d.nranks = px * py;
  d.exec_mode = DFFT_EXEC_FUSED;
  return d;
}
}  // namespace

NavierStokes::NavierStokes(int n, int px, int py, double lx, double nu, double dt, int form)
    : n_(n), px_(px), py_(py), P_(px * py), form_(form), lx_(lx) {
  if (n < 4 || n % 2 != 0) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "n must be even and >= 4");
  if (px < 1 || py < 1 || P_ > kMaxRanks) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "bad process grid");
  if (form != 0 && form != 1) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "form must be 0 or 1");
  g_.n = n;
  g_.cutoff = ((n / 2) * 2) / 3;  // main.cpp:166-169, integer arithmetic
  g_.kscale = 2.0 / lx;           // fftfreq(n, lx / (2 n)) = index * 2 / lx
  g_.nu = nu;
  g_.dt = dt;
  g_.coef = std::pow(lx * M_PI, 3) / std::pow(static_cast<double>(n), 6);
  nf_b_ = form == 0 ? 3 : 12;
  nf_f_ = form == 0 ? 6 : 3;
  plan_b_ = std::make_unique<Plan>(ns_desc(nf_b_, n, px, py));
  plan_f_ = std::make_unique<Plan>(ns_desc(nf_f_, n, px, py));
  plan_f3_ = nf_f_ == 3 ? nullptr : std::make_unique<Plan>(ns_desc(3, n, px, py));
  plan_b3_ = nf_b_ == 3 ? nullptr : std::make_unique<Plan>(ns_desc(3, n, px, py));
  for (int r = 0; r < P_; ++r) {
    Vec e, st;
    plan_f_->local_starts(r, 1, e, st);
    NsBlock sb{3, {int32_t(e[1]), int32_t(e[2]), int32_t(e[3])}, {int32_t(st[1]), int32_t(st[2]), int32_t(st[3])}};
    plan_f_->local_starts(r, 0, e, st);
    NsBlock pb{3, {int32_t(e[1]), int32_t(e[2]), int32_t(e[3])}, {int32_t(st[1]), int32_t(st[2]), int32_t(st[3])}};
    for (const NsBlock* blk : {&sb, &pb})
      if (int64_t(blk->ext[0]) * blk->ext[1] * blk->ext[2] >= (int64_t(1) << 31))
        throw Error(DFFT_ERROR_UNSUPPORTED, "Navier-Stokes rank block of 2^31 points or more");
    sb.set_divisors();
    pb.set_divisors();
    sblk_.push_back(sb);
    pblk_.push_back(pb);
  }
  y_ = alloc_set(true, 3);
  yc_ = alloc_set(true, 3);
  k1_ = alloc_set(true, 3);
  k2_ = alloc_set(true, 3);
  k3_ = alloc_set(true, 3);
  nl_ = alloc_set(true, nf_f_);
  bin_ = form == 1 ? alloc_set(true, 12) : std::vector<void*>(P_, nullptr);
  pb_ = alloc_set(false, std::max(nf_b_, 3));
  pf_ = alloc_set(false, nf_f_);
  hip_check(hipMalloc(&partials_, sizeof(double) * 256 * P_), "hipMalloc");
  hip_check(hipMalloc(&esum_, sizeof(double)), "hipMalloc");
}

NavierStokes::~NavierStokes() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (graph_stream_) (void)hipStreamDestroy(graph_stream_);
  for (void* p : all_) (void)hipFree(p);
  if (partials_) (void)hipFree(partials_);
  if (spec_partials_) (void)hipFree(spec_partials_);
  if (spec_bins_) (void)hipFree(spec_bins_);
  if (esum_) (void)hipFree(esum_);
}

int64_t NavierStokes::spec_pts(int r) const {
  return int64_t(sblk_[r].ext[0]) * sblk_[r].ext[1] * sblk_[r].ext[2];
}
int64_t NavierStokes::phys_pts(int r) const {
  return int64_t(pblk_[r].ext[0]) * pblk_[r].ext[1] * pblk_[r].ext[2];
}

void NavierStokes::check_fields(void* const* f3, bool spectral, const char* what) const {
  if (f3 == nullptr) throw Error(DFFT_ERROR_INVALID_ARGUMENT, std::string(what) + ": null pointer table");
  for (int r = 0; r < P_; ++r) {
    const std::size_t bytes = 3 * static_cast<std::size_t>(spectral ? spec_pts(r) * 16 : phys_pts(r) * 8);
    check_device_extent(f3[r], bytes, std::string(what) + " of rank " + std::to_string(r));
  }
}

std::vector<void*> NavierStokes::alloc_set(bool spectral, int nf) {
  std::vector<void*> out(P_, nullptr);
  for (int r = 0; r < P_; ++r) {
    const size_t bytes = size_t(nf) * (spectral ? spec_pts(r) * 16 : phys_pts(r) * 8);
    if (bytes == 0) continue;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) throw Error(DFFT_ERROR_OUT_OF_MEMORY, "hipMalloc of NS field failed");
    all_.push_back(p);
    out[r] = p;
  }
  return out;
}

void NavierStokes::forward_products(void* const* u, hipStream_t s) {
  if (fuse_products_ && plan_f_->products_supported()) {
    plan_f_->execute_forward_products(u, nl_.data(), s);
    return;
  }
  for (int r = 0; r < P_; ++r)
    hip_check(ns_products(static_cast<const double*>(u[r]), static_cast<double*>(pf_[r]), phys_pts(r), s),
              "ns_products");
  plan_f_->execute(pf_.data(), nl_.data(), DFFT_FORWARD, s);
}

void NavierStokes::set_option(const std::string& key, long value) {
  if (key == "fuse_products") {
    if (value != 0 && value != 1) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "fuse_products must be 0 or 1");
    if (graph_exec_) {  // the captured step holds the old launches
      hip_check(hipStreamSynchronize(graph_stream_), "hipStreamSynchronize");
      (void)hipGraphExecDestroy(graph_exec_);
      graph_exec_ = nullptr;
    }
    if (graph_) {
      (void)hipGraphDestroy(graph_);
      graph_ = nullptr;
    }
    fuse_products_ = value != 0;
    return;
  }
  throw Error(DFFT_ERROR_INVALID_ARGUMENT, "unknown NS option " + key);
}

long NavierStokes::get_option(const std::string& key) const {
  if (key == "fuse_products") return fuse_products_ ? 1 : 0;
  if (key == "products_fused") return (form_ == 0 && fuse_products_ && plan_f_->products_supported()) ? 1 : 0;
  throw Error(DFFT_ERROR_INVALID_ARGUMENT, "unknown NS option " + key);
}

void NavierStokes::rk_stage(int stage, hipStream_t s, void* const* u_keep) {
  if (form_ == 0) {
    void* const* u = u_keep ? u_keep : pb_.data();
    plan_b_->execute(y_.data(), u, DFFT_BACKWARD, s);
    forward_products(u, s);
  } else {
    for (int r = 0; r < P_; ++r)
      hip_check(ns_gradients(g_, sblk_[r], static_cast<double2*>(y_[r]), static_cast<double2*>(bin_[r]), s),
                "ns_gradients");
    plan_b_->execute(bin_.data(), pb_.data(), DFFT_BACKWARD, s);
    for (int r = 0; r < P_; ++r)
      hip_check(ns_advection(static_cast<double*>(pb_[r]), static_cast<double*>(pf_[r]), phys_pts(r), s),
                "ns_advection");
    plan_f_->execute(pf_.data(), nl_.data(), DFFT_FORWARD, s);
  }
  for (int r = 0; r < P_; ++r) {
    NsStage st{form_, stage, static_cast<const double2*>(nl_[r]), static_cast<double2*>(y_[r]),
               static_cast<double2*>(yc_[r]), static_cast<double2*>(k1_[r]), static_cast<double2*>(k2_[r]),
               static_cast<double2*>(k3_[r])};
    hip_check(ns_rhs_rk(g_, sblk_[r], st, s), "ns_rhs_rk");
  }
}

void NavierStokes::one_step(hipStream_t s) {
  for (int stage = 0; stage < 4; ++stage) rk_stage(stage, s);
}

void NavierStokes::step(int nsteps, bool use_graph, hipStream_t s) {
  TraceRange range("dfft::NavierStokes::step x%d%s", nsteps, use_graph ? " (graph)" : "");
  if (nsteps <= 0) return;
  if (!use_graph) {
    for (int i = 0; i < nsteps; ++i) one_step(s);
    return;
  }
  // Graph replay runs on a private stream ordered after / before the
  // caller's stream (the caller's may be the null stream, which cannot capture).
  if (!graph_exec_) {
    one_step(s);  // eager first step: one-time kernel attribute setup happens outside capture
    --nsteps;
    if (!graph_stream_) hip_check(hipStreamCreate(&graph_stream_), "hipStreamCreate");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    hip_check(hipStreamBeginCapture(graph_stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      one_step(graph_stream_);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(graph_stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hip_check(hipStreamEndCapture(graph_stream_, &graph_), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
  }
  hipEvent_t ev;
  hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventRecord(ev, s), "hipEventRecord");
  hip_check(hipStreamWaitEvent(graph_stream_, ev, 0), "hipStreamWaitEvent");
  for (int i = 0; i < nsteps; ++i) hip_check(hipGraphLaunch(graph_exec_, graph_stream_), "hipGraphLaunch");
  hip_check(hipEventRecord(ev, graph_stream_), "hipEventRecord");
  hip_check(hipStreamWaitEvent(s, ev, 0), "hipStreamWaitEvent");
  (void)hipEventDestroy(ev);
}

void NavierStokes::autotune(hipStream_t s) {
  if (graph_exec_) {  // the captured step holds the old launch shapes
    hip_check(hipStreamSynchronize(graph_stream_), "hipStreamSynchronize");
    (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
  }
  if (graph_) {
    (void)hipGraphDestroy(graph_);
    graph_ = nullptr;
  }
  // scratch only: yc_ and bin_ / nl_ / pb_ / pf_ are rewritten by every stage.
  // Plan::autotune(in, out) takes the plan's input (physical, real) first and
  // its output (spectral, complex) second, whatever direction the solver uses
  // the plan in: each buffer must hold the side it is passed as.
  if (form_ == 0) {
    plan_b_->autotune(pb_.data(), yc_.data(), s);  // 3 fields: pb_ holds >= 3 physical, yc_ 3 spectral
    plan_f_->autotune(pf_.data(), nl_.data(), s);  // 6 fields: pf_ physical, nl_ spectral
    if (plan_f3_) plan_f3_->autotune(pb_.data(), yc_.data(), s);
  } else {
    plan_b_->autotune(pb_.data(), bin_.data(), s);  // 12 fields: pb_ 12 physical, bin_ 12 spectral
    plan_f_->autotune(pf_.data(), nl_.data(), s);   // 3 fields
    if (plan_b3_) plan_b3_->autotune(pb_.data(), yc_.data(), s);
  }
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

void NavierStokes::set_physical(void* const* u3, hipStream_t s) {
  Plan* pf = plan_f3_ ? plan_f3_.get() : plan_f_.get();
  pf->execute(u3, y_.data(), DFFT_FORWARD, s);
  for (int r = 0; r < P_; ++r)
    hip_check(ns_dealias_project(g_, sblk_[r], static_cast<double2*>(y_[r]), s), "ns_dealias_project");
}

void NavierStokes::init_taylor_green(double v0, hipStream_t s) {
  const double h = lx_ * M_PI / n_;
  for (int r = 0; r < P_; ++r)
    hip_check(ns_taylor_green(pblk_[r], h, v0, static_cast<double*>(pb_[r]), s), "ns_taylor_green");
  set_physical(pb_.data(), s);
}

double NavierStokes::energy(hipStream_t s) {
  for (int r = 0; r < P_; ++r)
    hip_check(ns_energy_partials(g_, sblk_[r], static_cast<const double2*>(y_[r]), partials_ + 256 * r, 256, s),
              "ns_energy");
  hip_check(ns_sum(partials_, 256 * P_, esum_, s), "ns_sum");
  double e = 0.0;
  hip_check(hipMemcpyAsync(&e, esum_, sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return e;
}

void NavierStokes::get_state(void* const* dst, hipStream_t s) const {
  for (int r = 0; r < P_; ++r)
    if (spec_pts(r) > 0)
      hip_check(hipMemcpyAsync(dst[r], y_[r], size_t(3 * spec_pts(r)) * 16, hipMemcpyDeviceToDevice, s),
                "hipMemcpyAsync");
}

// The stored state is kept dealiased: the reference's rhs dealiases its
// argument in place first (main.cpp:576-578), so this changes no result.
void NavierStokes::set_state(void* const* src, hipStream_t s) {
  for (int r = 0; r < P_; ++r)
    hip_check(ns_dealias(g_, sblk_[r], static_cast<const double2*>(src[r]), static_cast<double2*>(y_[r]), s),
              "ns_dealias");
}

}  // namespace dfft
