// Tangent-linear and adjoint of the divergence-form Navier-Stokes right-hand
// side and of the RK4 step, and the checkpointed gradient of the final
// kinetic energy with respect to the initial physical velocity.  No
// counterpart in the reference (its NS example runs forward only).
//
// Inner products: spectral fields use <a,b>_h = sum_k h_k Re(conj(a_k) b_k)
// with h = 1 on the kz = 0 and kz = n/2 planes and 2 elsewhere, physical
// fields the plain sum; then FFT^* = n^3 IFFT and IFFT^* = FFT / n^3, the
// dealias mask D, the projection P and nu k^2 are self-adjoint and i k_j is
// anti-self-adjoint.  The right-hand side dealiases its argument first, as
// the reference's rhs does in place (main.cpp:576-578).  With
// N(z)_i = i k_j FFT(u_i u_j) and u = IFFT(D z):
//   R(z)      = P(-D N(z) - nu k^2 D z)
//   R'(z) d   = P(-D i k_j FFT(du_i u_j + u_i du_j) - nu k^2 D d),   du = IFFT(D d)
//   R'(z)^* l = -D FFT(c) - nu k^2 mu,  mu = D P l,
//               c_m = sum_j u_j S_mj,  S_mj = IFFT(-i (k_j mu_m + k_m mu_j)).
#include <algorithm>
#include <cmath>

#include "navier_stokes.hpp"

namespace dfft {

namespace {
void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(DFFT_ERROR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

void NavierStokes::ensure_aux() {
  if (!aux_s3a_.empty()) return;
  aux_s3a_ = alloc_set(true, 3);
  aux_s3b_ = alloc_set(true, 3);
  aux_p3_ = alloc_set(false, 3);
}

void NavierStokes::apply_rhs(void* const* y3, void* const* out3, hipStream_t s) {
  if (form_ != 0) throw Error(DFFT_ERROR_UNSUPPORTED, "apply_rhs needs the divergence form");
  ensure_aux();
  std::vector<void*>& ym = aux_s3a_;
  for (int r = 0; r < P_; ++r)
    hip_check(ns_dealias(g_, sblk_[r], static_cast<const double2*>(y3[r]), static_cast<double2*>(ym[r]), s),
              "dealias");
  plan_b_->execute(ym.data(), pb_.data(), DFFT_BACKWARD, s);
  for (int r = 0; r < P_; ++r)
    hip_check(ns_products(static_cast<double*>(pb_[r]), static_cast<double*>(pf_[r]), phys_pts(r), s), "products");
  plan_f_->execute(pf_.data(), nl_.data(), DFFT_FORWARD, s);
  for (int r = 0; r < P_; ++r)
    hip_check(ns_rhs_only(g_, sblk_[r], static_cast<const double2*>(nl_[r]), static_cast<const double2*>(ym[r]),
                          static_cast<double2*>(out3[r]), s),
              "rhs_only");
}

void NavierStokes::apply_rhs_tangent(void* const* y3, void* const* d3, void* const* out3, hipStream_t s) {
  if (form_ != 0) throw Error(DFFT_ERROR_UNSUPPORTED, "tangent needs the divergence form");
  ensure_aux();
  std::vector<void*>& dm = aux_s3a_;
  std::vector<void*>& ym = aux_s3b_;
  std::vector<void*>& du = aux_p3_;
  for (int r = 0; r < P_; ++r) {
    hip_check(ns_dealias(g_, sblk_[r], static_cast<const double2*>(d3[r]), static_cast<double2*>(dm[r]), s),
              "dealias");
    hip_check(ns_dealias(g_, sblk_[r], static_cast<const double2*>(y3[r]), static_cast<double2*>(ym[r]), s),
              "dealias");
  }
  plan_b_->execute(ym.data(), pb_.data(), DFFT_BACKWARD, s);
  plan_b_->execute(dm.data(), du.data(), DFFT_BACKWARD, s);
  for (int r = 0; r < P_; ++r)
    hip_check(ns_products_tangent(static_cast<double*>(pb_[r]), static_cast<double*>(du[r]),
                                  static_cast<double*>(pf_[r]), phys_pts(r), s),
              "products_tangent");
  plan_f_->execute(pf_.data(), nl_.data(), DFFT_FORWARD, s);
  for (int r = 0; r < P_; ++r)
    hip_check(ns_rhs_only(g_, sblk_[r], static_cast<const double2*>(nl_[r]), static_cast<const double2*>(dm[r]),
                          static_cast<double2*>(out3[r]), s),
              "rhs_only");
}

void NavierStokes::apply_rhs_adjoint(void* const* y3, void* const* l3, void* const* out3, hipStream_t s) {
  if (form_ != 0) throw Error(DFFT_ERROR_UNSUPPORTED, "adjoint needs the divergence form");
  ensure_aux();
  std::vector<void*>& mu = aux_s3a_;
  std::vector<void*>& ch = aux_s3b_;
  std::vector<void*>& c = aux_p3_;
  for (int r = 0; r < P_; ++r)
    hip_check(ns_adjoint_prepare(g_, sblk_[r], static_cast<const double2*>(l3[r]), static_cast<double2*>(mu[r]),
                                 static_cast<double2*>(nl_[r]), s),
              "adjoint_prepare");
  plan_f_->execute(nl_.data(), pf_.data(), DFFT_BACKWARD, s);  // S_mj (6 fields)
  for (int r = 0; r < P_; ++r)  // D y into ch, which is free until the forward transform of c
    hip_check(ns_dealias(g_, sblk_[r], static_cast<const double2*>(y3[r]), static_cast<double2*>(ch[r]), s),
              "dealias");
  plan_b_->execute(ch.data(), pb_.data(), DFFT_BACKWARD, s);  // u = IFFT(D y)
  for (int r = 0; r < P_; ++r)
    hip_check(ns_adjoint_contract(static_cast<double*>(pb_[r]), static_cast<double*>(pf_[r]),
                                  static_cast<double*>(c[r]), phys_pts(r), s),
              "adjoint_contract");
  plan_f3_->execute(c.data(), ch.data(), DFFT_FORWARD, s);
  for (int r = 0; r < P_; ++r)
    hip_check(ns_adjoint_finish(g_, sblk_[r], static_cast<const double2*>(ch[r]), static_cast<const double2*>(mu[r]),
                                static_cast<double2*>(out3[r]), s),
              "adjoint_finish");
}

// J(u0) = kinetic energy after `nsteps` RK4 steps from x0 = D P FFT(u0);
// writes dJ/du0 (plain inner product on the physical grid) to `grad`.
// Checkpoints every `ck` steps; each segment is recomputed once, keeping its
// RK stage values, and then reversed.
double NavierStokes::gradient(int nsteps, int ck, void* const* u0, void* const* grad, hipStream_t s) {
  if (form_ != 0) throw Error(DFFT_ERROR_UNSUPPORTED, "gradient needs the divergence form");
  if (nsteps < 0 || ck < 1) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "bad step counts");
  const double dt = g_.dt;
  std::vector<std::vector<void*>> owned;
  auto fresh = [&](bool spectral, int nf) {
    owned.push_back(alloc_set(spectral, nf));
    return owned.back();
  };
  auto copy3 = [&](const std::vector<void*>& dst, const std::vector<void*>& src) {
    for (int r = 0; r < P_; ++r)
      if (spec_pts(r) > 0)
        hip_check(hipMemcpyAsync(dst[r], src[r], size_t(3 * spec_pts(r)) * 16, hipMemcpyDeviceToDevice, s), "copy");
  };
  auto axpby = [&](double a, const std::vector<void*>& x, double b, const std::vector<void*>& y) {
    for (int r = 0; r < P_; ++r)
      hip_check(ns_axpby(3 * spec_pts(r), a, static_cast<double2*>(x[r]), b, static_cast<const double2*>(y[r]), s),
                "axpby");
  };
  double J = 0.0;
  try {
    set_physical(u0, s);
    const int nseg = (nsteps + ck - 1) / ck;
    std::vector<std::vector<void*>> ckpt;
    for (int k = 0; k < nseg; ++k) ckpt.push_back(fresh(true, 3));
    for (int st = 0; st < nsteps; ++st) {
      if (st % ck == 0) copy3(ckpt[st / ck], y_);
      one_step(s);
    }
    J = energy(s);
    std::vector<void*> lam = fresh(true, 3);
    copy3(lam, y_);
    axpby(g_.coef, lam, 0.0, lam);
    std::vector<std::vector<void*>> stage_states;
    for (int k = 0; k < 4 * std::min(ck, std::max(nsteps, 1)); ++k) stage_states.push_back(fresh(true, 3));
    std::vector<void*> fb[4] = {fresh(true, 3), fresh(true, 3), fresh(true, 3), fresh(true, 3)};
    std::vector<void*> yb = fresh(true, 3), t = fresh(true, 3);
    for (int seg = nseg - 1; seg >= 0; --seg) {
      const int first = seg * ck, last = std::min(nsteps, first + ck);
      copy3(y_, ckpt[seg]);
      for (int st = first; st < last; ++st) {
        for (int stage = 0; stage < 4; ++stage) {
          copy3(stage_states[4 * (st - first) + stage], y_);
          rk_stage(stage, s);
        }
      }
      for (int st = last - 1; st >= first; --st) {
        auto& y0 = stage_states[4 * (st - first) + 0];
        auto& y1 = stage_states[4 * (st - first) + 1];
        auto& y2 = stage_states[4 * (st - first) + 2];
        auto& y3 = stage_states[4 * (st - first) + 3];
        // seed: fb0 = dt/6 l, fb1 = dt/3 l, fb2 = dt/3 l, fb3 = dt/6 l, yb = l
        copy3(fb[0], lam);
        axpby(dt / 6.0, fb[0], 0.0, lam);
        copy3(fb[1], lam);
        axpby(dt / 3.0, fb[1], 0.0, lam);
        copy3(fb[2], lam);
        axpby(dt / 3.0, fb[2], 0.0, lam);
        copy3(fb[3], lam);
        axpby(dt / 6.0, fb[3], 0.0, lam);
        copy3(yb, lam);
        apply_rhs_adjoint(y3.data(), fb[3].data(), t.data(), s);
        axpby(1.0, yb, 1.0, t);
        axpby(1.0, fb[2], dt, t);
        apply_rhs_adjoint(y2.data(), fb[2].data(), t.data(), s);
        axpby(1.0, yb, 1.0, t);
        axpby(1.0, fb[1], dt / 2.0, t);
        apply_rhs_adjoint(y1.data(), fb[1].data(), t.data(), s);
        axpby(1.0, yb, 1.0, t);
        axpby(1.0, fb[0], dt / 2.0, t);
        apply_rhs_adjoint(y0.data(), fb[0].data(), t.data(), s);
        axpby(1.0, yb, 1.0, t);
        // the stage-0 input is dealiased in place (main.cpp:576-578)
        for (int r = 0; r < P_; ++r)
          hip_check(ns_dealias(g_, sblk_[r], static_cast<const double2*>(yb[r]), static_cast<double2*>(yb[r]), s),
                    "dealias");
        copy3(lam, yb);
      }
    }
    // x0 = D P FFT(u0)  =>  dJ/du0 = n^3 IFFT(D P l0)
    for (int r = 0; r < P_; ++r)
      hip_check(ns_dealias_project(g_, sblk_[r], static_cast<double2*>(lam[r]), s), "dealias_project");
    Plan* pb3 = plan_b3_ ? plan_b3_.get() : plan_b_.get();
    pb3->execute(lam.data(), grad, DFFT_BACKWARD, s);
    const double n3 = std::pow(static_cast<double>(n_), 3);
    for (int r = 0; r < P_; ++r) hip_check(ns_scale(3 * phys_pts(r), n3, static_cast<double*>(grad[r]), s), "scale");
    hip_check(hipStreamSynchronize(s), "sync");
  } catch (...) {
    (void)hipStreamSynchronize(s);
    for (auto& v : owned)
      for (void* p : v)
        if (p) {
          all_.erase(std::find(all_.begin(), all_.end(), p));
          (void)hipFree(p);
        }
    throw;
  }
  for (auto& v : owned)
    for (void* p : v)
      if (p) {
        all_.erase(std::find(all_.begin(), all_.end(), p));
        (void)hipFree(p);
      }
  return J;
}

}  // namespace dfft
