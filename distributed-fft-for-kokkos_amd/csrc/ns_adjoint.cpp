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
  forward_products(pb_.data(), s);
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
  AdjFinish f;
  f.mode = 0;
  f.out3 = out3;
  adjoint_core(y3, l3, nullptr, f, s);
}

void NavierStokes::adjoint_core(const void* const* y3, const void* const* l3, const void* const* u3,
                                const AdjFinish& f, hipStream_t s) {
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
  if (!u3) {
    for (int r = 0; r < P_; ++r)  // D y into ch, which is free until the forward transform of c
      hip_check(ns_dealias(g_, sblk_[r], static_cast<const double2*>(y3[r]), static_cast<double2*>(ch[r]), s),
                "dealias");
    plan_b_->execute(ch.data(), pb_.data(), DFFT_BACKWARD, s);  // u = IFFT(D y)
    u3 = pb_.data();
  }
  for (int r = 0; r < P_; ++r)
    hip_check(ns_adjoint_contract(static_cast<const double*>(u3[r]), static_cast<double*>(pf_[r]),
                                  static_cast<double*>(c[r]), phys_pts(r), s),
              "adjoint_contract");
  plan_f3_->execute(c.data(), ch.data(), DFFT_FORWARD, s);
  for (int r = 0; r < P_; ++r) {
    const auto* chr = static_cast<const double2*>(ch[r]);
    const auto* mur = static_cast<const double2*>(mu[r]);
    if (f.mode == 0)
      hip_check(ns_adjoint_finish(g_, sblk_[r], chr, mur, static_cast<double2*>(f.out3[r]), s), "adjoint_finish");
    else
      hip_check(ns_adjoint_finish_rk(g_, sblk_[r], chr, mur, static_cast<double2*>(f.acc[r]),
                                     static_cast<double2*>(f.lam[r]), f.fb ? static_cast<double2*>(f.fb[r]) : nullptr,
                                     f.a, f.c, f.mode, s),
                "adjoint_finish_rk");
  }
}

// J(u0) = kinetic energy after `nsteps` RK4 steps from x0 = D P FFT(u0);
// writes dJ/du0 (plain inner product on the physical grid) to `grad`.
// Checkpoints every `ck` steps.  A segment's forward recomputation keeps the
// physical velocity u = IFFT(y) of every RK stage (the only state the
// adjoint right-hand side needs); the last segment keeps them during the
// first forward sweep and is not recomputed.
//
// Reverse step (x the step input, x_k the stage inputs, yc = D x):
//   k_k = dt R(x_k), x_1 = yc + k_1/2, x_2 = yc + k_2/2, x_3 = yc + k_3,
//   x_new = yc + (k_1 + 2 k_2 + 2 k_3 + k_4)/6
// gives, with t_k = R'(x_k)^* of the stage seed fb_k:
//   fb_3 = dt lam/6,              t_3 = R'(x_3)^* fb_3
//   fb_2 = dt (lam/3 + t_3),      t_2 = R'(x_2)^* fb_2
//   fb_1 = dt (lam/3 + t_2/2),    t_1 = R'(x_1)^* fb_1
//   fb_0 = dt (lam/6 + t_1/2),    t_0 = R'(x_0)^* fb_0
//   lam_new = D (lam + t_3 + t_2 + t_1 + t_0).
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
  double J = 0.0;
  try {
    set_physical(u0, s);
    const int nseg = (nsteps + ck - 1) / ck;
    const int seg_len = std::min(ck, std::max(nsteps, 1));
    std::vector<std::vector<void*>> ckpt;
    for (int k = 0; k < nseg; ++k) ckpt.push_back(fresh(true, 3));
    std::vector<std::vector<void*>> stage_u;  // 3 physical fields per stage of a segment
    for (int k = 0; k < 4 * seg_len; ++k) stage_u.push_back(fresh(false, 3));
    const int last_first = nseg > 0 ? (nseg - 1) * ck : 0;
    for (int st = 0; st < nsteps; ++st) {
      if (st % ck == 0) copy3(ckpt[st / ck], y_);
      if (st >= last_first) {
        for (int stage = 0; stage < 4; ++stage) rk_stage(stage, s, stage_u[4 * (st - last_first) + stage].data());
      } else {
        one_step(s);
      }
    }
    J = energy(s);
    std::vector<void*> lam = fresh(true, 3), acc = fresh(true, 3), fb = fresh(true, 3);
    copy3(lam, y_);
    for (int r = 0; r < P_; ++r)
      hip_check(ns_axpby(3 * spec_pts(r), g_.coef, static_cast<double2*>(lam[r]), 0.0,
                         static_cast<const double2*>(lam[r]), s),
                "axpby");
    AdjFinish f;
    f.acc = acc.data();
    f.lam = lam.data();
    f.fb = fb.data();
    for (int seg = nseg - 1; seg >= 0; --seg) {
      const int first = seg * ck, last = std::min(nsteps, first + ck);
      if (seg != nseg - 1) {
        copy3(y_, ckpt[seg]);
        for (int st = first; st < last; ++st)
          for (int stage = 0; stage < 4; ++stage) rk_stage(stage, s, stage_u[4 * (st - first) + stage].data());
      }
      for (int st = last - 1; st >= first; --st) {
        const int base = 4 * (st - first);
        for (int r = 0; r < P_; ++r)
          hip_check(ns_rk_adjoint_init(3 * spec_pts(r), static_cast<const double2*>(lam[r]),
                                       static_cast<double2*>(fb[r]), static_cast<double2*>(acc[r]), dt / 6.0, s),
                    "rk_adjoint_init");
        const double a[3] = {1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0}, c[3] = {1.0, 0.5, 0.5};
        for (int k = 3; k >= 1; --k) {
          f.mode = 1;
          f.a = a[3 - k];
          f.c = c[3 - k];
          adjoint_core(nullptr, fb.data(), stage_u[base + k].data(), f, s);
        }
        f.mode = 2;
        adjoint_core(nullptr, fb.data(), stage_u[base].data(), f, s);
      }
    }
    // x0 = D P FFT(u0)  =>  dJ/du0 = n^3 IFFT(D P lam_0)
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
