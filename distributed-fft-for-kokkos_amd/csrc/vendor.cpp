#include "vendor.hpp"

#include <rocfft/rocfft.h>

#include "trace.hpp"

#include <algorithm>
#include <mutex>
#include <string>

namespace dfft {

namespace {

void rf_check(rocfft_status st, const char* what) {
  if (st != rocfft_status_success)
    throw Error(DFFT_ERROR_HIP, std::string("rocFFT ") + what + " failed (status " + std::to_string(st) + ")");
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(DFFT_ERROR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

std::size_t volume(const Vec& v) {
  std::size_t n = 1;
  for (auto e : v) n *= e;
  return n;
}

// one rocFFT plan of `type` over the transform lengths (fastest first) and
// `batch` contiguous transforms, scaled by `scale`
rocfft_plan make_plan(rocfft_transform_type type, rocfft_result_placement place, bool fp64, const Vec& lengths,
                      std::size_t batch, double scale) {
  rocfft_plan_description desc = nullptr;
  if (scale != 1.0) {
    rf_check(rocfft_plan_description_create(&desc), "plan description");
    rocfft_status st = rocfft_plan_description_set_scale_factor(desc, scale);
    if (st != rocfft_status_success) {
      (void)rocfft_plan_description_destroy(desc);
      rf_check(st, "scale factor");
    }
  }
  rocfft_plan plan = nullptr;
  rocfft_status st = rocfft_plan_create(&plan, place, type, fp64 ? rocfft_precision_double : rocfft_precision_single,
                                        lengths.size(), lengths.data(), batch, desc);
  if (desc) (void)rocfft_plan_description_destroy(desc);
  rf_check(st, "plan create");
  return plan;
}

}  // namespace

VendorFFT::VendorFFT(const Vec& gin, const Vec& gout, const Vec& axes, bool real, bool fp64, double fwd_scale,
                     double bwd_scale)
    : gin_(gin), gout_(gout), real_(real), fp64_(fp64) {
  static std::once_flag setup;
  std::call_once(setup, [] { (void)rocfft_setup(); });
  const std::size_t rank = gin.size();
  // storage order of the global arrays: the untransformed axes (the batch)
  // outermost, then the transform axes in the order given, so the R2C axis
  // is the contiguous one (rocFFT transforms the innermost dimensions)
  for (std::size_t a = 0; a < rank; ++a)
    if (std::find(axes.begin(), axes.end(), a) == axes.end()) map_.push_back(a);
  for (auto a : axes) map_.push_back(a);
  Vec lengths;  // fastest first, real lengths
  std::size_t batch = 1;
  for (std::size_t k = rank; k-- > 0;) {
    if (k >= rank - axes.size()) lengths.push_back(std::max(gin[map_[k]], gout[map_[k]]));
    else batch *= gin[map_[k]];
  }
  const std::size_t cs = fp64 ? 16 : 8;
  try {
    if (volume(gout) > 0) {
      hip_check(hipMalloc(&gout_buf_, volume(gout) * cs), "hipMalloc of the vendor plan's spectrum");
      bytes_ += volume(gout) * cs;
      if (real) {
        hip_check(hipMalloc(&gin_buf_, volume(gin) * cs / 2), "hipMalloc of the vendor plan's real array");
        bytes_ += volume(gin) * cs / 2;
      } else {
        gin_buf_ = gout_buf_;
      }
      rocfft_plan f, b;
      if (real) {
        f = make_plan(rocfft_transform_type_real_forward, rocfft_placement_notinplace, fp64, lengths, batch,
                      fwd_scale);
        plan_fwd_ = f;
        b = make_plan(rocfft_transform_type_real_inverse, rocfft_placement_notinplace, fp64, lengths, batch,
                      bwd_scale);
        plan_bwd_ = b;
      } else {
        f = make_plan(rocfft_transform_type_complex_forward, rocfft_placement_inplace, fp64, lengths, batch,
                      fwd_scale);
        plan_fwd_ = f;
        b = make_plan(rocfft_transform_type_complex_inverse, rocfft_placement_inplace, fp64, lengths, batch,
                      bwd_scale);
        plan_bwd_ = b;
      }
      std::size_t wf = 0, wb = 0;
      rf_check(rocfft_plan_get_work_buffer_size(f, &wf), "work buffer size");
      rf_check(rocfft_plan_get_work_buffer_size(b, &wb), "work buffer size");
      work_bytes_ = std::max(wf, wb);
      if (work_bytes_ > 0) {
        hip_check(hipMalloc(&work_, work_bytes_), "hipMalloc of the rocFFT work buffer");
        bytes_ += work_bytes_;
      }
      rocfft_execution_info info = nullptr;
      rf_check(rocfft_execution_info_create(&info), "execution info");
      info_ = info;
      if (work_bytes_ > 0) rf_check(rocfft_execution_info_set_work_buffer(info, work_, work_bytes_), "work buffer");
    }
  } catch (...) {
    release();
    throw;
  }
}

void VendorFFT::release() {
  if (info_) (void)rocfft_execution_info_destroy(static_cast<rocfft_execution_info>(info_));
  if (plan_fwd_) (void)rocfft_plan_destroy(static_cast<rocfft_plan>(plan_fwd_));
  if (plan_bwd_) (void)rocfft_plan_destroy(static_cast<rocfft_plan>(plan_bwd_));
  if (work_) (void)hipFree(work_);
  if (gin_buf_ && gin_buf_ != gout_buf_) (void)hipFree(gin_buf_);
  if (gout_buf_) (void)hipFree(gout_buf_);
  info_ = plan_fwd_ = plan_bwd_ = work_ = gin_buf_ = gout_buf_ = nullptr;
}

VendorFFT::~VendorFFT() { release(); }

DDesc VendorFFT::global_desc(const Vec& ext, void* p) const {
  DDesc d{};
  d.ndim = static_cast<int32_t>(ext.size());
  d.layout_left = 0;
  for (std::size_t a = 0; a < ext.size(); ++a) {
    d.ext[a] = static_cast<int32_t>(ext[a]);
    d.ntopo[a] = 1;
    d.rstride[a] = 0;
    d.base[a] = static_cast<int32_t>(ext[a]);
    d.rem[a] = 0;
    d.map[a] = static_cast<int32_t>(map_[a]);
    d.div_big[a] = FastDiv(static_cast<uint32_t>(ext[a] + 1));
    d.div_small[a] = FastDiv(static_cast<uint32_t>(std::max<std::size_t>(ext[a], 1)));
  }
  d.pitch = 0;
  d.ptr[0] = static_cast<char*>(p);
  return d;
}

void VendorFFT::run(const DDesc& src, const DDesc& dst, bool forward, hipStream_t s,
                    const std::function<void(hipStream_t)>& mark) {
  if (!info_) return;  // empty transform
  const std::size_t cs = fp64_ ? 16 : 8;
  // element sizes of the two sides of this direction
  const int in_es = static_cast<int>(forward && real_ ? cs / 2 : cs);
  const int out_es = static_cast<int>(!forward && real_ ? cs / 2 : cs);
  void* gsrc = forward ? gin_buf_ : gout_buf_;
  void* gdst = forward ? gout_buf_ : gin_buf_;
  auto remap = [&](const DDesc& from, const DDesc& to, int es) {
    RemapArgs ra{};
    ra.src = from;
    ra.dst = to;
    ra.total = 1;
    for (int k = 0; k < to.ndim; ++k) {
      // the destination's storage order, slowest first
      const int a = to.layout_left ? to.map[to.ndim - 1 - k] : to.map[k];
      ra.iter_axis[k] = a;
      ra.iter_ext[k] = to.ext[a];
      ra.iter_div[k] = FastDiv(static_cast<uint32_t>(std::max(to.ext[a], 1)));
      ra.total *= to.ext[a];
    }
    mark(s);
    hip_check(launch_remap(ra, es, s), "remap");
    mark(s);
  };
  // gather the caller's blocks into the global array
  remap(src, global_desc(forward ? gin_ : gout_, gsrc), in_es);
  // one vendor call over the whole array
  rocfft_execution_info info = static_cast<rocfft_execution_info>(info_);
  rf_check(rocfft_execution_info_set_stream(info, s), "set stream");
  void* ibuf[1] = {gsrc};
  void* obuf[1] = {gdst};
  {
    TraceRange range("dfft::rocFFT execute");
    mark(s);
    rf_check(rocfft_execute(static_cast<rocfft_plan>(forward ? plan_fwd_ : plan_bwd_), ibuf,
                            gsrc == gdst ? nullptr : obuf, info),
             "execute");
    mark(s);
  }
  // scatter to the output ranks
  remap(global_desc(forward ? gout_ : gin_, gdst), dst, out_es);
}

}  // namespace dfft
