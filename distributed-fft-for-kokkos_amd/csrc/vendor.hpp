// rocFFT behind the plan boundary (execution mode DFFT_EXEC_VENDOR): the
// one-GPU form of the reference's TplPlan (KokkosFFT_Distributed_TplPlan.hpp:
// 25-181), which hands a whole distributed transform to a vendor library
// (rocFFT-MPI, RocfftMPI_Plan.hpp:15-62).  Here the P virtual ranks' blocks
// are gathered into one global array (a remap), rocFFT transforms it in one
// call and a second remap scatters the result to the output ranks.  A
// comparator to run and profile beside the fused executor, not a fast path:
// it moves the array through HBM twice more than the fused passes.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>

#include "dfft.h"
#include "fft_kernels.hpp"
#include "planner.hpp"

namespace dfft {

class VendorFFT {
 public:
  // gin / gout: global extents; axes: transform axes (R2C along the last);
  // fwd_scale / bwd_scale: the plan's normalisation of each direction
  VendorFFT(const Vec& gin, const Vec& gout, const Vec& axes, bool real, bool fp64, double fwd_scale,
            double bwd_scale);
  ~VendorFFT();
  VendorFFT(const VendorFFT&) = delete;
  VendorFFT& operator=(const VendorFFT&) = delete;

  // src: the caller's input distribution, dst: the output distribution of
  // this direction; `mark` brackets each of the three steps (timing hooks)
  void run(const DDesc& src, const DDesc& dst, bool forward, hipStream_t s,
           const std::function<void(hipStream_t)>& mark);
  std::size_t workspace_bytes() const { return bytes_; }

 private:
  DDesc global_desc(const Vec& ext, void* p) const;
  void release();

  Vec gin_, gout_, map_;
  bool real_, fp64_;
  void* gin_buf_ = nullptr;   // global input-shaped array (R2C: real)
  void* gout_buf_ = nullptr;  // global spectrum (C2C: the same buffer)
  void* work_ = nullptr;
  std::size_t work_bytes_ = 0, bytes_ = 0;
  void* plan_fwd_ = nullptr;  // rocfft_plan
  void* plan_bwd_ = nullptr;
  void* info_ = nullptr;      // rocfft_execution_info
};

}  // namespace dfft
