// Plan runtime: the one-GPU replacement of the reference's SlabPlan /
// PencilPlan / SharedPlan behind the InternalPlan boundary
// (KokkosFFT_Distributed_InternalPlan.hpp:22-143).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "dfft.h"
#include "fft_kernels.hpp"
#include "planner.hpp"
#include "vendor.hpp"

namespace dfft {

struct PassSpec {
  int axis;
  int kind;  // LineKind
  int n;     // transform length
  // launch shape (0 / -1 = launcher heuristic); set by options or the autotuner
  int lines_per_block = 0;
  int values_per_thread = 0;
  int lines_fast = -1;
  int nontemporal = -1;  // -1: on for arrays of at least 1 GiB
  int paired = 1;        // R2C / C2R: two real lines per complex FFT
  int xcd = 0;           // XCD-contiguous block numbering
  int swz = 0;           // column-major walk of the block grid
  int split = 1;         // exchanges in real / imaginary halves (half the LDS; FP64, FP32 n >= 1024)
  int dual = 1;          // FP32 C2C: two adjacent lines per lane (16-byte accesses) where they apply
  int lin = 1;           // one-line C2C passes into odd unpadded rows: lines enumerated over real elements
};

// device table exp(-2 pi i e / n), e in [0, n): evaluated in long double,
// rounded once to FP64 (both precisions form butterfly powers from it in FP64)
void* upload_twiddles(int n);

class Plan {
 public:
  explicit Plan(const dfft_desc& d);
  ~Plan();
  Plan(const Plan&) = delete;
  Plan& operator=(const Plan&) = delete;

  void execute(void* const* in, void* const* out, int direction, hipStream_t stream);
  // throws DFFT_ERROR_MISMATCH unless every rank's in/out allocation can hold
  // the local array that execute(direction) reads/writes (checked at the C
  // boundary; internal callers pass the plan's own buffers)
  void check_buffers(void* const* in, void* const* out, int direction) const;
  // adjoint of execute(direction) under the real inner product Re sum conj(a) b
  // of each side: in has the shape of that direction's output, out of its input
  void execute_adjoint(void* const* in, void* const* out, int direction, hipStream_t stream);
  // Forward execute of a 6-field R2C plan over 4-D views (field axis first,
  // not distributed, transform axes 1..3) whose input fields are the products
  // u_i u_j of the 3 velocity fields in `u3` (per rank: 3 fields shaped like
  // the plan's input blocks), (i, j) = (0,0) (0,1) (0,2) (1,1) (1,2) (2,2):
  // the products are formed as the first (R2C) pass loads its lines, so the
  // 6 product fields are never stored.  products_supported() says whether
  // this plan can (FP64, fused executor, a paired R2C first pass of a
  // power-of-two length 8..1024).
  bool products_supported() const;
  void execute_forward_products(void* const* u3, void* const* out, hipStream_t stream);
  // execute with a HIP event pair around every launch; durations in ms
  std::vector<float> execute_timed(void* const* in, void* const* out, int direction, hipStream_t stream);
  int num_passes(int direction) const;
  // tuning knobs: "chunk" (planes per chunk of the first two passes, 0 = off),
  // "pad" (scratch rows padded to 128-byte multiples: -1 auto, 0 off, 1 on),
  // "fwd.<i>.lpb|vpt|lf|nt|pair|xcd|swz|split|dual|lin", "bwd.<i>.…" (lines
  // per block, values per thread, lanes-across-lines, nontemporal accesses,
  // two real lines per complex FFT in an R2C / C2R pass, XCD-contiguous
  // blocks, column-major block walk, LDS exchanges in real/imaginary halves,
  // two adjacent FP32 lines per lane, one-line passes into odd unpadded rows
  // enumerated over the real elements) of pass i
  void set_option(const std::string& key, long value);
  long get_option(const std::string& key) const;
  // measure candidate launch shapes and chunk sizes on the caller's arrays
  // (their contents are overwritten) and keep the fastest
  void autotune(void* const* in, void* const* out, hipStream_t stream);

  const std::string& name() const { return name_; }
  std::size_t nranks() const { return P_; }
  const Vec& in_local(std::size_t r) const { return in_local_.at(r); }
  const Vec& out_local(std::size_t r) const { return out_local_.at(r); }
  void local_starts(std::size_t r, int which, Vec& ext, Vec& st) const;
  const std::vector<BlockAnalyses>& analyses() const { return analyses_; }
  std::size_t max_buffer_size() const { return max_buffer_size_; }
  std::size_t workspace_bytes() const { return workspace_bytes_; }

 private:
  void allocate();
  void release();
  DDesc user_desc(const Vec& ext, const Vec& topo, Layout tl, void* const* ptrs) const;
  // the plan scratch holding an array of extents `ext` with elements of
  // `elem_bytes` bytes; LayoutRight rows are padded to 128-byte multiples
  DDesc scratch_desc(const Vec& ext, void* p, std::size_t elem_bytes) const;
  std::size_t scratch_elems(const Vec& ext, std::size_t elem_bytes) const;  // incl. padding
  LinePass make_pass(const PassSpec& s, const DDesc& in, const DDesc& out, bool inverse, double scale,
                     int slow_axis = -1) const;
  float time_direction(void* const* in, void* const* out, bool forward, hipStream_t s, int reps);
  void plan_fused();
  void run_fused(void* const* in, void* const* out, bool forward, hipStream_t s);
  void run_blockwalk(void* const* in, void* const* out, bool forward, hipStream_t s);
  const void* twiddles(int n) const;
  double norm_scale(bool forward) const;
  double exec_scale(bool forward) const { return adj_scale_ >= 0.0 ? adj_scale_ : norm_scale(forward); }

  dfft_desc d_{};
  TopologyType kind_ = TopologyType::Shared;
  std::string name_;
  std::size_t P_ = 1;
  std::size_t rank_ = 3;
  Vec gin_, gout_, axes_, in_topo_, out_topo_, fft_extents_;
  Layout in_tl_ = Layout::Right, out_tl_ = Layout::Right, view_layout_ = Layout::Right;
  bool real_ = false, fp64_ = true;
  int device_ = 0;
  std::vector<Vec> in_local_, out_local_;
  std::vector<BlockAnalyses> analyses_;
  std::size_t max_buffer_size_ = 0;
  void* scratch_ = nullptr;
  // lines longer than one LDS buffer: complex scratch of the two-launch
  // transform (nlines * n values of a pass)
  void* long_scratch_ = nullptr;
  std::size_t long_bytes_ = 0;
  // lengths with a prime factor beyond one LDS line: Bluestein through two
  // m-point transforms of an (nlines, m) scratch pair
  struct Bluestein {
    int m = 0, log2m = 0;
    void* chirp = nullptr;  // exp(-pi i e / n), e in [0, 2n), FP64
    void* filt = nullptr;   // m-point transform of the chirp filter / m, FP64
  };
  std::map<int, Bluestein> blue_;
  void* blue_s_[2] = {nullptr, nullptr};
  std::size_t blue_bytes_ = 0;
  void setup_bluestein(int n, std::size_t max_lines);
  LinePass dense_pass(int rows, int m, void* src, void* dst, bool inverse, bool fp64) const;
  // a line pass through the kernel its length needs (Bluestein included)
  hipError_t launch_pass(const LinePass& p, hipStream_t s);
  std::vector<void*> walk_bufs_[2];  // block-walk ping-pong local arrays per rank
  std::unique_ptr<VendorFFT> vendor_;  // DFFT_EXEC_VENDOR: rocFFT over the gathered global array
  std::size_t workspace_bytes_ = 0;
  std::map<int, void*> tw_;
  std::vector<PassSpec> fwd_, bwd_;
  int chunk_[2] = {0, 0};  // planes per chunk, forward / backward (0 = no chunking)
  int pad_ = -1;           // scratch row padding: -1 auto (<= 1/16 extra), 0 off, 1 on
  // set only while execute_adjoint runs: scale of the adjoint direction and
  // the interior half-spectrum weight of its R2C / C2R line pass
  double adj_scale_ = -1.0;
  double adj_herm_ = 0.0;
  // timing hooks (set only by execute_timed)
  std::vector<hipEvent_t>* timing_ = nullptr;
  void mark(hipStream_t s);
};

}  // namespace dfft
