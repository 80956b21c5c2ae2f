// Plan runtime.  Two strategies execute the same plan:
//
//  * fused (default): the transform is a sequence of batched 1-D line passes
//    over the global index space.  The first pass gathers lines straight out
//    of the P input-rank buffers, the last pass scatters them into the P
//    output-rank buffers, and the passes in between work on one plan-owned
//    global array.  The reference's pack -> MPI_Alltoall -> unpack of every
//    transpose block (TransBlock.hpp:76-107) therefore costs no extra sweep:
//    it is the addressing of the first and last pass.
//  * blockwalk: the reference's block sequence (SlabPlan.hpp:1143-1267,
//    PencilPlan.hpp:1444-1598): FFT blocks on each rank's mapped local array
//    and one exchange remap per transpose block.
#include "plan.hpp"
#include "trace.hpp"

#include "buffer_check.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace dfft {

namespace {
void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(DFFT_ERROR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

Vec to_vec(const size_t* a, std::size_t n) { return Vec(a, a + n); }

bool contains(const Vec& v, std::size_t x) { return std::find(v.begin(), v.end(), x) != v.end(); }
}  // namespace

void* upload_twiddles(int n) {
  const std::size_t es = 16;
  std::vector<unsigned char> host(es * static_cast<std::size_t>(n));
  for (int e = 0; e < n; ++e) {
    const long double a = 6.283185307179586476925286766559005768L * e / n;
    const long double c = cosl(a), s = -sinl(a);
    double v[2] = {static_cast<double>(c), static_cast<double>(s)};
    std::memcpy(&host[es * e], v, es);
  }
  void* p = nullptr;
  hipError_t err = hipMalloc(&p, std::max<std::size_t>(host.size(), 16));
  if (err != hipSuccess) throw Error(DFFT_ERROR_OUT_OF_MEMORY, "hipMalloc of twiddle table failed");
  err = hipMemcpy(p, host.data(), host.size(), hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    (void)hipFree(p);
    throw Error(DFFT_ERROR_HIP, std::string("twiddle upload: ") + hipGetErrorString(err));
  }
  return p;
}

Plan::Plan(const dfft_desc& d) : d_(d) {
  if (d.rank < 2 || d.rank > DFFT_MAX_RANK)
    throw Error(DFFT_ERROR_UNSUPPORTED, "view rank must be in [2, 4]");
  if (d.fft_rank < 1 || d.fft_rank > 3 || d.fft_rank > d.rank)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "Plan: the Rank of FFT axes must be between 1 and 3");
  if (d.transform != DFFT_C2C && d.transform != DFFT_R2C)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "transform must be DFFT_C2C or DFFT_R2C");
  if (d.precision != DFFT_FP32 && d.precision != DFFT_FP64)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "precision must be DFFT_FP32 or DFFT_FP64");
  if (d.norm < DFFT_NORM_BACKWARD || d.norm > DFFT_NORM_NONE)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "invalid normalization");
  if (d.exec_mode != DFFT_EXEC_FUSED && d.exec_mode != DFFT_EXEC_BLOCKWALK && d.exec_mode != DFFT_EXEC_VENDOR)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "invalid execution mode");
  rank_ = static_cast<std::size_t>(d.rank);
  real_ = d.transform == DFFT_R2C;
  fp64_ = d.precision == DFFT_FP64;
  view_layout_ = d.layout == DFFT_LAYOUT_LEFT ? Layout::Left : Layout::Right;
  in_tl_ = d.in_topology_layout == DFFT_LAYOUT_LEFT ? Layout::Left : Layout::Right;
  out_tl_ = d.out_topology_layout == DFFT_LAYOUT_LEFT ? Layout::Left : Layout::Right;
  gin_ = to_vec(d.in_extents, rank_);
  gout_ = to_vec(d.out_extents, rank_);
  in_topo_ = to_vec(d.in_topology, rank_);
  out_topo_ = to_vec(d.out_topology, rank_);
  axes_ = convert_negative_axes(std::vector<int>(d.axes, d.axes + d.fft_rank), rank_);

  // internal_plan_factory (Plan.hpp:46-113)
  if (total_size(in_topo_) != total_size(out_topo_))
    throw Error(DFFT_ERROR_INVALID_TOPOLOGY, "in_topology and out_topology must have the same size.");
  if (static_cast<int>(total_size(in_topo_)) != d.nranks)
    throw Error(DFFT_ERROR_INVALID_TOPOLOGY, "topology size must be identical to mpi size.");
  P_ = static_cast<std::size_t>(d.nranks);
  if (P_ > static_cast<std::size_t>(kMaxRanks))
    throw Error(DFFT_ERROR_UNSUPPORTED, "more than 64 virtual ranks per plan");
  are_valid_extents(gin_, gout_, axes_, real_, false, view_layout_);
  for (std::size_t a = 0; a < rank_; ++a)
    if (gin_[a] >= (1u << 30) || gout_[a] >= (1u << 30))
      throw Error(DFFT_ERROR_UNSUPPORTED, "extent too large");

  kind_ = get_common_topology_type(in_topo_, out_topo_);
  switch (kind_) {
    case TopologyType::Shared: name_ = "SharedPlan"; break;
    case TopologyType::Slab: name_ = "SlabPlan"; break;
    case TopologyType::Pencil:
      if (rank_ < 3) throw Error(DFFT_ERROR_UNSUPPORTED, "Pencil plan supports 3D or higher dimensional Views");
      name_ = "PencilPlan";
      break;
    case TopologyType::Empty:
      throw Error(DFFT_ERROR_UNSUPPORTED, "Empty topology is not supported for FFT plan.");
    case TopologyType::Brick:
      throw Error(DFFT_ERROR_UNSUPPORTED, "Brick topology is not supported for FFT plan.");
    default:
      throw Error(DFFT_ERROR_INVALID_TOPOLOGY, "Unsupported topology for FFT plan.");
  }

  Vec id = identity_map(rank_);
  for (std::size_t r = 0; r < P_; ++r) {
    in_local_.push_back(compute_local_extents(gin_, in_topo_, id, r, in_tl_));
    out_local_.push_back(compute_local_extents(gout_, out_topo_, id, r, out_tl_));
  }
  fft_extents_ = compute_fft_extents(gin_, gout_, axes_);

  if (kind_ == TopologyType::Slab || kind_ == TopologyType::Pencil) {
    for (std::size_t r = 0; r < P_; ++r) {
      analyses_.push_back(block_analyses(kind_, real_, view_layout_, in_local_[r], gin_, gout_, in_topo_,
                                         out_topo_, axes_, r, in_tl_, out_tl_));
      max_buffer_size_ = std::max(max_buffer_size_, analyses_.back().max_buffer_size);
    }
    const std::size_t nb = analyses_[0].blocks.size();
    const std::size_t maxb = kind_ == TopologyType::Slab ? 5 : 8;  // SlabPlan.hpp:905, PencilPlan.hpp:1106
    if (nb < 1 || nb > maxb)
      throw Error(DFFT_ERROR_INVALID_TOPOLOGY, "Number blocks must be in [1, " + std::to_string(maxb) + "]");
  }
  if (d.exec_mode == DFFT_EXEC_BLOCKWALK) {
    if (kind_ == TopologyType::Shared)
      throw Error(DFFT_ERROR_UNSUPPORTED, "block-walk execution needs a slab or pencil plan");
  }

  // the TplPlan role (TplPlan.hpp:25-181): the vendor library names the plan
  if (d.exec_mode == DFFT_EXEC_VENDOR) name_ = "rocFFTPlan";

  // every length runs: the fast set, one line in LDS, two such factors, or
  // Bluestein over a power-of-two convolution m >= 2n - 1
  for (std::size_t k = 0; k < axes_.size() && d.exec_mode != DFFT_EXEC_VENDOR; ++k) {
    const std::size_t n = fft_extents_[k];
    if (!line_length_supported(static_cast<int>(n), fp64_) && n > (std::size_t(1) << 27))
      throw Error(DFFT_ERROR_UNSUPPORTED, "transform length " + std::to_string(n) + " too long for Bluestein");
  }

  check_hip(hipGetDevice(&device_), "hipGetDevice");
  try {
    allocate();
  } catch (...) {
    release();
    throw;
  }
}

void Plan::release() {
  vendor_.reset();
  for (auto& kv : tw_) (void)hipFree(kv.second);
  tw_.clear();
  if (scratch_) (void)hipFree(scratch_);
  scratch_ = nullptr;
  if (long_scratch_) (void)hipFree(long_scratch_);
  long_scratch_ = nullptr;
  long_bytes_ = 0;
  for (auto& kv : blue_) {
    if (kv.second.chirp) (void)hipFree(kv.second.chirp);
    if (kv.second.filt) (void)hipFree(kv.second.filt);
  }
  blue_.clear();
  for (void*& b : blue_s_) {
    if (b) (void)hipFree(b);
    b = nullptr;
  }
  blue_bytes_ = 0;
  for (auto& v : walk_bufs_) {
    for (void* p : v) (void)hipFree(p);
    v.clear();
  }
}

// device memory of the plan: twiddle tables, then the fused executor's
// scratch array or the block walk's ping-pong buffers
void Plan::allocate() {
  if (d_.exec_mode == DFFT_EXEC_VENDOR) {
    vendor_ = std::make_unique<VendorFFT>(gin_, gout_, axes_, real_, fp64_, norm_scale(true), norm_scale(false));
    workspace_bytes_ += vendor_->workspace_bytes();
    return;
  }
  // twiddle tables exp(-2 pi i e / n), evaluated in long double and rounded
  // once to FP64 -- for both precisions: the kernels form the powers of a
  // butterfly in FP64 and round each once to the compute precision
  const std::size_t cs = fp64_ ? 16 : 8;
  // elements of the largest array a pass walks (padded scratch rows
  // included): every pass has at most elems / n lines of length n
  const std::size_t elems =
      std::max({scratch_elems(gout_, cs), scratch_elems(gin_, cs), scratch_elems(gin_, cs / 2)});
  auto add_twiddles = [&](int n) {
    const int n1 = long_line_split(n, fp64_), m1 = long_line_split(n, true);
    for (int len : {n, n1, n1 > 0 ? n / n1 : 0, m1, m1 > 0 ? n / m1 : 0}) {
      if (len <= 0 || tw_.count(len)) continue;
      tw_[len] = upload_twiddles(len);
      workspace_bytes_ += 16 * static_cast<std::size_t>(len);
    }
  };
  std::size_t long_need = 0;  // bytes of the long-line scratch
  std::vector<int> blue_lengths;
  for (std::size_t k = 0; k < axes_.size(); ++k) {
    const int n = static_cast<int>(fft_extents_[k]);
    add_twiddles(n);
    if (long_line_split(n, fp64_) > 0 && !bluestein_preferred(n, fp64_)) long_need = std::max(long_need, elems * cs);
    if (bluestein_preferred(n, fp64_) &&
        std::find(blue_lengths.begin(), blue_lengths.end(), n) == blue_lengths.end())
      blue_lengths.push_back(n);
  }
  for (int n : blue_lengths) {
    int m = 1;
    while (m < 2 * n - 1) m *= 2;
    add_twiddles(m);
    const std::size_t lines = (elems + n - 1) / n;
    blue_bytes_ = std::max(blue_bytes_, lines * static_cast<std::size_t>(m) * cs);
    if (long_line_split(m, fp64_) > 0) long_need = std::max(long_need, lines * static_cast<std::size_t>(m) * cs);
    // the FP64 transform of the filter at set-up
    if (long_line_split(m, true) > 0) long_need = std::max(long_need, static_cast<std::size_t>(m) * 16);
  }
  if (long_need > 0 && total_size(gout_) > 0) {
    long_bytes_ = long_need;
    hipError_t err = hipMalloc(&long_scratch_, long_bytes_);
    if (err != hipSuccess) throw Error(DFFT_ERROR_OUT_OF_MEMORY, "hipMalloc of the long-line scratch failed");
    workspace_bytes_ += long_bytes_;
  }
  if (!blue_lengths.empty() && total_size(gout_) > 0) {
    for (void*& b : blue_s_) {
      hipError_t err = hipMalloc(&b, blue_bytes_);
      if (err != hipSuccess) throw Error(DFFT_ERROR_OUT_OF_MEMORY, "hipMalloc of the Bluestein scratch failed");
      workspace_bytes_ += blue_bytes_;
    }
    for (int n : blue_lengths) setup_bluestein(n, (elems + n - 1) / n);
  }
  if (d_.exec_mode == DFFT_EXEC_FUSED) {
    plan_fused();
    // the spectrum (gout, complex) and, for an in-place backward call, the
    // input-shaped result (gin, real for R2C) both live in the scratch
    const std::size_t rs = real_ ? cs / 2 : cs;
    const std::size_t bytes = total_size(gout_) == 0
                                  ? 0
                                  : std::max(scratch_elems(gout_, cs) * cs, scratch_elems(gin_, rs) * rs);
    if (bytes > 0) {
      hipError_t err = hipMalloc(&scratch_, bytes);
      if (err != hipSuccess) throw Error(DFFT_ERROR_OUT_OF_MEMORY, "hipMalloc of plan scratch failed");
      workspace_bytes_ += bytes;
    }
  } else {
    const std::size_t bytes = std::max<std::size_t>(max_buffer_size_ * (cs / 2), 16);
    for (int k = 0; k < 2; ++k) {
      for (std::size_t r = 0; r < P_; ++r) {
        void* p = nullptr;
        hipError_t err = hipMalloc(&p, bytes);
        if (err != hipSuccess) throw Error(DFFT_ERROR_OUT_OF_MEMORY, "hipMalloc of exchange buffer failed");
        walk_bufs_[k].push_back(p);
        workspace_bytes_ += bytes;
      }
    }
  }
}

Plan::~Plan() { release(); }

void Plan::local_starts(std::size_t r, int which, Vec& ext, Vec& st) const {
  if (r >= P_) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "rank must be less than topology size.");
  if (which == 0) compute_local_extents_and_starts(gin_, in_topo_, r, in_tl_, ext, st);
  else compute_local_extents_and_starts(gout_, out_topo_, r, out_tl_, ext, st);
}

double Plan::norm_scale(bool forward) const {
  double n = 1.0;
  for (auto e : fft_extents_) n *= static_cast<double>(e);
  switch (d_.norm) {
    case DFFT_NORM_BACKWARD: return forward ? 1.0 : 1.0 / n;
    case DFFT_NORM_FORWARD: return forward ? 1.0 / n : 1.0;
    case DFFT_NORM_ORTHO: return 1.0 / std::sqrt(n);
    default: return 1.0;
  }
}

const void* Plan::twiddles(int n) const { return tw_.at(n); }

// Order of the line passes.  The first pass reads the caller's input
// distribution and the last writes the output distribution; a pass whose line
// axis is split across ranks on either side falls back to per-element
// addressing, so the order prefers axes that are whole on the side they touch.
void Plan::plan_fused() {
  const std::size_t last_axis = axes_.back();
  auto build = [&](const Vec& first_topo, const Vec& last_topo, bool forward) {
    std::vector<PassSpec> out;
    Vec rest = axes_;
    std::size_t first, last = rank_;
    bool have_last = false;
    if (real_ && forward) {
      first = last_axis;
    } else {
      first = last_axis;
      if (!(real_ && !forward && axes_.size() == 1)) {
        for (auto it = axes_.rbegin(); it != axes_.rend(); ++it) {
          if (real_ && !forward && *it == last_axis) continue;  // C2R must come last
          if (first_topo[*it] == 1) {
            first = *it;
            break;
          }
        }
        if (real_ && !forward && first == last_axis) {
          for (auto it = axes_.rbegin(); it != axes_.rend(); ++it)
            if (*it != last_axis) {
              first = *it;
              break;
            }
        }
      }
    }
    rest.erase(std::find(rest.begin(), rest.end(), first));
    if (!rest.empty()) {
      have_last = true;
      if (real_ && !forward) {
        last = last_axis;
      } else {
        last = rest.front();
        for (auto a : rest)
          if (last_topo[a] == 1) {
            last = a;
            break;
          }
      }
      rest.erase(std::find(rest.begin(), rest.end(), last));
    }
    std::vector<std::size_t> order{first};
    for (auto it = rest.rbegin(); it != rest.rend(); ++it) order.push_back(*it);
    if (have_last) order.push_back(last);
    for (auto a : order) {
      PassSpec s;
      s.axis = static_cast<int>(a);
      s.n = static_cast<int>(std::max(gin_[a], gout_[a]));
      s.kind = kC2C;
      if (real_ && a == last_axis) s.kind = forward ? kR2C : kC2R;
      out.push_back(s);
    }
    return out;
  };
  fwd_ = build(in_topo_, out_topo_, true);
  bwd_ = build(out_topo_, in_topo_, false);
}

DDesc Plan::user_desc(const Vec& ext, const Vec& topo, Layout tl, void* const* ptrs) const {
  DDesc d{};
  d.ndim = static_cast<int32_t>(rank_);
  d.layout_left = view_layout_ == Layout::Left ? 1 : 0;
  std::size_t stride = 1;
  // rank = sum coord[a]*rstride[a], matching rank_to_coord (MPI_Extents.hpp:492-529)
  if (tl == Layout::Right) {
    for (std::size_t a = rank_; a-- > 0;) {
      d.rstride[a] = static_cast<int32_t>(stride);
      stride *= topo[a];
    }
  } else {
    for (std::size_t a = 0; a < rank_; ++a) {
      d.rstride[a] = static_cast<int32_t>(stride);
      stride *= topo[a];
    }
  }
  for (std::size_t a = 0; a < rank_; ++a) {
    d.ext[a] = static_cast<int32_t>(ext[a]);
    d.ntopo[a] = static_cast<int32_t>(topo[a]);
    d.base[a] = static_cast<int32_t>(ext[a] / topo[a]);
    d.rem[a] = static_cast<int32_t>(ext[a] % topo[a]);
    d.map[a] = static_cast<int32_t>(a);
    d.div_big[a] = FastDiv(static_cast<uint32_t>(d.base[a] + 1));
    d.div_small[a] = FastDiv(static_cast<uint32_t>(d.base[a]));
  }
  for (std::size_t r = 0; r < P_; ++r) d.ptr[r] = static_cast<char*>(ptrs[r]);
  return d;
}

// Row pitch of the innermost storage axis of a LayoutRight scratch array:
// rows start on 128-byte boundaries (blocks of lines along the row axis then
// read and write whole cache lines; 129-element rows of an R2C spectrum cost
// ~15 % per pass otherwise).
// `mode`: 1 pads whenever the row is not aligned, 0 never, -1 only when the
// padding adds at most 1/16 to the row (129 -> 136 pays; 65 -> 72 does not).
static int32_t scratch_pitch(int32_t inner, std::size_t elem_bytes, bool layout_left, int mode) {
  if (mode == 0 || layout_left || inner <= 0) return 0;
  const int32_t align = static_cast<int32_t>(128 / elem_bytes);
  if (align <= 1 || inner % align == 0) return 0;
  const int32_t pitch = (inner + align - 1) / align * align;
  if (mode < 0 && (pitch - inner) * 16 > inner) return 0;
  return pitch;
}

std::size_t Plan::scratch_elems(const Vec& ext, std::size_t elem_bytes) const {
  // allocation: room for the padded rows whatever the option says later
  const int32_t pitch = scratch_pitch(static_cast<int32_t>(ext[rank_ - 1]), elem_bytes, view_layout_ == Layout::Left, 1);
  std::size_t n = 1;
  for (std::size_t a = 0; a + 1 < rank_; ++a) n *= ext[a];
  return n * (pitch > 0 ? static_cast<std::size_t>(pitch) : ext[rank_ - 1]);
}

DDesc Plan::scratch_desc(const Vec& ext, void* p, std::size_t elem_bytes) const {
  DDesc d{};
  d.ndim = static_cast<int32_t>(rank_);
  d.layout_left = view_layout_ == Layout::Left ? 1 : 0;
  for (std::size_t a = 0; a < rank_; ++a) {
    d.ext[a] = static_cast<int32_t>(ext[a]);
    d.ntopo[a] = 1;
    d.rstride[a] = 0;
    d.base[a] = static_cast<int32_t>(ext[a]);
    d.rem[a] = 0;
    d.map[a] = static_cast<int32_t>(a);
    d.div_big[a] = FastDiv(static_cast<uint32_t>(ext[a] + 1));
    d.div_small[a] = FastDiv(static_cast<uint32_t>(ext[a]));
  }
  d.pitch = scratch_pitch(static_cast<int32_t>(ext[rank_ - 1]), elem_bytes, d.layout_left != 0, pad_);
  d.ptr[0] = static_cast<char*>(p);
  return d;
}

LinePass Plan::make_pass(const PassSpec& s, const DDesc& in, const DDesc& out, bool inverse,
                         double scale, int slow_axis) const {
  LinePass p{};
  p.in = in;
  p.out = out;
  p.axis = s.axis;
  p.kind = s.kind;
  p.n = s.n;
  p.inverse = inverse ? 1 : 0;
  p.scale = scale;
  p.herm_weight = s.kind == kC2C ? 0.0 : adj_herm_;
  p.paired = s.kind == kC2C ? 0 : s.paired;
  p.xcd_remap = s.xcd;
  p.swizzle = s.swz;
  p.split_lds = s.split;
  p.dual_lines = s.dual;
  p.linear_lines = s.lin;
  p.tw = twiddles(s.n);
  p.lines_per_block = s.lines_per_block;
  p.values_per_thread = s.values_per_thread;
  // batch axes ordered slowest -> fastest in storage order
  std::vector<int> order;
  if (in.layout_left) {
    for (int a = static_cast<int>(rank_) - 1; a >= 0; --a) order.push_back(in.map[a]);
  } else {
    for (std::size_t k = 0; k < rank_; ++k) order.push_back(in.map[k]);
  }
  const int innermost = order.back();
  if (slow_axis >= 0 && slow_axis != innermost) {
    // a chunked pass enumerates the chunk axis slowest so a chunk is a contiguous range of lines
    order.erase(std::find(order.begin(), order.end(), slow_axis));
    order.insert(order.begin(), slow_axis);
  }
  p.nbatch = 0;
  int64_t nl = 1;
  for (int a : order) {
    if (a == s.axis) continue;
    p.batch_axis[p.nbatch] = a;
    p.batch_ext[p.nbatch] = in.ext[a];
    p.batch_div[p.nbatch] = FastDiv(static_cast<uint32_t>(in.ext[a]));
    nl *= in.ext[a];
    ++p.nbatch;
  }
  // a padded scratch row axis is enumerated over its pitch (padding lines are
  // skipped in the kernel) so each block of lines starts on an aligned row;
  // paired R2C / C2R passes keep the plain enumeration
  p.pad_axis = 0;
  p.pad_real = 0;
  const int32_t pitch = in.pitch > 0 ? in.pitch : out.pitch;
  const int inner = static_cast<int>(rank_) - 1;  // padded storage axis of the scratch (LayoutRight)
  const bool pairable = s.kind != kC2C && s.paired;
  if (pitch > 0 && s.axis != inner && !pairable) {
    for (int j = 0; j < p.nbatch; ++j) {
      if (p.batch_axis[j] != inner) continue;
      nl = nl / p.batch_ext[j] * pitch;
      p.pad_axis = inner;
      p.pad_real = p.batch_ext[j];
      p.batch_ext[j] = pitch;
      p.batch_div[j] = FastDiv(static_cast<uint32_t>(pitch));
    }
  }
  if (nl >= (int64_t(1) << 31)) throw Error(DFFT_ERROR_UNSUPPORTED, "too many lines in one pass");
  p.nlines = static_cast<int32_t>(nl);
  const int n1 = blue_.count(s.n) ? 0 : long_line_split(s.n, fp64_);
  if (n1 > 0) {
    const std::size_t need = static_cast<std::size_t>(nl) * static_cast<std::size_t>(s.n) * (fp64_ ? 16 : 8);
    if (!long_scratch_ || need > long_bytes_)
      throw Error(DFFT_ERROR_UNSUPPORTED, "long-line scratch too small for a pass of " + std::to_string(nl) + " lines");
    p.long_n1 = n1;
    p.long_tw1 = twiddles(n1);
    p.long_tw2 = twiddles(s.n / n1);
    p.long_scratch = long_scratch_;
  }
  p.line_begin = 0;
  p.lines_fast = s.lines_fast >= 0 ? s.lines_fast : (s.axis != innermost ? 1 : 0);
  if (s.nontemporal >= 0) {
    p.nontemporal = s.nontemporal;
  } else {
    // streaming accesses pay off once the array no longer fits the caches
    double elems = 1.0;
    for (std::size_t a = 0; a < rank_; ++a) elems *= static_cast<double>(in.ext[a]);
    const double es = (s.kind == kR2C ? 0.5 : 1.0) * (fp64_ ? 16.0 : 8.0);
    p.nontemporal = elems * es >= 1024.0 * 1024 * 1024 ? 1 : 0;
  }
  return p;
}

void Plan::run_fused(void* const* in, void* const* out, bool forward, hipStream_t stream) {
  const auto& passes = forward ? fwd_ : bwd_;
  const double scale = exec_scale(forward);
  // first and last descriptors
  DDesc src = forward ? user_desc(gin_, in_topo_, in_tl_, in) : user_desc(gout_, out_topo_, out_tl_, in);
  DDesc dst = forward ? user_desc(gout_, out_topo_, out_tl_, out) : user_desc(gin_, in_topo_, in_tl_, out);
  // the scratch array holds the complex spectrum shape (gout)
  DDesc scr = scratch_desc(gout_, scratch_, fp64_ ? 16 : 8);
  // element size of the scratch when it stands in for the output (alias case)
  const std::size_t alias_es = (forward || !real_) ? (fp64_ ? 16 : 8) : (fp64_ ? 8 : 4);
  // alias check: a single pass may not read and write the same memory
  bool alias = false;
  if (passes.size() == 1) {
    for (std::size_t r = 0; r < P_ && !alias; ++r)
      for (std::size_t q = 0; q < P_; ++q)
        if (in[r] != nullptr && in[r] == out[q]) alias = true;
  }
  // Chunking: with three passes, the first two touch each plane of the third
  // pass's axis independently, so they run plane-block by plane-block; a block
  // small enough for the 256 MB Infinity Cache is written by pass 1 and read
  // back by pass 2 without a round trip to HBM.
  const int chunk = chunk_[forward ? 0 : 1];
  std::size_t i0 = 0;
  LinePass p0{}, p1{};
  bool chunked = false;
  if (chunk > 0 && passes.size() == 3 && !alias) {
    const int cax = passes[2].axis;
    p0 = make_pass(passes[0], src, scr, !forward, 1.0, cax);
    p1 = make_pass(passes[1], scr, scr, !forward, 1.0, cax);
    // the chunk axis must be the slowest batch axis of both passes
    chunked = p0.nbatch > 0 && p1.nbatch > 0 && p0.batch_axis[0] == cax && p1.batch_axis[0] == cax &&
              p0.batch_ext[0] == p1.batch_ext[0];
  }
  if (chunked) {
    // the chunk axis is the slowest batch axis of both passes; its enumerated
    // extent (padded if it is the scratch row axis) splits the lines
    const int extc = p0.batch_ext[0];
    const int32_t per0 = p0.nlines / extc, per1 = p1.nlines / extc;
    for (int c0 = 0; c0 < extc; c0 += chunk) {
      const int nc = std::min(chunk, extc - c0);
      p0.line_begin = c0 * per0;
      p0.nlines = nc * per0;
      p1.line_begin = c0 * per1;
      p1.nlines = nc * per1;
      mark(stream);
      check_hip(launch_pass(p0, stream), "line pass");
      mark(stream);
      mark(stream);
      check_hip(launch_pass(p1, stream), "line pass");
      mark(stream);
    }
    i0 = 2;
  }
  for (std::size_t i = i0; i < passes.size(); ++i) {
    const bool first = i == 0, last = i + 1 == passes.size();
    DDesc a = first ? src : scr;
    DDesc b = last ? (alias ? scratch_desc(forward ? gout_ : gin_, scratch_, alias_es) : dst) : scr;
    LinePass p = make_pass(passes[i], a, b, !forward, last ? scale : 1.0);
    TraceRange range("dfft::%s pass %zu axis %d%s", forward ? "forward" : "backward", i, passes[i].axis,
                     first ? " (gather from the input ranks)" : last ? " (scatter to the output ranks)" : "");
    mark(stream);
    check_hip(launch_pass(p, stream), "line pass");
    mark(stream);
  }
  if (alias) {
    RemapArgs ra{};
    ra.src = scratch_desc(forward ? gout_ : gin_, scratch_, alias_es);
    ra.dst = dst;
    ra.total = 1;
    for (std::size_t k = 0; k < rank_; ++k) {
      ra.iter_axis[k] = static_cast<int32_t>(k);
      ra.iter_ext[k] = dst.ext[k];
      ra.iter_div[k] = FastDiv(static_cast<uint32_t>(dst.ext[k]));
      ra.total *= dst.ext[k];
    }
    const int es = (forward || !real_) ? (fp64_ ? 16 : 8) : (fp64_ ? 8 : 4);
    mark(stream);
    check_hip(launch_remap(ra, es, stream), "remap");
    mark(stream);
  }
}

bool Plan::products_supported() const {
  if (!fp64_ || !real_ || d_.exec_mode != DFFT_EXEC_FUSED || view_layout_ != Layout::Right) return false;
  if (rank_ != 4 || gin_[0] != 6 || in_topo_[0] != 1 || out_topo_[0] != 1) return false;
  if (fwd_.empty() || fwd_[0].kind != kR2C || fwd_[0].axis != 3 || fwd_[0].paired == 0) return false;
  for (int a : {1, 2, 3})
    if (std::find(axes_.begin(), axes_.end(), static_cast<std::size_t>(a)) == axes_.end()) return false;
  return products_pass_supported(fwd_[0].n);
}

void Plan::execute_forward_products(void* const* u3, void* const* out, hipStream_t stream) {
  if (!products_supported()) throw Error(DFFT_ERROR_UNSUPPORTED, "this plan cannot form the products in its R2C pass");
  if (u3 == nullptr || out == nullptr) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "null pointer table");
  int cur = -1;
  check_hip(hipGetDevice(&cur), "hipGetDevice");
  if (cur != device_) check_hip(hipSetDevice(device_), "hipSetDevice");
  if (total_size(gin_) == 0) return;
  TraceRange range("dfft::%s::execute forward (products of 3 fields)", name_.c_str());
  const auto& passes = fwd_;
  const double scale = exec_scale(true);
  // the 6-field input distribution (line enumeration) and the 3-field one
  // the products are read from: the field axis is not distributed, so every
  // rank's block of the 3-field array has the same strides but 3 fields
  DDesc src6 = user_desc(gin_, in_topo_, in_tl_, u3);
  DDesc src3 = src6;
  src3.ext[0] = 3;
  src3.base[0] = 3;
  src3.rem[0] = 0;
  src3.div_big[0] = FastDiv(4u);
  src3.div_small[0] = FastDiv(3u);
  DDesc dst = user_desc(gout_, out_topo_, out_tl_, out);
  DDesc scr = scratch_desc(gout_, scratch_, 16);
  for (std::size_t i = 0; i < passes.size(); ++i) {
    const bool first = i == 0, last = i + 1 == passes.size();
    DDesc b = last ? dst : scr;
    if (first) {
      LinePass p = make_pass(passes[0], src6, b, false, last ? scale : 1.0);
      p.in = src3;
      p.products = 1;
      p.prod_axis = 0;
      // enumerate the field axis fastest: the six product lines of one
      // position sit in neighbouring pairs and read the same three velocity
      // lines, which the caches serve after the first read
      int fj = -1;
      for (int j = 0; j < p.nbatch; ++j)
        if (p.batch_axis[j] == 0) fj = j;
      if (fj >= 0) {
        const int32_t ax = p.batch_axis[fj], ex = p.batch_ext[fj];
        const FastDiv dv = p.batch_div[fj];
        for (int j = fj; j + 1 < p.nbatch; ++j) {
          p.batch_axis[j] = p.batch_axis[j + 1];
          p.batch_ext[j] = p.batch_ext[j + 1];
          p.batch_div[j] = p.batch_div[j + 1];
        }
        p.batch_axis[p.nbatch - 1] = ax;
        p.batch_ext[p.nbatch - 1] = ex;
        p.batch_div[p.nbatch - 1] = dv;
      }
      TraceRange pr("dfft::forward pass 0 axis %d (products of the velocity lines)", passes[0].axis);
      mark(stream);
      check_hip(launch_products_pass(p, stream), "products pass");
      mark(stream);
      continue;
    }
    LinePass p = make_pass(passes[i], scr, b, false, last ? scale : 1.0);
    TraceRange pr("dfft::forward pass %zu axis %d%s", i, passes[i].axis, last ? " (scatter to the output ranks)" : "");
    mark(stream);
    check_hip(launch_pass(p, stream), "line pass");
    mark(stream);
  }
}

// An (rows, m) complex array transformed along its rows (length m).
LinePass Plan::dense_pass(int rows, int m, void* src, void* dst, bool inverse, bool fp64) const {
  auto dense = [&](void* ptr) {
    DDesc d{};
    d.ndim = 2;
    const int32_t ext[2] = {rows, m};
    for (int a = 0; a < 2; ++a) {
      d.ext[a] = ext[a];
      d.ntopo[a] = 1;
      d.base[a] = ext[a];
      d.map[a] = a;
      d.div_big[a] = FastDiv(static_cast<uint32_t>(ext[a] + 1));
      d.div_small[a] = FastDiv(static_cast<uint32_t>(ext[a]));
    }
    d.ptr[0] = static_cast<char*>(ptr);
    return d;
  };
  LinePass q{};
  q.in = dense(src);
  q.out = dense(dst);
  q.axis = 1;
  q.nbatch = 1;
  q.batch_axis[0] = 0;
  q.batch_ext[0] = rows;
  q.batch_div[0] = FastDiv(static_cast<uint32_t>(rows));
  q.nlines = rows;
  q.kind = kC2C;
  q.n = m;
  q.inverse = inverse ? 1 : 0;
  q.tw = twiddles(m);
  q.scale = 1.0;
  q.split_lds = 1;
  const int m1 = long_line_split(m, fp64);
  if (m1 > 0) {
    const std::size_t need = static_cast<std::size_t>(rows) * m * (fp64 ? 16 : 8);
    if (!long_scratch_ || need > long_bytes_) throw Error(DFFT_ERROR_UNSUPPORTED, "long-line scratch too small");
    q.long_n1 = m1;
    q.long_tw1 = twiddles(m1);
    q.long_tw2 = twiddles(m / m1);
    q.long_scratch = long_scratch_;
  }
  return q;
}

// Tables of the Bluestein transform of length n: the chirp exp(-pi i e / n)
// (e = j^2 mod 2n, evaluated in long double, rounded once to FP64) and the
// FP64 m-point transform of the filter b_j = conj(chirp(j^2)), |j| < n
// (b at m - j for negative j), divided by m.
void Plan::setup_bluestein(int n, std::size_t max_lines) {
  Bluestein b;
  b.m = 1;
  while (b.m < 2 * n - 1) {
    b.m *= 2;
    ++b.log2m;
  }
  if (max_lines * static_cast<std::size_t>(b.m) >= (std::size_t(1) << 31))
    throw Error(DFFT_ERROR_UNSUPPORTED, "too many Bluestein lines of length " + std::to_string(n));
  std::vector<double> chirp(4 * static_cast<std::size_t>(n));
  for (int e = 0; e < 2 * n; ++e) {
    const long double a = 3.141592653589793238462643383279502884L * e / n;
    chirp[2 * e] = static_cast<double>(cosl(a));
    chirp[2 * e + 1] = static_cast<double>(-sinl(a));
  }
  std::vector<double> filt(2 * static_cast<std::size_t>(b.m), 0.0);
  for (int64_t j = 0; j < n; ++j) {
    const int64_t e = (j * j) % (2 * static_cast<int64_t>(n));
    const double re = chirp[2 * e], im = -chirp[2 * e + 1];
    filt[2 * j] = re;
    filt[2 * j + 1] = im;
    if (j > 0) {
      filt[2 * (b.m - j)] = re;
      filt[2 * (b.m - j) + 1] = im;
    }
  }
  const std::size_t cb = chirp.size() * sizeof(double), fb = filt.size() * sizeof(double);
  check_hip(hipMalloc(&b.chirp, cb), "hipMalloc of the Bluestein chirp");
  blue_[n] = b;  // released with the plan from here on
  check_hip(hipMalloc(&blue_[n].filt, fb), "hipMalloc of the Bluestein filter");
  struct Tmp {
    void* p = nullptr;
    ~Tmp() {
      if (p) (void)hipFree(p);
    }
  } tmp;
  check_hip(hipMalloc(&tmp.p, fb), "hipMalloc");
  hipError_t err = hipMemcpy(b.chirp, chirp.data(), cb, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(tmp.p, filt.data(), fb, hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    LinePass q = dense_pass(1, b.m, tmp.p, blue_[n].filt, false, true);
    q.scale = 1.0 / b.m;
    err = launch_line_pass(q, true, nullptr);
  }
  if (err == hipSuccess) err = hipStreamSynchronize(nullptr);
  check_hip(err, "Bluestein filter transform");
  workspace_bytes_ += cb + fb;
}

hipError_t Plan::launch_pass(const LinePass& p, hipStream_t s) {
  auto it = blue_.find(p.n);
  if (it == blue_.end()) return launch_line_pass(p, fp64_, s);
  if (p.nlines <= 0) return hipSuccess;
  const Bluestein& b = it->second;
  const std::size_t need = static_cast<std::size_t>(p.nlines) * b.m * (fp64_ ? 16 : 8);
  if (need > blue_bytes_) return hipErrorInvalidValue;
  hipError_t e = launch_bluestein_pre(p, b.log2m, b.chirp, blue_s_[0], fp64_, s);
  if (e == hipSuccess) e = launch_line_pass(dense_pass(p.nlines, b.m, blue_s_[0], blue_s_[1], false, fp64_), fp64_, s);
  if (e == hipSuccess)
    e = launch_bluestein_mul(static_cast<int64_t>(p.nlines) << b.log2m, b.log2m, b.filt, blue_s_[1], fp64_, s);
  if (e == hipSuccess) e = launch_line_pass(dense_pass(p.nlines, b.m, blue_s_[1], blue_s_[0], true, fp64_), fp64_, s);
  if (e == hipSuccess) e = launch_bluestein_post(p, b.log2m, b.chirp, blue_s_[0], fp64_, s);
  return e;
}

void Plan::execute(void* const* in, void* const* out, int direction, hipStream_t stream) {
  if (in == nullptr || out == nullptr) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "null pointer table");
  if (direction != DFFT_FORWARD && direction != DFFT_BACKWARD)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "Invalid FFT direction specified.");
  const bool forward = direction == DFFT_FORWARD;
  for (std::size_t r = 0; r < P_; ++r) {
    const Vec& ie = forward ? in_local_[r] : out_local_[r];
    const Vec& oe = forward ? out_local_[r] : in_local_[r];
    if (total_size(ie) > 0 && in[r] == nullptr) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "null input pointer");
    if (total_size(oe) > 0 && out[r] == nullptr) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "null output pointer");
  }
  int cur = -1;
  check_hip(hipGetDevice(&cur), "hipGetDevice");
  if (cur != device_) check_hip(hipSetDevice(device_), "hipSetDevice");
  if (total_size(gin_) == 0) return;
  TraceRange range("dfft::%s::execute %s", name_.c_str(), forward ? "forward" : "backward");
  if (d_.exec_mode == DFFT_EXEC_VENDOR) {
    if (adj_scale_ >= 0.0) throw Error(DFFT_ERROR_UNSUPPORTED, "the vendor plan has no adjoint");
    DDesc src = forward ? user_desc(gin_, in_topo_, in_tl_, in) : user_desc(gout_, out_topo_, out_tl_, in);
    DDesc dst = forward ? user_desc(gout_, out_topo_, out_tl_, out) : user_desc(gin_, in_topo_, in_tl_, out);
    vendor_->run(src, dst, forward, stream, [this](hipStream_t s) { mark(s); });
    return;
  }
  if (d_.exec_mode == DFFT_EXEC_BLOCKWALK) run_blockwalk(in, out, forward, stream);
  else run_fused(in, out, forward, stream);
}

void Plan::check_buffers(void* const* in, void* const* out, int direction) const {
  if (in == nullptr || out == nullptr) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "null pointer table");
  if (direction != DFFT_FORWARD && direction != DFFT_BACKWARD)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "Invalid FFT direction specified.");
  const bool forward = direction == DFFT_FORWARD;
  const std::size_t cs = fp64_ ? 16 : 8, rs = real_ ? cs / 2 : cs;
  for (std::size_t r = 0; r < P_; ++r) {
    const std::size_t nin = total_size(in_local_[r]) * rs, nout = total_size(out_local_[r]) * cs;
    const std::string who = " of rank " + std::to_string(r);
    check_device_extent(in[r], forward ? nin : nout, "input View" + who);
    check_device_extent(out[r], forward ? nout : nin, "output View" + who);
  }
}

// The adjoint of a transform is the opposite-direction pass sequence with
// the transform's own scale: (s F)^* = s n F^{-1} for C2C.  For R2C the
// half-spectrum side counts each stored entry once, so
//   R2C^* y = C2R(w y),  w = 1 at k = 0 and k = n/2, 1/2 in between,
//   C2R^* x = w' R2C(x), w' = 1 at k = 0 and k = n/2, 2 in between,
// with the weight applied in the R2C / C2R line pass (it commutes with the
// transforms along the other axes).
void Plan::execute_adjoint(void* const* in, void* const* out, int direction, hipStream_t stream) {
  if (direction != DFFT_FORWARD && direction != DFFT_BACKWARD)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "Invalid FFT direction specified.");
  const bool forward = direction == DFFT_FORWARD;
  struct Reset {
    Plan* p;
    ~Reset() {
      p->adj_scale_ = -1.0;
      p->adj_herm_ = 0.0;
    }
  } reset{this};
  adj_scale_ = norm_scale(forward);
  adj_herm_ = real_ ? (forward ? 0.5 : 2.0) : 0.0;
  execute(in, out, forward ? DFFT_BACKWARD : DFFT_FORWARD, stream);
}

void Plan::mark(hipStream_t s) {
  if (!timing_) return;
  hipEvent_t e;
  check_hip(hipEventCreate(&e), "hipEventCreate");
  timing_->push_back(e);
  check_hip(hipEventRecord(e, s), "hipEventRecord");
}

std::vector<float> Plan::execute_timed(void* const* in, void* const* out, int direction, hipStream_t stream) {
  std::vector<hipEvent_t> ev;
  timing_ = &ev;
  std::vector<float> ms;
  try {
    execute(in, out, direction, stream);
  } catch (...) {
    timing_ = nullptr;
    for (auto e : ev) (void)hipEventDestroy(e);
    throw;
  }
  timing_ = nullptr;
  hipError_t err = hipStreamSynchronize(stream);
  for (std::size_t i = 0; err == hipSuccess && i + 1 < ev.size(); i += 2) {
    float t = 0.f;
    err = hipEventElapsedTime(&t, ev[i], ev[i + 1]);
    ms.push_back(t);
  }
  for (auto e : ev) (void)hipEventDestroy(e);
  check_hip(err, "timed execute");
  return ms;
}

void Plan::set_option(const std::string& key, long value) {
  if (key == "chunk") {
    chunk_[0] = chunk_[1] = static_cast<int>(std::max(0L, value));
    return;
  }
  if (key == "fwd.chunk" || key == "bwd.chunk") {
    chunk_[key[0] == 'f' ? 0 : 1] = static_cast<int>(std::max(0L, value));
    return;
  }
  if (key == "pad") {
    if (value < -1 || value > 1) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "pad must be -1, 0 or 1");
    pad_ = static_cast<int>(value);
    return;
  }
  // "<dir>.<i>.<field>"
  const auto d1 = key.find('.'), d2 = key.rfind('.');
  if (d1 == std::string::npos || d2 == d1) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "unknown option " + key);
  const std::string dir = key.substr(0, d1), field = key.substr(d2 + 1);
  auto& passes = dir == "fwd" ? fwd_ : bwd_;
  if (dir != "fwd" && dir != "bwd") throw Error(DFFT_ERROR_INVALID_ARGUMENT, "unknown option " + key);
  const std::size_t i = static_cast<std::size_t>(std::stoul(key.substr(d1 + 1, d2 - d1 - 1)));
  if (i >= passes.size()) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "pass index out of range in " + key);
  if (field == "lpb") passes[i].lines_per_block = static_cast<int>(value);
  else if (field == "vpt") passes[i].values_per_thread = static_cast<int>(value);
  else if (field == "lf") passes[i].lines_fast = static_cast<int>(value);
  else if (field == "nt") passes[i].nontemporal = static_cast<int>(value);
  else if (field == "pair") passes[i].paired = static_cast<int>(value);
  else if (field == "xcd") passes[i].xcd = static_cast<int>(value);
  else if (field == "swz") passes[i].swz = static_cast<int>(value);
  else if (field == "split") passes[i].split = static_cast<int>(value);
  else if (field == "dual") passes[i].dual = static_cast<int>(value);
  else if (field == "lin") passes[i].lin = static_cast<int>(value);
  else throw Error(DFFT_ERROR_INVALID_ARGUMENT, "unknown option " + key);
}

long Plan::get_option(const std::string& key) const {
  if (key == "chunk" || key == "fwd.chunk") return chunk_[0];
  if (key == "bwd.chunk") return chunk_[1];
  if (key == "pad") return pad_;
  const auto d1 = key.find('.'), d2 = key.rfind('.');
  if (d1 == std::string::npos || d2 == d1) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "unknown option " + key);
  const std::string dir = key.substr(0, d1), field = key.substr(d2 + 1);
  const auto& passes = dir == "fwd" ? fwd_ : bwd_;
  const std::size_t i = static_cast<std::size_t>(std::stoul(key.substr(d1 + 1, d2 - d1 - 1)));
  if (i >= passes.size()) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "pass index out of range in " + key);
  if (field == "lpb") return passes[i].lines_per_block;
  if (field == "vpt") return passes[i].values_per_thread;
  if (field == "lf") return passes[i].lines_fast;
  if (field == "nt") return passes[i].nontemporal;
  if (field == "pair") return passes[i].paired;
  if (field == "xcd") return passes[i].xcd;
  if (field == "swz") return passes[i].swz;
  if (field == "split") return passes[i].split;
  if (field == "dual") return passes[i].dual;
  if (field == "lin") return passes[i].lin;
  throw Error(DFFT_ERROR_INVALID_ARGUMENT, "unknown option " + key);
}

float Plan::time_direction(void* const* in, void* const* out, bool forward, hipStream_t s, int reps) {
  hipEvent_t e0, e1;
  check_hip(hipEventCreate(&e0), "hipEventCreate");
  check_hip(hipEventCreate(&e1), "hipEventCreate");
  execute(in, out, forward ? DFFT_FORWARD : DFFT_BACKWARD, s);
  check_hip(hipEventRecord(e0, s), "hipEventRecord");
  for (int k = 0; k < reps; ++k) execute(in, out, forward ? DFFT_FORWARD : DFFT_BACKWARD, s);
  check_hip(hipEventRecord(e1, s), "hipEventRecord");
  check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
  float ms = 0.f;
  check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

// Autotuner: per pass, every compile-time launch shape (values per thread x
// lines per block x lane order) is timed on the pass's own descriptors; then
// the chunk size of the first two passes.  Shapes the launcher rejects
// (too many threads or too much LDS) are skipped before launch.
void Plan::autotune(void* const* in, void* const* out, hipStream_t stream) {
  if (d_.exec_mode != DFFT_EXEC_FUSED) return;
  // the tuner writes both sides: refuse arrays that cannot hold them
  check_buffers(in, out, DFFT_FORWARD);
  check_hip(hipSetDevice(device_), "hipSetDevice");
  hipEvent_t e0, e1;
  check_hip(hipEventCreate(&e0), "hipEventCreate");
  check_hip(hipEventCreate(&e1), "hipEventCreate");
  const int vpts[] = {4, 8, 16};
  const int lpbs[] = {1, 2, 4, 8, 16, 32};
  for (int dir = 0; dir < 2; ++dir) {
    const bool forward = dir == 0;
    auto& passes = forward ? fwd_ : bwd_;
    void* const* a_ptr = forward ? in : out;
    void* const* b_ptr = forward ? out : in;
    DDesc src = forward ? user_desc(gin_, in_topo_, in_tl_, a_ptr) : user_desc(gout_, out_topo_, out_tl_, a_ptr);
    DDesc dst = forward ? user_desc(gout_, out_topo_, out_tl_, b_ptr) : user_desc(gin_, in_topo_, in_tl_, b_ptr);
    DDesc scr = scratch_desc(gout_, scratch_, fp64_ ? 16 : 8);
    for (std::size_t i = 0; i < passes.size(); ++i) {
      if (!has_fast_line_kernel(passes[i].n)) continue;
      DDesc a = i == 0 ? src : scr;
      DDesc b = i + 1 == passes.size() ? dst : scr;
      // one timing of 3 launches per shape, then the six fastest shapes are
      // timed three more times and the best single timing wins (one noisy
      // sample does not decide)
      auto time_spec = [&](const PassSpec& s) -> float {
        LinePass p = make_pass(s, a, b, !forward, 1.0);
        if (launch_pass(p, stream) != hipSuccess) {
          (void)hipGetLastError();
          return -1.f;
        }
        check_hip(hipEventRecord(e0, stream), "hipEventRecord");
        for (int k = 0; k < 3; ++k) check_hip(launch_pass(p, stream), "line pass");
        check_hip(hipEventRecord(e1, stream), "hipEventRecord");
        check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ms;
      };
      std::vector<std::pair<float, PassSpec>> timed;
      for (int vpt : vpts) {
        for (int lpb : lpbs) {
          // the column-major block walk ("swz") is left out of the search: it
          // measured slower on every pass shape tried (512^3 x pass +36 %)
          // bit 3: two real lines per complex FFT (R2C / C2R) or two adjacent
          // lines per lane (FP32 C2C lanes across lines)
          const bool c2c = passes[i].kind == kC2C;
          for (int lfnt = 0; lfnt < (c2c && fp64_ ? 8 : 16); ++lfnt) {
            if (c2c && (lfnt >> 3) && !(lfnt & 1)) continue;  // dual lines need lanes across lines
            PassSpec s = passes[i];
            s.values_per_thread = vpt;
            s.lines_per_block = lpb;
            s.lines_fast = lfnt & 1;
            s.nontemporal = (lfnt >> 1) & 1;
            s.xcd = (lfnt >> 2) & 1;
            if (c2c) {
              s.dual = lfnt >> 3;
            } else {
              s.paired = lfnt >> 3;
            }
            const float ms = time_spec(s);
            if (ms >= 0.f) timed.emplace_back(ms, s);
          }
        }
      }
      if (timed.empty()) continue;
      std::stable_sort(timed.begin(), timed.end(),
                       [](const auto& x, const auto& y) { return x.first < y.first; });
      const std::size_t finalists = std::min<std::size_t>(6, timed.size());
      float best = 1e30f;
      PassSpec best_spec = passes[i];
      for (std::size_t f = 0; f < finalists; ++f) {
        // each finalist also with the other LDS exchange form (split halves
        // or complex; kernels without a split form ignore the option)
        for (int flip = 0; flip < 2; ++flip) {
          PassSpec s = timed[f].second;
          if (flip) s.split = s.split ? 0 : 1;
          float ms = flip ? 1e30f : timed[f].first;
          for (int trial = 0; trial < 3; ++trial) {
            const float m = time_spec(s);
            if (m >= 0.f) ms = std::min(ms, m);
          }
          if (ms < best) {
            best = ms;
            best_spec = s;
          }
        }
      }
      passes[i] = best_spec;
    }
    if (passes.size() == 3) {
      // a chunk size must beat the unchunked passes by 3 % (timing noise
      // alone has picked chunks that are slower on a re-measurement)
      chunk_[dir] = 0;
      float best = 0.97f * time_direction(a_ptr, b_ptr, forward, stream, 3);
      int best_chunk = 0;
      for (int c : {4, 8, 16, 32, 64}) {
        chunk_[dir] = c;
        const float ms = time_direction(a_ptr, b_ptr, forward, stream, 3);
        if (ms < best) {
          best = ms;
          best_chunk = c;
        }
      }
      chunk_[dir] = best_chunk;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

int Plan::num_passes(int direction) const {
  if (d_.exec_mode == DFFT_EXEC_VENDOR) return 3;  // gather, vendor call, scatter
  if (d_.exec_mode == DFFT_EXEC_FUSED) {
    const auto& p = direction == DFFT_FORWARD ? fwd_ : bwd_;
    return static_cast<int>(p.size());
  }
  return -1;  // data dependent for the block walk; execute_timed reports it
}

// ---------------------------------------------------------------------------
// Block walk: SlabInternalPlan/PencilInternalPlan forward_impl/backward_impl.
// ---------------------------------------------------------------------------
void Plan::run_blockwalk(void* const* in, void* const* out, bool forward, hipStream_t stream) {
  const auto& blocks = analyses_[0].blocks;
  const auto& topos = analyses_[0].topologies;
  const auto& lays = analyses_[0].layouts;
  const std::size_t nb = blocks.size();
  const std::size_t cs = fp64_ ? 16 : 8, rs = fp64_ ? 8 : 4;

  struct State {
    std::vector<void*> ptr;
    Vec map, topo, ext;
    Layout tl;
    bool is_real;
  };
  auto desc_of = [&](const State& s) {
    DDesc d = user_desc(s.ext, s.topo, s.tl, s.ptr.data());
    for (std::size_t a = 0; a < rank_; ++a) d.map[a] = static_cast<int32_t>(s.map[a]);
    return d;
  };
  auto remap = [&](const State& from, State& to) {
    RemapArgs ra{};
    ra.src = desc_of(from);
    ra.dst = desc_of(to);
    ra.total = 1;
    for (std::size_t k = 0; k < rank_; ++k) {
      const int a = static_cast<int>(to.map[k]);
      ra.iter_axis[k] = a;
      ra.iter_ext[k] = static_cast<int32_t>(to.ext[a]);
      ra.iter_div[k] = FastDiv(static_cast<uint32_t>(to.ext[a]));
      ra.total *= static_cast<int64_t>(to.ext[a]);
    }
    TraceRange range("dfft::TransBlock (pack + all2all + unpack)");
    mark(stream);
    check_hip(launch_remap(ra, from.is_real ? static_cast<int>(rs) : static_cast<int>(cs), stream), "remap");
    mark(stream);
  };
  int pp = 0;  // ping-pong selector of the next free buffer set
  auto fresh = [&](const Vec& map, const Vec& topo, const Vec& ext, Layout tl, bool is_real) {
    State s;
    s.ptr.assign(walk_bufs_[pp].begin(), walk_bufs_[pp].end());
    pp ^= 1;
    s.map = map;
    s.topo = topo;
    s.ext = ext;
    s.tl = tl;
    s.is_real = is_real;
    return s;
  };
  auto topo_layout = [&](std::size_t ti) {
    if (kind_ == TopologyType::Slab) return Layout::Right;
    return lays.at(ti) ? Layout::Right : Layout::Left;
  };
  // number of transpose blocks before block bi = index of the topology in use
  auto topo_index = [&](std::size_t bi) {
    std::size_t c = 0;
    for (std::size_t k = 0; k < bi; ++k)
      if (blocks[k].block_type == BlockType::Transpose) ++c;
    return c;
  };
  double scale = exec_scale(forward);
  bool scaled = false;
  Vec ident = identity_map(rank_);

  auto fft_block = [&](State& st, const BlockInfo& b, bool fwd) {
    // axes of the block: the innermost fft_dim storage axes of the mapped
    // array -- the last ones of a LayoutRight array, the first ones of a
    // LayoutLeft array -- listed so that the contiguous (R2C) axis is last
    Vec ax;
    if (view_layout_ == Layout::Right) {
      ax.assign(b.in_map.end() - static_cast<long>(b.fft_dim), b.in_map.end());
    } else {
      for (std::size_t k = b.fft_dim; k-- > 0;) ax.push_back(b.in_map[k]);
    }
    const bool r2c = real_ && b.block_idx == 0;
    std::vector<std::size_t> order;
    if (fwd) {
      for (auto it = ax.rbegin(); it != ax.rend(); ++it) order.push_back(*it);
    } else {
      for (auto a : ax) order.push_back(a);
    }
    for (std::size_t k = 0; k < order.size(); ++k) {
      const std::size_t a = order[k];
      PassSpec s;
      s.axis = static_cast<int>(a);
      s.n = static_cast<int>(std::max(gin_[a], gout_[a]));
      s.kind = kC2C;
      State dst = st;
      if (r2c && a == ax.back()) {
        s.kind = fwd ? kR2C : kC2R;
        Vec next_ext = st.ext;
        next_ext[a] = fwd ? gout_[a] : gin_[a];
        dst = fresh(st.map, st.topo, next_ext, st.tl, !fwd);
      }
      LinePass p = make_pass(s, desc_of(st), desc_of(dst), !fwd, scaled ? 1.0 : scale);
      scaled = true;
      TraceRange range("dfft::FFT block %zu axis %zu", static_cast<std::size_t>(b.block_idx), a);
      mark(stream);
      check_hip(launch_pass(p, stream), "line pass");
      mark(stream);
      st = dst;
    }
  };

  if (forward) {
    State st;
    st.ptr.assign(in, in + P_);
    st.map = ident;
    st.topo = in_topo_;
    st.ext = gin_;
    st.tl = in_tl_;
    st.is_real = real_;
    bool user_input = true;
    for (std::size_t bi = 0; bi < nb; ++bi) {
      const BlockInfo& b = blocks[bi];
      const std::size_t ti = topo_index(bi);
      if (b.block_type == BlockType::FFT) {
        if (st.map != b.in_map || user_input) {
          State t = fresh(b.in_map, st.topo, st.ext, st.tl, st.is_real);
          remap(st, t);
          st = t;
          user_input = false;
        }
        fft_block(st, b, true);
      } else {
        State t = fresh(b.out_map, topos.at(ti + 1), st.ext, topo_layout(ti + 1), st.is_real);
        remap(st, t);
        st = t;
        user_input = false;
      }
    }
    State o;
    o.ptr.assign(out, out + P_);
    o.map = ident;
    o.topo = out_topo_;
    o.ext = gout_;
    o.tl = out_tl_;
    o.is_real = false;
    remap(st, o);
  } else {
    State st;
    st.ptr.assign(in, in + P_);
    st.map = ident;
    st.topo = out_topo_;
    st.ext = gout_;
    st.tl = out_tl_;
    st.is_real = false;
    bool user_input = true;
    for (std::size_t bi = nb; bi-- > 0;) {
      const BlockInfo& b = blocks[bi];
      const std::size_t ti = topo_index(bi);
      if (b.block_type == BlockType::FFT) {
        if (st.map != b.out_map || user_input) {
          State t = fresh(b.out_map, st.topo, st.ext, st.tl, st.is_real);
          remap(st, t);
          st = t;
          user_input = false;
        }
        fft_block(st, b, false);
      } else {
        State t = fresh(b.in_map, topos.at(ti), st.ext, topo_layout(ti), st.is_real);
        remap(st, t);
        st = t;
        user_input = false;
      }
    }
    State o;
    o.ptr.assign(out, out + P_);
    o.map = ident;
    o.topo = in_topo_;
    o.ext = gin_;
    o.tl = in_tl_;
    o.is_real = real_;
    remap(st, o);
  }
  (void)contains;
}

}  // namespace dfft
