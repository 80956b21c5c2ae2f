#include "group.hpp"

#include <algorithm>
#include <cmath>
#include <string>

#include "buffer_check.hpp"
#include "plan.hpp"
#include "trace.hpp"

namespace dfft {

namespace {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(DFFT_ERROR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// one single-GPU LayoutRight array
DDesc array_desc(const Vec& ext, void* p) {
  DDesc d{};
  d.ndim = static_cast<int32_t>(ext.size());
  d.layout_left = 0;
  for (std::size_t a = 0; a < ext.size(); ++a) {
    d.ext[a] = static_cast<int32_t>(ext[a]);
    d.ntopo[a] = 1;
    d.rstride[a] = 0;
    d.base[a] = static_cast<int32_t>(ext[a]);
    d.rem[a] = 0;
    d.map[a] = static_cast<int32_t>(a);
    d.div_big[a] = FastDiv(static_cast<uint32_t>(ext[a] + 1));
    d.div_small[a] = FastDiv(static_cast<uint32_t>(std::max<std::size_t>(ext[a], 1)));
  }
  d.ptr[0] = static_cast<char*>(p);
  return d;
}

std::size_t volume(const Vec& v) {
  std::size_t n = 1;
  for (auto e : v) n *= e;
  return n;
}

}  // namespace

Group::Group(int count, const dfft_desc* descs) {
  if (count < 1 || count > 4096) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "group size must be in [1, 4096]");
  if (!descs) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "null descriptor array");
  check(hipGetDevice(&device_), "hipGetDevice");
  fp64_ = descs[0].precision == DFFT_FP64;
  try {
    for (int i = 0; i < count; ++i) {
      const dfft_desc& d = descs[i];
      const std::string who = "group member " + std::to_string(i) + ": ";
      if (d.rank < 1 || d.rank > DFFT_MAX_RANK) throw Error(DFFT_ERROR_INVALID_ARGUMENT, who + "bad rank");
      if (d.fft_rank < 1 || d.fft_rank > 3 || d.fft_rank > d.rank)
        throw Error(DFFT_ERROR_INVALID_ARGUMENT, who + "bad number of FFT axes");
      if ((d.precision == DFFT_FP64) != fp64_)
        throw Error(DFFT_ERROR_MISMATCH, who + "all members must share one precision");
      if (d.transform != DFFT_C2C && d.transform != DFFT_R2C)
        throw Error(DFFT_ERROR_INVALID_ARGUMENT, who + "bad transform type");
      if (d.layout != DFFT_LAYOUT_RIGHT) throw Error(DFFT_ERROR_UNSUPPORTED, who + "members must be LayoutRight");
      if (d.norm < DFFT_NORM_BACKWARD || d.norm > DFFT_NORM_NONE)
        throw Error(DFFT_ERROR_INVALID_ARGUMENT, who + "invalid normalization");
      if (d.nranks != 1) throw Error(DFFT_ERROR_INVALID_TOPOLOGY, who + "members are single-GPU transforms");
      Member m;
      m.rank = d.rank;
      m.real = d.transform == DFFT_R2C;
      for (int a = 0; a < d.rank; ++a) {
        if (d.in_topology[a] != 1 || d.out_topology[a] != 1)
          throw Error(DFFT_ERROR_INVALID_TOPOLOGY, who + "members are single-GPU transforms");
        m.gin.push_back(d.in_extents[a]);
        m.gout.push_back(d.out_extents[a]);
      }
      std::vector<int> axes;
      for (int k = 0; k < d.fft_rank; ++k) {
        const int a = d.axes[k] < 0 ? d.axes[k] + d.rank : d.axes[k];
        if (a < 0 || a >= d.rank || std::find(axes.begin(), axes.end(), a) != axes.end())
          throw Error(DFFT_ERROR_INVALID_ARGUMENT, who + "bad FFT axes");
        axes.push_back(a);
      }
      for (int a = 0; a < d.rank; ++a) {
        std::size_t want = m.gin[a];
        if (m.real && a == axes.back()) want = m.gin[a] / 2 + 1;
        if (m.gout[a] != want || m.gin[a] == 0)
          throw Error(DFFT_ERROR_INVALID_EXTENTS, who + "input and output extents do not match");
      }
      // forward: the R2C axis first, then the other axes innermost first;
      // backward: the reverse, so a C2R pass comes last
      std::vector<int> rest;
      for (int a : axes)
        if (!(m.real && a == axes.back())) rest.push_back(a);
      std::sort(rest.begin(), rest.end(), [](int x, int y) { return x > y; });
      if (m.real) m.fwd_axes.push_back(axes.back());
      for (int a : rest) m.fwd_axes.push_back(a);
      m.bwd_axes.assign(m.fwd_axes.rbegin(), m.fwd_axes.rend());
      double n = 1.0;
      for (int a : axes) n *= static_cast<double>(m.gin[a]);
      switch (d.norm) {
        case DFFT_NORM_BACKWARD: m.scale_f = 1.0; m.scale_b = 1.0 / n; break;
        case DFFT_NORM_FORWARD: m.scale_f = 1.0 / n; m.scale_b = 1.0; break;
        case DFFT_NORM_ORTHO: m.scale_f = m.scale_b = 1.0 / std::sqrt(n); break;
        default: break;
      }
      for (int a : axes) {
        const int len = static_cast<int>(m.gin[a]);
        if (!line_length_single_pass(len, fp64_))
          throw Error(DFFT_ERROR_UNSUPPORTED, who + "no line kernel for transform length " + std::to_string(len));
      }
      for (int a : axes) {
        const int len = static_cast<int>(m.gin[a]);
        if (!tw_.count(len)) tw_[len] = upload_twiddles(len);
      }
      if (m.real && m.fwd_axes.size() > 1) {
        const std::size_t bytes = volume(m.gout) * (fp64_ ? 16 : 8);
        if (hipMalloc(&m.scratch, bytes) != hipSuccess)
          throw Error(DFFT_ERROR_OUT_OF_MEMORY, "hipMalloc of a group member's scratch failed");
      }
      max_passes_ = std::max(max_passes_, static_cast<int>(m.fwd_axes.size()));
      members_.push_back(m);
    }
  } catch (...) {
    for (auto& m : members_)
      if (m.scratch) (void)hipFree(m.scratch);
    for (auto& kv : tw_) (void)hipFree(kv.second);
    throw;
  }
}

Group::~Group() {
  for (auto& ds : dir_) {
    if (ds.exec) (void)hipGraphExecDestroy(ds.exec);
    if (ds.graph) (void)hipGraphDestroy(ds.graph);
    for (auto* p : ds.items) (void)hipFree(p);
    for (auto* p : ds.first) (void)hipFree(p);
  }
  if (graph_stream_) (void)hipStreamDestroy(graph_stream_);
  for (auto& m : members_)
    if (m.scratch) (void)hipFree(m.scratch);
  for (auto& kv : tw_) (void)hipFree(kv.second);
}

// Line pass k of every member that has one (members in order).
std::vector<LinePass> Group::build_pass(const std::vector<void*>& in, const std::vector<void*>& out, bool forward,
                                        int k) const {
  std::vector<LinePass> passes;
  for (std::size_t i = 0; i < members_.size(); ++i) {
    const Member& m = members_[i];
    const std::vector<int>& order = forward ? m.fwd_axes : m.bwd_axes;
    const int np = static_cast<int>(order.size());
    if (k >= np) continue;
    const bool first = k == 0, last = k + 1 == np;
    const int axis = order[k];
    int kind = kC2C;
    if (m.real && forward && first) kind = kR2C;
    if (m.real && !forward && last) kind = kC2R;
    // source / destination arrays of this pass
    DDesc src, dst;
    if (forward) {
      src = first ? array_desc(m.real ? m.gin : m.gout, in[i]) : array_desc(m.gout, out[i]);
      dst = array_desc(m.gout, out[i]);
    } else if (!m.real) {
      src = array_desc(m.gout, first ? in[i] : out[i]);
      dst = array_desc(m.gout, out[i]);
    } else {
      src = array_desc(m.gout, first ? in[i] : m.scratch);
      dst = last ? array_desc(m.gin, out[i]) : array_desc(m.gout, m.scratch);
    }
    LinePass p{};
    p.in = src;
    p.out = dst;
    p.axis = axis;
    p.kind = kind;
    p.n = static_cast<int32_t>(m.gin[axis]);
    p.inverse = forward ? 0 : 1;
    p.scale = last ? (forward ? m.scale_f : m.scale_b) : 1.0;
    p.herm_weight = 0.0;
    p.tw = tw_.at(p.n);
    p.nbatch = 0;
    int64_t nl = 1;
    for (int a = 0; a < m.rank; ++a) {
      if (a == axis) continue;
      p.batch_axis[p.nbatch] = a;
      p.batch_ext[p.nbatch] = src.ext[a];
      p.batch_div[p.nbatch] = FastDiv(static_cast<uint32_t>(std::max(src.ext[a], 1)));
      nl *= src.ext[a];
      ++p.nbatch;
    }
    if (nl >= (int64_t(1) << 31)) throw Error(DFFT_ERROR_UNSUPPORTED, "too many lines in one pass");
    p.nlines = static_cast<int32_t>(nl);
    p.line_begin = 0;
    p.lines_fast = axis != m.rank - 1 ? 1 : 0;
    int B = 0;
    std::size_t lds = 0;
    p.values_per_thread = 8;
    p.lines_per_block = group_line_shape(p.n, fp64_, lds_cap_, kGroupThreads, &B, &lds) ? B : 0;
    if (!fp64_ && dual_ && kind == kC2C && p.lines_fast && group_dual_shape(p.n, lds_cap_, &B, &lds)) {
      LinePass q = p;
      q.dual_lines = 1;
      if (dual_lines_apply(q, B)) {
        p = q;
        p.lines_per_block = B;
      }
    }
    passes.push_back(p);
  }
  return passes;
}

void Group::build_tables(const std::vector<void*>& in, const std::vector<void*>& out, int dir, Tables& t) const {
  const std::size_t np = static_cast<std::size_t>(max_passes_);
  const std::size_t ns = 2 * np;  // slots: (pass, one or two lines per lane)
  t.items.assign(ns, {});
  t.first.assign(ns, {});
  t.count.assign(ns, 0);
  t.total.assign(ns, 0);
  t.grid.assign(ns, 0);
  t.nt.assign(ns, 0);
  t.lds.assign(ns, 0);
  t.lanes.assign(ns, kGroupThreads);
  t.single.assign(np, {});
  for (std::size_t k = 0; k < np; ++k) {
    std::vector<LinePass> all = build_pass(in, out, dir == 0, static_cast<int>(k));
    double bytes = 0.0;  // array bytes the pass streams through, all members
    for (const LinePass& p : all) {
      double e = 1.0;
      for (int a = 0; a < p.in.ndim; ++a) e *= static_cast<double>(p.in.ext[a]);
      bytes += e * (p.kind == kR2C ? 0.5 : 1.0) * (fp64_ ? 16.0 : 8.0);
    }
    const int nt = nt_ >= 0 ? nt_ : (bytes >= 1024.0 * 1024 * 1024 ? 1 : 0);
    for (LinePass& p : all) p.nontemporal = nt;
    std::vector<LinePass> grouped[2];
    std::vector<int32_t> first[2] = {{0}, {0}};
    std::size_t lds[2] = {0, 0};
    bool contiguous = true;  // every one-line member walks its lines with lanes along the line
    for (const LinePass& p : all) {
      int B = 0;
      std::size_t l = 0;
      if (p.nlines == 0) continue;
      const int v = p.dual_lines ? 1 : 0;
      const bool shaped = v ? group_dual_shape(p.n, lds_cap_, &B, &l)
                            : group_line_shape(p.n, fp64_, lds_cap_, kGroupThreads, &B, &l);
      if (shaped) {
        grouped[v].push_back(p);
        if (v == 0 && p.lines_fast) contiguous = false;
      } else {
        LinePass q = p;
        q.lines_per_block = 0;  // launcher heuristic
        q.values_per_thread = 0;
        q.dual_lines = 0;
        t.single[k].push_back(q);
      }
    }
    for (int v = 0; v < 2; ++v) {
      // a one-line launch of contiguous lines only takes lanes_ lanes per block;
      // strided lines keep full blocks (more adjacent lines per access)
      const int lanes = v == 0 && contiguous ? lanes_ : kGroupThreads;
      for (LinePass& p : grouped[v]) {
        int B = 0;
        std::size_t l = 0;
        if (v) group_dual_shape(p.n, lds_cap_, &B, &l);
        else group_line_shape(p.n, fp64_, lds_cap_, lanes, &B, &l);
        p.lines_per_block = B;
        const int64_t nb = first[v].back() + (static_cast<int64_t>(p.nlines) + B - 1) / B;
        if (nb >= (int64_t(1) << 31)) throw Error(DFFT_ERROR_UNSUPPORTED, "too many blocks in one grouped pass");
        first[v].push_back(static_cast<int32_t>(nb));
        lds[v] = std::max(lds[v], l);
      }
      const std::size_t slot = 2 * k + static_cast<std::size_t>(v);
      t.nt[slot] = v == 0 && contiguous && contig_nt_ >= 0 ? contig_nt_ : nt;
      t.lanes[slot] = lanes;
      t.count[slot] = static_cast<int>(grouped[v].size());
      t.total[slot] = first[v].back();
      t.lds[slot] = lds[v];
      if (grouped[v].empty()) continue;
      const int resident = group_grid(lds[v], fp64_, v == 1, lanes);
      if (resident <= 0) throw Error(DFFT_ERROR_HIP, "occupancy query of the grouped kernel failed");
      t.grid[slot] = grid_mult_ > 0 ? resident * grid_mult_ : first[v].back();
      t.items[slot] = std::move(grouped[v]);
      t.first[slot] = std::move(first[v]);
    }
  }
}

void Group::upload(const std::vector<void*>& in, const std::vector<void*>& out, int dir, hipStream_t s) {
  DirState& ds = dir_[dir];
  check(hipStreamSynchronize(s), "hipStreamSynchronize");
  build_tables(in, out, dir, ds.tables);
  const std::size_t ns = ds.tables.items.size();
  ds.items.resize(ns, nullptr);
  ds.first.resize(ns, nullptr);
  for (std::size_t slot = 0; slot < ns; ++slot) {
    if (ds.tables.count[slot] == 0) continue;
    if (!ds.items[slot]) {
      check(hipMalloc(&ds.items[slot], sizeof(LinePass) * members_.size()), "hipMalloc");
      check(hipMalloc(&ds.first[slot], sizeof(int32_t) * (members_.size() + 1)), "hipMalloc");
    }
    check(hipMemcpy(ds.items[slot], ds.tables.items[slot].data(), sizeof(LinePass) * ds.tables.items[slot].size(),
                    hipMemcpyHostToDevice),
          "hipMemcpy");
    check(hipMemcpy(ds.first[slot], ds.tables.first[slot].data(), sizeof(int32_t) * ds.tables.first[slot].size(),
                    hipMemcpyHostToDevice),
          "hipMemcpy");
  }
  ds.in = in;
  ds.out = out;
  ds.built = true;
}

void Group::launch_tables(const Tables& t, LinePass* const* items, int32_t* const* first, int dir,
                          hipStream_t s) const {
  TraceRange range("dfft::Group %s (%zu members)", dir == 0 ? "forward" : "backward", members_.size());
  for (int k = 0; k < max_passes_; ++k) {
    for (int v = 0; v < 2; ++v) {
      const int slot = 2 * k + v;
      if (t.count[slot] > 0)
        check(launch_group_pass(items[slot], first[slot], t.count[slot], t.total[slot], t.grid[slot], t.lds[slot],
                                fp64_, t.nt[slot] != 0, v == 1, t.lanes[slot], stage_ != 0, s),
              "grouped line pass");
    }
    for (const LinePass& p : t.single[k]) check(launch_line_pass(p, fp64_, s), "line pass");
  }
}

void Group::launch_all(int dir, hipStream_t s) {
  const DirState& ds = dir_[dir];
  launch_tables(ds.tables, ds.items.data(), ds.first.data(), dir, s);
}

int Group::num_launches(int direction) const {
  const DirState& ds = dir_[direction == DFFT_FORWARD ? 0 : 1];
  if (ds.built) {
    // the launches of the tables in use: grouped slots plus single passes
    int n = 0;
    for (int c : ds.tables.count) n += c > 0 ? 1 : 0;
    for (const auto& v : ds.tables.single) n += static_cast<int>(v.size());
    return n;
  }
  // before the first execute: one grouped launch per pass (two where some of
  // its members take two lines per lane) plus the members without a kernel
  int n = 0;
  for (int k = 0; k < max_passes_; ++k) {
    bool grouped = false;
    for (const Member& m : members_) {
      if (k >= static_cast<int>(m.fwd_axes.size())) continue;
      const std::vector<int>& order = direction == DFFT_FORWARD ? m.fwd_axes : m.bwd_axes;
      int B = 0;
      std::size_t l = 0;
      if (group_line_shape(static_cast<int>(m.gin[order[k]]), fp64_, lds_cap_, kGroupThreads, &B, &l)) grouped = true;
      else ++n;
    }
    if (grouped) ++n;
  }
  return n;
}

void Group::invalidate() {
  for (auto& ds : dir_) {
    if (ds.exec) {
      if (graph_stream_) (void)hipStreamSynchronize(graph_stream_);
      (void)hipGraphExecDestroy(ds.exec);
      ds.exec = nullptr;
    }
    if (ds.graph) {
      (void)hipGraphDestroy(ds.graph);
      ds.graph = nullptr;
    }
    ds.built = false;
  }
}

void Group::set_option(const std::string& key, long value) {
  if (key == "lds_cap") {
    if (value < 1024 || value > static_cast<long>(kGroupLds))
      throw Error(DFFT_ERROR_INVALID_ARGUMENT, "lds_cap must be in [1024, 65536]");
    lds_cap_ = static_cast<std::size_t>(value);
  } else if (key == "nt") {
    if (value < -1 || value > 1) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "nt must be -1, 0 or 1");
    nt_ = static_cast<int>(value);
  } else if (key == "grid_mult") {
    if (value < 0 || value > 64) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "grid_mult must be in [0, 64]");
    grid_mult_ = static_cast<int>(value);
  } else if (key == "dual") {
    if (value < 0 || value > 1) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "dual must be 0 or 1");
    dual_ = static_cast<int>(value);
  } else if (key == "contig_nt") {
    if (value < -1 || value > 1) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "contig_nt must be -1, 0 or 1");
    contig_nt_ = static_cast<int>(value);
  } else if (key == "stage") {
    if (value < 0 || value > 1) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "stage must be 0 or 1");
    stage_ = static_cast<int>(value);
  } else if (key == "lanes") {
    if (value != 64 && value != 128 && value != kGroupThreads)
      throw Error(DFFT_ERROR_INVALID_ARGUMENT, "lanes must be 64, 128 or 256");
    lanes_ = static_cast<int>(value);
  } else {
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "unknown group option: " + key);
  }
  invalidate();
}

long Group::get_option(const std::string& key) const {
  if (key == "lds_cap") return static_cast<long>(lds_cap_);
  if (key == "grid_mult") return grid_mult_;
  if (key == "nt") return nt_;
  if (key == "dual") return dual_;
  if (key == "lanes") return lanes_;
  if (key == "stage") return stage_;
  if (key == "contig_nt") return contig_nt_;
  throw Error(DFFT_ERROR_INVALID_ARGUMENT, "unknown group option: " + key);
}

void Group::autotune(void* const* in, void* const* out, hipStream_t s) {
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "hipEventCreate");
  check(hipEventCreate(&e1), "hipEventCreate");
  float best = -1.0f;
  std::size_t best_cap = lds_cap_;
  int best_mult = grid_mult_, best_nt = nt_, best_dual = dual_, best_lanes = lanes_, best_stage = stage_;
  int best_cnt = contig_nt_;
  // one configuration: warm-up round trip (builds the tables), then the best
  // of two trials of three round trips (single trials are noisy)
  auto time_config = [&]() {
    execute(in, out, DFFT_FORWARD, false, s);
    execute(out, in, DFFT_BACKWARD, false, s);
    float ms = -1.0f;
    for (int trial = 0; trial < 2; ++trial) {
      check(hipEventRecord(e0, s), "hipEventRecord");
      for (int rep = 0; rep < 3; ++rep) {
        execute(in, out, DFFT_FORWARD, false, s);
        execute(out, in, DFFT_BACKWARD, false, s);
      }
      check(hipEventRecord(e1, s), "hipEventRecord");
      check(hipEventSynchronize(e1), "hipEventSynchronize");
      float t = 0.0f;
      check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
      if (ms < 0.0f || t < ms) ms = t;
    }
    return ms;
  };
  try {
    set_option("lanes", kGroupThreads);
    set_option("stage", 1);
    set_option("contig_nt", -1);
    for (std::size_t cap : {std::size_t(65536), std::size_t(49152), std::size_t(32768), std::size_t(24576)}) {
      for (int multnt = 0; multnt < (fp64_ ? 8 : 16); ++multnt) {
        const int mult = (multnt & 3) == 3 ? 0 : (1 << (multnt & 3));
        const int nt = (multnt >> 2) & 1;
        const int dual = fp64_ ? dual_ : (multnt >> 3);
        set_option("lds_cap", static_cast<long>(cap));
        set_option("grid_mult", mult);
        set_option("nt", nt);
        set_option("dual", dual);
        const float ms = time_config();
        if (best < 0.0f || ms < best) {
          best = ms;
          best_cap = cap;
          best_mult = mult;
          best_nt = nt;
          best_dual = dual;
        }
      }
    }
    // then the launches of contiguous lines: block size, FP32 staging and
    // their own access kind, with every grid size (the LDS cap, the dual
    // choice and the other launches' access kind found above)
    set_option("lds_cap", static_cast<long>(best_cap));
    set_option("dual", best_dual);
    set_option("nt", best_nt);
    best_lanes = kGroupThreads;
    best_stage = 1;
    best_cnt = -1;
    for (int stage = fp64_ ? 1 : 0; stage < 2; ++stage) {
      for (int lanes : {256, 128, 64}) {
        for (int multnt = 0; multnt < 8; ++multnt) {
          const int mult = (multnt & 3) == 3 ? 0 : (1 << (multnt & 3));
          const int cnt = multnt >> 2;
          set_option("stage", stage);
          set_option("lanes", lanes);
          set_option("grid_mult", mult);
          set_option("contig_nt", cnt);
          const float ms = time_config();
          if (ms < best) {
            best = ms;
            best_stage = stage;
            best_lanes = lanes;
            best_mult = mult;
            best_cnt = cnt;
          }
        }
      }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  set_option("lds_cap", static_cast<long>(best_cap));
  set_option("grid_mult", best_mult);
  set_option("nt", best_nt);
  set_option("dual", best_dual);
  set_option("lanes", best_lanes);
  set_option("stage", best_stage);
  set_option("contig_nt", best_cnt);
}

int Group::check_call(const std::vector<void*>& vin, const std::vector<void*>& vout, int direction) const {
  if (direction != DFFT_FORWARD && direction != DFFT_BACKWARD)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "Invalid FFT direction specified.");
  const int dir = direction == DFFT_FORWARD ? 0 : 1;
  for (std::size_t i = 0; i < members_.size(); ++i) {
    if (!vin[i] || !vout[i]) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "null member array");
    if (members_[i].real && vin[i] == vout[i])
      throw Error(DFFT_ERROR_ALIAS, "an R2C member cannot run in place");
  }
  // every member's allocations must hold its arrays
  const std::size_t cs = fp64_ ? 16 : 8;
  for (std::size_t i = 0; i < members_.size(); ++i) {
    const Member& m = members_[i];
    const std::size_t nreal = volume(m.gin) * (m.real ? cs / 2 : cs), ncplx = volume(m.gout) * cs;
    const std::string who = "group member " + std::to_string(i);
    check_device_extent(vin[i], dir == 0 ? nreal : ncplx, who + " input");
    check_device_extent(vout[i], dir == 0 ? ncplx : nreal, who + " output");
  }
  return dir;
}

void Group::execute(void* const* in, void* const* out, int direction, bool use_graph, hipStream_t s) {
  if (!in || !out) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "null pointer table");
  if (direction != DFFT_FORWARD && direction != DFFT_BACKWARD)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "Invalid FFT direction specified.");
  const int dir = direction == DFFT_FORWARD ? 0 : 1;
  std::vector<void*> vin(in, in + members_.size()), vout(out, out + members_.size());
  int cur = -1;
  check(hipGetDevice(&cur), "hipGetDevice");
  if (cur != device_) check(hipSetDevice(device_), "hipSetDevice");
  DirState& ds = dir_[dir];
  if (!ds.built || ds.in != vin || ds.out != vout) {
    // a new pointer set: checked once, then the tables are rebuilt
    check_call(vin, vout, direction);
    if (ds.exec) {
      check(hipStreamSynchronize(graph_stream_), "hipStreamSynchronize");
      (void)hipGraphExecDestroy(ds.exec);
      ds.exec = nullptr;
    }
    if (ds.graph) {
      (void)hipGraphDestroy(ds.graph);
      ds.graph = nullptr;
    }
    upload(vin, vout, dir, s);
  }
  if (!use_graph) {
    launch_all(dir, s);
    return;
  }
  if (!ds.exec) {
    // eager run for this call; the capture below records the same launches
    launch_all(dir, s);
    if (!graph_stream_) check(hipStreamCreate(&graph_stream_), "hipStreamCreate");
    check(hipStreamSynchronize(s), "hipStreamSynchronize");
    check(hipStreamBeginCapture(graph_stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      launch_all(dir, graph_stream_);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(graph_stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    check(hipStreamEndCapture(graph_stream_, &ds.graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&ds.exec, ds.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    return;
  }
  hipEvent_t ev;
  check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
  check(hipEventRecord(ev, s), "hipEventRecord");
  check(hipStreamWaitEvent(graph_stream_, ev, 0), "hipStreamWaitEvent");
  check(hipGraphLaunch(ds.exec, graph_stream_), "hipGraphLaunch");
  check(hipEventRecord(ev, graph_stream_), "hipEventRecord");
  check(hipStreamWaitEvent(s, ev, 0), "hipStreamWaitEvent");
  (void)hipEventDestroy(ev);
}

}  // namespace dfft
