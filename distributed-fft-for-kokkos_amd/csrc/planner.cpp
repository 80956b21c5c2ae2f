// Host planner: see planner.hpp.  Each function cites the reference function
// it restates.
#include "planner.hpp"

#include <algorithm>
#include <set>

#include "dfft.h"

namespace dfft {

namespace {
[[noreturn]] void fail(int code, const std::string& msg) { throw Error(code, msg); }
std::string str(const Vec& v) {
  std::string s = "(";
  for (std::size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s + ")";
}
bool contains(const Vec& v, std::size_t x) { return std::find(v.begin(), v.end(), x) != v.end(); }
std::size_t count_non_ones(const Vec& v) {
  return std::count_if(v.begin(), v.end(), [](std::size_t x) { return x != 1; });
}
Vec extract_different_indices(const Vec& a, const Vec& b) {  // ContainerAnalyses.hpp:37-54
  if (a.size() != b.size()) fail(DFFT_ERROR_INVALID_TOPOLOGY, "Containers must have the same size.");
  Vec d;
  for (std::size_t i = 0; i < a.size(); ++i)
    if (a[i] != b[i]) d.push_back(i);
  return d;
}
std::size_t different_value_set_size(const Vec& a, const Vec& b) {  // :63-78
  std::set<std::size_t> s;
  for (std::size_t i = 0; i < a.size(); ++i) {
    s.insert(a[i]);
    s.insert(b[i]);
  }
  return s.size();
}
Vec extract_non_one_indices(const Vec& a, const Vec& b) {  // :89-104
  Vec d;
  for (std::size_t i = 0; i < a.size(); ++i)
    if (a[i] != 1 || b[i] != 1) d.push_back(i);
  return d;
}
Vec extract_non_one_values(const Vec& a) {  // :113-125
  Vec d;
  for (auto x : a)
    if (x != 1) d.push_back(x);
  return d;
}
bool has_identical_non_ones(const Vec& v) {  // :133-145
  if (count_non_ones(v) < 2) return false;
  return v.size() == 2 && v[0] == v[1];
}
Vec swap_elements(Vec v, std::size_t i, std::size_t j) {  // :157-164
  std::swap(v.at(i), v.at(j));
  return v;
}
Vec reversed(Vec v) {
  std::reverse(v.begin(), v.end());
  return v;
}
}  // namespace

bool BlockInfo::operator==(const BlockInfo& o) const {
  return in_topology == o.in_topology && out_topology == o.out_topology &&
         in_extents == o.in_extents && out_extents == o.out_extents &&
         buffer_extents == o.buffer_extents && in_map == o.in_map && out_map == o.out_map &&
         in_axis == o.in_axis && out_axis == o.out_axis && fft_dim == o.fft_dim &&
         comm_axis == o.comm_axis && block_type == o.block_type && block_idx == o.block_idx;
}

std::size_t total_size(const Vec& v) {
  std::size_t s = 1;
  for (auto x : v) s *= x;
  return s;
}

Vec identity_map(std::size_t n) {
  Vec v(n);
  for (std::size_t i = 0; i < n; ++i) v[i] = i;
  return v;
}

std::size_t get_index(const Vec& v, std::size_t value) {
  for (std::size_t i = 0; i < v.size(); ++i)
    if (v[i] == value) return i;
  fail(DFFT_ERROR_INVALID_TOPOLOGY, "value is not included in the container");
}

Vec convert_negative_axes(const std::vector<int>& axes, std::size_t rank) {
  Vec out;
  for (int a : axes) {
    int r = static_cast<int>(rank);
    if (a < -r || a >= r) fail(DFFT_ERROR_INVALID_ARGUMENT, "axis out of range");
    out.push_back(static_cast<std::size_t>(a < 0 ? a + r : a));
  }
  return out;
}

// kokkos-fft get_map_axes: FFT axes innermost.
void get_map_axes(Layout layout, std::size_t rank, const Vec& axes, Vec& map, Vec& map_inv) {
  map.clear();
  if (layout == Layout::Right) {
    for (std::size_t i = 0; i < rank; ++i)
      if (!contains(axes, i)) map.push_back(i);
    for (auto a : axes) map.push_back(a);
  } else {
    for (auto it = axes.rbegin(); it != axes.rend(); ++it) map.push_back(*it);
    for (std::size_t i = 0; i < rank; ++i)
      if (!contains(axes, i)) map.push_back(i);
  }
  map_inv.assign(rank, 0);
  for (std::size_t i = 0; i < rank; ++i) map_inv[i] = get_index(map, i);
}

Vec compute_mapped_extents(const Vec& e, const Vec& m) {
  Vec out(m.size());
  for (std::size_t i = 0; i < m.size(); ++i) out[i] = e.at(m[i]);
  return out;
}

// ContainerAnalyses.hpp:181-226
Vec merge_topology(const Vec& in, const Vec& out) {
  if (total_size(in) != total_size(out))
    fail(DFFT_ERROR_INVALID_TOPOLOGY, "Input and output topologies must have the same size.");
  if (total_size(in) == 1) return in;
  if (extract_different_indices(in, out).size() != 2)
    fail(DFFT_ERROR_INVALID_TOPOLOGY,
         "Input and output topologies must differ exactly two positions: in_topology " + str(in) +
             ", out_topology " + str(out));
  Vec m(in.size());
  for (std::size_t i = 0; i < in.size(); ++i) m[i] = std::max(in[i], out[i]);
  return m;
}

// ContainerAnalyses.hpp:243-257
std::size_t diff_topology(const Vec& in, const Vec& out) {
  if (total_size(in) == 1 && total_size(out) == 1) return 1;
  auto d = extract_different_indices(in, out);
  if (d.size() != 1)
    fail(DFFT_ERROR_INVALID_TOPOLOGY, "Input and output topologies must differ at exactly one position.");
  return std::max(in[d[0]], out[d[0]]);
}

// Topologies.hpp:25-40
TopologyType to_topology_type(const Vec& t) {
  if (total_size(t) == 0) return TopologyType::Empty;
  switch (count_non_ones(t)) {
    case 0: return TopologyType::Shared;
    case 1: return TopologyType::Slab;
    case 2: return TopologyType::Pencil;
    case 3: return TopologyType::Brick;
    default: return TopologyType::Invalid;
  }
}

// Topologies.hpp:65-89
TopologyType get_common_topology_type(const Vec& a, const Vec& b) {
  auto ta = to_topology_type(a), tb = to_topology_type(b);
  if (ta == TopologyType::Empty || tb == TopologyType::Empty) return TopologyType::Empty;
  if (ta == tb && ta != TopologyType::Invalid) return ta;
  return TopologyType::Invalid;
}

// Topologies.hpp:108-132
void slab_in_out_axes(const Vec& in, const Vec& out, std::size_t& in_axis, std::size_t& out_axis) {
  if (total_size(in) != total_size(out))
    fail(DFFT_ERROR_INVALID_TOPOLOGY, "Input and output topologies must have the same size.");
  if (to_topology_type(in) != TopologyType::Slab || to_topology_type(out) != TopologyType::Slab)
    fail(DFFT_ERROR_INVALID_TOPOLOGY, "Input and output topologies must be slab topologies.");
  in_axis = out_axis = 0;
  for (std::size_t i = 0; i < in.size(); ++i) {
    if (in[i] > 1 && out[i] == 1) out_axis = i;
    if (in[i] == 1 && out[i] > 1) in_axis = i;
  }
}

// Topologies.hpp:150-175
void pencil_in_out_axes(const Vec& in, const Vec& out, std::size_t& in_axis, std::size_t& out_axis) {
  if (total_size(in) == 1 || total_size(out) == 1)
    fail(DFFT_ERROR_INVALID_TOPOLOGY,
         "Input and output topologies must have at least one non-trivial dimension.");
  if (total_size(in) != total_size(out))
    fail(DFFT_ERROR_INVALID_TOPOLOGY, "Input and output topologies must have the same size.");
  in_axis = out_axis = 0;
  for (std::size_t i = 0; i < in.size(); ++i) {
    if (in[i] != out[i]) {
      if (in[i] == 1) in_axis = i;
      if (out[i] == 1) out_axis = i;
    }
  }
}

// Topologies.hpp:190-242
Vec propose_mid_array(const Vec& in, const Vec& out) {
  auto diff = extract_different_indices(in, out);
  auto nvals = different_value_set_size(in, out);
  auto non_one = extract_non_one_indices(in, out);
  if (non_one.size() < 3)
    fail(DFFT_ERROR_INVALID_TOPOLOGY,
         "The total number of non-one elements either in Input and output topologies must be three.");
  if (diff.size() < 3 && nvals == 3)
    fail(DFFT_ERROR_INVALID_TOPOLOGY, "Input and output topologies must differ exactly three positions.");
  Vec in_trim(in.size(), 0), out_trim(in.size(), 0);
  for (auto d : diff) {
    in_trim[d] = in[d];
    out_trim[d] = out[d];
  }
  std::size_t one_in = get_index(in_trim, 1), one_out = get_index(out_trim, 1);
  for (std::size_t i = 0; i < non_one.size(); ++i) {
    for (std::size_t j = i + 1; j < non_one.size(); ++j) {
      Vec mid = swap_elements(in, non_one[i], non_one[j]);
      std::size_t one_mid = get_index(mid, 1);
      auto mi = extract_different_indices(mid, in);
      auto mo = extract_different_indices(mid, out);
      if (mi.size() == 2 && mo.size() == 2 && !(one_mid == one_in || one_mid == one_out)) {
        if ((mid[mi[0]] == 1 || mid[mi[1]] == 1) && (mid[mo[0]] == 1 || mid[mo[1]] == 1))
          return mid;
      }
    }
  }
  return out;
}

// Topologies.hpp:258-331
std::vector<Vec> decompose_axes(const std::vector<Vec>& topologies, const Vec& axes) {
  Vec rev = reversed(axes);
  std::vector<Vec> all;
  for (const auto& t : topologies) {
    Vec ready;
    for (auto a : rev) {
      if (t.at(a) > 1) break;
      ready.push_back(a);
    }
    all.push_back(reversed(ready));
    for (auto a : ready) {
      auto it = std::find(rev.begin(), rev.end(), a);
      if (it != rev.end()) rev.erase(it);
    }
  }
  std::size_t total = 0;
  for (const auto& r : all) total += r.size();
  if (total != axes.size()) fail(DFFT_ERROR_INVALID_TOPOLOGY, "Axes are not decomposed correctly");
  return all;
}

// Topologies.hpp:351-399
std::size_t compute_trans_axis(const Vec& in, const Vec& out, std::size_t first_non_one) {
  auto in_non = extract_non_one_values(in), out_non = extract_non_one_values(out);
  if (in_non.size() != 2 || out_non.size() != 2)
    fail(DFFT_ERROR_INVALID_TOPOLOGY, "Input and output topologies must have exactly two non-one elements.");
  if (has_identical_non_ones(in_non) || has_identical_non_ones(out_non))
    fail(DFFT_ERROR_INVALID_TOPOLOGY,
         "Input and output topologies must not have identical non-one elements.");
  auto d = extract_different_indices(in, out);
  if (d.size() != 2)
    fail(DFFT_ERROR_INVALID_TOPOLOGY, "Input and output topologies must differ exactly two positions");
  std::size_t ex = 0;
  for (auto i : d)
    if (in[i] > 1) {
      ex = in[i];
      break;
    }
  return ex == first_non_one ? 0 : 1;
}

// Topologies.hpp:420-618
std::vector<Vec> get_all_slab_topologies(const Vec& in, const Vec& out, const Vec& axes) {
  const std::size_t dim = in.size(), fdim = axes.size();
  if (fdim < 1 || fdim > 3 || dim < 2 || dim < fdim)
    fail(DFFT_ERROR_INVALID_ARGUMENT, "invalid dimensions");
  if (to_topology_type(in) != TopologyType::Slab || to_topology_type(out) != TopologyType::Slab)
    fail(DFFT_ERROR_INVALID_TOPOLOGY, "Input and output topologies must be slab topologies.");
  std::vector<Vec> topos{in};
  const std::size_t p = total_size(in);
  auto add_at = [&](std::size_t axis) {
    Vec t(dim, 1);
    t[axis] = p;
    if (topos.back() != t) topos.push_back(t);
  };
  auto add_at_first_one = [&](const Vec& topo) {
    for (std::size_t i = 0; i < dim; ++i)
      if (topo[i] == 1) {
        add_at(i);
        break;
      }
  };
  auto finalize = [&]() {
    if (topos.back() != out) topos.push_back(out);
    return topos;
  };
  if (dim == 2 && fdim == 2) {
    if (in == out) {
      add_at_first_one(in);
    } else if (in[axes.back()] != 1) {
      add_at_first_one(in);
      if (topos.back() != in) topos.push_back(in);
    }
    return finalize();
  }
  if (dim == 3 && fdim == 3) {
    if (in == out) {
      if (in[axes.back()] != 1) add_at(axes[0]);
      else add_at(axes.back());
    } else {
      if (in[axes.back()] != 1) {
        add_at(axes[0]);
        if (out[axes.front()] != 1) add_at_first_one(Vec(topos.back()));
      } else if (in[axes[1]] != 1) {
        add_at(axes.back());
      } else if (out[axes.front()] != 1) {
        add_at_first_one(Vec(topos.back()));
      }
    }
    return finalize();
  }
  Vec rest = axes;
  auto is_ready = [&](const Vec& t) {
    for (auto a : rest)
      if (t[a] > 1) return false;
    return true;
  };
  if (is_ready(in) || is_ready(out)) return finalize();
  if (fdim == 1) {
    for (std::size_t i = 0; i < dim; ++i)
      if (!contains(rest, i)) {
        add_at(i);
        break;
      }
    return finalize();
  }
  if (dim > 3 && fdim == 3) {
    Vec sorted_axes = axes;
    std::sort(sorted_axes.begin(), sorted_axes.end());
    std::reverse(rest.begin(), rest.end());
    Vec ready;
    for (auto a : rest) {
      if (in[a] > 1) break;
      ready.push_back(a);
    }
    if (ready.empty()) {
      add_at(rest.back());
      for (auto a : rest) {
        if (topos.back()[a] > 1) break;
        ready.push_back(a);
      }
    }
    for (auto a : ready) {
      auto it = std::find(rest.begin(), rest.end(), a);
      if (it != rest.end()) rest.erase(it);
    }
    if (!is_ready(out)) {
      for (auto a : sorted_axes)
        if (!contains(rest, a)) {
          add_at(a);
          break;
        }
    }
    return finalize();
  }
  std::reverse(rest.begin(), rest.end());
  Vec ready;
  for (auto a : rest) {
    if (in[a] > 1) break;
    ready.push_back(a);
  }
  for (auto a : ready) {
    auto it = std::find(rest.begin(), rest.end(), a);
    if (it != rest.end()) rest.erase(it);
  }
  if (!is_ready(out)) {
    for (std::size_t i = 0; i < dim; ++i)
      if (!contains(rest, i)) {
        add_at(i);
        break;
      }
  }
  return finalize();
}

// Topologies.hpp:642-811
void get_all_pencil_topologies(const Vec& in_topology, const Vec& out_topology, const Vec& axes,
                               Layout in_layout, Layout out_layout, std::vector<Vec>& topos,
                               Vec& trans_axes, Vec& layouts) {
  const std::size_t dim = in_topology.size(), fdim = axes.size();
  if (fdim < 1 || fdim > 3 || dim < 3 || dim < fdim)
    fail(DFFT_ERROR_INVALID_ARGUMENT, "invalid dimensions");
  if (to_topology_type(in_topology) != TopologyType::Pencil ||
      to_topology_type(out_topology) != TopologyType::Pencil)
    fail(DFFT_ERROR_INVALID_TOPOLOGY, "Input and output topologies must be pencil topologies.");
  Vec rest = axes;
  Vec non_ones = extract_non_one_values(in_topology);
  bool same = has_identical_non_ones(non_ones);
  Vec in_t = in_topology, out_t = out_topology;
  if (same) {
    Vec dummies{2, 3};
    std::size_t c = 0;
    for (std::size_t i = 0; i < dim; ++i)
      if (in_t[i] > 1) in_t[i] = dummies[c++];
    c = 0;
    if (in_layout != out_layout) std::reverse(dummies.begin(), dummies.end());
    for (std::size_t i = 0; i < dim; ++i)
      if (out_t[i] > 1) out_t[i] = dummies[c++];
  }
  Vec nz = extract_non_one_values(in_t);
  std::size_t first_non_one = in_layout == Layout::Right ? nz.at(0) : nz.at(1);
  topos.assign(1, in_t);
  trans_axes.clear();
  auto get_layout = [&](const Vec& t) -> std::size_t {
    return extract_non_one_values(t).at(0) == first_non_one ? 1 : 0;
  };
  layouts.assign(1, get_layout(in_t));
  auto add = [&](const Vec& t) {
    if (topos.back() != t) {
      trans_axes.push_back(compute_trans_axis(topos.back(), t, first_non_one));
      topos.push_back(t);
      layouts.push_back(get_layout(t));
    }
  };
  auto try_add_mid = [&]() {
    try {
      add(propose_mid_array(topos.back(), out_t));
    } catch (const Error&) {
    }
  };
  auto to_original = [&]() {
    if (same) {
      for (auto& t : topos)
        for (auto& x : t)
          if (x > 1) x = non_ones[0];
    }
  };
  auto is_ready = [&](const Vec& t, const Vec& cur) {
    for (auto a : cur)
      if (t[a] != 1) return false;
    return true;
  };
  if (is_ready(in_t, rest) || is_ready(out_t, rest)) {
    try_add_mid();
    add(out_t);
    to_original();
    return;
  }
  std::reverse(rest.begin(), rest.end());
  Vec shuffled = in_t;
  std::size_t last_axis = rest.front(), first_axis = rest.back();
  if (in_t[last_axis] == 1) rest.erase(rest.begin());
  if (out_t[first_axis] == 1 && !rest.empty()) rest.pop_back();
  for (auto axis : rest) {
    std::size_t swap_idx = 0;
    if (in_t == out_t) {
      for (std::size_t idx = 0; idx < dim; ++idx)
        if (shuffled[idx] == 1 && idx != axis) {
          swap_idx = idx;
          break;
        }
    } else {
      for (auto d : extract_non_one_indices(in_t, out_t))
        if (shuffled[d] == 1 && d != axis) {
          swap_idx = d;
          break;
        }
    }
    shuffled = swap_elements(shuffled, axis, swap_idx);
    add(shuffled);
  }
  if (topos.back() == out_t) {
    to_original();
    return;
  }
  try_add_mid();
  add(out_t);
  to_original();
}

// Mapping.hpp:39-85
Vec permute_map_by_axes(Layout layout, const Vec& src_map, const Vec& axes) {
  Vec m;
  if (layout == Layout::Right) {
    for (auto s : src_map)
      if (!contains(axes, s)) m.push_back(s);
    for (auto a : axes) m.push_back(a);
  } else {
    Vec rev = reversed(axes);
    for (auto a : rev) m.push_back(a);
    for (auto s : src_map)
      if (!contains(rev, s)) m.push_back(s);
  }
  m.resize(src_map.size());
  return m;
}

// Extents.hpp:35-60
Vec compute_buffer_extents(Layout layout, const Vec& e, const Vec& in, const Vec& out) {
  Vec merged = merge_topology(in, out);
  std::size_t p0 = diff_topology(merged, in);
  Vec b;
  if (layout == Layout::Right) b.push_back(p0);
  for (std::size_t i = 0; i < e.size(); ++i) b.push_back((e[i] - 1) / merged[i] + 1);
  if (layout == Layout::Left) b.push_back(p0);
  return b;
}

// Extents.hpp:116-137
Vec compute_fft_extents(const Vec& gin, const Vec& gout, const Vec& axes) {
  Vec f;
  for (auto a : axes) f.push_back(std::max(gin.at(a), gout.at(a)));
  return f;
}

// MPI_Extents.hpp:492-529
Vec rank_to_coord(const Vec& t, std::size_t rank, Layout layout) {
  if (rank >= total_size(t)) fail(DFFT_ERROR_INVALID_ARGUMENT, "rank must be less than topology size.");
  Vec c(t.size(), 0);
  std::size_t r = rank;
  if (layout == Layout::Right) {
    for (std::size_t k = t.size(); k-- > 0;) {
      c[k] = r % t[k];
      r /= t[k];
    }
  } else {
    for (std::size_t k = 0; k < t.size(); ++k) {
      c[k] = r % t[k];
      r /= t[k];
    }
  }
  return c;
}

// MPI_Extents.hpp:195-235
Vec compute_local_extents(const Vec& e, const Vec& t, const Vec& map, std::size_t rank, Layout layout) {
  Vec local = e;
  Vec c = rank_to_coord(t, rank, layout);
  for (std::size_t i = 0; i < e.size(); ++i) {
    if (t[i] > 1) {
      std::size_t q = e[i] / t[i], r = e[i] % t[i];
      local[i] = c[i] < r ? q + 1 : q;
    }
  }
  return compute_mapped_extents(local, map);
}

// MPI_Extents.hpp:546-577 (starts: accumulate_extents :55-107 over the
// virtual ranks instead of MPI_Allgather)
void compute_local_extents_and_starts(const Vec& e, const Vec& t, std::size_t rank, Layout layout,
                                      Vec& local, Vec& starts) {
  Vec id = identity_map(e.size());
  local = compute_local_extents(e, t, id, rank, layout);
  Vec c = rank_to_coord(t, rank, layout);
  starts.assign(e.size(), 0);
  for (std::size_t i = 0; i < e.size(); ++i) {
    std::size_t q = e[i] / t[i], r = e[i] % t[i];
    starts[i] = c[i] * q + std::min(c[i], r);
  }
}

// MPI_Extents.hpp:322-417
void are_valid_extents(const Vec& gin, const Vec& gout, const Vec& axes, bool real_in, bool real_out,
                       Layout layout) {
  const std::size_t rank = gin.size();
  if (real_in && real_out) fail(DFFT_ERROR_INVALID_EXTENTS, "real to real transform is not supported");
  for (std::size_t i = 0; i < axes.size(); ++i) {
    if (axes[i] >= rank) fail(DFFT_ERROR_INVALID_EXTENTS, "input axes are not valid for the view");
    for (std::size_t j = 0; j < i; ++j)
      if (axes[i] == axes[j]) fail(DFFT_ERROR_INVALID_EXTENTS, "input axes are not valid for the view");
  }
  std::size_t inner = layout == Layout::Left ? 0 : rank - 1;
  Vec m, mi;
  get_map_axes(layout, rank, axes, m, mi);
  Vec ine = compute_mapped_extents(gin, m), oute = compute_mapped_extents(gout, m);
  for (std::size_t i = 0; i < rank; ++i) {
    if (i == inner) continue;
    if (ine[i] != oute[i])
      fail(DFFT_ERROR_INVALID_EXTENTS,
           "input and output extents must be the same except for the transform axis: " + str(ine) +
               ", " + str(oute));
  }
  if (!real_in && !real_out && ine[inner] != oute[inner])
    fail(DFFT_ERROR_INVALID_EXTENTS, "input and output extents must be the same for C2C transform: " +
                                         str(ine) + ", " + str(oute));
  if (real_in && oute[inner] != ine[inner] / 2 + 1)
    fail(DFFT_ERROR_INVALID_EXTENTS,
         "For R2C, the 'output extent' of transform must be equal to 'input extent'/2 + 1: " + str(ine) +
             ", " + str(oute));
  if (real_out && ine[inner] != oute[inner] / 2 + 1)
    fail(DFFT_ERROR_INVALID_EXTENTS,
         "For C2R, the 'input extent' of transform must be equal to 'output extent' / 2 + 1: " +
             str(ine) + ", " + str(oute));
}

namespace {
BlockInfo fft_block(const Vec& map, const Vec& in_e, const Vec& out_e, std::size_t fft_dim,
                    std::size_t idx, std::size_t& bufsize) {  // BlockAnalyses.hpp:57-74
  BlockInfo b;
  b.block_type = BlockType::FFT;
  b.fft_dim = fft_dim;
  b.in_extents = in_e;
  b.out_extents = out_e;
  b.in_map = map;
  b.out_map = map;
  b.block_idx = idx;
  b.in_topology.assign(map.size(), 0);
  b.out_topology.assign(map.size(), 0);
  b.buffer_extents.assign(map.size() + 1, 0);
  bufsize = total_size(out_e) * 2;
  return b;
}

BlockInfo transpose_block(Layout layout, const Vec& map, const Vec& in_t, const Vec& out_t,
                          const Vec& in_e, const Vec& gext, const Vec& axes, std::size_t rank,
                          bool is_lr, bool is_last, std::size_t in_axis, std::size_t out_axis,
                          std::size_t comm_axis, std::size_t idx, std::size_t size_factor,
                          std::size_t& bufsize) {  // BlockAnalyses.hpp:95-138
  BlockInfo b;
  b.block_type = BlockType::Transpose;
  b.in_map = map;
  b.in_axis = in_axis;
  b.out_axis = out_axis;
  b.comm_axis = comm_axis;
  b.in_topology = in_t;
  b.out_topology = out_t;
  b.in_extents = in_e;
  b.block_idx = idx;
  if (is_last) b.out_map = identity_map(map.size());
  else b.out_map = permute_map_by_axes(layout, map, axes.empty() ? Vec{out_axis} : axes);
  b.out_extents = compute_local_extents(gext, out_t, b.out_map, rank, is_lr ? Layout::Right : Layout::Left);
  b.buffer_extents = compute_buffer_extents(layout, gext, in_t, out_t);
  bufsize = std::max({total_size(b.in_extents), total_size(b.buffer_extents), total_size(b.out_extents)}) *
            size_factor;
  return b;
}
}  // namespace

// BlockAnalyses.hpp:158-293 (slab) and :311-395 (pencil); :415-506
BlockAnalyses block_analyses(TopologyType kind, bool real_value, Layout layout, const Vec& in_extents,
                             const Vec& gin, const Vec& gout, const Vec& in_topology,
                             const Vec& out_topology, const Vec& axes, std::size_t rank,
                             Layout in_layout, Layout out_layout) {
  BlockAnalyses res;
  const std::size_t dim = in_extents.size();
  Vec map, map_inv;
  get_map_axes(layout, dim, axes, map, map_inv);
  if (kind == TopologyType::Slab) {
    res.topologies = get_all_slab_topologies(in_topology, out_topology, axes);
  } else {
    get_all_pencil_topologies(in_topology, out_topology, axes, in_layout, out_layout, res.topologies,
                              res.trans_axes, res.layouts);
  }
  res.all_axes = decompose_axes(res.topologies, axes);
  auto& blocks = res.blocks;
  std::size_t maxbuf = 0;
  const auto& topos = res.topologies;
  for (std::size_t bi = 0; bi < topos.size(); ++bi) {
    auto make_transpose = [&](const std::vector<BlockInfo>& bl, std::size_t size_factor) {
      bool is_last = bi == topos.size() - 1;
      std::size_t nb_trans = std::count_if(bl.begin(), bl.end(), [](const BlockInfo& b) {
        return b.block_type == BlockType::Transpose;
      });
      const Vec& cur = topos.at(nb_trans);
      const Vec& nxt = topos.at(nb_trans + 1);
      std::size_t in_axis, out_axis, comm_axis = 0;
      bool is_lr = true;
      if (kind == TopologyType::Slab) {
        slab_in_out_axes(cur, nxt, in_axis, out_axis);
      } else {
        pencil_in_out_axes(cur, nxt, in_axis, out_axis);
        comm_axis = res.trans_axes.at(nb_trans);
        is_lr = res.layouts.at(nb_trans + 1) != 0;
      }
      Vec next_axes = is_last ? Vec{} : res.all_axes.at(bi + 1);
      std::size_t sz = 0;
      BlockInfo b;
      if (bl.empty())
        b = transpose_block(layout, identity_map(dim), cur, nxt, in_extents, gin, next_axes, rank, is_lr,
                            false, in_axis, out_axis, comm_axis, 0, size_factor, sz);
      else
        b = transpose_block(layout, bl.back().out_map, cur, nxt, bl.back().out_extents, gout, next_axes,
                            rank, is_lr, is_last, in_axis, out_axis, comm_axis, nb_trans, 2, sz);
      maxbuf = std::max(maxbuf, sz);
      return b;
    };
    const Vec& fax = res.all_axes.at(bi);
    std::size_t fft_dim = fax.size();
    if (fft_dim == 0) {
      blocks.push_back(make_transpose(blocks, real_value ? 1 : 2));
      continue;
    }
    Layout lay = Layout::Right;
    if (kind == TopologyType::Pencil) lay = res.layouts.at(bi) ? Layout::Right : Layout::Left;
    std::size_t done = 0;
    for (const auto& b : blocks)
      if (b.block_type == BlockType::FFT) done += b.fft_dim;
    bool is_last_fft = done + fft_dim == axes.size();
    std::size_t sz = 0;
    if (blocks.empty()) {
      Vec fin = compute_mapped_extents(in_extents, map);
      Vec fout = compute_local_extents(gout, topos.at(bi), map, rank, lay);
      blocks.push_back(fft_block(map, fin, fout, fft_dim, 0, sz));
    } else {
      std::size_t nb_fft = std::count_if(blocks.begin(), blocks.end(), [](const BlockInfo& b) {
        return b.block_type == BlockType::FFT;
      });
      Vec in_map = blocks.back().out_map;
      Vec fin = blocks.back().out_extents;
      if (nb_fft == 0) {
        Vec fout = compute_local_extents(gout, blocks.back().out_topology, in_map, rank, lay);
        blocks.push_back(fft_block(in_map, fin, fout, fft_dim, 0, sz));
      } else {
        blocks.push_back(fft_block(in_map, fin, fin, fft_dim, nb_fft, sz));
      }
    }
    maxbuf = std::max(maxbuf, sz);
    if (!is_last_fft) blocks.push_back(make_transpose(blocks, 2));
  }
  res.max_buffer_size = maxbuf;
  return res;
}

}  // namespace dfft
