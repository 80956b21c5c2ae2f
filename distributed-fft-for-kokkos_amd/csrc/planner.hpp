// Host-side planner of the distributed FFT: topology chain, axis
// decomposition, block list, local extents.  Integer logic only.
//
// Restates, function by function, the reference's planning headers:
//   KokkosFFT_Distributed_Topologies.hpp, _BlockAnalyses.hpp, _Mapping.hpp,
//   _Extents.hpp, _MPI_Extents.hpp, _ContainerAnalyses.hpp, _Types.hpp
// with the MPI collectives of the setup phase (Allgather of extents,
// Allreduce of buffer sizes) replaced by loops over the P virtual ranks that
// one process holds.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace dfft {

using Vec = std::vector<std::size_t>;

// Errors carry one of the status codes of the C ABI (dfft.h).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum class TopologyType { Empty, Shared, Slab, Pencil, Brick, Invalid };
enum class BlockType { Transpose, FFT };
enum class Layout { Right, Left };

// Types.hpp:61-154
struct BlockInfo {
  Vec in_topology, out_topology, in_extents, out_extents, buffer_extents;
  Vec in_map, out_map;
  std::size_t in_axis = 0, out_axis = 0, fft_dim = 0, comm_axis = 0;
  BlockType block_type = BlockType::FFT;
  std::size_t block_idx = 0;
  bool operator==(const BlockInfo& o) const;
};

struct BlockAnalyses {
  std::vector<BlockInfo> blocks;
  std::size_t max_buffer_size = 0;  // this rank, scalar units
  std::vector<Vec> topologies;
  Vec trans_axes, layouts;
  std::vector<Vec> all_axes;
};

std::size_t total_size(const Vec& v);
Vec identity_map(std::size_t n);
std::size_t get_index(const Vec& v, std::size_t value);  // throws when absent
Vec convert_negative_axes(const std::vector<int>& axes, std::size_t rank);
void get_map_axes(Layout layout, std::size_t rank, const Vec& axes, Vec& map, Vec& map_inv);
Vec compute_mapped_extents(const Vec& extents, const Vec& map);

// ContainerAnalyses.hpp
Vec merge_topology(const Vec& in, const Vec& out);
std::size_t diff_topology(const Vec& in, const Vec& out);

// Topologies.hpp
TopologyType to_topology_type(const Vec& t);
TopologyType get_common_topology_type(const Vec& a, const Vec& b);
void slab_in_out_axes(const Vec& in, const Vec& out, std::size_t& in_axis, std::size_t& out_axis);
void pencil_in_out_axes(const Vec& in, const Vec& out, std::size_t& in_axis, std::size_t& out_axis);
Vec propose_mid_array(const Vec& in, const Vec& out);
std::vector<Vec> decompose_axes(const std::vector<Vec>& topologies, const Vec& axes);
std::size_t compute_trans_axis(const Vec& in, const Vec& out, std::size_t first_non_one);
std::vector<Vec> get_all_slab_topologies(const Vec& in, const Vec& out, const Vec& axes);
void get_all_pencil_topologies(const Vec& in, const Vec& out, const Vec& axes, Layout in_layout,
                               Layout out_layout, std::vector<Vec>& topologies, Vec& trans_axes,
                               Vec& layouts);

// Mapping.hpp / Extents.hpp
Vec permute_map_by_axes(Layout layout, const Vec& src_map, const Vec& axes);
Vec compute_buffer_extents(Layout layout, const Vec& extents, const Vec& in, const Vec& out);
Vec compute_fft_extents(const Vec& gin, const Vec& gout, const Vec& axes);

// MPI_Extents.hpp
Vec rank_to_coord(const Vec& topology, std::size_t rank, Layout layout);
Vec compute_local_extents(const Vec& extents, const Vec& topology, const Vec& map,
                          std::size_t rank, Layout layout);
void compute_local_extents_and_starts(const Vec& extents, const Vec& topology, std::size_t rank,
                                      Layout layout, Vec& local, Vec& starts);
void are_valid_extents(const Vec& gin, const Vec& gout, const Vec& axes, bool real_in,
                       bool real_out, Layout layout);

// BlockAnalyses.hpp:415-506 for one rank
BlockAnalyses block_analyses(TopologyType kind, bool real_value, Layout layout,
                             const Vec& in_extents, const Vec& gin, const Vec& gout,
                             const Vec& in_topology, const Vec& out_topology, const Vec& axes,
                             std::size_t rank, Layout in_layout, Layout out_layout);

}  // namespace dfft
