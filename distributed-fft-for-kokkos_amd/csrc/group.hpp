// A group of independent single-GPU transforms of possibly different
// extents executed together: line pass k of every member is ONE grouped
// kernel launch (the reference would loop execute() over separate plans,
// Plan.hpp:212-238; its batched NS example batches only equal-size fields
// through a leading axis, navier-stokes-MPI-batched/main.cpp:102-107).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "dfft.h"
#include "fft_kernels.hpp"
#include "planner.hpp"

namespace dfft {

class Group {
 public:
  Group(int count, const dfft_desc* descs);
  ~Group();
  Group(const Group&) = delete;
  Group& operator=(const Group&) = delete;

  // in[i] / out[i]: member i's arrays (inputs are not modified)
  void execute(void* const* in, void* const* out, int direction, bool use_graph, hipStream_t s);
  // execute() without any device allocation: the pass tables go into a
  // static device array of kStaticTableBytes (group_static.hip) shared by
  // every group on the device; no graph.  Throws DFFT_ERROR_UNSUPPORTED when
  // this call's tables do not fit.
  void execute_static(void* const* in, void* const* out, int direction, hipStream_t s);
  static constexpr std::size_t kStaticTableBytes = std::size_t(4) << 20;
  int count() const { return static_cast<int>(members_.size()); }
  // launches one execute of `direction` issues
  int num_launches(int direction) const;
  // "lds_cap": LDS bytes per block of the grouped kernel (lines per block
  // shrink to fit; 1024..65536, default 65536); "grid_mult": resident-block
  // multiples of the persistent grid (0: one block per virtual block);
  // "nt": nontemporal global accesses (-1: on when a pass moves >= 1 GiB)
  void set_option(const std::string& key, long value);
  long get_option(const std::string& key) const;
  // time LDS caps and grid sizes on the caller's arrays (contents are
  // overwritten by forward + backward round trips) and keep the fastest
  void autotune(void* const* in, void* const* out, hipStream_t s);

 private:
  struct Member {
    int rank = 3;
    bool real = false;
    Vec gin, gout;
    std::vector<int> fwd_axes, bwd_axes;  // pass order
    double scale_f = 1.0, scale_b = 1.0;
    void* scratch = nullptr;  // complex spectrum for the backward R2C passes
  };
  std::vector<LinePass> build_pass(const std::vector<void*>& in, const std::vector<void*>& out, bool forward,
                                   int k) const;
  // host form of one direction's launches: slot 2k + v holds the grouped
  // members of pass k taken one (v = 0) or two (v = 1) lines per lane
  struct Tables {
    std::vector<std::vector<LinePass>> items;
    std::vector<std::vector<int32_t>> first;  // virtual-block prefix per slot
    std::vector<int> count, total, grid, nt, lanes;
    std::vector<size_t> lds;
    std::vector<std::vector<LinePass>> single;  // per pass: individually launched passes
  };
  void build_tables(const std::vector<void*>& in, const std::vector<void*>& out, int dir, Tables& t) const;
  // argument checks of execute(): pointers, direction, R2C aliasing, and that
  // every member allocation holds its arrays; returns the direction index
  int check_call(const std::vector<void*>& vin, const std::vector<void*>& vout, int direction) const;
  void upload(const std::vector<void*>& in, const std::vector<void*>& out, int dir, hipStream_t s);
  void launch_all(int dir, hipStream_t s);
  // launches of one direction from tables whose slots sit at items[slot] /
  // first[slot] on the device
  void launch_tables(const Tables& t, LinePass* const* items, int32_t* const* first, int dir, hipStream_t s) const;
  void invalidate();

  std::vector<Member> members_;
  bool fp64_ = true;
  int device_ = 0;
  int max_passes_ = 0;
  std::size_t lds_cap_ = kGroupLds;
  int grid_mult_ = 1;
  int nt_ = -1;
  int dual_ = 1;  // FP32: two adjacent lines per lane where they apply
  int lanes_ = kGroupThreads;  // lanes per block of launches that hold only contiguous lines
  int stage_ = 1;  // FP32: contiguous aligned lines 16 bytes per lane through LDS
  int contig_nt_ = -1;  // nontemporal accesses of launches of contiguous lines (-1: as nt_)
  std::map<int, void*> tw_;
  // per direction: the pointer set the device tables were built for, the
  // tables of every pass (LinePass items + virtual-block prefix), members
  // whose length has no grouped kernel (launched one by one), and a graph
  struct DirState {
    std::vector<void*> in, out;
    bool built = false;
    Tables tables;                    // host form of the launches
    std::vector<LinePass*> items;     // device copies of tables.items, one array per slot
    std::vector<int32_t*> first;      // device copies of tables.first
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  } dir_[2];
  hipStream_t graph_stream_ = nullptr;
};

}  // namespace dfft
