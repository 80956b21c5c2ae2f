// A compiled consumer of include/dfft.h and include/dfft.hpp (the C++ mirror
// of KokkosFFT::Distributed).  Built and run by tests/test_capi_consumer.py:
//
//   capi_consumer layout   struct sizes and field offsets of the C ABI (JSON),
//                          compared there with the ctypes declarations
//   capi_consumer host     host-only calls: planner queries and the status
//                          codes of invalid plan descriptions (no GPU needed)
//   capi_consumer gpu      (built with -DDFFT_CONSUMER_GPU) slab and pencil
//                          round trips through dfft::Plan / dfft::execute on
//                          hipMalloc'd views, and the execute-time extent check
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dfft.h"
#include "dfft.hpp"

#ifdef DFFT_CONSUMER_GPU
#include <hip/hip_runtime_api.h>
#endif

namespace {

#define FIELD(S, F) std::printf("\"%s.%s\": %zu, ", #S, #F, offsetof(S, F))

int layout() {
  std::printf("{");
  std::printf("\"sizeof.dfft_desc\": %zu, \"sizeof.dfft_block_info\": %zu, \"sizeof.dfft_ns_desc\": %zu, ",
              sizeof(dfft_desc), sizeof(dfft_block_info), sizeof(dfft_ns_desc));
  FIELD(dfft_desc, rank);
  FIELD(dfft_desc, fft_rank);
  FIELD(dfft_desc, axes);
  FIELD(dfft_desc, transform);
  FIELD(dfft_desc, precision);
  FIELD(dfft_desc, layout);
  FIELD(dfft_desc, in_extents);
  FIELD(dfft_desc, out_extents);
  FIELD(dfft_desc, in_topology);
  FIELD(dfft_desc, out_topology);
  FIELD(dfft_desc, in_topology_layout);
  FIELD(dfft_desc, out_topology_layout);
  FIELD(dfft_desc, norm);
  FIELD(dfft_desc, nranks);
  FIELD(dfft_desc, exec_mode);
  FIELD(dfft_block_info, in_topology);
  FIELD(dfft_block_info, out_topology);
  FIELD(dfft_block_info, in_extents);
  FIELD(dfft_block_info, out_extents);
  FIELD(dfft_block_info, buffer_extents);
  FIELD(dfft_block_info, in_map);
  FIELD(dfft_block_info, out_map);
  FIELD(dfft_block_info, in_axis);
  FIELD(dfft_block_info, out_axis);
  FIELD(dfft_block_info, fft_dim);
  FIELD(dfft_block_info, comm_axis);
  FIELD(dfft_block_info, block_type);
  FIELD(dfft_block_info, block_idx);
  FIELD(dfft_ns_desc, n);
  FIELD(dfft_ns_desc, px);
  FIELD(dfft_ns_desc, py);
  FIELD(dfft_ns_desc, lx);
  FIELD(dfft_ns_desc, nu);
  FIELD(dfft_ns_desc, dt);
  FIELD(dfft_ns_desc, form);
  std::printf("\"end\": 0}\n");
  return 0;
}

dfft_desc desc3(std::array<size_t, 3> gin, std::array<size_t, 3> gout, std::array<size_t, 3> tin,
                std::array<size_t, 3> tout, int transform, int precision, int nranks) {
  dfft_desc d{};
  d.rank = 3;
  d.fft_rank = 3;
  d.axes[0] = 0;
  d.axes[1] = 1;
  d.axes[2] = 2;
  d.transform = transform;
  d.precision = precision;
  d.layout = DFFT_LAYOUT_RIGHT;
  for (int a = 0; a < 3; ++a) {
    d.in_extents[a] = gin[a];
    d.out_extents[a] = gout[a];
    d.in_topology[a] = tin[a];
    d.out_topology[a] = tout[a];
  }
  d.norm = DFFT_NORM_BACKWARD;
  d.nranks = nranks;
  d.exec_mode = DFFT_EXEC_FUSED;
  return d;
}

int host() {
  std::printf("{");
  // the reference's factory and extent checks (Plan.hpp:62-112, MPI_Extents.hpp:322-417)
  struct Case {
    const char* name;
    dfft_desc d;
  };
  std::vector<Case> cases;
  cases.push_back({"topology_sizes_differ", desc3({8, 7, 5}, {8, 7, 5}, {2, 1, 1}, {1, 3, 1}, DFFT_C2C, DFFT_FP64, 2)});
  cases.push_back({"topology_size_not_nranks", desc3({8, 7, 5}, {8, 7, 5}, {2, 1, 1}, {1, 2, 1}, DFFT_C2C, DFFT_FP64, 4)});
  cases.push_back({"c2c_extents_differ", desc3({8, 7, 5}, {8, 7, 6}, {2, 1, 1}, {1, 2, 1}, DFFT_C2C, DFFT_FP64, 2)});
  cases.push_back({"r2c_out_not_half", desc3({8, 7, 6}, {8, 7, 6}, {2, 1, 1}, {1, 2, 1}, DFFT_R2C, DFFT_FP64, 2)});
  cases.push_back({"brick", desc3({8, 8, 8}, {8, 8, 8}, {2, 2, 2}, {2, 2, 2}, DFFT_C2C, DFFT_FP64, 8)});
  cases.push_back({"empty", desc3({8, 8, 8}, {8, 8, 8}, {0, 1, 1}, {0, 1, 1}, DFFT_C2C, DFFT_FP64, 0)});
  cases.push_back({"slab_to_pencil", desc3({8, 8, 8}, {8, 8, 8}, {4, 1, 1}, {1, 2, 2}, DFFT_C2C, DFFT_FP64, 4)});
  cases.push_back({"extent_too_large", desc3({1u << 30, 4, 4}, {1u << 30, 4, 4}, {1, 2, 1}, {1, 1, 2}, DFFT_C2C, DFFT_FP64, 2)});
  {
    dfft_desc d = desc3({8, 7, 5}, {8, 7, 5}, {2, 1, 1}, {1, 2, 1}, DFFT_C2C, DFFT_FP64, 2);
    d.axes[2] = 1;  // repeated axis
    cases.push_back({"repeated_axis", d});
  }
  {
    dfft_desc d = desc3({8, 7, 5}, {8, 7, 5}, {2, 1, 1}, {1, 2, 1}, DFFT_C2C, DFFT_FP64, 2);
    d.fft_rank = 4;
    cases.push_back({"fft_rank_4", d});
  }
  {
    dfft_desc d{};
    d.rank = 2;
    d.fft_rank = 2;
    d.axes[0] = 0;
    d.axes[1] = 1;
    d.transform = DFFT_C2C;
    d.precision = DFFT_FP64;
    d.in_extents[0] = d.out_extents[0] = 8;
    d.in_extents[1] = d.out_extents[1] = 8;
    d.in_topology[0] = d.in_topology[1] = 2;
    d.out_topology[0] = d.out_topology[1] = 2;
    d.nranks = 4;
    cases.push_back({"pencil_rank2", d});
  }
  {
    dfft_desc d = desc3({8, 7, 5}, {8, 7, 5}, {2, 1, 1}, {1, 2, 1}, DFFT_C2C, DFFT_FP64, 2);
    d.norm = 7;
    cases.push_back({"bad_norm", d});
  }
  for (const Case& c : cases) {
    dfft_plan* p = nullptr;
    const int st = dfft_plan_create(&c.d, &p);
    if (st == DFFT_SUCCESS) dfft_plan_destroy(p);
    std::printf("\"%s\": %d, ", c.name, st);
  }
  // planner queries through the C++ mirror's types
  const dfft::Topology<3> tin{4, 1, 1}, tout{1, 4, 1};
  size_t topos[16 * 3];
  int ntopo = 16;
  const int axes[3] = {0, 1, 2};
  dfft::check(dfft_get_all_slab_topologies(3, tin.data(), tout.data(), 3, axes, topos, &ntopo));
  std::printf("\"slab_topologies\": [");
  for (int i = 0; i < ntopo; ++i)
    std::printf("%s[%zu, %zu, %zu]", i ? ", " : "", topos[3 * i], topos[3 * i + 1], topos[3 * i + 2]);
  std::printf("], ");
  const auto le = dfft::compute_local_extents_and_starts<3>({129, 64, 10}, dfft::Topology<3>{2, 1, 1}, 1);
  std::printf("\"local_extents_129_r1\": [%zu, %zu, %zu], \"local_starts_129_r1\": [%zu, %zu, %zu], ", le.first[0],
              le.first[1], le.first[2], le.second[0], le.second[1], le.second[2]);
  const auto c4 = dfft::rank_to_coord(dfft::Topology<3, dfft::LayoutLeft>{2, 3, 1}, 4);
  std::printf("\"coord_left_231_r4\": [%zu, %zu, %zu], ", c4[0], c4[1], c4[2]);
  bool threw = false;
  try {
    (void)dfft::rank_to_coord(dfft::Topology<3>{2, 1, 1}, 2);
  } catch (const dfft::Error& e) {
    threw = e.code == DFFT_ERROR_INVALID_ARGUMENT;
  }
  std::printf("\"rank_to_coord_out_of_range_throws\": %s, ", threw ? "true" : "false");
  dfft_block_info infos[16];
  int nb = 16;
  size_t mx = 0;
  const dfft_desc a = desc3({64, 64, 64}, {64, 64, 64}, {2, 1, 1}, {1, 2, 1}, DFFT_C2C, DFFT_FP64, 2);
  dfft::check(dfft_block_analyses(&a, 0, infos, &nb, &mx));
  std::printf("\"config_a_blocks\": %d, \"config_a_max_buffer\": %zu, \"config_a_block1_buffer\": [%zu, %zu, %zu, %zu]",
              nb, mx, infos[1].buffer_extents[0], infos[1].buffer_extents[1], infos[1].buffer_extents[2],
              infos[1].buffer_extents[3]);
  std::printf("}\n");
  return 0;
}

#ifdef DFFT_CONSUMER_GPU
void hip_ok(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

template <typename T>
std::vector<dfft::View<T, 3>> alloc_views(const std::array<size_t, 3>& g, const dfft::Topology<3>& t) {
  std::vector<dfft::View<T, 3>> v;
  for (size_t r = 0; r < t.total(); ++r) {
    const auto e = dfft::compute_local_extents_and_starts(g, t, r).first;
    T* p = nullptr;
    hip_ok(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * e[0] * e[1] * e[2]));
    v.emplace_back(p, e);
  }
  return v;
}

template <typename T>
void fill(const std::vector<dfft::View<T, 3>>& v, unsigned seed) {
  for (const auto& x : v) {
    std::vector<T> h(x.size());
    for (size_t i = 0; i < h.size(); ++i) {
      seed = seed * 1664525u + 1013904223u;
      const double a = (seed >> 8) * (1.0 / 16777216.0) * 2.0 - 1.0;
      if constexpr (dfft::is_complex<T>::value) {
        seed = seed * 1664525u + 1013904223u;
        const double b = (seed >> 8) * (1.0 / 16777216.0) * 2.0 - 1.0;
        h[i] = T(a, b);
      } else {
        h[i] = T(a);
      }
    }
    hip_ok(hipMemcpy(x.data(), h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  }
}

// relative L2 distance between two sets of views
template <typename T>
double rel_l2(const std::vector<dfft::View<T, 3>>& a, const std::vector<dfft::View<T, 3>>& b) {
  double num = 0.0, den = 0.0;
  for (size_t r = 0; r < a.size(); ++r) {
    std::vector<T> ha(a[r].size()), hb(b[r].size());
    hip_ok(hipMemcpy(ha.data(), a[r].data(), sizeof(T) * ha.size(), hipMemcpyDeviceToHost));
    hip_ok(hipMemcpy(hb.data(), b[r].data(), sizeof(T) * hb.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ha.size(); ++i) {
      num += std::norm(std::complex<double>(ha[i]) - std::complex<double>(hb[i]));
      den += std::norm(std::complex<double>(ha[i]));
    }
  }
  return std::sqrt(num / den);
}

template <typename V>
void release(const V& v) {
  for (const auto& x : v) (void)hipFree(x.data());
}

int gpu() {
  std::printf("{");
  using C = std::complex<double>;
  {
    // config A shape at 16 x 12 x 10: slab (2,1,1) -> (1,2,1), C2C FP64
    const std::array<size_t, 3> g{16, 12, 10};
    const dfft::Topology<3> tin{2, 1, 1}, tout{1, 2, 1};
    auto in = alloc_views<C>(g, tin), back = alloc_views<C>(g, tin);
    auto out = alloc_views<C>(g, tout);
    fill(in, 1u);
    dfft::Plan<C, C, 3, 3> plan(in, out, {0, 1, 2}, tin, tout);
    dfft::execute(plan, in, out, dfft::Direction::forward);
    dfft::execute(plan, out, back, dfft::Direction::backward);
    hip_ok(hipDeviceSynchronize());
    std::printf("\"slab_name\": \"%s\", \"slab_c2c_roundtrip\": %.3e, ", plan.name().c_str(), rel_l2(in, back));
    // good(): views of other extents are refused before anything runs
    auto wrong = alloc_views<C>({16, 12, 10}, dfft::Topology<3>{1, 2, 1});
    int code = 0;
    try {
      dfft::execute(plan, wrong, out, dfft::Direction::forward);
    } catch (const dfft::Error& e) {
      code = e.code;
    }
    std::printf("\"good_mismatch_code\": %d, ", code);
    release(in);
    release(back);
    release(out);
    release(wrong);
  }
  {
    // pencil (2,2,1) -> (1,2,2), R2C FP64 at 16^3
    const std::array<size_t, 3> g{16, 16, 16}, gh{16, 16, 9};
    const dfft::Topology<3> tin{2, 2, 1}, tout{1, 2, 2};
    auto in = alloc_views<double>(g, tin), back = alloc_views<double>(g, tin);
    auto out = alloc_views<C>(gh, tout);
    fill(in, 2u);
    dfft::Plan<double, C, 3, 3> plan(in, out, {0, 1, 2}, tin, tout);
    dfft::execute(plan, in, out, dfft::Direction::forward);
    dfft::execute(plan, out, back, dfft::Direction::backward);
    hip_ok(hipDeviceSynchronize());
    std::printf("\"pencil_name\": \"%s\", \"pencil_r2c_roundtrip\": %.3e, ", plan.name().c_str(), rel_l2(in, back));
    release(in);
    release(back);
    release(out);
  }
  {
    // the allocation guard of the C ABI: a rank buffer smaller than its local array
    dfft_desc d = desc3({16, 12, 10}, {16, 12, 10}, {2, 1, 1}, {1, 2, 1}, DFFT_C2C, DFFT_FP64, 2);
    dfft_plan* p = nullptr;
    dfft::check(dfft_plan_create(&d, &p));
    // rank 0's input starts 4 KiB before the end of a 1 MiB allocation; its
    // local array needs 8 x 12 x 10 x 16 = 15360 bytes
    const size_t need = 8 * 12 * 10 * 16, whole = 1 << 20;
    char* base = nullptr;
    void* other[3] = {nullptr, nullptr, nullptr};
    hip_ok(hipMalloc(reinterpret_cast<void**>(&base), whole));
    for (auto& q : other) hip_ok(hipMalloc(&q, whole));
    void* ins[2] = {base + whole - 4096, other[0]};
    void* outs[2] = {other[1], other[2]};
    // the guard reads the allocation range from the runtime; run the call
    // only where the runtime reports it (else the guard lets it through)
    hipDeviceptr_t b = nullptr;
    size_t sz = 0;
    if (hipMemGetAddressRange(&b, &sz, ins[0]) == hipSuccess && static_cast<char*>(b) + sz < base + whole - 4096 + need) {
      const int st = dfft_execute(p, ins, outs, DFFT_FORWARD, nullptr);
      std::printf("\"short_buffer_code\": %d", st);
    } else {
      std::printf("\"short_buffer_code\": \"untestable\"");
    }
    dfft_plan_destroy(p);
    (void)hipFree(base);
    for (auto q : other) (void)hipFree(q);
  }
  std::printf("}\n");
  return 0;
}
#endif

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "layout";
  try {
    if (mode == "layout") return layout();
    if (mode == "host") return host();
#ifdef DFFT_CONSUMER_GPU
    if (mode == "gpu") return gpu();
#endif
  } catch (const std::exception& e) {
    std::fprintf(stderr, "capi_consumer %s: %s\n", mode.c_str(), e.what());
    return 1;
  }
  std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
  return 2;
}
