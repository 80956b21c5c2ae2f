// C++ mirror of the reference's public API (KokkosFFT::Distributed) over the
// C ABI of libdfft.so (dfft.h).  Header-only; needs no HIP or Kokkos headers.
//
//   reference (KokkosFFT_Distributed_Plan.hpp:124-238)      here
//   Plan plan(exec, in, out, axes, tin, tout, comm, norm);  dfft::Plan<...> plan(in, out, axes, tin, tout, norm);
//   execute(plan, in, out, Direction::forward);             dfft::execute(plan, in, out, dfft::Direction::forward);
//   plan.name();                                            plan.name();
//
// One process drives the P virtual ranks of a topology, so where the
// reference takes the calling rank's Kokkos::View, this mirror takes a
// vector of P views: dfft::View is a device pointer plus the rank's local
// extents (the view's data() and extent(i)).  Errors throw dfft::Error (a
// std::runtime_error, as the reference throws) carrying the C ABI status.
#ifndef DFFT_HPP
#define DFFT_HPP

#include <array>
#include <complex>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "dfft.h"

namespace dfft {

// Kokkos::LayoutRight / Kokkos::LayoutLeft tags
struct LayoutRight {
  static constexpr int value = DFFT_LAYOUT_RIGHT;
};
struct LayoutLeft {
  static constexpr int value = DFFT_LAYOUT_LEFT;
};

// KokkosFFT::Direction, KokkosFFT::Normalization
enum class Direction : int { forward = DFFT_FORWARD, backward = DFFT_BACKWARD };
enum class Normalization : int {
  backward = DFFT_NORM_BACKWARD,
  forward = DFFT_NORM_FORWARD,
  ortho = DFFT_NORM_ORTHO,
  none = DFFT_NORM_NONE
};

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& msg) : std::runtime_error(msg), code(status) {}
  int code;
};

inline void check(int status) {
  if (status != DFFT_SUCCESS) throw Error(status, dfft_last_error());
}

// Topology<std::size_t, N, Layout> (KokkosFFT_Distributed_Types.hpp:174-428):
// the number of ranks along each axis plus the layout of rank numbering.
template <std::size_t N, typename LayoutType = LayoutRight>
class Topology {
 public:
  using value_type = std::size_t;
  using layout_type = LayoutType;
  Topology() = default;
  Topology(const std::array<std::size_t, N>& a) : m_data(a) {}
  Topology(std::initializer_list<std::size_t> init) {
    if (init.size() != N) throw std::length_error("Initializer list size must match array size");
    std::size_t i = 0;
    for (auto v : init) m_data[i++] = v;
  }
  std::size_t operator[](std::size_t i) const { return m_data[i]; }
  std::size_t& operator[](std::size_t i) { return m_data[i]; }
  static constexpr std::size_t size() { return N; }
  const std::array<std::size_t, N>& array() const { return m_data; }
  const std::size_t* data() const { return m_data.data(); }
  std::size_t total() const {
    std::size_t p = 1;
    for (auto v : m_data) p *= v;
    return p;
  }

 private:
  std::array<std::size_t, N> m_data{};
};

// One rank's local array: device pointer + extents, stored LayoutRight
// (row-major) or LayoutLeft (column-major) -- a Kokkos::View stand-in.
template <typename T, std::size_t N, typename LayoutType = LayoutRight>
struct View {
  using value_type = T;
  using layout_type = LayoutType;
  static constexpr std::size_t rank() { return N; }
  T* ptr = nullptr;
  std::array<std::size_t, N> ext{};
  View() = default;
  View(T* p, const std::array<std::size_t, N>& e) : ptr(p), ext(e) {}
  T* data() const { return ptr; }
  std::size_t extent(std::size_t i) const { return ext[i]; }
  std::size_t size() const {
    std::size_t s = 1;
    for (auto v : ext) s *= v;
    return s;
  }
};

template <typename T> struct is_complex : std::false_type {};
template <typename R> struct is_complex<std::complex<R>> : std::true_type {};
template <typename T> struct real_of { using type = T; };
template <typename R> struct real_of<std::complex<R>> { using type = R; };

// rank_to_coord (MPI_Extents.hpp:492-529): LayoutRight topologies number
// ranks with the last axis fastest, LayoutLeft with the first.
template <std::size_t N, typename L>
std::array<std::size_t, N> rank_to_coord(const Topology<N, L>& t, std::size_t rank) {
  if (rank >= t.total()) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "rank must be less than topology size.");
  std::array<std::size_t, N> c{};
  std::size_t r = rank;
  if (std::is_same<L, LayoutRight>::value) {
    for (std::size_t i = N; i-- > 0;) {
      c[i] = r % t[i];
      r /= t[i];
    }
  } else {
    for (std::size_t i = 0; i < N; ++i) {
      c[i] = r % t[i];
      r /= t[i];
    }
  }
  return c;
}

// compute_local_extents_and_starts (MPI_Extents.hpp:546-577): the first
// N % t ranks along an axis hold one element more.
template <std::size_t N, typename L>
std::pair<std::array<std::size_t, N>, std::array<std::size_t, N>> compute_local_extents_and_starts(
    const std::array<std::size_t, N>& global, const Topology<N, L>& t, std::size_t rank) {
  const auto c = rank_to_coord(t, rank);
  std::array<std::size_t, N> ext{}, start{};
  for (std::size_t i = 0; i < N; ++i) {
    const std::size_t q = global[i] / t[i], rem = global[i] % t[i];
    ext[i] = q + (c[i] < rem ? 1 : 0);
    start[i] = c[i] * q + (c[i] < rem ? c[i] : rem);
  }
  return {ext, start};
}

namespace detail {
struct PlanDeleter {
  void operator()(dfft_plan* p) const { (void)dfft_plan_destroy(p); }
};

// compute_global_extents (MPI_Extents.hpp:114-129) over the ranks' views:
// the sum of the local extents along each axis's coordinate line
template <std::size_t N, typename L, typename V>
std::array<std::size_t, N> global_extents(const std::vector<V>& views, const Topology<N, L>& t) {
  std::array<std::size_t, N> g{};
  for (std::size_t a = 0; a < N; ++a) {
    for (std::size_t r = 0; r < views.size(); ++r) {
      const auto c = rank_to_coord(t, r);
      bool on_line = true;
      for (std::size_t b = 0; b < N; ++b)
        if (b != a && c[b] != 0) on_line = false;
      if (on_line) g[a] += views[r].extent(a);
    }
  }
  return g;
}

template <std::size_t N>
std::string extents_string(const std::array<std::size_t, N>& e) {
  std::string s = "(";
  for (std::size_t i = 0; i < N; ++i) s += (i ? "," : "") + std::to_string(e[i]);
  return s + ")";
}
}  // namespace detail

// Plan<InT, OutT, N, DIM, InLayout, OutLayout, ViewLayout>: the reference's
// Plan<ExecSpace, InView, OutView, DIM, InLayout, OutLayout> with the view
// types spelled as value type + rank + view layout.
template <typename InT, typename OutT, std::size_t N, std::size_t DIM = 1, typename InLayoutType = LayoutRight,
          typename OutLayoutType = LayoutRight, typename ViewLayout = LayoutRight>
class Plan {
  static_assert(DIM >= 1 && DIM <= 3, "Plan: the Rank of FFT axes must be between 1 and 3");
  static_assert(is_complex<OutT>::value,
                "Plan: the output type must be complex, while the input type can be either real or complex");
  static_assert(std::is_same<typename real_of<InT>::type, typename real_of<OutT>::type>::value,
                "Plan: input and output must share one precision");

 public:
  using in_view_type = View<InT, N, ViewLayout>;
  using out_view_type = View<OutT, N, ViewLayout>;
  using extents_type = std::array<std::size_t, N>;

  Plan(const std::vector<in_view_type>& in, const std::vector<out_view_type>& out, const std::array<int, DIM>& axes,
       const Topology<N, InLayoutType>& in_topology, const Topology<N, OutLayoutType>& out_topology,
       Normalization norm = Normalization::backward, int exec_mode = DFFT_EXEC_FUSED) {
    const std::size_t P = in_topology.total();
    if (in.size() != P || out.size() != P)
      throw Error(DFFT_ERROR_INVALID_TOPOLOGY, "topology size must be identical to the number of views");
    dfft_desc d{};
    d.rank = static_cast<int>(N);
    d.fft_rank = static_cast<int>(DIM);
    for (std::size_t i = 0; i < DIM; ++i) d.axes[i] = axes[i];
    d.transform = is_complex<InT>::value ? DFFT_C2C : DFFT_R2C;
    d.precision = std::is_same<typename real_of<OutT>::type, double>::value ? DFFT_FP64 : DFFT_FP32;
    d.layout = ViewLayout::value;
    const auto gin = detail::global_extents(in, in_topology);
    const auto gout = detail::global_extents(out, out_topology);
    for (std::size_t a = 0; a < N; ++a) {
      d.in_extents[a] = gin[a];
      d.out_extents[a] = gout[a];
      d.in_topology[a] = in_topology[a];
      d.out_topology[a] = out_topology[a];
    }
    d.in_topology_layout = InLayoutType::value;
    d.out_topology_layout = OutLayoutType::value;
    d.norm = static_cast<int>(norm);
    d.nranks = static_cast<int>(P);
    d.exec_mode = exec_mode;
    dfft_plan* p = nullptr;
    check(dfft_plan_create(&d, &p));
    m_plan.reset(p);
    for (std::size_t r = 0; r < P; ++r) {
      extents_type ei{}, eo{};
      check(dfft_plan_local_extents(p, static_cast<int>(r), 0, ei.data(), nullptr));
      check(dfft_plan_local_extents(p, static_cast<int>(r), 1, eo.data(), nullptr));
      if (ei != in[r].ext || eo != out[r].ext)
        throw Error(DFFT_ERROR_INVALID_EXTENTS, "Extents are not valid: rank " + std::to_string(r) + " holds " +
                                                    detail::extents_string(in[r].ext) + " -> " +
                                                    detail::extents_string(out[r].ext) + ", the topologies give " +
                                                    detail::extents_string(ei) + " -> " +
                                                    detail::extents_string(eo));
      m_in_extents.push_back(ei);
      m_out_extents.push_back(eo);
    }
  }

  std::string name() const {
    char buf[64];
    check(dfft_plan_name(m_plan.get(), buf, sizeof buf));
    return buf;
  }
  dfft_plan* native() const { return m_plan.get(); }
  std::size_t nranks() const { return m_in_extents.size(); }

  void forward_impl(const std::vector<in_view_type>& in, const std::vector<out_view_type>& out,
                    void* stream = nullptr) const {
    good(in, m_in_extents, out, m_out_extents);
    run(in, out, DFFT_FORWARD, stream);
  }
  void backward_impl(const std::vector<out_view_type>& out, const std::vector<in_view_type>& in,
                     void* stream = nullptr) const {
    good(out, m_out_extents, in, m_in_extents);
    run(out, in, DFFT_BACKWARD, stream);
  }

 private:
  // InternalPlan::good (KokkosFFT_Distributed_InternalPlan.hpp:109-142)
  template <typename A, typename B>
  static void good(const std::vector<A>& a, const std::vector<extents_type>& ea, const std::vector<B>& b,
                   const std::vector<extents_type>& eb) {
    if (a.size() != ea.size() || b.size() != eb.size())
      throw Error(DFFT_ERROR_MISMATCH, "one view per rank is required");
    for (std::size_t r = 0; r < ea.size(); ++r) {
      if (a[r].ext != ea[r])
        throw Error(DFFT_ERROR_MISMATCH, "extents of input View for plan and execution are not identical: View " +
                                             detail::extents_string(a[r].ext) + ", Plan " +
                                             detail::extents_string(ea[r]));
      if (b[r].ext != eb[r])
        throw Error(DFFT_ERROR_MISMATCH, "extents of output View for plan and execution are not identical: View " +
                                             detail::extents_string(b[r].ext) + ", Plan " +
                                             detail::extents_string(eb[r]));
    }
  }
  template <typename A, typename B>
  void run(const std::vector<A>& a, const std::vector<B>& b, int dir, void* stream) const {
    std::vector<void*> pa, pb;
    for (const auto& v : a) pa.push_back(const_cast<void*>(static_cast<const void*>(v.data())));
    for (const auto& v : b) pb.push_back(static_cast<void*>(v.data()));
    check(dfft_execute(m_plan.get(), pa.data(), pb.data(), dir, stream));
  }

  std::unique_ptr<dfft_plan, detail::PlanDeleter> m_plan;
  std::vector<extents_type> m_in_extents, m_out_extents;
};

// execute(plan, in, out, direction) (KokkosFFT_Distributed_Plan.hpp:212-238)
template <typename PlanType, typename InViews, typename OutViews>
void execute(const PlanType& plan, const InViews& in, const OutViews& out, Direction direction,
             void* stream = nullptr) {
  using in_value = typename InViews::value_type::value_type;
  using out_value = typename OutViews::value_type::value_type;
  switch (direction) {
    case Direction::forward:
      if constexpr (is_complex<out_value>::value) {
        plan.forward_impl(in, out, stream);
      } else {
        throw Error(DFFT_ERROR_MISMATCH, "Forward FFT operation requires complex output type.");
      }
      break;
    case Direction::backward:
      if constexpr (is_complex<in_value>::value) {
        plan.backward_impl(in, out, stream);
      } else {
        throw Error(DFFT_ERROR_MISMATCH, "Backward FFT operation requires complex input type.");
      }
      break;
    default:
      throw Error(DFFT_ERROR_INVALID_ARGUMENT, "Invalid FFT direction specified.");
  }
}

}  // namespace dfft

#endif  // DFFT_HPP
