// extern "C" boundary (include/dfft.h).  Every entry point converts C++
// exceptions into a status code and a thread-local message.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "dfft.h"
#include "fft_kernels.hpp"
#include "group.hpp"
#include "navier_stokes.hpp"
#include "plan.hpp"
#include "trace.hpp"
#include "planner.hpp"

struct dfft_plan {
  dfft::Plan* impl;
};

struct dfft_ns {
  dfft::NavierStokes* impl;
};

struct dfft_group {
  dfft::Group* impl;
};

namespace {
thread_local std::string g_last_error;

template <typename F> int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return DFFT_SUCCESS;
  } catch (const dfft::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return DFFT_ERROR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return DFFT_ERROR_INVALID_ARGUMENT;
  }
}

void need(bool ok, const char* what) {
  if (!ok) throw dfft::Error(DFFT_ERROR_INVALID_ARGUMENT, what);
}
void need(bool ok, const std::string& what) {
  if (!ok) throw dfft::Error(DFFT_ERROR_INVALID_ARGUMENT, what);
}

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw dfft::Error(DFFT_ERROR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void fill_info(const dfft::BlockInfo& b, dfft_block_info* o) {
  std::memset(o, 0, sizeof(*o));
  auto cp = [](const dfft::Vec& v, size_t* dst, std::size_t cap) {
    for (std::size_t i = 0; i < v.size() && i < cap; ++i) dst[i] = v[i];
  };
  cp(b.in_topology, o->in_topology, DFFT_MAX_RANK);
  cp(b.out_topology, o->out_topology, DFFT_MAX_RANK);
  cp(b.in_extents, o->in_extents, DFFT_MAX_RANK);
  cp(b.out_extents, o->out_extents, DFFT_MAX_RANK);
  cp(b.buffer_extents, o->buffer_extents, DFFT_MAX_RANK + 1);
  cp(b.in_map, o->in_map, DFFT_MAX_RANK);
  cp(b.out_map, o->out_map, DFFT_MAX_RANK);
  o->in_axis = b.in_axis;
  o->out_axis = b.out_axis;
  o->fft_dim = b.fft_dim;
  o->comm_axis = b.comm_axis;
  o->block_type = b.block_type == dfft::BlockType::FFT ? 1 : 0;
  o->block_idx = b.block_idx;
}

dfft::Vec vec(const size_t* p, int n) { return dfft::Vec(p, p + n); }
}  // namespace

extern "C" {

const char* dfft_last_error(void) { return g_last_error.c_str(); }

int dfft_version(int* major, int* minor) {
  return guarded([&] {
    need(major && minor, "null pointer");
    *major = 0;
    *minor = 1;
  });
}

int dfft_plan_create(const dfft_desc* desc, dfft_plan** plan) {
  return guarded([&] {
    need(desc && plan, "null pointer");
    *plan = nullptr;
    auto* p = new dfft_plan{nullptr};
    try {
      p->impl = new dfft::Plan(*desc);
    } catch (...) {
      delete p;
      throw;
    }
    *plan = p;
  });
}

int dfft_plan_destroy(dfft_plan* plan) {
  return guarded([&] {
    if (!plan) return;
    delete plan->impl;
    delete plan;
  });
}

int dfft_plan_name(const dfft_plan* plan, char* buf, size_t len) {
  return guarded([&] {
    need(plan && buf && len > 0, "null pointer");
    std::strncpy(buf, plan->impl->name().c_str(), len - 1);
    buf[len - 1] = '\0';
  });
}

int dfft_plan_local_extents(const dfft_plan* plan, int vrank, int which, size_t* extents, size_t* starts) {
  return guarded([&] {
    need(plan && extents, "null pointer");
    need(vrank >= 0, "negative rank");
    dfft::Vec e, s;
    plan->impl->local_starts(static_cast<std::size_t>(vrank), which, e, s);
    for (std::size_t i = 0; i < e.size(); ++i) {
      extents[i] = e[i];
      if (starts) starts[i] = s[i];
    }
  });
}

int dfft_plan_num_blocks(const dfft_plan* plan, int* nblocks) {
  return guarded([&] {
    need(plan && nblocks, "null pointer");
    const auto& a = plan->impl->analyses();
    *nblocks = a.empty() ? 0 : static_cast<int>(a[0].blocks.size());
  });
}

int dfft_plan_block_info(const dfft_plan* plan, int vrank, int idx, dfft_block_info* info) {
  return guarded([&] {
    need(plan && info, "null pointer");
    const auto& a = plan->impl->analyses();
    need(vrank >= 0 && static_cast<std::size_t>(vrank) < a.size(), "rank out of range");
    const auto& blocks = a[static_cast<std::size_t>(vrank)].blocks;
    need(idx >= 0 && static_cast<std::size_t>(idx) < blocks.size(), "block index out of range");
    fill_info(blocks[static_cast<std::size_t>(idx)], info);
  });
}

int dfft_plan_max_buffer_size(const dfft_plan* plan, size_t* size) {
  return guarded([&] {
    need(plan && size, "null pointer");
    *size = plan->impl->max_buffer_size();
  });
}

int dfft_plan_workspace_bytes(const dfft_plan* plan, size_t* bytes) {
  return guarded([&] {
    need(plan && bytes, "null pointer");
    *bytes = plan->impl->workspace_bytes();
  });
}

int dfft_execute(dfft_plan* plan, void* const* in, void* const* out, int direction, void* stream) {
  return guarded([&] {
    need(plan, "null plan");
    plan->impl->check_buffers(in, out, direction);
    plan->impl->execute(in, out, direction, static_cast<hipStream_t>(stream));
  });
}

int dfft_execute_adjoint(dfft_plan* plan, void* const* in, void* const* out, int direction, void* stream) {
  return guarded([&] {
    need(plan, "null plan");
    need(direction == DFFT_FORWARD || direction == DFFT_BACKWARD, "Invalid FFT direction specified.");
    // the adjoint of `direction` reads and writes what the opposite transform does
    plan->impl->check_buffers(in, out, direction == DFFT_FORWARD ? DFFT_BACKWARD : DFFT_FORWARD);
    plan->impl->execute_adjoint(in, out, direction, static_cast<hipStream_t>(stream));
  });
}

int dfft_plan_num_passes(const dfft_plan* plan, int direction, int* npasses) {
  return guarded([&] {
    need(plan && npasses, "null pointer");
    *npasses = plan->impl->num_passes(direction);
  });
}

int dfft_execute_timed(dfft_plan* plan, void* const* in, void* const* out, int direction, void* stream,
                       float* ms_per_pass, int capacity, int* npasses) {
  return guarded([&] {
    need(plan && npasses, "null pointer");
    plan->impl->check_buffers(in, out, direction);
    auto ms = plan->impl->execute_timed(in, out, direction, static_cast<hipStream_t>(stream));
    *npasses = static_cast<int>(ms.size());
    for (std::size_t i = 0; i < ms.size() && static_cast<int>(i) < capacity; ++i) ms_per_pass[i] = ms[i];
  });
}

int dfft_plan_set_option(dfft_plan* plan, const char* key, long value) {
  return guarded([&] {
    need(plan && key, "null pointer");
    plan->impl->set_option(key, value);
  });
}

int dfft_plan_get_option(const dfft_plan* plan, const char* key, long* value) {
  return guarded([&] {
    need(plan && key && value, "null pointer");
    *value = plan->impl->get_option(key);
  });
}

int dfft_plan_autotune(dfft_plan* plan, void* const* in, void* const* out, void* stream) {
  return guarded([&] {
    need(plan && in && out, "null pointer");
    plan->impl->autotune(in, out, static_cast<hipStream_t>(stream));
  });
}

namespace {

// Pack/unpack arguments of one rank's local array and its exchange buffer;
// `layout` is the view layout: LayoutRight buffers are (nprocs, e_0..e_{D-1}),
// LayoutLeft buffers (e_0..e_{D-1}, nprocs) (Extents.hpp:44-60,
// PackUnpack.hpp:176-178).
dfft::PackArgs pack_args(const size_t* arr_extents, const size_t* arr_map, int rank, const size_t* buffer_extents,
                         int nprocs, int axis, int layout, const char* what) {
  need(arr_extents && arr_map && buffer_extents, "null pointer");
  need(rank >= 2 && rank <= DFFT_MAX_RANK, std::string(what) + ": View rank must be in [2, 4]");
  need(axis >= 0 && axis < rank, "axis out of range");
  need(nprocs >= 1, "nprocs must be positive");
  need(layout == DFFT_LAYOUT_RIGHT || layout == DFFT_LAYOUT_LEFT, "layout must be DFFT_LAYOUT_RIGHT or LEFT");
  const bool left = layout == DFFT_LAYOUT_LEFT;
  const int pdim = left ? rank : 0;
  if (static_cast<int>(buffer_extents[pdim]) != nprocs)
    throw dfft::Error(DFFT_ERROR_MISMATCH, std::string("buffer extent ") + std::to_string(pdim) +
                                               " must equal the number of processes");
  dfft::PackArgs a{};
  a.ndim = rank;
  a.layout_left = left ? 1 : 0;
  a.nprocs = nprocs;
  a.axis = axis;
  std::size_t pos = dfft::get_index(vec(arr_map, rank), static_cast<std::size_t>(axis));
  a.merged_size = static_cast<int32_t>(arr_extents[pos]);
  a.total = 1;
  for (int k = 0; k < rank; ++k) {
    a.map[k] = static_cast<int32_t>(arr_map[k]);
    a.local_ext[k] = static_cast<int64_t>(arr_extents[k]);
  }
  for (int k = 0; k <= rank; ++k) {
    a.buf_ext[k] = static_cast<int64_t>(buffer_extents[k]);
    a.total *= a.buf_ext[k];
  }
  return a;
}

}  // namespace

int dfft_pack(const void* src, const size_t* src_extents, const size_t* src_map, int rank, void* buffer,
              const size_t* buffer_extents, int nprocs, int axis, int layout, int elem_size, void* stream) {
  return guarded([&] {
    dfft::PackArgs a = pack_args(src_extents, src_map, rank, buffer_extents, nprocs, axis, layout, "pack");
    std::size_t nsrc = 1;
    for (int k = 0; k < rank; ++k) nsrc *= src_extents[k];
    need(a.total == 0 || buffer, "null buffer");
    need(nsrc == 0 || src, "null source");
    // TransBlock::good (TransBlock.hpp:175-185): no aliasing between view and buffer
    if (src == buffer && a.total > 0)
      throw dfft::Error(DFFT_ERROR_ALIAS, "Input view must not be aliasing with send buffer");
    a.src = static_cast<const char*>(src);
    a.dst = static_cast<char*>(buffer);
    dfft::TraceRange range("dfft::pack");
    hip_ok(dfft::launch_pack(a, elem_size, false, static_cast<hipStream_t>(stream)), "pack");
  });
}

int dfft_unpack(const void* buffer, const size_t* buffer_extents, int nprocs, void* dst,
                const size_t* dst_extents, const size_t* dst_map, int rank, int axis, int layout, int elem_size,
                void* stream) {
  return guarded([&] {
    dfft::PackArgs a = pack_args(dst_extents, dst_map, rank, buffer_extents, nprocs, axis, layout, "unpack");
    if (dst == buffer && a.total > 0)
      throw dfft::Error(DFFT_ERROR_ALIAS, "Output view must not be aliasing with receive buffer");
    a.src = static_cast<const char*>(buffer);
    a.dst = static_cast<char*>(dst);
    dfft::TraceRange range("dfft::unpack");
    hip_ok(dfft::launch_pack(a, elem_size, true, static_cast<hipStream_t>(stream)), "unpack");
  });
}

int dfft_all2all(void* const* send, void* const* recv, int nprocs, size_t block_bytes, void* stream) {
  return guarded([&] {
    need(send && recv, "null pointer table");
    need(nprocs >= 1 && nprocs <= DFFT_MAX_VRANKS, "nprocs out of range");
    dfft::All2AllArgs a{};
    a.nprocs = nprocs;
    a.block_bytes = static_cast<int64_t>(block_bytes);
    for (int p = 0; p < nprocs; ++p) {
      for (int q = 0; q < nprocs; ++q)
        if (send[p] == recv[q] && block_bytes > 0)
          throw dfft::Error(DFFT_ERROR_ALIAS, "send and receive buffers must not alias");
      a.send[p] = static_cast<const char*>(send[p]);
      a.recv[p] = static_cast<char*>(recv[p]);
    }
    dfft::TraceRange range("dfft::all2all");
    hip_ok(dfft::launch_all2all(a, static_cast<hipStream_t>(stream)), "all2all");
  });
}

int dfft_exchange(void* const* in, void* const* out, int rank, const size_t* global_extents,
                  const size_t* in_topology, int in_topology_layout, const size_t* in_map,
                  const size_t* out_topology, int out_topology_layout, const size_t* out_map, int elem_size,
                  void* stream) {
  return guarded([&] {
    need(in && out && global_extents && in_topology && out_topology && in_map && out_map, "null pointer");
    need(rank >= 2 && rank <= DFFT_MAX_RANK, "rank must be in [2, 4]");
    dfft::Vec g = vec(global_extents, rank), ti = vec(in_topology, rank), to = vec(out_topology, rank);
    if (dfft::total_size(ti) != dfft::total_size(to))
      throw dfft::Error(DFFT_ERROR_INVALID_TOPOLOGY, "Input and output topologies must have the same size.");
    const std::size_t P = dfft::total_size(ti);
    need(P >= 1 && P <= static_cast<std::size_t>(DFFT_MAX_VRANKS), "number of ranks out of range");
    auto make = [&](const dfft::Vec& t, int tl, const size_t* map, void* const* ptrs) {
      dfft::DDesc d{};
      d.ndim = rank;
      std::size_t stride = 1;
      if (tl == DFFT_LAYOUT_LEFT) {
        for (int a = 0; a < rank; ++a) {
          d.rstride[a] = static_cast<int32_t>(stride);
          stride *= t[a];
        }
      } else {
        for (int a = rank; a-- > 0;) {
          d.rstride[a] = static_cast<int32_t>(stride);
          stride *= t[a];
        }
      }
      for (int a = 0; a < rank; ++a) {
        need(t[a] >= 1, "topology entries must be positive");
        d.ext[a] = static_cast<int32_t>(g[a]);
        d.ntopo[a] = static_cast<int32_t>(t[a]);
        d.base[a] = static_cast<int32_t>(g[a] / t[a]);
        d.rem[a] = static_cast<int32_t>(g[a] % t[a]);
        d.map[a] = static_cast<int32_t>(map[a]);
        d.div_big[a] = dfft::FastDiv(static_cast<uint32_t>(d.base[a] + 1));
        d.div_small[a] = dfft::FastDiv(static_cast<uint32_t>(d.base[a]));
      }
      for (std::size_t r = 0; r < P; ++r) d.ptr[r] = static_cast<char*>(ptrs[r]);
      return d;
    };
    dfft::RemapArgs ra{};
    ra.src = make(ti, in_topology_layout, in_map, in);
    ra.dst = make(to, out_topology_layout, out_map, out);
    ra.total = 1;
    for (int k = 0; k < rank; ++k) {
      const int a = static_cast<int>(out_map[k]);
      ra.iter_axis[k] = a;
      ra.iter_ext[k] = static_cast<int32_t>(g[a]);
      ra.iter_div[k] = dfft::FastDiv(static_cast<uint32_t>(g[a]));
      ra.total *= static_cast<int64_t>(g[a]);
    }
    for (std::size_t r = 0; r < P; ++r)
      for (std::size_t q = 0; q < P; ++q)
        if (in[r] == out[q] && in[r] != nullptr)
          throw dfft::Error(DFFT_ERROR_ALIAS, "input and output local arrays must not alias");
    dfft::TraceRange range("dfft::exchange (pack + all2all + unpack)");
    hip_ok(dfft::launch_remap(ra, elem_size, static_cast<hipStream_t>(stream)), "exchange");
  });
}

int dfft_group_create(int count, const dfft_desc* descs, dfft_group** group) {
  return guarded([&] {
    need(descs && group, "null pointer");
    *group = nullptr;
    auto* h = new dfft_group{nullptr};
    try {
      h->impl = new dfft::Group(count, descs);
    } catch (...) {
      delete h;
      throw;
    }
    *group = h;
  });
}

int dfft_group_destroy(dfft_group* group) {
  return guarded([&] {
    if (!group) return;
    delete group->impl;
    delete group;
  });
}

int dfft_group_execute(dfft_group* group, void* const* in, void* const* out, int direction, int use_graph,
                       void* stream) {
  return guarded([&] {
    need(group, "null group");
    group->impl->execute(in, out, direction, use_graph != 0, static_cast<hipStream_t>(stream));
  });
}

int dfft_group_execute_noalloc(dfft_group* group, void* const* in, void* const* out, int direction, void* stream) {
  return guarded([&] {
    need(group, "null group");
    group->impl->execute_static(in, out, direction, static_cast<hipStream_t>(stream));
  });
}

int dfft_group_num_launches(const dfft_group* group, int direction, int* nlaunches) {
  return guarded([&] {
    need(group && nlaunches, "null pointer");
    *nlaunches = group->impl->num_launches(direction);
  });
}

int dfft_group_set_option(dfft_group* group, const char* key, long value) {
  return guarded([&] {
    need(group && key, "null pointer");
    group->impl->set_option(key, value);
  });
}

int dfft_group_get_option(const dfft_group* group, const char* key, long* value) {
  return guarded([&] {
    need(group && key && value, "null pointer");
    *value = group->impl->get_option(key);
  });
}

int dfft_group_autotune(dfft_group* group, void* const* in, void* const* out, void* stream) {
  return guarded([&] {
    need(group && in && out, "null pointer");
    group->impl->autotune(in, out, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_create(const dfft_ns_desc* d, dfft_ns** ns) {
  return guarded([&] {
    need(d && ns, "null pointer");
    *ns = nullptr;
    auto* h = new dfft_ns{nullptr};
    try {
      h->impl = new dfft::NavierStokes(d->n, d->px, d->py, d->lx, d->nu, d->dt, d->form);
    } catch (...) {
      delete h;
      throw;
    }
    *ns = h;
  });
}

int dfft_ns_destroy(dfft_ns* ns) {
  return guarded([&] {
    if (!ns) return;
    delete ns->impl;
    delete ns;
  });
}

int dfft_ns_local_extents(const dfft_ns* ns, int vrank, int which, size_t* extents, size_t* starts) {
  return guarded([&] {
    need(ns && extents, "null pointer");
    need(vrank >= 0 && vrank < ns->impl->nranks(), "rank out of range");
    const dfft::NsBlock& b = which == 0 ? ns->impl->physical_block(vrank) : ns->impl->spectral_block(vrank);
    for (int a = 0; a < 3; ++a) {
      extents[a] = static_cast<size_t>(b.ext[a]);
      if (starts) starts[a] = static_cast<size_t>(b.start[a]);
    }
  });
}

int dfft_ns_init_taylor_green(dfft_ns* ns, double v0, void* stream) {
  return guarded([&] {
    need(ns, "null pointer");
    ns->impl->init_taylor_green(v0, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_set_physical(dfft_ns* ns, void* const* u3, void* stream) {
  return guarded([&] {
    need(ns && u3, "null pointer");
    ns->impl->check_fields(u3, false, "physical velocity");
    ns->impl->set_physical(u3, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_set_state(dfft_ns* ns, void* const* y3, void* stream) {
  return guarded([&] {
    need(ns && y3, "null pointer");
    ns->impl->check_fields(y3, true, "state");
    ns->impl->set_state(y3, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_get_state(dfft_ns* ns, void* const* y3, void* stream) {
  return guarded([&] {
    need(ns && y3, "null pointer");
    ns->impl->check_fields(y3, true, "state");
    ns->impl->get_state(y3, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_step(dfft_ns* ns, int nsteps, int use_graph, void* stream) {
  return guarded([&] {
    need(ns, "null pointer");
    ns->impl->step(nsteps, use_graph != 0, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_autotune(dfft_ns* ns, void* stream) {
  return guarded([&] {
    need(ns, "null pointer");
    ns->impl->autotune(static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_set_option(dfft_ns* ns, const char* key, long value) {
  return guarded([&] {
    need(ns && key, "null pointer");
    ns->impl->set_option(key, value);
  });
}

int dfft_ns_get_option(const dfft_ns* ns, const char* key, long* value) {
  return guarded([&] {
    need(ns && key && value, "null pointer");
    *value = ns->impl->get_option(key);
  });
}

int dfft_ns_energy(dfft_ns* ns, double* energy, void* stream) {
  return guarded([&] {
    need(ns && energy, "null pointer");
    *energy = ns->impl->energy(static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_velocity(dfft_ns* ns, void* const* u3, void* stream) {
  return guarded([&] {
    need(ns && u3, "null pointer");
    ns->impl->check_fields(u3, false, "velocity");
    ns->impl->velocity(u3, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_vorticity(dfft_ns* ns, void* const* w3, void* stream) {
  return guarded([&] {
    need(ns && w3, "null pointer");
    ns->impl->check_fields(w3, false, "vorticity");
    ns->impl->vorticity(w3, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_energy_spectrum(dfft_ns* ns, int num_bins, double* e_spec, double* k_vals, void* stream) {
  return guarded([&] {
    need(ns && e_spec, "null pointer");
    ns->impl->energy_spectrum(num_bins, e_spec, k_vals, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_diag(dfft_ns* ns, const char* out_dir, int diag_it, double* energy, void* stream) {
  return guarded([&] {
    need(ns && out_dir, "null pointer");
    const double e = ns->impl->diag(out_dir, diag_it, static_cast<hipStream_t>(stream));
    if (energy) *energy = e;
  });
}

int dfft_ns_rhs(dfft_ns* ns, void* const* y3, void* const* out3, void* stream) {
  return guarded([&] {
    need(ns && y3 && out3, "null pointer");
    ns->impl->check_fields(y3, true, "state");
    ns->impl->check_fields(out3, true, "right-hand side");
    ns->impl->apply_rhs(y3, out3, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_rhs_tangent(dfft_ns* ns, void* const* y3, void* const* d3, void* const* out3, void* stream) {
  return guarded([&] {
    need(ns && y3 && d3 && out3, "null pointer");
    ns->impl->check_fields(y3, true, "state");
    ns->impl->check_fields(d3, true, "direction");
    ns->impl->check_fields(out3, true, "tangent");
    ns->impl->apply_rhs_tangent(y3, d3, out3, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_rhs_adjoint(dfft_ns* ns, void* const* y3, void* const* l3, void* const* out3, void* stream) {
  return guarded([&] {
    need(ns && y3 && l3 && out3, "null pointer");
    ns->impl->check_fields(y3, true, "state");
    ns->impl->check_fields(l3, true, "adjoint variable");
    ns->impl->check_fields(out3, true, "adjoint");
    ns->impl->apply_rhs_adjoint(y3, l3, out3, static_cast<hipStream_t>(stream));
  });
}

int dfft_ns_gradient(dfft_ns* ns, int nsteps, int checkpoint_every, void* const* u0, void* const* grad, double* J,
                     void* stream) {
  return guarded([&] {
    need(ns && u0 && grad && J, "null pointer");
    ns->impl->check_fields(u0, false, "initial velocity");
    ns->impl->check_fields(grad, false, "gradient");
    *J = ns->impl->gradient(nsteps, checkpoint_every, u0, grad, static_cast<hipStream_t>(stream));
  });
}

int dfft_get_all_slab_topologies(int rank, const size_t* in_topology, const size_t* out_topology, int fft_rank,
                                 const int* axes, size_t* topologies, int* ntopo) {
  return guarded([&] {
    need(in_topology && out_topology && axes && topologies && ntopo, "null pointer");
    need(rank >= 1 && rank <= 8 && fft_rank >= 1 && fft_rank <= 3, "bad rank");
    auto ax = dfft::convert_negative_axes(std::vector<int>(axes, axes + fft_rank), rank);
    auto t = dfft::get_all_slab_topologies(vec(in_topology, rank), vec(out_topology, rank), ax);
    need(static_cast<int>(t.size()) <= *ntopo, "output capacity too small");
    for (std::size_t i = 0; i < t.size(); ++i)
      for (int k = 0; k < rank; ++k) topologies[i * rank + k] = t[i][k];
    *ntopo = static_cast<int>(t.size());
  });
}

int dfft_get_all_pencil_topologies(int rank, const size_t* in_topology, int in_layout, const size_t* out_topology,
                                   int out_layout, int fft_rank, const int* axes, size_t* topologies,
                                   size_t* trans_axes, size_t* layouts, int* ntopo) {
  return guarded([&] {
    need(in_topology && out_topology && axes && topologies && trans_axes && layouts && ntopo, "null pointer");
    need(rank >= 1 && rank <= 8 && fft_rank >= 1 && fft_rank <= 3, "bad rank");
    auto ax = dfft::convert_negative_axes(std::vector<int>(axes, axes + fft_rank), rank);
    std::vector<dfft::Vec> t;
    dfft::Vec ta, la;
    dfft::get_all_pencil_topologies(vec(in_topology, rank), vec(out_topology, rank), ax,
                                    in_layout == DFFT_LAYOUT_LEFT ? dfft::Layout::Left : dfft::Layout::Right,
                                    out_layout == DFFT_LAYOUT_LEFT ? dfft::Layout::Left : dfft::Layout::Right, t,
                                    ta, la);
    need(static_cast<int>(t.size()) <= *ntopo, "output capacity too small");
    for (std::size_t i = 0; i < t.size(); ++i) {
      for (int k = 0; k < rank; ++k) topologies[i * rank + k] = t[i][k];
      layouts[i] = la[i];
      if (i + 1 < t.size()) trans_axes[i] = ta[i];
    }
    *ntopo = static_cast<int>(t.size());
  });
}

int dfft_block_analyses(const dfft_desc* desc, int vrank, dfft_block_info* infos, int* nblocks,
                        size_t* max_buffer_size) {
  return guarded([&] {
    need(desc && infos && nblocks, "null pointer");
    const int rank = desc->rank;
    need(rank >= 2 && rank <= DFFT_MAX_RANK, "rank must be in [2, 4]");
    dfft::Vec gin = vec(desc->in_extents, rank), gout = vec(desc->out_extents, rank);
    dfft::Vec ti = vec(desc->in_topology, rank), to = vec(desc->out_topology, rank);
    auto ax = dfft::convert_negative_axes(std::vector<int>(desc->axes, desc->axes + desc->fft_rank), rank);
    auto il = desc->in_topology_layout == DFFT_LAYOUT_LEFT ? dfft::Layout::Left : dfft::Layout::Right;
    auto ol = desc->out_topology_layout == DFFT_LAYOUT_LEFT ? dfft::Layout::Left : dfft::Layout::Right;
    auto kind = dfft::get_common_topology_type(ti, to);
    if (kind != dfft::TopologyType::Slab && kind != dfft::TopologyType::Pencil)
      throw dfft::Error(DFFT_ERROR_INVALID_TOPOLOGY, "block analyses need slab or pencil topologies");
    need(vrank >= 0 && static_cast<std::size_t>(vrank) < dfft::total_size(ti), "rank out of range");
    dfft::Vec local = dfft::compute_local_extents(gin, ti, dfft::identity_map(rank), vrank, il);
    auto vl = desc->layout == DFFT_LAYOUT_LEFT ? dfft::Layout::Left : dfft::Layout::Right;
    auto res = dfft::block_analyses(kind, desc->transform == DFFT_R2C, vl, local, gin, gout, ti, to, ax,
                                    static_cast<std::size_t>(vrank), il, ol);
    need(static_cast<int>(res.blocks.size()) <= *nblocks, "output capacity too small");
    for (std::size_t i = 0; i < res.blocks.size(); ++i) fill_info(res.blocks[i], &infos[i]);
    *nblocks = static_cast<int>(res.blocks.size());
    if (max_buffer_size) *max_buffer_size = res.max_buffer_size;
  });
}

int dfft_rank_to_coord(int rank, const size_t* topology, int topology_layout, int vrank, size_t* coord) {
  return guarded([&] {
    need(topology && coord, "null pointer");
    need(rank >= 1 && rank <= DFFT_MAX_RANK, "rank must be in [1, 4]");
    need(vrank >= 0, "rank must be non-negative");
    auto tl = topology_layout == DFFT_LAYOUT_LEFT ? dfft::Layout::Left : dfft::Layout::Right;
    dfft::Vec c = dfft::rank_to_coord(vec(topology, rank), static_cast<std::size_t>(vrank), tl);
    for (int k = 0; k < rank; ++k) coord[k] = c[k];
  });
}

int dfft_compute_local_extents_and_starts(int rank, const size_t* global_extents, const size_t* topology,
                                          int topology_layout, int vrank, size_t* extents, size_t* starts) {
  return guarded([&] {
    need(global_extents && topology && extents && starts, "null pointer");
    need(rank >= 1 && rank <= DFFT_MAX_RANK, "rank must be in [1, 4]");
    need(vrank >= 0, "rank must be non-negative");
    auto tl = topology_layout == DFFT_LAYOUT_LEFT ? dfft::Layout::Left : dfft::Layout::Right;
    dfft::Vec local, st;
    dfft::compute_local_extents_and_starts(vec(global_extents, rank), vec(topology, rank),
                                           static_cast<std::size_t>(vrank), tl, local, st);
    for (int k = 0; k < rank; ++k) {
      extents[k] = local[k];
      starts[k] = st[k];
    }
  });
}

}  // extern "C"
