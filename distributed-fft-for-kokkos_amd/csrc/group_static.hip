// Group execution without device allocation (dfft_group_execute_noalloc):
// the per-call pass tables (LinePass records and virtual-block prefixes of
// every grouped launch) live in one static device array per GPU instead of
// hipMalloc'ed tables, for callers that cannot allocate (code under stream
// capture, serving loops, harnesses that forbid hipMalloc).
//
// The array is shared by every group and every call on a device and keeps
// the last call's tables; calls that overlap on one device race (each call
// rewrites the tables the other's kernels read).  Nothing in it needs to
// start at zero: every call writes the tables it launches with.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "group.hpp"

namespace dfft {

// the static table area: at most Group::kStaticTableBytes per device
static __device__ __align__(256) unsigned char scratch[Group::kStaticTableBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(DFFT_ERROR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// device address of `scratch` on `device` (the current device), looked up
// once per device: host code never uses the array's name as an address
char* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(DFFT_ERROR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress of the static table area");
    if (!p) throw Error(DFFT_ERROR_HIP, "hipGetSymbolAddress returned a null address");
    g_scratch_ptr[device] = p;
  }
  return static_cast<char*>(g_scratch_ptr[device]);
}

std::size_t align256(std::size_t n) { return (n + 255) / 256 * 256; }

}  // namespace

void Group::execute_static(void* const* in, void* const* out, int direction, hipStream_t s) {
  if (!in || !out) throw Error(DFFT_ERROR_INVALID_ARGUMENT, "null pointer table");
  if (direction != DFFT_FORWARD && direction != DFFT_BACKWARD)
    throw Error(DFFT_ERROR_INVALID_ARGUMENT, "Invalid FFT direction specified.");
  std::vector<void*> vin(in, in + members_.size()), vout(out, out + members_.size());
  int cur = -1;
  check(hipGetDevice(&cur), "hipGetDevice");
  if (cur != device_) check(hipSetDevice(device_), "hipSetDevice");
  const int dir = check_call(vin, vout, direction);
  Tables t;
  build_tables(vin, vout, dir, t);
  // place every slot's records and prefix in the static area (256-byte aligned)
  const std::size_t ns = t.items.size();
  std::vector<std::size_t> off_items(ns, 0), off_first(ns, 0);
  std::size_t used = 0;
  for (std::size_t slot = 0; slot < ns; ++slot) {
    if (t.count[slot] == 0) continue;
    off_items[slot] = used;
    used = align256(used + sizeof(LinePass) * t.items[slot].size());
    off_first[slot] = used;
    used = align256(used + sizeof(int32_t) * t.first[slot].size());
  }
  if (used > kStaticTableBytes)
    throw Error(DFFT_ERROR_UNSUPPORTED, "the pass tables of this call need " + std::to_string(used) +
                                            " bytes; the static table area holds " +
                                            std::to_string(kStaticTableBytes));
  if (used == 0) {  // nothing grouped: every pass is launched on its own
    std::vector<LinePass*> none(ns, nullptr);
    std::vector<int32_t*> nonef(ns, nullptr);
    launch_tables(t, none.data(), nonef.data(), dir, s);
    return;
  }
  std::vector<unsigned char> host(used, 0);
  for (std::size_t slot = 0; slot < ns; ++slot) {
    if (t.count[slot] == 0) continue;
    std::memcpy(host.data() + off_items[slot], t.items[slot].data(), sizeof(LinePass) * t.items[slot].size());
    std::memcpy(host.data() + off_first[slot], t.first[slot].data(), sizeof(int32_t) * t.first[slot].size());
  }
  char* base = scratch_on(device_);
  // ordered after earlier work on the stream (which may still read the area);
  // the host copy is released once the stream has taken it
  check(hipMemcpyAsync(base, host.data(), used, hipMemcpyHostToDevice, s), "hipMemcpyAsync of the pass tables");
  check(hipStreamSynchronize(s), "hipStreamSynchronize");
  std::vector<LinePass*> items(ns, nullptr);
  std::vector<int32_t*> first(ns, nullptr);
  for (std::size_t slot = 0; slot < ns; ++slot) {
    if (t.count[slot] == 0) continue;
    items[slot] = reinterpret_cast<LinePass*>(base + off_items[slot]);
    first[slot] = reinterpret_cast<int32_t*>(base + off_first[slot]);
  }
  launch_tables(t, items.data(), first.data(), dir, s);
}

}  // namespace dfft
