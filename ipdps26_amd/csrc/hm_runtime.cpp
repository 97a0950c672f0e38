// Error reporting and device selection for libhmotif.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <shared_mutex>
#include <vector>
#include <string>
#include <unordered_map>

#include "hm_internal.hpp"
#include "hm_tables.hpp"

namespace hm {

static thread_local std::string t_error;
static thread_local int t_device = 0;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_error = buf;
}

void clear_error() { t_error.clear(); }

int current_device() { return t_device; }

hm_status check_hip(hipError_t e, const char* what, const char* file, int line) {
  if (e == hipSuccess) return HM_OK;
  set_error("%s failed at %s:%d: %s", what, file, line, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? HM_ERR_OOM : HM_ERR_DEVICE;
}

hm_status allow_lds(const void* kernel, size_t bytes) {
  if (bytes <= 65536) return HM_OK;
  static std::mutex mu;
  static std::unordered_map<const void*, size_t> granted;
  std::lock_guard<std::mutex> lk(mu);
  size_t& g = granted[kernel];
  if (bytes <= g) return HM_OK;
  hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return check_hip(e, "hipFuncSetAttribute", __FILE__, __LINE__);
  g = bytes;
  return HM_OK;
}

namespace {
std::shared_mutex g_capture_mu;
thread_local bool t_capturing = false;
thread_local bool t_no_alloc = false;
}  // namespace

bool alloc_forbidden() { return t_no_alloc; }
NoAllocScope::NoAllocScope() : prev_(t_no_alloc) { t_no_alloc = true; }
NoAllocScope::~NoAllocScope() { t_no_alloc = prev_; }

AllocGuard::AllocGuard() : held_(!t_capturing) {
  if (held_) g_capture_mu.lock_shared();
}
AllocGuard::~AllocGuard() {
  if (held_) g_capture_mu.unlock_shared();
}
CaptureGuard::CaptureGuard() {
  g_capture_mu.lock();
  t_capturing = true;
}
CaptureGuard::~CaptureGuard() {
  t_capturing = false;
  g_capture_mu.unlock();
}

hm_status end_capture(hipStream_t st, CapturedGraph* out) {
  out->reset();
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(st, &graph);
  if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture", __FILE__, __LINE__);
  size_t n = 0;
  e = hipGraphGetNodes(graph, nullptr, &n);
  std::vector<hipGraphNode_t> nodes(n);
  if (e == hipSuccess && n) e = hipGraphGetNodes(graph, nodes.data(), &n);
  // only kernel nodes are replayed: a memset node has given wrong results on replays, and a copy
  // node would refer to host memory of the capturing call
  bool other_node = false;
  for (size_t k = 0; e == hipSuccess && k < n; ++k) {
    hipGraphNodeType t;
    e = hipGraphNodeGetType(nodes[k], &t);
    other_node = other_node || (e == hipSuccess && t != hipGraphNodeTypeKernel && t != hipGraphNodeTypeEmpty);
  }
  hipGraphExec_t exec = nullptr;
  if (e == hipSuccess && !other_node) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess || !exec) {
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    return e != hipSuccess ? check_hip(e, "captured graph", __FILE__, __LINE__) : HM_OK;
  }
  out->graph = graph;
  out->exec = exec;
  return HM_OK;
}

}  // namespace hm

extern "C" {

const char* hm_status_string(hm_status st) {
  switch (st) {
    case HM_OK: return "ok";
    case HM_ERR_OOM: return "out of memory";
    case HM_ERR_OVER_BUDGET: return "over budget";
    case HM_ERR_BAD_INPUT: return "bad input";
    case HM_ERR_DEVICE: return "device error";
    case HM_ERR_UNSUPPORTED: return "unsupported input";
    case HM_ERR_BUFFER: return "buffer too small";
  }
  return "unknown status";
}

const char* hm_last_error(void) { return hm::t_error.c_str(); }

hm_status hm_bin_table(int32_t which, int8_t out[128]) {
  hm::clear_error();
  if (!out || (which != 0 && which != 1)) return hm::fail(HM_ERR_BAD_INPUT, "hm_bin_table: bad argument");
  const int8_t* t = which == 0 ? hm::kStaticBin : hm::kTemporalBin;
  for (int i = 0; i < 128; ++i) out[i] = t[i];
  return HM_OK;
}

hm_status hm_init(int32_t device) {
  hm::clear_error();
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return hm::fail(HM_ERR_DEVICE, "hm_init: no HIP device visible");
  if (device < 0 || device >= n) return hm::fail(HM_ERR_BAD_INPUT, "hm_init: device index out of range");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess)
    return hm::fail(HM_ERR_DEVICE, "hm_init: hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    hm::set_error("hm_init: device %d is %s; libhmotif is built for gfx950 only", device,
                  prop.gcnArchName);
    return HM_ERR_DEVICE;
  }
  hm::t_device = device;
  return HM_OK;
}

}  // extern "C"
