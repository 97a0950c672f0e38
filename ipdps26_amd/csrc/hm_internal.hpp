// Internal helpers shared by the libhmotif translation units: error reporting and checks.
#pragma once
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "hmotif.h"

namespace hm {

// Per-thread message for hm_last_error(); every C entry point clears it on entry.
void set_error(const char* fmt, ...);
void clear_error();

inline hm_status fail(hm_status st, const char* msg) {
  set_error("%s", msg);
  return st;
}

// Device selected by hm_init on this thread.
int current_device();

// Device and pinned allocations and frees (AllocGuard, shared) never overlap a stream capture on
// another thread (CaptureGuard, exclusive): a free synchronizes the device, which a capture in
// progress must not see. A thread holding CaptureGuard passes its own AllocGuards through.
class AllocGuard {
 public:
  AllocGuard();
  ~AllocGuard();
  AllocGuard(const AllocGuard&) = delete;
  AllocGuard& operator=(const AllocGuard&) = delete;

 private:
  bool held_;
};
class CaptureGuard {
 public:
  CaptureGuard();
  ~CaptureGuard();
  CaptureGuard(const CaptureGuard&) = delete;
  CaptureGuard& operator=(const CaptureGuard&) = delete;
};

// While a NoAllocScope lives on a thread, growing a device or pinned buffer on that thread fails
// with HM_ERR_OVER_BUDGET instead of allocating (the entry points that promise no allocation run
// the shared pipeline under one).
bool alloc_forbidden();
class NoAllocScope {
 public:
  NoAllocScope();
  ~NoAllocScope();
  NoAllocScope(const NoAllocScope&) = delete;
  NoAllocScope& operator=(const NoAllocScope&) = delete;

 private:
  bool prev_;
};

// Device bytes g holds now: its store, count scratch, dynamic scratch and captured-update input
// (the capacities of its buffers, headroom included). g must not be in use by another thread.
int64_t graph_device_bytes(const hm_graph* g);

// Validates a caller CSR (offsets[0] == 0, non-decreasing, node ids >= 0) and returns its rows
// sorted with repeated nodes removed; *vmax_out = largest node id (-1 if none).
hm_status prepare_host_rows(const int64_t* offsets, const int32_t* nodes, int32_t E,
                            std::vector<int64_t>* off, std::vector<int32_t>* out, int32_t* vmax_out);

}  // namespace hm

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
namespace hm {
hm_status check_hip(hipError_t e, const char* what, const char* file, int line);
// Raises a kernel's dynamic-LDS limit to at least `bytes` (once per kernel and size; the
// attribute call is not repeated on every launch).
hm_status allow_lds(const void* kernel, size_t bytes);
// An instantiated stream capture and the graph it was instantiated from. The template lives as
// long as the executable graph (reset() destroys both), so nothing that inspects a launch, such
// as a profiler's graph hooks, can meet a destroyed template.
struct CapturedGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};
// Ends the stream capture on st and instantiates the captured graph into *out (reset first).
// out->exec stays null (HM_OK) when the capture holds any node other than kernels (memset nodes
// are not trusted on replays, copy nodes would refer to the capturing call's host memory), so
// such a sequence always runs eagerly.
hm_status end_capture(hipStream_t st, CapturedGraph* out);
}
// Evaluates a HIP call; on failure records the message and returns from the enclosing function.
#define HM_HIP(call)                                                          \
  do {                                                                        \
    hm_status _st = hm::check_hip((call), #call, __FILE__, __LINE__);         \
    if (_st != HM_OK) return _st;                                             \
  } while (0)
#define HM_TRY(expr)                  \
  do {                                \
    hm_status _st = (expr);           \
    if (_st != HM_OK) return _st;     \
  } while (0)
#endif
