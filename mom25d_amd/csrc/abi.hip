// C-ABI plumbing: thread-local error text, launch checks, version query.
#include "common.h"

#include <cstring>
#include <map>
#include <mutex>

namespace mom {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

int fail(const std::string& msg) {
  g_error = msg;
  return -1;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

static std::mutex g_lds_mu;
static std::map<const void*, size_t> g_lds;

int reserve_lds(const void* kernel, size_t bytes, const char* what) {
  std::lock_guard<std::mutex> lock(g_lds_mu);
  size_t& have = g_lds[kernel];
  if (bytes <= have) return 0;
  const hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return fail(std::string(what) + ": cannot reserve " + std::to_string(bytes) + " bytes of LDS: " +
                                   hipGetErrorString(e));
  have = bytes;
  return 0;
}

}  // namespace mom

extern "C" const char* mom_last_error(void) { return mom::g_error.c_str(); }

extern "C" int mom_abi_version(void) { return 1; }

// HIP graphs of library calls: the caller issues calls on `stream` between
// mom_graph_begin and mom_graph_end (nothing else may be enqueued on that
// stream meanwhile, and the calls must not allocate); the graph replays them
// with the same arguments.  Work the calls fork onto other streams (the LU
// look-ahead) joins back into the graph through its events.
extern "C" int mom_graph_begin(void* stream) {
  if (!stream) return mom::fail("mom_graph_begin: the default stream cannot be captured");
  MOM_TRY(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed));
  return 0;
}

extern "C" int mom_graph_end(void* stream, void** graph_exec) {
  if (!graph_exec) return mom::fail("mom_graph_end: null output");
  hipGraph_t g = nullptr;
  MOM_TRY(hipStreamEndCapture((hipStream_t)stream, &g));
  hipGraphExec_t e = nullptr;
  const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (rc != hipSuccess) return mom::fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(rc));
  *graph_exec = (void*)e;
  return 0;
}

extern "C" int mom_graph_launch(void* graph_exec, void* stream) {
  if (!graph_exec) return mom::fail("mom_graph_launch: null graph");
  MOM_TRY(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
  return 0;
}

extern "C" int mom_graph_destroy(void* graph_exec) {
  if (graph_exec) MOM_TRY(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
  return 0;
}
