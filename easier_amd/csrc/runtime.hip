// Library-level entry points: error reporting, device query, runtime
// compilation (hipRTC) of fused kernels, HIP graph capture and replay.
#include <hip/hiprtc.h>

#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "esr_common.h"

namespace esr {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[4096];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

} // namespace esr

ESR_API int esr_abi_version(void) { return 1; }

ESR_API const char *esr_last_error(void) { return esr::g_last_error.c_str(); }

ESR_API int esr_device_arch(char *buf, int buflen) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(buf != nullptr && buflen > 0, "esr_device_arch: empty buffer");
  int dev = 0;
  ESR_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  ESR_HIP(hipGetDeviceProperties(&prop, dev));
  std::snprintf(buf, (size_t)buflen, "%s", prop.gcnArchName);
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_stream_sync(void *stream) {
  ESR_HIP(hipStreamSynchronize(esr::as_stream(stream)));
  return ESR_OK;
}

// ---------------------------------------------------------------------------
// hipRTC: one program per esr.compile() holds every fused kernel of it.
// ---------------------------------------------------------------------------
ESR_API int esr_rtc_compile(const char *source, const char *name,
                            const char *options, void **module_out) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(source && module_out, "esr_rtc_compile: null argument");
  *module_out = nullptr;
  hiprtcProgram prog;
  hiprtcResult rc =
      hiprtcCreateProgram(&prog, source, name ? name : "esr_fused.hip", 0,
                          nullptr, nullptr);
  if (rc != HIPRTC_SUCCESS) {
    esr::set_error("hiprtcCreateProgram: %s", hiprtcGetErrorString(rc));
    return ESR_ERR_RTC;
  }
  std::vector<std::string> opts;
  if (options) {
    std::istringstream ss(options);
    std::string tok;
    while (ss >> tok) opts.push_back(tok);
  }
  std::vector<const char *> copts;
  for (auto &o : opts) copts.push_back(o.c_str());
  rc = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    esr::set_error("hipRTC compile failed: %s\n%s", hiprtcGetErrorString(rc),
                   log.c_str());
    return ESR_ERR_RTC;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  hipModule_t mod;
  ESR_HIP(hipModuleLoadData(&mod, code.data()));
  *module_out = reinterpret_cast<void *>(mod);
  return ESR_OK;
  ESR_GUARD_END
}

// Compile without loading (no device needed): returns ESR_ERR_RTC with the
// compiler log on failure, and the code object size in *code_bytes.
ESR_API int esr_rtc_check(const char *source, const char *options, int64_t *code_bytes) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(source && code_bytes, "esr_rtc_check: null argument");
  hiprtcProgram prog;
  hiprtcResult rc = hiprtcCreateProgram(&prog, source, "esr_check.hip", 0, nullptr, nullptr);
  if (rc != HIPRTC_SUCCESS) {
    esr::set_error("hiprtcCreateProgram: %s", hiprtcGetErrorString(rc));
    return ESR_ERR_RTC;
  }
  std::vector<std::string> opts;
  if (options) {
    std::istringstream ss(options);
    std::string tok;
    while (ss >> tok) opts.push_back(tok);
  }
  std::vector<const char *> copts;
  for (auto &o : opts) copts.push_back(o.c_str());
  rc = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    esr::set_error("hipRTC compile failed: %s\n%s", hiprtcGetErrorString(rc), log.c_str());
    return ESR_ERR_RTC;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  hiprtcDestroyProgram(&prog);
  *code_bytes = (int64_t)cs;
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_rtc_function(void *module, const char *kernel_name,
                             void **fn_out) {
  ESR_REQUIRE(module && kernel_name && fn_out, "esr_rtc_function: null arg");
  hipFunction_t fn;
  ESR_HIP(hipModuleGetFunction(&fn, reinterpret_cast<hipModule_t>(module),
                               kernel_name));
  *fn_out = reinterpret_cast<void *>(fn);
  return ESR_OK;
}

ESR_API int esr_rtc_unload(void *module) {
  if (!module) return ESR_OK;
  ESR_HIP(hipModuleUnload(reinterpret_cast<hipModule_t>(module)));
  return ESR_OK;
}

ESR_API int esr_launch(void *fn, uint32_t gx, uint32_t gy, uint32_t gz,
                       uint32_t bx, uint32_t by, uint32_t bz, uint32_t shmem,
                       void *stream, void *argbuf, size_t argbytes) {
  ESR_REQUIRE(fn != nullptr, "esr_launch: null function");
  ESR_REQUIRE(gx > 0 && gy > 0 && gz > 0 && bx > 0 && by > 0 && bz > 0,
              "esr_launch: empty launch %u,%u,%u x %u,%u,%u", gx, gy, gz, bx,
              by, bz);
  ESR_REQUIRE((uint64_t)bx * by * bz <= 1024, "esr_launch: block too large");
  size_t sz = argbytes;
  void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, argbuf,
                    HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz, HIP_LAUNCH_PARAM_END};
  ESR_HIP(hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(fn), gx, gy,
                                gz, bx, by, bz, shmem, esr::as_stream(stream),
                                nullptr, config));
  return ESR_OK;
}

// ---------------------------------------------------------------------------
// HIP graphs
// ---------------------------------------------------------------------------
ESR_API int esr_graph_begin(void *stream) {
  ESR_REQUIRE(stream != nullptr, "esr_graph_begin: capture needs a non-null stream");
  ESR_HIP(hipStreamBeginCapture(esr::as_stream(stream),
                                hipStreamCaptureModeThreadLocal));
  return ESR_OK;
}

ESR_API int esr_graph_end(void *stream, void **exec_out) {
  ESR_REQUIRE(stream != nullptr && exec_out != nullptr, "esr_graph_end: null arg");
  hipGraph_t graph = nullptr;
  ESR_HIP(hipStreamEndCapture(esr::as_stream(stream), &graph));
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) {
    esr::set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
    return ESR_ERR_HIP;
  }
  *exec_out = reinterpret_cast<void *>(exec);
  return ESR_OK;
}

ESR_API int esr_graph_launch(void *exec, void *stream) {
  ESR_REQUIRE(exec != nullptr, "esr_graph_launch: null graph");
  ESR_HIP(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(exec),
                         esr::as_stream(stream)));
  return ESR_OK;
}

ESR_API int esr_graph_destroy(void *exec) {
  if (!exec) return ESR_OK;
  ESR_HIP(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(exec)));
  return ESR_OK;
}
