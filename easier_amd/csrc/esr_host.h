// Host-only helpers (no HIP headers): error reporting for the host entry
// points of libesrhip.so.
#pragma once

#include <cstdint>
#include <exception>

#include "esr_hip.h"

namespace esr {
void set_error(const char *fmt, ...);
}

#define ESR_HAPI extern "C" __attribute__((visibility("default")))

#define ESR_HREQUIRE(cond, ...)                                                \
  do {                                                                         \
    if (!(cond)) {                                                             \
      esr::set_error(__VA_ARGS__);                                             \
      return ESR_ERR_INVALID;                                                  \
    }                                                                          \
  } while (0)

#define ESR_HGUARD_BEGIN try {
#define ESR_HGUARD_END                                                         \
  }                                                                            \
  catch (const std::exception &ex) {                                           \
    esr::set_error("internal error: %s", ex.what());                           \
    return ESR_ERR_INTERNAL;                                                   \
  }                                                                            \
  catch (...) {                                                                \
    esr::set_error("internal error");                                          \
    return ESR_ERR_INTERNAL;                                                   \
  }
