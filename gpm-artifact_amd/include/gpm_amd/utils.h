// Shared enums and the abort convention of the reference host API
// (src/utils/utils.h:23-25, src/utils/logging.h), over the gpm C ABI.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../../include/gpm.h"

namespace project_AntiRF {

enum algorithms { TC, SM, CF, MC, FSM };
enum modes { v_centric, e_centric };
typedef unsigned long long AccType;

// The reference aborts through LOG(FATAL) on every error
// (src/utils/cuda_utils.h:21-28); the shim keeps that behaviour and prints
// the library's message first.
inline void check_status(int st, const char* what) {
  if (st != GPM_OK) {
    std::fprintf(stderr, "F %s failed (status %d): %s\n", what, st, gpm_last_error());
    std::abort();
  }
}

[[noreturn]] inline void fatal(const std::string& msg) {
  std::fprintf(stderr, "F %s\n", msg.c_str());
  std::abort();
}

}  // namespace project_AntiRF
