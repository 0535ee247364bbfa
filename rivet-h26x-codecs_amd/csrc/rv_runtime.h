// Library state shared by the host translation units (runtime.hip owns it).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

namespace rv {
std::mutex& lib_mutex();           // serialises launches and per-entry calls
hipStream_t lib_stream();          // the kernel stream (null before rv_init)
bool lib_ready();                  // rv_init succeeded
int lib_fail(int code, const std::string& msg);  // sets rv_last_error, returns code
}  // namespace rv
