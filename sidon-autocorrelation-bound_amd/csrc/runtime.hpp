// runtime.hpp -- host-side helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <cstdint>

#include "../../include/sidon.h"

namespace sidon {

// Per-device attributes the launches need (read once by sidon_init or on
// first use on a device; nothing is allocated).  `device` is stored last,
// with release order, once the other fields are complete; readers that see
// it equal to their device may use the rest.
struct DeviceState {
  std::atomic<int> device{-1};
  int lds_max = 0;        // max dynamic LDS bytes per workgroup
  int cu_count = 0;
  void* scratch = nullptr;  // this device's address of the no-alloc polish scratch
};

// Bytes of the static device array behind sidon_polish_f64_noalloc.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;

// Device address of that array on the current device (hipGetSymbolAddress;
// polish.hip, where the array lives).
int noalloc_scratch_address(void** out);

// Returns the state of the current device (initialising it on first use), or
// nullptr with *rc set on failure.
DeviceState* device_state(int* rc);

inline int hip_status(hipError_t e) { return e == hipSuccess ? SIDON_OK : SIDON_ERR_HIP; }

// Records the last HIP error string for sidon_last_error().
void set_last_error(const char* where, hipError_t e);
void set_last_error_msg(const char* msg);

#define SIDON_HIP_TRY(call)                                    \
  do {                                                         \
    hipError_t _e = (call);                                    \
    if (_e != hipSuccess) {                                    \
      ::sidon::set_last_error(#call, _e);                      \
      return SIDON_ERR_HIP;                                    \
    }                                                          \
  } while (0)

// Sets the caller-owned device counter *counter to zero on `stream` with a
// one-thread kernel (no hipMemsetAsync, so the sequence is safe inside
// captured graphs).
int reset_counter(int* counter, hipStream_t stream);

}  // namespace sidon
