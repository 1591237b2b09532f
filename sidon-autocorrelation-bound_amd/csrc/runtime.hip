// runtime.hip -- per-device attributes, error reporting and queries of the C ABI.
#include <cstdio>
#include <cstring>
#include <mutex>

#include "common.hpp"
#include "runtime.hpp"

extern "C" long long sidon_ls_lds_bytes_impl(int kernel, int n, int block);
extern "C" long long sidon_eval_lds_bytes_impl(int n);
extern "C" long long sidon_fw_lds_bytes_impl(int n, int block);

namespace sidon {
namespace {

constexpr int kMaxDevices = 64;
DeviceState g_states[kMaxDevices];
std::mutex g_mu;
thread_local char g_err[512] = "no error";

__global__ void k_reset_counter(int* p) { *p = 0; }

}  // namespace

void set_last_error(const char* where, hipError_t e) {
  std::snprintf(g_err, sizeof g_err, "%s: %s", where, hipGetErrorString(e));
}
void set_last_error_msg(const char* msg) { std::snprintf(g_err, sizeof g_err, "%s", msg); }

// Fills g_states[dev] with `dev` current.
static int init_current(int dev) {
  DeviceState& s = g_states[dev];
  int lds_block = 0, lds_optin = 0, cus = 0;
  SIDON_HIP_TRY(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess)
    lds_optin = 0;
  (void)hipGetLastError();
  SIDON_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  void* scratch = nullptr;
  if (int rc = noalloc_scratch_address(&scratch)) return rc;
  s.lds_max = lds_optin > lds_block ? lds_optin : lds_block;
  s.cu_count = cus;
  s.scratch = scratch;
  s.device.store(dev, std::memory_order_release);
  return SIDON_OK;
}

// Called with g_mu held.  Leaves the caller's current device as it was.
static int init_device(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return SIDON_ERR_ARG;
  if (g_states[dev].device.load(std::memory_order_acquire) == dev) return SIDON_OK;
  int cur = 0;
  SIDON_HIP_TRY(hipGetDevice(&cur));
  if (cur != dev) SIDON_HIP_TRY(hipSetDevice(dev));
  const int rc = init_current(dev);
  if (cur != dev) {
    const hipError_t e = hipSetDevice(cur);
    if (rc == SIDON_OK && e != hipSuccess) {
      set_last_error("hipSetDevice(restore)", e);
      return SIDON_ERR_HIP;
    }
  }
  return rc;
}

DeviceState* device_state(int* rc) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_last_error("hipGetDevice", e);
    *rc = SIDON_ERR_HIP;
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) { *rc = SIDON_ERR_ARG; return nullptr; }
  if (g_states[dev].device.load(std::memory_order_acquire) != dev) {
    std::lock_guard<std::mutex> lk(g_mu);
    const int r = init_device(dev);
    if (r) { *rc = r; return nullptr; }
  }
  *rc = SIDON_OK;
  return &g_states[dev];
}

int reset_counter(int* counter, hipStream_t stream) {
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, stream, counter);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}

}  // namespace sidon

using namespace sidon;

extern "C" int sidon_init(int device) {
  int dev = device;
  if (dev < 0) {
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { set_last_error("hipGetDevice", e); return SIDON_ERR_HIP; }
  }
  std::lock_guard<std::mutex> lk(g_mu);
  return init_device(dev);
}

extern "C" const char* sidon_last_error(void) { return g_err; }

extern "C" const char* sidon_status_string(int status) {
  switch (status) {
    case SIDON_OK: return "ok";
    case SIDON_ERR_ARG: return "invalid argument";
    case SIDON_ERR_LDS: return "length needs more LDS than one workgroup may use";
    case SIDON_ERR_HIP: return "HIP runtime error";
    case SIDON_ERR_UNSUPPORTED: return "unsupported configuration";
    default: return "unknown status";
  }
}

extern "C" long long sidon_lds_bytes(int kernel, int n, int block) {
  switch (kernel) {
    case 0: case 1: case 4: return sidon_ls_lds_bytes_impl(kernel, n, block);
    case 2: return sidon_eval_lds_bytes_impl(n);
    case 3: return sidon_fw_lds_bytes_impl(n, block);
    default: return -1;
  }
}

extern "C" int sidon_lds_limit(void) {
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  return ds ? ds->lds_max : -rc;
}
