// Fixed device workspace of the allocation-free occupied-factor J/K entry
// point (vibeqc_fock_plan_evaluate_occupied_noalloc).
//
// The workspace is one file-scope device array, `scratch`, of
// kNoAllocScratchBytes.  Host code never passes the array's name to a kernel
// (in host code it is a host-side stand-in, not a device address): the first
// use on a device resolves the device's own copy with hipGetSymbolAddress,
// with that device current, and every later call on the device uses that
// address.  Every call on a device shares the device's array, which keeps
// the previous call's contents; calls that overlap on one device race.
#include <array>
#include <mutex>

#include "noalloc.hpp"

namespace vqc {
namespace {

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_lock;
std::array<void*, kMaxDevices> scratch_address{};

}  // namespace

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index out of range");
  std::lock_guard<std::mutex> lock(scratch_lock);
  if (!scratch_address[device]) {
    int current = 0;
    VQC_HIP(hipGetDevice(&current));
    VQC_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    (void)hipSetDevice(current);
    hip_check(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) throw DeviceError("no device address for the allocation-free scratch array");
    scratch_address[device] = ptr;
  }
  return scratch_address[device];
}

}  // namespace vqc
