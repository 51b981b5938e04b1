// The device context shared by every entry point of the C ABI: one device, one stream.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

struct pbrt_ctx {
  int device = 0;
  int cu_count = 0;
  std::string name;
  hipStream_t stream = nullptr;
  // Device address of this device's copy of the no-alloc entry point's spill array (api.cpp).
  void* scratch = nullptr;
};
