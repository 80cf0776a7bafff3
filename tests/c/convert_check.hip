// Device check of the f32 -> integer conversions of include/st_math.h: the
// one-instruction GPU forms against the portable definitions compiled for the
// same device, over every f32 bit pattern.  Built by the package build into
// tests/c/libconvert_check.so and run by tests/test_gpu_convert.py.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "st_math.h"

__global__ void __launch_bounds__(256) k_convert_check(unsigned long long* bad, uint32_t* first) {
  uint32_t local = 0, firstBad = 0xffffffffu;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
    float x = st_u2f(uint32_t(i));
    bool ok = st_f2i_sat(x) == st_f2i_sat_portable(x) && st_f2u_sat(x) == st_f2u_sat_portable(x);
    if (!ok) {
      local++;
      firstBad = min(firstBad, uint32_t(i));
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    local += uint32_t(__shfl_xor(int(local), off));
    firstBad = min(firstBad, uint32_t(__shfl_xor(int(firstBad), off)));
  }
  if ((threadIdx.x & 63u) == 0u && local) {
    atomicAdd(bad, (unsigned long long)local);
    atomicMin(first, firstBad);
  }
}

extern "C" int st_check_conversions(unsigned long long* mismatches, uint32_t* first_mismatch) {
  unsigned long long* dBad = nullptr;
  uint32_t* dFirst = nullptr;
  const unsigned long long zero = 0;
  const uint32_t none = 0xffffffffu;
  int rc = 1;
  if (hipMalloc((void**)&dBad, sizeof *dBad) != hipSuccess) goto done;
  if (hipMalloc((void**)&dFirst, sizeof *dFirst) != hipSuccess) goto done;
  if (hipMemcpy(dBad, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess) goto done;
  if (hipMemcpy(dFirst, &none, sizeof none, hipMemcpyHostToDevice) != hipSuccess) goto done;
  hipLaunchKernelGGL(k_convert_check, dim3(4096), dim3(256), 0, 0, dBad, dFirst);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) goto done;
  if (hipMemcpy(mismatches, dBad, sizeof *dBad, hipMemcpyDeviceToHost) != hipSuccess) goto done;
  if (hipMemcpy(first_mismatch, dFirst, sizeof *dFirst, hipMemcpyDeviceToHost) != hipSuccess) goto done;
  rc = 0;
done:  // every exit frees what was allocated
  if (dBad) (void)hipFree(dBad);
  if (dFirst) (void)hipFree(dFirst);
  return rc;
}
