// Measurement helper: the VALU's 32-bit integer issue rate on this device, the peak against which
// the VALU-bound kernels (Blake2s Merkle hashing, the M31 butterflies of the FFT) are judged.
//
// Every lane runs 8 independent chains of v_add_u32 / v_xor_b32 / v_alignbit_b32 (the operations
// of a Blake2s round), written one instruction per asm statement so that the count is exact:
// 24 VALU instructions per loop iteration and lane. Enough waves fill every SIMD of every CU.
#include "field.h"
#include "kernels.h"

namespace sh {

__global__ void __launch_bounds__(256) int32_issue_kernel(uint32_t iters, uint32_t seed, uint32_t* out) {
  uint32_t a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = seed * (threadIdx.x + 1) + k;
  // loop-invariant operands stay in vector registers, so the loop body holds only the 24 timed
  // instructions (the loop counter lives in scalar registers)
  const uint32_t b = seed ^ threadIdx.x, c = seed + threadIdx.x, r = 7 + (threadIdx.x & 1);
  for (uint32_t it = 0; it < iters; ++it) {
    // the chains interleaved, so that no instruction waits on the one just before it
#pragma unroll
    for (int k = 0; k < 8; ++k) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
#pragma unroll
    for (int k = 0; k < 8; ++k) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a[k]) : "v"(c));
#pragma unroll
    for (int k = 0; k < 8; ++k) asm volatile("v_alignbit_b32 %0, %0, %0, %1" : "+v"(a[k]) : "v"(r));
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) s ^= a[k];
  if (s == 0x12345678u) out[blockIdx.x] = s;  // keeps the chains live
}

double measure_int32_issue(uint32_t* scratch, hipStream_t stream) {
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const uint32_t blocks = (uint32_t)cus * 16, threads = 256, iters = 4096;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(int32_issue_kernel, dim3(blocks), dim3(threads), 0, stream, 64u, 3u, scratch);  // warm-up
  (void)hipEventRecord(e0, stream);
  hipLaunchKernelGGL(int32_issue_kernel, dim3(blocks), dim3(threads), 0, stream, iters, 5u, scratch);
  (void)hipEventRecord(e1, stream);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double lane_instr = (double)blocks * threads * iters * 24.0;
  return ms > 0 ? lane_instr / (ms * 1e-3) : 0.0;
}

}  // namespace sh
