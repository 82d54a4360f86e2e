// Measurement helpers: ceilings of the VALU on this device, against which the VALU-bound kernels
// are judged.
//
// int32_issue_kernel: every lane runs 8 independent chains of v_add_u32 / v_xor_b32 /
// v_alignbit_b32 (the operations of a Blake2s round), written one instruction per asm statement so
// that the count is exact: 24 VALU instructions per loop iteration and lane.
//
// butterfly_kernel: every lane holds 32 M31 values in registers and runs rounds of 5 butterfly
// stages on them (80 butterflies per round), with the FFT's own butterfly code (butterfly.h) and
// twiddles held in registers: the rate of the transform's arithmetic with no memory traffic, the
// ceiling of the FFT kernels' butterflies per second.
//
// compression_kernel: every lane holds a Blake2s chaining state and a 64-byte message block in
// registers and compresses with the Merkle kernels' own one-lane code (blake2s.h), each result
// folded into the next block: Blake2s compressions per second with no memory traffic, the ceiling
// of the Merkle layers' hashing.
//
// All launch 16 workgroups of 256 threads per CU, enough waves to fill every SIMD.
#include "blake2s.h"
#include "butterfly.h"
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

template <bool INV>
__global__ void __launch_bounds__(256) butterfly_kernel(uint32_t iters, uint32_t seed, uint32_t* out) {
  uint32_t v[32], tw[16];
#pragma unroll
  for (int k = 0; k < 32; ++k) v[k] = (seed * (threadIdx.x + 1) + 7919u * k) & P;
#pragma unroll
  for (int k = 0; k < 16; ++k) tw[k] = 2u * ((seed * 2654435761u + 40503u * k + threadIdx.x) & 0x3fffffffu);
  for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
    for (int kk = 0; kk < 5; ++kk) {
      const int k = INV ? kk : 4 - kk;  // the FFT's stage order within a round
#pragma unroll
      for (int u = 0; u < (32 >> (k + 1)); ++u) {
        const uint32_t t = tw[(u + 3 * k) & 15];
#pragma unroll
        for (int l = 0; l < (1 << k); ++l) {
          const int q = (u << (k + 1)) | l;
          bfly<INV>(v[q], v[q | (1 << k)], t);
        }
      }
    }
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < 32; ++k) s ^= v[k];
  if (s == 0x12345678u) out[blockIdx.x] = s;  // keeps the values live
}

double measure_butterfly_rate(bool inverse, uint32_t* scratch, hipStream_t stream) {
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const uint32_t blocks = (uint32_t)cus * 16, threads = 256, iters = 256;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  auto launch = [&](uint32_t n, uint32_t seed) {
    if (inverse) hipLaunchKernelGGL(butterfly_kernel<true>, dim3(blocks), dim3(threads), 0, stream, n, seed, scratch);
    else hipLaunchKernelGGL(butterfly_kernel<false>, dim3(blocks), dim3(threads), 0, stream, n, seed, scratch);
  };
  launch(8u, 3u);  // warm-up
  (void)hipEventRecord(e0, stream);
  launch(iters, 5u);
  (void)hipEventRecord(e1, stream);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double butterflies = (double)blocks * threads * iters * 80.0;
  return ms > 0 ? butterflies / (ms * 1e-3) : 0.0;
}

__global__ void __launch_bounds__(256) compression_kernel(uint32_t iters, uint32_t seed, uint32_t* out) {
  uint32_t h[8], m[16];
#pragma unroll
  for (int k = 0; k < 8; ++k) h[k] = seed * (threadIdx.x + 1) + 0x9E3779B9u * k;
#pragma unroll
  for (int k = 0; k < 16; ++k) m[k] = (seed ^ blockIdx.x) + 0x85EBCA6Bu * k + threadIdx.x;
  for (uint32_t it = 0; it < iters; ++it) {
    blake2s_compress(h, m, 64, true);
    m[it & 15] ^= h[it & 7];  // the next block depends on this one's result
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) s ^= h[k];
  if (s == 0x12345678u) out[blockIdx.x] = s;  // keeps the state live
}

double measure_compression_rate(uint32_t* scratch, hipStream_t stream) {
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const uint32_t blocks = (uint32_t)cus * 16, threads = 256, iters = 128;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(compression_kernel, dim3(blocks), dim3(threads), 0, stream, 4u, 3u, scratch);  // warm-up
  (void)hipEventRecord(e0, stream);
  hipLaunchKernelGGL(compression_kernel, dim3(blocks), dim3(threads), 0, stream, iters, 5u, scratch);
  (void)hipEventRecord(e1, stream);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double compressions = (double)blocks * threads * iters;
  return ms > 0 ? compressions / (ms * 1e-3) : 0.0;
}

}  // namespace sh
