// FRI folds, DEEP quotients, accumulation, out-of-domain sampling and gathers.
//
//   fold_circle_into_line  prover/fri.rs:46-66; core/fri.rs:751-780
//   fold_line              prover/fri.rs:31-44; core/fri.rs:702-725
//   accumulate_numerators  prover/pcs/quotient_ops.rs:19-43; core/pcs/quotients.rs:200-298
//   compute_quotients_and_combine  prover/backend/cpu/quotients.rs:45-80
//   lift_and_accumulate    prover/backend/cpu/accumulation.rs (lifted index (i >> (r+1) << 1) + (i & 1))
//   eval_at_point          prover/backend/cpu/circle.rs:82-99; core/poly/utils.rs:27-70
// All kernels are HBM streaming kernels: one lane per row, coordinate columns read coalesced.
#include "field.h"
#include "kernels.h"

namespace sh {

__device__ __forceinline__ qm31 load_q(const uint32_t* const c[4], uint32_t i) {
  return {gmem(c[0])[i], gmem(c[1])[i], gmem(c[2])[i], gmem(c[3])[i]};
}
__device__ __forceinline__ void store_q(uint32_t* const c[4], uint32_t i, qm31 v) {
  gmem(c[0])[i] = v.a; gmem(c[1])[i] = v.b; gmem(c[2])[i] = v.c; gmem(c[3])[i] = v.d;
}

// Inverse circle twiddle of pair h: line layer 0 pairs [x, y] -> [y, -y, -x, x].
__device__ __forceinline__ uint32_t circle_itw(const uint32_t* __restrict__ lt0, uint32_t h) {
  uint32_t x = lt0[(h >> 2) * 2], y = lt0[(h >> 2) * 2 + 1];
  switch (h & 3) {
    case 0: return y;
    case 1: return m31_neg(y);
    case 2: return m31_neg(x);
    default: return x;
  }
}

__global__ void __launch_bounds__(256) fold_circle_kernel(uint32_t n_out, CCols4 src, Cols4 dst, qm31 alpha,
                                                          qm31 alpha2, const uint32_t* __restrict__ lt0,
                                                          const qm31* __restrict__ alpha_dev, uint32_t dst_zero) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  if (alpha_dev) {  // folding coefficient drawn on the device
    alpha = *alpha_dev;
    alpha2 = qm_mul(alpha, alpha);
  }
  qm31 fp = load_q(src.c, 2 * i), fn = load_q(src.c, 2 * i + 1);
  uint32_t t = circle_itw(lt0, i);  // 1 / p.y for p = domain.at(bitrev(2i))
  qm31 f0 = qm_add(fp, fn);
  qm31 f1 = qm_mul_m31(qm_sub(fp, fn), t);
  qm31 fr = qm_add(f0, qm_mul(alpha, f1));
  if (!dst_zero) fr = qm_add(qm_mul(load_q((const uint32_t* const*)dst.c, i), alpha2), fr);
  store_q(dst.c, i, fr);
}

// Vector form: lane i produces outputs 4i..4i+3 from source words 8i..8i+7 of each coordinate (two
// 16-byte loads per coordinate column). The four outputs share one pair [x, y] of line layer 0:
// their inverse twiddles are y, -y, -x, x, so the odd output of each pair takes its difference
// the other way round instead of a negated twiddle.
__device__ __forceinline__ u32x4 ld4(const uint32_t* c, uint32_t i) {
  return *reinterpret_cast<const GLOBAL_AS u32x4*>(gmem(c) + i);
}
__device__ __forceinline__ void st4(uint32_t* c, uint32_t i, u32x4 v) { *reinterpret_cast<GLOBAL_AS u32x4*>(gmem(c) + i) = v; }
__device__ __forceinline__ void fold_circle_quad(uint32_t i, const CCols4& src, const Cols4& dst, qm31 alpha,
                                                 qm31 alpha2, const uint32_t* __restrict__ lt0, uint32_t dst_zero) {
  const uint32_t x = lt0[2 * i], y = lt0[2 * i + 1];
  u32x4 lo[4], hi[4], d[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    lo[k] = ld4(src.c[k], 8 * i);
    hi[k] = ld4(src.c[k], 8 * i + 4);
  }
  if (!dst_zero)
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = ld4(dst.c[k], 4 * i);
  u32x4 o[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    // source pair (2q, 2q+1) of the lane's eight words
    qm31 fp, fn;
    if (q < 2) {
      fp = {lo[0][2 * q], lo[1][2 * q], lo[2][2 * q], lo[3][2 * q]};
      fn = {lo[0][2 * q + 1], lo[1][2 * q + 1], lo[2][2 * q + 1], lo[3][2 * q + 1]};
    } else {
      fp = {hi[0][2 * q - 4], hi[1][2 * q - 4], hi[2][2 * q - 4], hi[3][2 * q - 4]};
      fn = {hi[0][2 * q - 3], hi[1][2 * q - 3], hi[2][2 * q - 3], hi[3][2 * q - 3]};
    }
    const uint32_t t = q < 2 ? y : x;
    const bool swap = q == 1 || q == 2;  // twiddles -y and -x
    qm31 f0 = qm_add(fp, fn);
    qm31 f1 = qm_mul_m31(swap ? qm_sub(fn, fp) : qm_sub(fp, fn), t);
    qm31 fr = qm_add(f0, qm_mul(alpha, f1));
    if (!dst_zero) fr = qm_add(qm_mul(qm31{d[0][q], d[1][q], d[2][q], d[3][q]}, alpha2), fr);
    o[0][q] = fr.a;
    o[1][q] = fr.b;
    o[2][q] = fr.c;
    o[3][q] = fr.d;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) st4(dst.c[k], 4 * i, o[k]);
}
__global__ void __launch_bounds__(256) fold_circle_vec_kernel(uint32_t n_quads, CCols4 src, Cols4 dst, qm31 alpha,
                                                              qm31 alpha2, const uint32_t* __restrict__ lt0,
                                                              uint32_t dst_zero) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_quads) return;
  fold_circle_quad(i, src, dst, alpha, alpha2, lt0, dst_zero);
}
// Grouped form: secure column g of the launch is blockIdx.y, its coordinate columns 4g..4g+3 of the
// address tables.
__global__ void __launch_bounds__(256) fold_circle_grouped_kernel(uint32_t n_quads, const uint32_t* const* srcs,
                                                                  uint32_t* const* dsts, qm31 alpha, qm31 alpha2,
                                                                  const uint32_t* __restrict__ lt0) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_quads) return;
  const uint32_t g = blockIdx.y;
  const CCols4 s = {{srcs[4 * g], srcs[4 * g + 1], srcs[4 * g + 2], srcs[4 * g + 3]}};
  const Cols4 d = {{dsts[4 * g], dsts[4 * g + 1], dsts[4 * g + 2], dsts[4 * g + 3]}};
  fold_circle_quad(i, s, d, alpha, alpha2, lt0, 0u);
}

static bool aligned16(const uint32_t* const* c, int n) {
  for (int k = 0; k < n; ++k)
    if ((reinterpret_cast<uintptr_t>(c[k]) & 15) != 0) return false;
  return true;
}

void launch_fold_circle_into_line(uint32_t log_src, CCols4 src, Cols4 dst, qm31 alpha, const uint32_t* itree,
                                  uint32_t tree_len, hipStream_t stream) {
  uint32_t n = 1u << (log_src - 1);
  const uint32_t* lt0 = itree + tree_len - (1u << (log_src - 1));
  if (log_src >= 3 && aligned16(src.c, 4) && aligned16((const uint32_t* const*)dst.c, 4)) {
    const uint32_t nq = n / 4;
    hipLaunchKernelGGL(fold_circle_vec_kernel, dim3((nq + 255) / 256), dim3(256), 0, stream, nq, src, dst, alpha,
                       qm_mul(alpha, alpha), lt0, 0u);
    return;
  }
  hipLaunchKernelGGL(fold_circle_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, src, dst, alpha,
                     qm_mul(alpha, alpha), lt0, (const qm31*)nullptr, 0u);
}
void launch_fold_circle_into_line_grouped(uint32_t log_src, uint32_t n_groups, const uint32_t* const* srcs,
                                          uint32_t* const* dsts, qm31 alpha, const uint32_t* itree, uint32_t tree_len,
                                          hipStream_t stream) {
  const uint32_t nq = (1u << (log_src - 1)) / 4;
  const uint32_t* lt0 = itree + tree_len - (1u << (log_src - 1));
  hipLaunchKernelGGL(fold_circle_grouped_kernel, dim3((nq + 255) / 256, n_groups), dim3(256), 0, stream, nq, srcs, dsts,
                     alpha, qm_mul(alpha, alpha), lt0);
}
void launch_fold_circle_into_line_zero(uint32_t log_src, CCols4 src, Cols4 dst, qm31 alpha, const uint32_t* itree,
                                       uint32_t tree_len, hipStream_t stream) {
  uint32_t n = 1u << (log_src - 1);
  const uint32_t* lt0 = itree + tree_len - (1u << (log_src - 1));
  if (log_src >= 3 && aligned16(src.c, 4) && aligned16((const uint32_t* const*)dst.c, 4)) {
    const uint32_t nq = n / 4;
    hipLaunchKernelGGL(fold_circle_vec_kernel, dim3((nq + 255) / 256), dim3(256), 0, stream, nq, src, dst, alpha,
                       qm_mul(alpha, alpha), lt0, 1u);
    return;
  }
  hipLaunchKernelGGL(fold_circle_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, src, dst, alpha,
                     qm_mul(alpha, alpha), lt0, (const qm31*)nullptr, 1u);
}
void launch_fold_circle_into_line_dev(uint32_t log_src, CCols4 src, Cols4 dst, const qm31* alpha_dev, bool dst_zero,
                                      const uint32_t* itree, uint32_t tree_len, hipStream_t stream) {
  uint32_t n = 1u << (log_src - 1);
  const uint32_t* lt0 = itree + tree_len - (1u << (log_src - 1));
  hipLaunchKernelGGL(fold_circle_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, src, dst, qm_zero(),
                     qm_zero(), lt0, alpha_dev, (uint32_t)dst_zero);
}

__global__ void __launch_bounds__(256) fold_line_kernel(uint32_t n_out, CCols4 src, Cols4 dst, qm31 alpha,
                                                        const uint32_t* __restrict__ lt,
                                                        const qm31* __restrict__ alpha_dev, uint32_t n_squarings) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  if (alpha_dev) {  // folding coefficient drawn on the device, squared once per earlier fold step
    alpha = *alpha_dev;
    for (uint32_t k = 0; k < n_squarings; ++k) alpha = qm_mul(alpha, alpha);
  }
  qm31 fx = load_q(src.c, 2 * i), fnx = load_q(src.c, 2 * i + 1);
  uint32_t xi = lt[i];  // 1 / x for x = domain.at(bitrev(2i))
  qm31 f0 = qm_add(fx, fnx);
  qm31 f1 = qm_mul_m31(qm_sub(fx, fnx), xi);
  store_q(dst.c, i, qm_add(f0, qm_mul(alpha, f1)));
}

// Vector form: lane i produces outputs 4i..4i+3 from source words 8i..8i+7 of each coordinate and
// the four inverse twiddles lt[4i..4i+3] (one 16-byte load each).
__global__ void __launch_bounds__(256) fold_line_vec_kernel(uint32_t n_quads, CCols4 src, Cols4 dst, qm31 alpha,
                                                            const uint32_t* __restrict__ lt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_quads) return;
  const u32x4 t = ld4(lt, 4 * i);
  u32x4 lo[4], hi[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    lo[k] = ld4(src.c[k], 8 * i);
    hi[k] = ld4(src.c[k], 8 * i + 4);
  }
  u32x4 o[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    qm31 fx, fnx;
    if (q < 2) {
      fx = {lo[0][2 * q], lo[1][2 * q], lo[2][2 * q], lo[3][2 * q]};
      fnx = {lo[0][2 * q + 1], lo[1][2 * q + 1], lo[2][2 * q + 1], lo[3][2 * q + 1]};
    } else {
      fx = {hi[0][2 * q - 4], hi[1][2 * q - 4], hi[2][2 * q - 4], hi[3][2 * q - 4]};
      fnx = {hi[0][2 * q - 3], hi[1][2 * q - 3], hi[2][2 * q - 3], hi[3][2 * q - 3]};
    }
    qm31 f0 = qm_add(fx, fnx);
    qm31 f1 = qm_mul_m31(qm_sub(fx, fnx), t[q]);
    qm31 fr = qm_add(f0, qm_mul(alpha, f1));
    o[0][q] = fr.a;
    o[1][q] = fr.b;
    o[2][q] = fr.c;
    o[3][q] = fr.d;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) st4(dst.c[k], 4 * i, o[k]);
}

void launch_fold_line(uint32_t log_src, CCols4 src, Cols4 dst, qm31 alpha, const uint32_t* itree, uint32_t tree_len,
                      hipStream_t stream) {
  uint32_t n = 1u << (log_src - 1);
  const uint32_t* lt = itree + tree_len - (1u << log_src);
  if (log_src >= 3 && aligned16(src.c, 4) && aligned16((const uint32_t* const*)dst.c, 4) &&
      (reinterpret_cast<uintptr_t>(lt) & 15) == 0) {
    const uint32_t nq = n / 4;
    hipLaunchKernelGGL(fold_line_vec_kernel, dim3((nq + 255) / 256), dim3(256), 0, stream, nq, src, dst, alpha, lt);
    return;
  }
  hipLaunchKernelGGL(fold_line_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, src, dst, alpha, lt,
                     (const qm31*)nullptr, 0u);
}
void launch_fold_line_dev(uint32_t log_src, CCols4 src, Cols4 dst, const qm31* alpha_dev, uint32_t n_squarings,
                          const uint32_t* itree, uint32_t tree_len, hipStream_t stream) {
  uint32_t n = 1u << (log_src - 1);
  const uint32_t* lt = itree + tree_len - (1u << log_src);
  hipLaunchKernelGGL(fold_line_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, src, dst, qm_zero(), lt,
                     alpha_dev, n_squarings);
}

// partial_b[row] = sum_{j in batch b} c_j * col_j[row] - sum_j b_j.
//
// Wide form (2^k >= 256 rows, 16-byte aligned columns): a block of 8 waves covers 256 rows, lane l
// owns rows 4l..4l+3 (one 16-byte load per column), and wave w takes every 8th entry of a batch,
// 8 entries per step with their loads issued together; the 8 partial sums meet in LDS.
constexpr uint32_t NUM_WAVES = 8;
constexpr uint32_t NUM_UNROLL = 8;
__device__ __forceinline__ void mac4(qacc* acc, qm31 c, u32x4 v) {
  qacc_mac(acc[0], c, v.x);
  qacc_mac(acc[1], c, v.y);
  qacc_mac(acc[2], c, v.z);
  qacc_mac(acc[3], c, v.w);
}
// The same without folding: canonical coefficients and values make each product below 2^62, so
// three of them may follow a fold (fold4), which leaves every coordinate below 2^34.
__device__ __forceinline__ void mad4(qacc* acc, qm31 c, u32x4 v) {
  qacc_mad(acc[0], c, v.x);
  qacc_mad(acc[1], c, v.y);
  qacc_mad(acc[2], c, v.z);
  qacc_mad(acc[3], c, v.w);
}
__device__ __forceinline__ void fold4(qacc* acc) {
#pragma unroll
  for (int r = 0; r < 4; ++r) qacc_fold(acc[r]);
}
__global__ void __launch_bounds__(512) accumulate_numerators_wide_kernel(const NumEntry* __restrict__ entries,
                                                                         const uint32_t* __restrict__ batch_start,
                                                                         const qm31* __restrict__ batch_b,
                                                                         uint32_t n_batches,
                                                                         uint32_t* const* __restrict__ out) {
  __shared__ uint32_t part[NUM_WAVES][4][256];
  // the wave index is uniform across the wave: readfirstlane lets the entries (column address and
  // coefficient) load through the scalar cache into scalar registers
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t row0 = blockIdx.x * 256 + lane * 4;
  for (uint32_t b = 0; b < n_batches; ++b) {
    qacc acc[4] = {qacc_zero(), qacc_zero(), qacc_zero(), qacc_zero()};
    const uint32_t e_end = batch_start[b + 1];
    uint32_t e = batch_start[b] + wave;
    for (; e + (NUM_UNROLL - 1) * NUM_WAVES < e_end; e += NUM_UNROLL * NUM_WAVES) {
      const uint32_t* col[NUM_UNROLL];
      qm31 c[NUM_UNROLL];
      u32x4 v[NUM_UNROLL];
#pragma unroll
      for (uint32_t k = 0; k < NUM_UNROLL; ++k) {
        col[k] = entries[e + k * NUM_WAVES].col;
        c[k] = entries[e + k * NUM_WAVES].c;
      }
#pragma unroll
      for (uint32_t k = 0; k < NUM_UNROLL; ++k) v[k] = *reinterpret_cast<const GLOBAL_AS u32x4*>(gmem(col[k]) + row0);
#pragma unroll
      for (uint32_t k = 0; k < NUM_UNROLL; ++k) {
        mad4(acc, c[k], v[k]);
        if (k % 3 == 2 || k == NUM_UNROLL - 1) fold4(acc);
      }
    }
    for (; e < e_end; e += NUM_WAVES) {
      const NumEntry en = entries[e];
      mac4(acc, en.c, *reinterpret_cast<const GLOBAL_AS u32x4*>(gmem(en.col) + row0));
    }
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r) {
      qm31 q = qacc_reduce(acc[r]);
      part[wave][0][lane * 4 + r] = q.a;
      part[wave][1][lane * 4 + r] = q.b;
      part[wave][2][lane * 4 + r] = q.c;
      part[wave][3][lane * 4 + r] = q.d;
    }
    __syncthreads();
    {
      const uint32_t r = threadIdx.x & 255, c0 = (threadIdx.x >> 8) * 2;
      const qm31 bb = batch_b[b];
      const uint32_t bw[4] = {bb.a, bb.b, bb.c, bb.d};
#pragma unroll
      for (uint32_t cc = c0; cc < c0 + 2; ++cc) {
        uint32_t s = 0;
#pragma unroll
        for (uint32_t w = 0; w < NUM_WAVES; ++w) s = m31_add(s, part[w][cc][r]);
        gmem(out[4 * b + cc])[blockIdx.x * 256 + r] = m31_sub(s, bw[cc]);
      }
    }
    __syncthreads();
  }
}

// Any size: one row per lane, 4 waves splitting each batch's entries.
constexpr uint32_t NUM_SLICES = 4;
__global__ void __launch_bounds__(256) accumulate_numerators_kernel(uint32_t n_rows,
                                                                    const NumEntry* __restrict__ entries,
                                                                    const uint32_t* __restrict__ batch_start,
                                                                    const qm31* __restrict__ batch_b,
                                                                    uint32_t n_batches, uint32_t* const* __restrict__ out) {
  __shared__ qm31 part[NUM_SLICES][64];
  const uint32_t lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
  const uint32_t row = blockIdx.x * 64 + lane;
  const bool active = row < n_rows;
  for (uint32_t b = 0; b < n_batches; ++b) {
    qacc acc = qacc_zero();
    uint32_t e_end = batch_start[b + 1];
    if (active)
      for (uint32_t e = batch_start[b] + slice; e < e_end; e += NUM_SLICES) {
        NumEntry en = entries[e];
        qacc_mac(acc, en.c, en.col[row]);
      }
    part[slice][lane] = qacc_reduce(acc);
    __syncthreads();
    if (slice == 0 && active) {
      qm31 r = qm_add(qm_add(part[0][lane], part[1][lane]), qm_add(part[2][lane], part[3][lane]));
      r = qm_sub(r, batch_b[b]);
      out[4 * b][row] = r.a;
      out[4 * b + 1][row] = r.b;
      out[4 * b + 2][row] = r.c;
      out[4 * b + 3][row] = r.d;
    }
    __syncthreads();
  }
}

void launch_accumulate_numerators(uint32_t log_size, const NumEntry* entries, const uint32_t* batch_start,
                                  const qm31* batch_b, uint32_t n_batches, uint32_t* const* out, bool aligned16,
                                  hipStream_t stream) {
  uint32_t n = 1u << log_size;
  if (aligned16 && n >= 256)
    hipLaunchKernelGGL(accumulate_numerators_wide_kernel, dim3(n / 256), dim3(512), 0, stream, entries, batch_start,
                       batch_b, n_batches, out);
  else
    hipLaunchKernelGGL(accumulate_numerators_kernel, dim3((n + 63) / 64), dim3(256), 0, stream, n, entries,
                       batch_start, batch_b, n_batches, out);
}

// q[row] = sum_a (sum_parts lift(partial)[row] - first_linear_a * y) / den_a(row),
// den_a = (Re p.x - x) Im p.y - (Re p.y - y) Im p.x  (a CM31 value; core/pcs/quotients.rs:200-226)
__global__ void __launch_bounds__(256) quotients_kernel(uint32_t lifting_log, const uint32_t* __restrict__ domain_xy,
                                                        const QuotPoint* __restrict__ points, uint32_t n_points,
                                                        const QuotPart* __restrict__ parts, Cols4 out) {
  uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= (1u << lifting_log)) return;
  uint32_t x = domain_xy[2 * row], y = domain_xy[2 * row + 1];
  qm31 q = qm_zero();
  for (uint32_t a = 0; a < n_points; ++a) {
    QuotPoint pt = points[a];
    cm31 dx = {m31_sub(pt.prx.a, x), pt.prx.b};
    cm31 dy = {m31_sub(pt.pry.a, y), pt.pry.b};
    cm31 den = cm_sub(cm_mul(dx, pt.piy), cm_mul(dy, pt.pix));
    cm31 dinv = cm_inv(den);
    qm31 num = qm_zero(), lin = pt.first_linear;
    for (uint32_t p = pt.part_begin; p < pt.part_end; ++p) {
      const QuotPart& pa = parts[p];
      uint32_t r = lifting_log - pa.log_size;
      uint32_t li = ((row >> (r + 1)) << 1) + (row & 1);
      num = qm_add(num, load_q(pa.c, li));
      lin = qm_add(lin, pa.a);
    }
    num = qm_sub(num, qm_mul_m31(lin, y));
    q = qm_add(q, qm_mul_cm31(num, dinv));
  }
  store_q(out.c, row, q);
}

// Two launches: one lane per entry computes r = qrc^q, the entry's coefficient and its two terms
// r v and r (conj(v) - v) (into scratch); then one workgroup per batch sums its entries' terms.
__device__ __forceinline__ qm31 qm_pow_small(qm31 b, uint32_t e) {
  qm31 r = qm_one();
  while (e) {
    if (e & 1) r = qm_mul(r, b);
    b = qm_sqr(b);
    e >>= 1;
  }
  return r;
}
__global__ void __launch_bounds__(256) quotient_coeff_entries_kernel(qm31 qrc, const qm31* __restrict__ vals,
                                                                     const QCoefEntry* __restrict__ entries,
                                                                     uint32_t n_entries,
                                                                     const QCoefBatch* __restrict__ batches,
                                                                     NumEntry* __restrict__ out_entries,
                                                                     qm31* __restrict__ terms) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_entries) return;
  const QCoefEntry en = entries[e];
  const qm31 r = qm_pow_small(qrc, en.q), v = vals[en.vi];
  NumEntry ne{};
  ne.col = en.col;
  ne.c = qm_mul(r, batches[en.batch].cy);
  out_entries[e] = ne;
  terms[2 * e] = qm_mul(r, v);
  terms[2 * e + 1] = qm_mul(r, qm_sub(qm_conj(v), v));
}
__global__ void __launch_bounds__(256) quotient_coeff_batches_kernel(const QCoefBatch* __restrict__ batches,
                                                                     const qm31* __restrict__ terms,
                                                                     qm31* __restrict__ batch_b,
                                                                     QuotPart* __restrict__ parts) {
  __shared__ qm31 red[2][256];
  const QCoefBatch bt = batches[blockIdx.x];
  const uint32_t t = threadIdx.x;
  qm31 srv = qm_zero(), sra = qm_zero();
  for (uint32_t k = t; k < bt.count; k += 256) {
    srv = qm_add(srv, terms[2 * (bt.first + k)]);
    sra = qm_add(sra, terms[2 * (bt.first + k) + 1]);
  }
  red[0][t] = srv;
  red[1][t] = sra;
  __syncthreads();
  for (uint32_t h = 128; h > 0; h >>= 1) {
    if (t < h) {
      red[0][t] = qm_add(red[0][t], red[0][t + h]);
      red[1][t] = qm_add(red[1][t], red[1][t + h]);
    }
    __syncthreads();
  }
  if (t == 0) {
    batch_b[blockIdx.x] = qm_sub(qm_mul(bt.cy, red[0][0]), qm_mul(bt.py, red[1][0]));
    parts[bt.part].a = red[1][0];
  }
}

void launch_quotient_coeffs(qm31 qrc, const qm31* vals, const QCoefEntry* entries, uint32_t n_entries,
                            const QCoefBatch* batches, uint32_t n_batches, NumEntry* out_entries, qm31* batch_b,
                            QuotPart* parts, qm31* scratch, hipStream_t stream) {
  if (!n_batches) return;
  if (n_entries)
    hipLaunchKernelGGL(quotient_coeff_entries_kernel, dim3((n_entries + 255) / 256), dim3(256), 0, stream, qrc, vals,
                       entries, n_entries, batches, out_entries, scratch);
  hipLaunchKernelGGL(quotient_coeff_batches_kernel, dim3(n_batches), dim3(256), 0, stream, batches,
                     (const qm31*)scratch, batch_b, parts);
}

void launch_quotients(uint32_t lifting_log, const uint32_t* domain_xy, const QuotPoint* points, uint32_t n_points,
                      const QuotPart* parts, Cols4 out, hipStream_t stream) {
  uint32_t n = 1u << lifting_log;
  hipLaunchKernelGGL(quotients_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, lifting_log, domain_xy, points,
                     n_points, parts, out);
}

// dst[i] += src[(i >> (r+1) << 1) + (i & 1)], r = log_dst - log_src (r = 0 is a plain add)
__global__ void __launch_bounds__(256) accumulate_lifted_kernel(uint32_t log_dst, Cols4 dst, uint32_t log_src,
                                                                CCols4 src) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1u << log_dst)) return;
  uint32_t r = log_dst - log_src;
  uint32_t li = r ? ((i >> (r + 1)) << 1) + (i & 1) : i;
  qm31 v = qm_add(load_q((const uint32_t* const*)dst.c, i), load_q(src.c, li));
  store_q(dst.c, i, v);
}

void launch_accumulate_lifted(uint32_t log_dst, Cols4 dst, uint32_t log_src, CCols4 src, hipStream_t stream) {
  uint32_t n = 1u << log_dst;
  hipLaunchKernelGGL(accumulate_lifted_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, log_dst, dst, log_src,
                     src);
}

// ---- barycentric evaluation (PolyOps::barycentric_weights / barycentric_eval_at_point) --------
// w_i = V_n(p) / (S_i V_i(p)) over the bit-reversed CanonicCoset(k) circle domain
// (prover/backend/cpu/circle.rs:101-141): S_i = -2 y_i 4^(k-1) prod_{j=1}^{k-1} V_{CanonicCoset(j)}(x_i)
// (coset_vanishing_derivative), V_i(p) = h.y / (1 + h.x) with h = p - x_i (point_vanishing).
// shift[j-1] = half step of CanonicCoset(j) minus its initial point, so that
// V_{CanonicCoset(j)}(q) = x-coordinate of q + shift[j-1] doubled j - 1 times.
struct BaryShifts { cpm31 s[31]; };
__global__ void __launch_bounds__(256) barycentric_weights_kernel(uint32_t k, cpqm31 p, qm31 vn, BaryShifts sh,
                                                                  uint32_t four_pow, Cols4 out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1u << k)) return;
  // domain point bitrev(i) of CanonicCoset(k).circle_domain(): half_odds(k - 1) and its negation
  const uint32_t idx = bitrev(i, k), half = 1u << (k - 1);
  const uint32_t init = 1u << (31 - (k + 1)), step = 1u << (31 - (k - 1));
  uint32_t pi = idx < half ? (init + idx * step) : (0x80000000u - (init + (idx - half) * step));
  const cpm31 x = index_to_point(pi & 0x7fffffffu);
  uint32_t van = 1;
  for (uint32_t j = 1; j < k; ++j) {
    uint32_t v = cp_add(x, sh.s[j - 1]).x;
    for (uint32_t r = 1; r < j; ++r) v = m31_double_x(v);
    van = m31_mul(van, v);
  }
  const uint32_t si = m31_mul(m31_mul(m31_neg(m31_add(x.y, x.y)), four_pow), van);
  const cpqm31 h = cpq_add(p, cpq_from(cp_conj(x)));
  const qm31 w = qm_mul(qm_mul(vn, qm_add(qm_one(), h.x)), qm_inv(qm_mul_m31(h.y, si)));
  store_q(out.c, i, w);
}

void launch_barycentric_weights(uint32_t log_size, cpqm31 p, qm31 vn, const cpm31* shifts, Cols4 out,
                                hipStream_t stream) {
  BaryShifts sh{};
  for (uint32_t j = 1; j < log_size; ++j) sh.s[j - 1] = shifts[j - 1];
  uint32_t four_pow = 1;
  for (uint32_t j = 1; j < log_size; ++j) four_pow = m31_mul(four_pow, 4);
  const uint32_t n = 1u << log_size;
  hipLaunchKernelGGL(barycentric_weights_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, log_size, p, vn, sh,
                     four_pow, out);
}

// partial[c * gridDim.x + b] = sum over the 2048 rows of block b of evals_c[i] * w_i.
__global__ void __launch_bounds__(256) barycentric_dot_kernel(uint32_t n, const uint32_t* const* __restrict__ evals,
                                                              CCols4 w, qm31* __restrict__ partial) {
  __shared__ qm31 red[256];
  const uint32_t c = blockIdx.y;
  const uint32_t* __restrict__ e = evals[c];
  qacc acc = qacc_zero();
  for (uint32_t i = blockIdx.x * 2048 + threadIdx.x; i < n && i < (blockIdx.x + 1) * 2048; i += 256)
    qacc_mac(acc, load_q(w.c, i), e[i]);
  red[threadIdx.x] = qacc_reduce(acc);
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = qm_add(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[c * gridDim.x + blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(64) sum_partials_kernel(const qm31* __restrict__ partial, uint32_t n_parts,
                                                          qm31* __restrict__ out) {
  if (threadIdx.x != 0) return;
  qm31 s = qm_zero();
  for (uint32_t b = 0; b < n_parts; ++b) s = qm_add(s, partial[blockIdx.x * n_parts + b]);
  out[blockIdx.x] = s;
}

void launch_barycentric_eval(uint32_t log_size, const uint32_t* const* evals, uint32_t n_cols, CCols4 weights,
                             qm31* partial, qm31* out, hipStream_t stream) {
  const uint32_t n = 1u << log_size, blocks = (n + 2047) / 2048;
  hipLaunchKernelGGL(barycentric_dot_kernel, dim3(blocks, n_cols), dim3(256), 0, stream, n, evals, weights, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(n_cols), dim3(64), 0, stream, partial, blocks, out);
}

// ---- out-of-domain sampling on coefficients ------------------------------------------------
// f(P) = sum_n c_n prod_b f_b^(n_b), f_0 = y, f_1 = x, f_b = pi^(b-1)(x). Tasks that share a
// point and a size share a weight table W[i] = prod_{b < T} f_b^(i_b) over the low
// T = min(k, 16) index bits (1 MiB per group, read through L2 by every column of the group). A
// block takes 8192 coefficients of one column: each lane accumulates W[i] c_i for its 32
// coefficients with a fold after every 2 products, the block sums its lanes,
// and for columns longer than 2^16 the block sum is multiplied by the factors of its index bits
// >= 16.
// A block of 256 threads writes 256 consecutive entries: the weights of their low 8 index bits are
// built in LDS by doubling (step b multiplies the first 2^b entries by f_b), the product of the
// factors of the block's bits 8.. by a 3-level tree in the same steps, and each entry is one
// product of the two (QM31 products are exact: the order of the factors does not matter).
__global__ void __launch_bounds__(256) oods_weights_kernel(const qm31* __restrict__ factors,
                                                           const uint32_t* __restrict__ group_log,
                                                           qm31* __restrict__ W) {
  __shared__ qm31 lo[256];
  __shared__ qm31 hi[8];
  const uint32_t g = blockIdx.y;
  const uint32_t k = group_log[g];
  const uint32_t T = k < OODS_TABLE_LOG ? k : OODS_TABLE_LOG;
  const uint32_t base = blockIdx.x * 256;  // blockDim.x is 256
  if (base >= (1u << T)) return;  // uniform per block
  const qm31* f = factors + (size_t)g * 32;
  const uint32_t t = threadIdx.x;
  const uint32_t tl = T < 8 ? T : 8;  // index bits below the block's
  if (t == 0) lo[0] = qm_one();
  if (t >= 128 && t < 136) {
    const uint32_t b = 8 + (t - 128);
    hi[t - 128] = (b < T && ((blockIdx.x >> (b - 8)) & 1)) ? f[b] : qm_one();
  }
  __syncthreads();
  for (uint32_t b = 0; b < 8; ++b) {
    qm31 v;
    const bool dbl = b < tl && t < (1u << b);
    const bool tree = b < 3 && t >= 128 && t < 128 + (4u >> b);
    if (dbl) v = qm_mul(lo[t], f[b]);
    if (tree) v = qm_mul(hi[2 * (t - 128)], hi[2 * (t - 128) + 1]);
    __syncthreads();  // every read of the step is done
    if (dbl) lo[t + (1u << b)] = v;
    if (tree) hi[t - 128] = v;
    __syncthreads();
  }
  const uint32_t i = base + t;
  if (i < (1u << T)) W[((size_t)g << OODS_TABLE_LOG) + i] = T > 8 ? qm_mul(lo[t], hi[0]) : lo[t];
}

// A block takes 2^OODS_BLOCK_LOG coefficients of up to OODS_BATCH columns of one group (tasks
// batch_start[b] .. batch_start[b+1]), so that each weight it reads from L2 serves every column of
// the batch; small blocks keep many of them resident to hide the loads.
__global__ void __launch_bounds__(256) oods_chunk_kernel(const OodsTask* __restrict__ tasks,
                                                         const uint32_t* __restrict__ task_group,
                                                         const uint32_t* __restrict__ batch_start,
                                                         const qm31* __restrict__ factors,
                                                         const qm31* __restrict__ Wall, qm31* __restrict__ partials,
                                                         uint32_t max_chunks) {
  __shared__ uint32_t red[4 * OODS_BATCH][256 + 8];  // padded rows: the 8 words a wave sums at once sit in distinct banks
  __shared__ uint32_t fin[4 * OODS_BATCH];
  const uint32_t t0 = batch_start[blockIdx.y], cnt = batch_start[blockIdx.y + 1] - t0;
  const uint32_t grp = task_group[t0];
  const uint32_t k = tasks[t0].log_size;
  const uint32_t n = 1u << k;
  const uint32_t blk = blockIdx.x;
  if (((size_t)blk << OODS_BLOCK_LOG) >= n) return;  // uniform per block
  const uint32_t T = k < OODS_TABLE_LOG ? k : OODS_TABLE_LOG;
  const uint32_t tmask = (1u << T) - 1;
  const qm31* W = Wall + ((size_t)grp << OODS_TABLE_LOG);
  const uint32_t t = threadIdx.x;
  const uint32_t base = blk << OODS_BLOCK_LOG;
  const GLOBAL_AS uint32_t* col[OODS_BATCH];
#pragma unroll
  for (uint32_t c = 0; c < OODS_BATCH; ++c) col[c] = gmem(tasks[t0 + (c < cnt ? c : 0)].coeffs);
  qacc acc[OODS_BATCH];
#pragma unroll
  for (uint32_t c = 0; c < OODS_BATCH; ++c) acc[c] = qacc_zero();
  if (n >= 4) {
    for (uint32_t r = 0; r < (1u << OODS_BLOCK_LOG) / 1024; ++r) {
      const uint32_t i = base + 4 * (t + 256 * r);
      if (i < n) {
        const uint32_t wi = i & tmask;
        const qm31 w0 = W[wi], w1 = W[wi + 1], w2 = W[wi + 2], w3 = W[wi + 3];
        // every column slot loads and accumulates (slots past cnt repeat column 0 and are never
        // written): loads behind per-column branches would each wait on their own
        u32x4 v[OODS_BATCH];
#pragma unroll
        for (uint32_t c = 0; c < OODS_BATCH; ++c) v[c] = *reinterpret_cast<const GLOBAL_AS u32x4*>(col[c] + i);
#pragma unroll
        for (uint32_t c = 0; c < OODS_BATCH; ++c) {
          qacc_mad(acc[c], w0, v[c].x);
          qacc_mad(acc[c], w1, v[c].y);
          qacc_fold(acc[c]);
          qacc_mad(acc[c], w2, v[c].z);
          qacc_mad(acc[c], w3, v[c].w);
          qacc_fold(acc[c]);
        }
      }
    }
  } else if (t < n) {
#pragma unroll
    for (uint32_t c = 0; c < OODS_BATCH; ++c)
      if (c < cnt) qacc_mad(acc[c], W[t], col[c][t]);
  }
  // Block sum of each of the 4 OODS_BATCH accumulator words: every lane folds its words below
  // 2^31 + 8 into LDS (word-major), then 8 lanes per word add 32 values each in 64 bits and meet
  // by shuffles (a shuffle reduction of every word over the whole block cost more than the
  // accumulation itself).
#pragma unroll
  for (uint32_t c = 0; c < OODS_BATCH; ++c) {
    red[4 * c + 0][t] = (uint32_t)fold31(fold31(acc[c].a));
    red[4 * c + 1][t] = (uint32_t)fold31(fold31(acc[c].b));
    red[4 * c + 2][t] = (uint32_t)fold31(fold31(acc[c].c));
    red[4 * c + 3][t] = (uint32_t)fold31(fold31(acc[c].d));
  }
  __syncthreads();
  {
    const uint32_t w = t >> 3, j = t & 7;  // accumulator word w = 4 c + coordinate
    uint64_t sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < 32; ++i) sum += red[w][(i << 3) | j];
#pragma unroll
    for (uint32_t off = 4; off > 0; off >>= 1) {
      const uint32_t lo = (uint32_t)__shfl_down((int)(uint32_t)sum, off, 8);
      const uint32_t hi = (uint32_t)__shfl_down((int)(uint32_t)(sum >> 32), off, 8);
      sum += ((uint64_t)hi << 32) | lo;
    }
    if (j == 0) fin[w] = m31_reduce64(sum);
  }
  __syncthreads();
  if (t < cnt) {
    const qm31* f = factors + (size_t)grp * 32;
    qm31 s = {fin[4 * t], fin[4 * t + 1], fin[4 * t + 2], fin[4 * t + 3]};
    for (uint32_t b = OODS_TABLE_LOG; b < k; ++b)
      if ((base >> b) & 1) s = qm_mul(s, f[b]);
    partials[(size_t)(t0 + t) * max_chunks + blk] = s;
  }
}

__global__ void oods_reduce_kernel(const OodsTask* __restrict__ tasks, uint32_t n_tasks,
                                   const qm31* __restrict__ partials, uint32_t max_chunks, qm31* __restrict__ out) {
  uint32_t tsk = blockIdx.x * blockDim.x + threadIdx.x;
  if (tsk >= n_tasks) return;
  uint32_t k = tasks[tsk].log_size;
  uint32_t n_chunks = k > OODS_BLOCK_LOG ? 1u << (k - OODS_BLOCK_LOG) : 1u;
  qm31 s = qm_zero();
  for (uint32_t i = 0; i < n_chunks; ++i) s = qm_add(s, partials[(size_t)tsk * max_chunks + i]);
  out[tasks[tsk].out_index] = s;
}

void launch_oods(const OodsTask* tasks, const uint32_t* task_group, uint32_t n_tasks, const uint32_t* batch_start,
                 uint32_t n_batches, uint32_t max_log, const qm31* group_factors, const uint32_t* group_log,
                 uint32_t n_groups, qm31* weights, qm31* partials, qm31* out, hipStream_t stream) {
  if (n_tasks == 0) return;
  const uint32_t max_chunks = oods_max_chunks(max_log);
  const uint32_t tab = max_log < OODS_TABLE_LOG ? max_log : OODS_TABLE_LOG;
  hipLaunchKernelGGL(oods_weights_kernel, dim3(((1u << tab) + 255) / 256, n_groups), dim3(256), 0, stream,
                     group_factors, group_log, weights);
  hipLaunchKernelGGL(oods_chunk_kernel, dim3(max_chunks, n_batches), dim3(256), 0, stream, tasks, task_group,
                     batch_start, group_factors, weights, partials, max_chunks);
  hipLaunchKernelGGL(oods_reduce_kernel, dim3((n_tasks + 255) / 256), dim3(256), 0, stream, tasks, n_tasks, partials,
                     max_chunks, out);
}

__global__ void gather_words_kernel(const uint32_t* const* __restrict__ bases, const uint32_t* __restrict__ base_idx,
                                    const uint32_t* __restrict__ offsets, uint32_t n, uint32_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = bases[base_idx[i]][offsets[i]];
}

void launch_gather_words(const uint32_t* const* bases, const uint32_t* base_idx, const uint32_t* offsets, uint32_t n,
                         uint32_t* out, hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(gather_words_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, bases, base_idx, offsets, n,
                     out);
}

// HOST: out is coherent host memory, read by the host once the stream has synchronised; the
// system-scope fence orders the store before the kernel's completion
template <bool HOST>
__global__ void gather_args_kernel(GatherArgs g, uint32_t n_cols, uint32_t n, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = g.cols[i % n_cols][g.rows[i / n_cols]];
  if (HOST) __threadfence_system();
}
__global__ void gather_rows_kernel(const uint32_t* const* __restrict__ cols, uint32_t n_cols,
                                   const uint32_t* __restrict__ rows, uint32_t n, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = cols[i % n_cols][rows[i / n_cols]];
}
void launch_gather_args(const GatherArgs& g, uint32_t n_cols, uint32_t n_rows, uint32_t* out, bool out_is_host,
                        hipStream_t stream) {
  if (n_cols > GATHER_ARG_COLS || n_rows > GATHER_ARG_ROWS)
    throw std::invalid_argument("launch_gather_args: more columns or rows than the argument block holds");
  const uint32_t n = n_cols * n_rows;
  if (!n) return;
  if (out_is_host)
    hipLaunchKernelGGL(gather_args_kernel<true>, dim3((n + 255) / 256), dim3(256), 0, stream, g, n_cols, n, out);
  else
    hipLaunchKernelGGL(gather_args_kernel<false>, dim3((n + 255) / 256), dim3(256), 0, stream, g, n_cols, n, out);
}
void launch_gather_rows(const uint32_t* const* cols, uint32_t n_cols, const uint32_t* rows, uint32_t n_rows,
                        uint32_t* out, hipStream_t stream) {
  const uint64_t n = (uint64_t)n_cols * n_rows;
  if (!n) return;
  if (n > 0xFFFFFFFFull - 255) throw std::invalid_argument("launch_gather_rows: gather too large");
  hipLaunchKernelGGL(gather_rows_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, cols, n_cols, rows,
                     (uint32_t)n, out);
}

struct DecQueries { uint32_t first, n, p[DEC_MAX_QUERIES]; };

__device__ __forceinline__ uint32_t dec_pos(uint32_t mode, uint32_t a, uint32_t p) {
  return mode == 0 ? ((p >> 1) << a) + (p & 1) : mode == 1 ? ((p >> a) << 1) + (p & 1) : p >> a;
}

// One workgroup per (item, query). Each thread reads up to 4 words per step with the loads issued
// together (indices clamped into range, out-of-level nodes masked afterwards).
__global__ void __launch_bounds__(256) decommit_gather_kernel(const DecItem* __restrict__ items, DecQueries q,
                                                              uint32_t* __restrict__ out) {
  const DecItem it = items[blockIdx.x];
  const uint32_t i = blockIdx.y;
  const uint32_t pos = dec_pos(it.pos_mode, it.pos_arg, q.p[i]);
  const uint32_t node0 = (pos >> it.shift) << it.rows_log;
  const uint32_t level_n = 1u << it.level_log;
  const uint32_t total = it.width << it.rows_log;
  uint32_t* o = out + it.out + (q.first + i) * it.stride;
  for (uint32_t e0 = 0; e0 < total; e0 += 4 * 256) {
    uint32_t v[4], dst[4];
    bool ok[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t e = e0 + u * 256 + threadIdx.x;
      const uint32_t ec = e < total ? e : total - 1;
      const uint32_t r = ec / it.width, c = ec - r * it.width;
      const uint32_t node = node0 + r;
      ok[u] = e < total;
      dst[u] = r * it.row_width + c;
      const uint32_t nd = node < level_n ? node : 0;
      const uint32_t* src = it.kind == DEC_COLS ? it.cols[c] + nd : it.layer + (size_t)nd * 8 + c;
      v[u] = *src;
      if (node >= level_n) v[u] = 0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (ok[u]) o[dst[u]] = v[u];
  }
}

void launch_decommit_gather(const DecItem* items, uint32_t n_items, const uint32_t* positions, uint32_t n_queries,
                            uint32_t* out, hipStream_t stream) {
  if (!n_items) return;
  for (uint32_t f = 0; f < n_queries; f += DEC_MAX_QUERIES) {
    DecQueries q{};
    q.first = f;
    q.n = n_queries - f < DEC_MAX_QUERIES ? n_queries - f : DEC_MAX_QUERIES;
    for (uint32_t i = 0; i < q.n; ++i) q.p[i] = positions[f + i];
    hipLaunchKernelGGL(decommit_gather_kernel, dim3(n_items, q.n), dim3(256), 0, stream, items, q, out);
  }
}

__global__ void fill_zero_kernel(uint32_t* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
void launch_fill_zero(uint32_t* p, size_t n, hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n);
}
__global__ void copy_words_kernel(uint32_t* dst, const uint32_t* src, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
void launch_copy_words(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(copy_words_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, dst, src, n);
}

}  // namespace sh
