// Link-prediction evaluation on the device-resident model.
//
// Scores are dot products of input-embedding rows (Wih) of vertex pairs, the
// evaluation the DistGER paper uses for link prediction on held-out edges
// (held-out split as in src/tools/gtrain.cpp:300-343; the reference has no
// scoring code of its own).  The ROC AUC is the Mann-Whitney statistic:
// for every positive pair, the negatives scoring below it count one and
// those scoring equal count one half, over npos * nneg.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>

#include "sgns_internal.h"

namespace dger {
namespace {

typedef unsigned long long u64;

__device__ __forceinline__ float bf16_bits_to_f32(unsigned h) { return __uint_as_float(h << 16); }

template <typename T>
__device__ __forceinline__ T coherent_load(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One wavefront per pair: lanes read element pairs of both rows (coalesced),
// FMA, then a wavefront sum.
template <typename R>
__global__ __launch_bounds__(256) void k_link_scores(const void* wih, const int32_t* id_rank, uint32_t V, int D,
                                                     const int32_t* pairs, u64 n, float* scores) {
  const u64 p = ((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p >= n) return;
  const int32_t u = pairs[2 * p], v = pairs[2 * p + 1];
  const bool ok = u >= 0 && v >= 0 && (uint32_t)u < V && (uint32_t)v < V;
  float acc = 0.f;
  if (ok) {
    const u64 ru = (u64)id_rank[u], rv = (u64)id_rank[v];
    for (int k = 2 * lane; k < D; k += 128) {
      float a0, a1, b0, b1;
      // Model rows are read with agent-scope loads, like the trainer's.
      if constexpr (sizeof(R) == 4) {
        const u64 a = coherent_load(reinterpret_cast<const u64*>(static_cast<const float*>(wih) + ru * D + k));
        const u64 b = coherent_load(reinterpret_cast<const u64*>(static_cast<const float*>(wih) + rv * D + k));
        a0 = __uint_as_float((unsigned)a); a1 = __uint_as_float((unsigned)(a >> 32));
        b0 = __uint_as_float((unsigned)b); b1 = __uint_as_float((unsigned)(b >> 32));
      } else {
        const unsigned a =
            coherent_load(reinterpret_cast<const unsigned*>(static_cast<const uint16_t*>(wih) + ru * D + k));
        const unsigned b =
            coherent_load(reinterpret_cast<const unsigned*>(static_cast<const uint16_t*>(wih) + rv * D + k));
        a0 = bf16_bits_to_f32(a & 0xFFFFu); a1 = bf16_bits_to_f32(a >> 16);
        b0 = bf16_bits_to_f32(b & 0xFFFFu); b1 = bf16_bits_to_f32(b >> 16);
      }
      acc = __builtin_fmaf(a0, b0, acc);
      acc = __builtin_fmaf(a1, b1, acc);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) scores[p] = ok ? acc : 0.f;
}

// Splits scores by label: positives and negatives, in input order.
__global__ __launch_bounds__(256) void k_split(const float* scores, const int8_t* labels, u64 n, float* pos,
                                               float* neg, unsigned long long* counts) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (labels[i]) pos[atomicAdd(&counts[0], 1ull)] = scores[i];
  else neg[atomicAdd(&counts[1], 1ull)] = scores[i];
}

// For each positive score: negatives below it (lower bound) and equal to it
// (upper minus lower bound) in the sorted negatives; twice the U count is
// accumulated in integers, so the sum does not depend on the order.
__global__ __launch_bounds__(256) void k_rank_sum(const float* pos, u64 npos, const float* neg_sorted, u64 nneg,
                                                  unsigned long long* twice_u) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  u64 t = 0;
  if (i < npos) {
    const float s = pos[i];
    u64 lo = 0, hi = nneg;
    while (lo < hi) {
      const u64 mid = (lo + hi) >> 1;
      if (neg_sorted[mid] < s) lo = mid + 1; else hi = mid;
    }
    const u64 less = lo;
    hi = nneg;
    while (lo < hi) {
      const u64 mid = (lo + hi) >> 1;
      if (neg_sorted[mid] <= s) lo = mid + 1; else hi = mid;
    }
    t = 2 * less + (lo - less);
  }
  if (t) atomicAdd(twice_u, t);
}

}  // namespace

hipError_t launch_link_scores(const void* wih, bool bf16_rows, const int32_t* id_rank, uint32_t V, int D,
                              const int32_t* pairs, uint64_t n, float* scores, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const unsigned blocks = (unsigned)((n + 3) / 4);
  if (bf16_rows)
    hipLaunchKernelGGL(k_link_scores<uint16_t>, dim3(blocks), dim3(256), 0, s, wih, id_rank, V, D, pairs, (u64)n,
                       scores);
  else
    hipLaunchKernelGGL(k_link_scores<float>, dim3(blocks), dim3(256), 0, s, wih, id_rank, V, D, pairs, (u64)n,
                       scores);
  return hipGetLastError();
}

size_t link_auc_scratch_bytes(uint64_t n) {
  size_t sort_bytes = 0;
  (void)hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, static_cast<const float*>(nullptr),
                                          static_cast<float*>(nullptr), (int)n);
  return 3 * n * sizeof(float) + 256 + sort_bytes;
}

hipError_t launch_link_auc(const float* scores, const int8_t* labels, uint64_t n, void* scratch,
                           unsigned long long* counts3, uint64_t* npos_out, uint64_t* nneg_out,
                           unsigned long long* twice_u_out, hipStream_t s) {
  float* pos = static_cast<float*>(scratch);
  float* neg = pos + n;
  float* neg_sorted = neg + n;
  void* sort_tmp = reinterpret_cast<char*>(neg_sorted + n) + 256 - ((uintptr_t)(neg_sorted + n) & 255);
  size_t sort_bytes = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, neg, neg_sorted, (int)n, 0, 32, s);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(counts3, 0, 3 * sizeof(unsigned long long), s);
  if (e != hipSuccess) return e;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_split, dim3(blocks), dim3(256), 0, s, scores, labels, (u64)n, pos, neg, counts3);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  unsigned long long c[2];
  if ((e = hipMemcpyAsync(c, counts3, sizeof(c), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  *npos_out = c[0];
  *nneg_out = c[1];
  if (c[0] == 0 || c[1] == 0) return hipSuccess;
  e = hipcub::DeviceRadixSort::SortKeys(sort_tmp, sort_bytes, neg, neg_sorted, (int)c[1], 0, 32, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_rank_sum, dim3((unsigned)((c[0] + 255) / 256)), dim3(256), 0, s, pos, (u64)c[0],
                     neg_sorted, (u64)c[1], counts3 + 2);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(twice_u_out, counts3 + 2, sizeof(unsigned long long), hipMemcpyDeviceToHost, s)) !=
      hipSuccess)
    return e;
  return hipStreamSynchronize(s);
}

}  // namespace dger
