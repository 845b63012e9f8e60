// gfx950 kernels of the SGNS trainer (DistGER DSGL, include/dsgl.hpp).
//
// Trainer layout: one 64-lane wavefront per Hogwild worker.  Worker k plays
// compute thread k+1 of the reference (dsgl.hpp:786-1408): same corpus slice,
// same LCG seed, same lazily interleaved two-sentence schedule, same buffer
// write-back semantics.  Inside a worker the sequential pieces of the
// reference are spread over the 64 lanes:
//   * sentence packing: 64 tokens per coalesced load; each lane draws its own
//     subsampling number through LCG jump-ahead (lane l holds the l+1-step
//     affine map), kept tokens are compacted with ballot/popcount into LDS;
//   * the shared negative batch: K*max(L1,L2) draws, 64 per round, same
//     jump-ahead;
//   * the step: contexts and output rows staged in LDS, the three
//     contractions (dsgl.hpp:1241,1290,1309) with lanes over contexts
//     (GEMM1) or over embedding pairs (GEMM2, GEMM3), each sum in sequential
//     order with one FMA per term;
//   * scatter: duplicate contexts of a step are reduced in LDS in context
//     order (the order the reference adds them into its sentence buffer),
//     then one store per distinct buffer row; target rows take agent-scope
//     atomic adds, so concurrent workers do not lose each other's updates.
// Rows of Wih/Woh are float or bf16 (Row<> below); all arithmetic and the
// per-worker working buffers are FP32.

#include <hip/hip_runtime.h>

#include "dger_sgns.h"
#include "sgns_internal.h"

namespace dger {
namespace {

typedef unsigned long long u64;

__device__ __forceinline__ u64 readlane64(u64 v, int l) {
  unsigned lo = __builtin_amdgcn_readlane((unsigned)v, l);
  unsigned hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), l);
  return ((u64)hi << 32) | lo;
}

__device__ __forceinline__ int readlane32(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

__device__ __forceinline__ u64 lanemask_lt(int lane) { return (1ull << lane) - 1ull; }

__device__ __forceinline__ float bf16_to_f32(unsigned h) { return __uint_as_float(h << 16); }

// Round-to-nearest-even float -> bf16 (finite inputs).
__device__ __forceinline__ unsigned f32_to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}

// Row storage traits.  A lane works on pairs of adjacent elements.
template <typename R>
struct Row;

template <>
struct Row<float> {
  __device__ static float2 load(const void* base, size_t row, int D, int pr) {
    const u64* p = reinterpret_cast<const u64*>(static_cast<const float*>(base) + row * (size_t)D) + pr;
    u64 v = __hip_atomic_load(const_cast<u64*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_float2(__uint_as_float((unsigned)v), __uint_as_float((unsigned)(v >> 32)));
  }
  __device__ static void store(void* base, size_t row, int D, int pr, float2 v) {
    float2* p = reinterpret_cast<float2*>(static_cast<float*>(base) + row * (size_t)D) + pr;
    *p = v;
  }
  __device__ static void atomic_add(void* base, size_t row, int D, int pr, float2 v) {
    float* p = static_cast<float*>(base) + row * (size_t)D + 2 * pr;
    unsafeAtomicAdd(p, v.x);
    unsafeAtomicAdd(p + 1, v.y);
  }
};

template <>
struct Row<uint16_t> {
  __device__ static float2 load(const void* base, size_t row, int D, int pr) {
    const unsigned* p =
        reinterpret_cast<const unsigned*>(static_cast<const uint16_t*>(base) + row * (size_t)D) + pr;
    unsigned v = __hip_atomic_load(const_cast<unsigned*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_float2(bf16_to_f32(v & 0xFFFFu), bf16_to_f32(v >> 16));
  }
  __device__ static void store(void* base, size_t row, int D, int pr, float2 v) {
    unsigned* p = reinterpret_cast<unsigned*>(static_cast<uint16_t*>(base) + row * (size_t)D) + pr;
    *p = f32_to_bf16(v.x) | (f32_to_bf16(v.y) << 16);
  }
  __device__ static void atomic_add(void* base, size_t row, int D, int pr, float2 v) {
    unsigned* p = reinterpret_cast<unsigned*>(static_cast<uint16_t*>(base) + row * (size_t)D) + pr;
    unsigned old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
      float x = bf16_to_f32(old & 0xFFFFu) + v.x;
      float y = bf16_to_f32(old >> 16) + v.y;
      unsigned nv = f32_to_bf16(x) | (f32_to_bf16(y) << 16);
      unsigned seen = old;
      if (__hip_atomic_compare_exchange_strong(p, &seen, nv, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT))
        break;
      old = seen;
    }
  }
};

// Table-quantised sigmoid gradient, dsgl.hpp:1249-1270.
__device__ __forceinline__ float sig_grad(float f, int label, float alpha, const float* exp_table) {
  if (f > 6.0f) return (float)(label - 1) * alpha;
  if (f < -6.0f) return (float)label * alpha;
  int idx = (int)((f + 6.0f) * 83.0f);
  return ((float)label - exp_table[idx]) * alpha;
}

// LDS carve-up of one worker (all offsets 16-byte aligned).
struct Lds {
  int* tok[2];    // Lmax ranks per sentence
  int* slot[2];   // Lmax: first position holding the same rank
  int* negw;      // K*Lmax negative draws
  int* negs;      // K*Lmax: first draw holding the same rank
  int* cpos;      // 4W context positions
  int* csent;     // 4W context sentence (0/1)
  int* crep;      // 4W first context of the step with the same buffer row
  int* oidx;      // K+2 output ranks
  int* ometa;     // K+2 multiplicities
  int* oslot;     // K+2 negative-buffer rows
  int* misc;      // 8 scalars
  float* O;       // (K+2) x OST staged output rows
  float* I;       // 4W x IST staged context rows
  float* C;       // (K+2) x 4W gradient coefficients
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }
__host__ __device__ inline int o_stride(int D) { return (D + 3) & ~3; }
__host__ __device__ inline int i_stride(int D) { return D + 1; }

__host__ __device__ inline size_t lds_layout(int Lmax, int K, int W, int D, size_t* off) {
  const int KO = K + 2, MC = 4 * W;
  size_t o = 0;
  off[0] = o; o = align16(o + 4 * (size_t)Lmax * 4);      // tok0, tok1, slot0, slot1
  off[1] = o; o = align16(o + 2 * (size_t)K * Lmax * 4);  // negw, negs
  off[2] = o; o = align16(o + 3 * (size_t)MC * 4);        // cpos, csent, crep
  off[3] = o; o = align16(o + 3 * (size_t)KO * 4 + 8 * 4);// oidx, ometa, oslot, misc
  off[4] = o; o = align16(o + (size_t)KO * o_stride(D) * 4);
  off[5] = o; o = align16(o + (size_t)MC * i_stride(D) * 4);
  off[6] = o; o = align16(o + (size_t)KO * MC * 4);
  return o;
}

__device__ inline Lds carve(unsigned char* smem, int Lmax, int K, int W, int D) {
  size_t off[7];
  lds_layout(Lmax, K, W, D, off);
  const int KO = K + 2, MC = 4 * W;
  Lds s;
  int* b0 = reinterpret_cast<int*>(smem + off[0]);
  s.tok[0] = b0; s.tok[1] = b0 + Lmax; s.slot[0] = b0 + 2 * Lmax; s.slot[1] = b0 + 3 * Lmax;
  int* b1 = reinterpret_cast<int*>(smem + off[1]);
  s.negw = b1; s.negs = b1 + K * Lmax;
  int* b2 = reinterpret_cast<int*>(smem + off[2]);
  s.cpos = b2; s.csent = b2 + MC; s.crep = b2 + 2 * MC;
  int* b3 = reinterpret_cast<int*>(smem + off[3]);
  s.oidx = b3; s.ometa = b3 + KO; s.oslot = b3 + 2 * KO; s.misc = b3 + 3 * KO;
  s.O = reinterpret_cast<float*>(smem + off[4]);
  s.I = reinterpret_cast<float*>(smem + off[5]);
  s.C = reinterpret_cast<float*>(smem + off[6]);
  return s;
}

// first[p] = smallest q <= p with v[q] == v[p], for p < n (values in LDS).
__device__ inline void first_occurrence(const int* v, int* first, int n, int lane) {
  for (int c0 = 0; c0 < n; c0 += 64) {
    int p = c0 + lane;
    if (p < n) {
      int x = v[p];
      int f = p;
      for (int q = 0; q < p; q++) {
        if (v[q] == x) { f = q; break; }
      }
      first[p] = f;
    }
  }
}

// GEMM1 + sigmoid (dsgl.hpp:1241-1272): lanes over contexts j; for each
// output row i, f = sum_k O[i][k]*I[j][k] in sequential k.  C[i][j] = g * meta.
__device__ inline void step_gemm1(const Lds& s, int D, int MC, int os, int nin, int fin, float alpha,
                                  const float* exp_table, int lane) {
  const int OST = o_stride(D), IST = i_stride(D);
  for (int j = lane; j < nin; j += 64) {
    const float* x = s.I + (size_t)j * IST;
    for (int i0 = 0; i0 < os; i0 += 8) {
      const int ni = min(8, os - i0);
      float acc[8];
#pragma unroll
      for (int ii = 0; ii < 8; ii++) acc[ii] = 0.f;
      for (int k = 0; k < D; k++) {
        const float xv = x[k];
#pragma unroll
        for (int ii = 0; ii < 8; ii++)
          if (ii < ni) acc[ii] = __builtin_fmaf(s.O[(i0 + ii) * OST + k], xv, acc[ii]);
      }
#pragma unroll
      for (int ii = 0; ii < 8; ii++) {
        if (ii < ni) {
          const int i = i0 + ii;
          const int label = (j < fin) ? (i == 0 ? 1 : 0) : (i == os - 1 ? 1 : 0);
          const float g = sig_grad(acc[ii], label, alpha, exp_table);
          s.C[i * MC + j] = g * (float)s.ometa[i];
        }
      }
    }
  }
}

// GEMM2 for output row i (dsgl.hpp:1290): md = sum_j C[i][j] * I[j][:] in
// sequential j, for the element pair pr of this lane.
__device__ inline float2 step_gemm2_row(const Lds& s, int D, int MC, int i, int nin, int pr) {
  const int IST = i_stride(D);
  float ax = 0.f, ay = 0.f;
  for (int j = 0; j < nin; j++) {
    const float c = s.C[i * MC + j];
    ax = __builtin_fmaf(c, s.I[(size_t)j * IST + 2 * pr], ax);
    ay = __builtin_fmaf(c, s.I[(size_t)j * IST + 2 * pr + 1], ay);
  }
  return make_float2(ax, ay);
}

// GEMM3 for context j (dsgl.hpp:1309): gin = sum_i C[i][j] * O[i][:] in
// sequential i, for the element pair pr of this lane.
__device__ inline float2 step_gemm3_col(const Lds& s, int D, int MC, int j, int os, int pr) {
  const int OST = o_stride(D);
  float ax = 0.f, ay = 0.f;
  for (int i = 0; i < os; i++) {
    const float c = s.C[i * MC + j];
    ax = __builtin_fmaf(c, s.O[i * OST + 2 * pr], ax);
    ay = __builtin_fmaf(c, s.O[i * OST + 2 * pr + 1], ay);
  }
  return make_float2(ax, ay);
}

__device__ __forceinline__ float2 ld2(const float* p) { return *reinterpret_cast<const float2*>(p); }
__device__ __forceinline__ void st2(float* p, float2 v) { *reinterpret_cast<float2*>(p) = v; }

template <typename R, int NP>
__global__ __launch_bounds__(64) void k_train(TrainArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x;
  const int wk = blockIdx.x;
  const int D = a.D, W = a.W, K = a.K, Lmax = a.Lmax;
  const int DP = D >> 1;
  const int MC = 4 * W, IST = i_stride(D), OST = o_stride(D);
  const Lds s = carve(smem, Lmax, K, W, D);

  // Corpus slice and seed of compute thread wk+1 (dsgl.hpp:791, :811-821).
  const u64 N = a.n_tokens;
  const u64 nparts = (u64)a.workers;
  const u64 per = N / nparts;
  const u64 start = per * (u64)wk;
  u64 cnt = per;
  if (N % nparts > 0 && wk == a.workers - 1) cnt = N - (nparts - 1) * per;
  const u64 end = start + cnt;
  u64 rng = (u64)wk;

  // Lane l holds the (l+1)-step LCG map s -> jA*s + jC.
  u64 jA = 1, jC = 0;
  for (int i = 0; i <= lane; i++) { jA = jA * kLcgMul; jC = jC * kLcgMul + kLcgAdd; }

  float* const buf[2] = {a.fscratch + (size_t)wk * worker_scratch_floats(Lmax, K, D),
                         a.fscratch + (size_t)wk * worker_scratch_floats(Lmax, K, D) + (size_t)Lmax * D};
  float* const nbuf = buf[1] + (size_t)Lmax * D;

  const float a0 = a.alpha0;
  float alpha = a0;
  int iters_left = a.iter;
  u64 wc = 0, last_wc = 0, cur = start;
  int len[2] = {0, 0}, pos[2] = {0, 0};
  int done_sent = 0, refresh = 1, neg_count = 0, nbn = 1;
  u64 steps = 0, raw = 0, sentences = 0;
  const int stride = record_stride(W, K);
  int32_t* const rec = a.rec ? a.rec + (size_t)wk * (size_t)a.rec_cap * stride : nullptr;

  for (;;) {
    // Learning-rate decay in >10000-token chunks (dsgl.hpp:886-897).
    if (wc - last_wc > 10000) {
      const u64 diff = wc - last_wc;
      u64 old = 0;
      if (lane == 0) old = atomicAdd(&a.counters[kWordCountActual], diff);
      const u64 total = readlane64(old, 0) + diff;
      last_wc = wc;
      alpha = a0 * (1 - (float)total / a.alpha_den);
      if (alpha < a0 * 0.0001f) alpha = a0 * 0.0001f;
    }

    // Sentence loads (dsgl.hpp:902-965, :968-1030).
    for (int q = 0; q < 2; q++) {
      while (len[q] == 0) {
        if (cur >= end) break;
        int L = 0;
        for (;;) {
          const u64 idx = cur + (u64)lane;
          const int id = (idx < end) ? a.corpus[idx] : -1;
          const bool term = (id == -1) || (idx + 1 >= end);
          const u64 tm = __ballot(term);
          const int tl = tm ? (__ffsll((long long)tm) - 1) : 64;
          const bool regular = lane < tl;
          int r = 0;
          bool keep = regular;
          if (regular) {
            r = a.id_rank[id];
            if (a.sample_on) {
              const u64 sr = jA * rng + jC;
              keep = !(a.keep_thr[r] < (float)(unsigned)(sr & 0xFFFFu) / 65536.f);
            }
          }
          u64 km = __ballot(keep);
          int nk = __popcll(km);
          const int room = kMaxSentence - L;
          int consumed, draws;
          bool stop;
          if (nk >= room) {
            const u64 cm = __ballot(keep && (__popcll(km & lanemask_lt(lane)) + 1 == room));
            const int cl = __ffsll((long long)cm) - 1;
            keep = keep && (lane <= cl);
            consumed = cl + 1; draws = cl + 1; stop = true; nk = room;
          } else if (tl < 64) {
            consumed = tl + 1; draws = tl; stop = true;
          } else {
            consumed = 64; draws = 64; stop = false;
          }
          if (keep) s.tok[q][L + __popcll(km & lanemask_lt(lane))] = r;
          L += nk;
          if (a.sample_on && draws > 0) rng = readlane64(jA, draws - 1) * rng + readlane64(jC, draws - 1);
          cur += (u64)consumed;
          wc += (u64)consumed;
          if (stop) break;
        }
        pos[q] = 0;
        if (L <= 1) { len[q] = 0; continue; }
        len[q] = L;
        __syncthreads();
        first_occurrence(s.tok[q], s.slot[q], L, lane);
        __syncthreads();
        // Working copy of the sentence's Wih rows, one per distinct rank.
        for (int p = 0; p < L; p++) {
          if (s.slot[q][p] != p) continue;
          const size_t row = (size_t)s.tok[q][p];
#pragma unroll
          for (int v = 0; v < NP; v++) {
            const int pr = lane + 64 * v;
            if (pr < DP) st2(buf[q] + (size_t)p * D + 2 * pr, Row<R>::load(a.wih, row, D, pr));
          }
        }
      }
    }

    // End of the slice: in-flight sentences are dropped (dsgl.hpp:1031-1050).
    if (cur >= end) {
      if (lane == 0) atomicAdd(&a.counters[kWordCountActual], wc - last_wc);
      raw += wc;
      if (--iters_left == 0) break;
      wc = 0; cur = start; last_wc = 0;
      len[0] = 0; len[1] = 0;
      continue;
    }

    // Shared negative batch (dsgl.hpp:1054-1088).
    if (refresh) {
      nbn = max(len[0], len[1]);
      const int NN = nbn * K;
      u64 s0 = rng;
      for (int base = 0; base < NN; base += 64) {
        const int c = min(64, NN - base);
        if (lane < c) {
          const u64 sr = jA * s0 + jC;
          int v = a.table[(sr >> 16) % a.table_size];
          if (!v) v = (int)(sr % (u64)(a.V - 1) + 1);
          s.negw[base + lane] = v;
        }
        s0 = readlane64(jA, c - 1) * s0 + readlane64(jC, c - 1);
      }
      rng = s0;
      __syncthreads();
      first_occurrence(s.negw, s.negs, NN, lane);
      __syncthreads();
      for (int n = 0; n < NN; n++) {
        if (s.negs[n] != n) continue;
        const size_t row = (size_t)s.negw[n];
#pragma unroll
        for (int v = 0; v < NP; v++) {
          const int pr = lane + 64 * v;
          if (pr < DP) st2(nbuf + (size_t)n * D + 2 * pr, Row<R>::load(a.woh, row, D, pr));
        }
      }
      refresh = 0;
    }

    // ---- one step (dsgl.hpp:1089-1368) ----
    const int t1 = s.tok[0][pos[0]];
    const int t2 = s.tok[1][pos[1]];
    rng = rng * kLcgMul + kLcgAdd;
    const int b = (int)(rng % (u64)W);
    const int hw = W - b;
    const int lo0 = max(0, pos[0] - hw), hi0 = min(len[0] - 1, pos[0] + hw);
    const int lo1 = max(0, pos[1] - hw), hi1 = min(len[1] - 1, pos[1] + hw);
    const int n0 = (pos[0] - lo0) + (hi0 - pos[0]);
    const int n1 = (pos[1] - lo1) + (hi1 - pos[1]);
    const int nin = n0 + n1, fin = n0;
    for (int j = lane; j < nin; j += 64) {
      const int q = (j < n0) ? 0 : 1;
      const int jj = q ? j - n0 : j;
      const int lo = q ? lo1 : lo0, pp = pos[q];
      s.cpos[j] = (jj < pp - lo) ? lo + jj : pp + 1 + (jj - (pp - lo));
      s.csent[j] = q;
    }
    neg_count++;
    const int grp = neg_count % nbn;
    if (lane == 0) {
      int os = 1;
      s.oidx[0] = t1;
      s.ometa[0] = 1;
      for (int k = 0; k < K; k++) {
        const int v = s.negw[grp * K + k];
        int hit = -1;
        for (int i = 0; i < os; i++)
          if (s.oidx[i] == v) { hit = i; break; }
        if (hit < 0) {
          s.oidx[os] = v; s.ometa[os] = 1; s.oslot[os] = s.negs[grp * K + k]; os++;
        } else {
          s.ometa[hit]++;
        }
      }
      s.ometa[0] = 1;  // a negative equal to the target is dropped
      s.oidx[os] = t2;
      s.ometa[os] = 1;
      os++;
      s.misc[0] = os;
    }
    __syncthreads();
    const int os = s.misc[0];
    // Repeat contexts of a step share one buffer row: crep[j] = first such j.
    for (int j = lane; j < nin; j += 64) {
      const int q = s.csent[j];
      const int sl = s.slot[q][s.cpos[j]];
      int r = j;
      for (int j2 = (q ? n0 : 0); j2 < j; j2++)
        if (s.slot[q][s.cpos[j2]] == sl) { r = j2; break; }
      s.crep[j] = r;
    }
    // Stage output rows (targets from Woh, negatives from the negative buffer)
    // and context rows (from the sentence buffers).
    for (int i = 0; i < os; i++) {
#pragma unroll
      for (int v = 0; v < NP; v++) {
        const int pr = lane + 64 * v;
        if (pr < DP) {
          float2 x;
          if (i == 0) x = Row<R>::load(a.woh, (size_t)t1, D, pr);
          else if (i == os - 1) x = Row<R>::load(a.woh, (size_t)t2, D, pr);
          else x = ld2(nbuf + (size_t)s.oslot[i] * D + 2 * pr);
          st2(s.O + i * OST + 2 * pr, x);
        }
      }
    }
    __syncthreads();
    for (int j = 0; j < nin; j++) {
      const int q = s.csent[j];
      const float* src = buf[q] + (size_t)s.slot[q][s.cpos[j]] * D;
#pragma unroll
      for (int v = 0; v < NP; v++) {
        const int pr = lane + 64 * v;
        if (pr < DP) {
          const float2 x = ld2(src + 2 * pr);
          s.I[(size_t)j * IST + 2 * pr] = x.x;
          s.I[(size_t)j * IST + 2 * pr + 1] = x.y;
        }
      }
    }
    __syncthreads();

    if (rec && (int64_t)steps < a.rec_cap && lane == 0) {
      int32_t* r = rec + (size_t)steps * stride;
      for (int z = 0; z < stride; z++) r[z] = -1;
      r[0] = t1; r[1] = t2; r[2] = b; r[3] = fin; r[4] = nin; r[5] = os;
      r[6] = __float_as_int(alpha);
      r[7] = grp;
      for (int j = 0; j < nin; j++) r[8 + j] = s.tok[s.csent[j]][s.cpos[j]];
      for (int i = 0; i < os; i++) { r[8 + MC + i] = s.oidx[i]; r[8 + MC + K + 2 + i] = s.ometa[i]; }
    }

    step_gemm1(s, D, MC, os, nin, fin, alpha, a.exp_table, lane);
    __syncthreads();

    // GEMM2 and the output-row scatter: target rows straight into Woh
    // (row 0 first, then row os-1, as dsgl.hpp:1352-1356), negatives into the
    // negative buffer (dsgl.hpp:1358-1366).
    for (int i = 0; i < os; i++) {
#pragma unroll
      for (int v = 0; v < NP; v++) {
        const int pr = lane + 64 * v;
        if (pr < DP) {
          const float2 md = step_gemm2_row(s, D, MC, i, nin, pr);
          if (i == 0) Row<R>::atomic_add(a.woh, (size_t)t1, D, pr, md);
          else if (i == os - 1) Row<R>::atomic_add(a.woh, (size_t)t2, D, pr, md);
          else {
            float* dst = nbuf + (size_t)s.oslot[i] * D + 2 * pr;
            float2 x = ld2(dst);
            x.x += md.x; x.y += md.y;
            st2(dst, x);
          }
        }
      }
    }
    // GEMM3 and the context scatter: updates reduced into the first context
    // row of each buffer row in context order (dsgl.hpp:1314-1333).
    for (int j = 0; j < nin; j++) {
      const int r = s.crep[j];
#pragma unroll
      for (int v = 0; v < NP; v++) {
        const int pr = lane + 64 * v;
        if (pr < DP) {
          const float2 g = step_gemm3_col(s, D, MC, j, os, pr);
          s.I[(size_t)r * IST + 2 * pr] += g.x;
          s.I[(size_t)r * IST + 2 * pr + 1] += g.y;
        }
      }
    }
    for (int j = 0; j < nin; j++) {
      if (s.crep[j] != j) continue;
      const int q = s.csent[j];
      float* dst = buf[q] + (size_t)s.slot[q][s.cpos[j]] * D;
#pragma unroll
      for (int v = 0; v < NP; v++) {
        const int pr = lane + 64 * v;
        if (pr < DP)
          st2(dst + 2 * pr, make_float2(s.I[(size_t)j * IST + 2 * pr], s.I[(size_t)j * IST + 2 * pr + 1]));
      }
    }
    steps++;

    // Write-backs (dsgl.hpp:1377-1407).
    pos[0]++;
    pos[1]++;
    for (int q = 0; q < 2; q++) {
      if (pos[q] >= len[q]) {
        for (int p = 0; p < len[q]; p++) {
          if (s.slot[q][p] != p) continue;
          const size_t row = (size_t)s.tok[q][p];
#pragma unroll
          for (int v = 0; v < NP; v++) {
            const int pr = lane + 64 * v;
            if (pr < DP) Row<R>::store(a.wih, row, D, pr, ld2(buf[q] + (size_t)p * D + 2 * pr));
          }
        }
        len[q] = 0;
        done_sent++;
        sentences++;
      }
    }
    if (done_sent >= 2) {
      const int NN = nbn * K;
      for (int n = 0; n < NN; n++) {
        if (s.negs[n] != n) continue;
        const size_t row = (size_t)s.negw[n];
#pragma unroll
        for (int v = 0; v < NP; v++) {
          const int pr = lane + 64 * v;
          if (pr < DP) Row<R>::store(a.woh, row, D, pr, ld2(nbuf + (size_t)n * D + 2 * pr));
        }
      }
      refresh = 1;
      done_sent = 0;
    }
    __syncthreads();
  }

  if (lane == 0) {
    a.worker_steps[wk] = steps;
    a.worker_alpha[wk] = alpha;
    atomicAdd(&a.counters[kSteps], steps);
    atomicAdd(&a.counters[kRawWords], raw);
    atomicAdd(&a.counters[kSentences], sentences);
  }
}

// InitNet (dsgl.hpp:433-457): element e of Wih takes LCG state e+1 from seed
// 1.  Lane l of a block covers elements e0+l, e0+l+64, ... using the 64-step
// map, so stores are coalesced.
constexpr int kInitChunk = 32;

__global__ __launch_bounds__(256) void k_init_net(float* wih, u64 n, int D) {
  const u64 wave = ((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const u64 e0 = wave * 64 * kInitChunk + (u64)lane;
  if (e0 >= n) return;
  // Map for e0+1 steps from seed 1, by repeated squaring.
  u64 m = e0 + 1, a = kLcgMul, c = kLcgAdd, ra = 1, rc = 0;
  while (m) {
    if (m & 1) { ra = a * ra; rc = a * rc + c; }
    c = a * c + c;
    a = a * a;
    m >>= 1;
  }
  u64 st = ra * 1ULL + rc;
  // 64-step map
  u64 a64 = 1, c64 = 0;
  for (int i = 0; i < 64; i++) { a64 = a64 * kLcgMul; c64 = c64 * kLcgMul + kLcgAdd; }
  const float fd = (float)D;
  for (int i = 0; i < kInitChunk; i++) {
    const u64 e = e0 + (u64)i * 64;
    if (e >= n) break;
    wih[e] = (((float)(unsigned)(st & 0xFFFFu) / 65536.f) - 0.5f) / fd;
    st = a64 * st + c64;
  }
}

// Noise-table fill: table[a] = largest k with starts[k] <= a.
__global__ __launch_bounds__(256) void k_table_fill(int32_t* table, u64 T, const int64_t* starts, int n) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < T; x += stride) {
    int lo = 0, hi = n - 1;  // starts[lo] <= x always (starts[0] = 0)
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if ((u64)starts[mid] <= x) lo = mid; else hi = mid - 1;
    }
    table[x] = lo;
  }
}

__global__ __launch_bounds__(256) void k_f32_to_bf16(const float* src, uint16_t* dst, u64 n) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride)
    dst[x] = (uint16_t)f32_to_bf16(src[x]);
}

__global__ __launch_bounds__(256) void k_bf16_to_f32(const uint16_t* src, float* dst, u64 n) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride)
    dst[x] = bf16_to_f32(src[x]);
}

// One step on given rows, through the trainer's device code.
template <int NP>
__global__ __launch_bounds__(64) void k_step(int D, int os, int nin, int fin, const int32_t* meta,
                                             float alpha, const float* exp_table, const float* O,
                                             const float* I, float* corr, float* md, float* gin) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x;
  const int K = os - 2, W = (nin + 3) / 4 > 0 ? (nin + 3) / 4 : 1;
  const int MC = 4 * W, DP = D >> 1, OST = o_stride(D), IST = i_stride(D);
  const Lds s = carve(smem, 1, K, W, D);
  for (int i = lane; i < os; i += 64) s.ometa[i] = meta[i];
  for (int i = 0; i < os; i++)
    for (int d = lane; d < D; d += 64) s.O[i * OST + d] = O[(size_t)i * D + d];
  for (int j = 0; j < nin; j++)
    for (int d = lane; d < D; d += 64) s.I[(size_t)j * IST + d] = I[(size_t)j * D + d];
  __syncthreads();
  step_gemm1(s, D, MC, os, nin, fin, alpha, exp_table, lane);
  __syncthreads();
  for (int i = 0; i < os; i++)
    for (int j = lane; j < nin; j += 64) corr[(size_t)i * nin + j] = s.C[i * MC + j];
  for (int i = 0; i < os; i++) {
#pragma unroll
    for (int v = 0; v < NP; v++) {
      const int pr = lane + 64 * v;
      if (pr < DP) st2(md + (size_t)i * D + 2 * pr, step_gemm2_row(s, D, MC, i, nin, pr));
    }
  }
  for (int j = 0; j < nin; j++) {
#pragma unroll
    for (int v = 0; v < NP; v++) {
      const int pr = lane + 64 * v;
      if (pr < DP) st2(gin + (size_t)j * D + 2 * pr, step_gemm3_col(s, D, MC, j, os, pr));
    }
  }
}

// ---- SGNS logistic loss and adjoint (exact sigmoid) -------------------------
template <typename T>
__device__ __forceinline__ T softplus(T x) {
  // log(1 + e^x) = max(x, 0) + log1p(e^-|x|)
  const T ax = x < (T)0 ? -x : x;
  return (x > (T)0 ? x : (T)0) + (T)log1p(exp(-ax));
}

template <typename T>
__device__ __forceinline__ T sigmoid(T x) {
  if (x >= (T)0) return (T)1 / ((T)1 + (T)exp(-x));
  const T e = (T)exp(x);
  return e / ((T)1 + e);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One block (64 lanes) per minibatch; lanes over embedding elements.
template <typename T>
__global__ __launch_bounds__(64) void k_loss_fwd(const T* syn0, const T* syn1, int D, const int32_t* ctx,
                                                 const int32_t* out, int n_ctx, int n_out, T* loss) {
  const int m = blockIdx.x, lane = threadIdx.x;
  T acc = (T)0;
  for (int c = 0; c < n_ctx; c++) {
    const T* x = syn0 + (size_t)ctx[(size_t)m * n_ctx + c] * D;
    for (int o = 0; o < n_out; o++) {
      const T* y = syn1 + (size_t)out[(size_t)m * n_out + o] * D;
      T part = (T)0;
      for (int d = lane; d < D; d += 64) part += x[d] * y[d];
      const T f = wave_sum(part);
      acc += (o == 0) ? softplus(-f) : softplus(f);
    }
  }
  if (lane == 0) atomicAdd(loss, acc);
}

template <typename T>
__global__ __launch_bounds__(64) void k_loss_bwd(const T* syn0, const T* syn1, int D, const int32_t* ctx,
                                                 const int32_t* out, int n_ctx, int n_out, T scale,
                                                 T* g0, T* g1) {
  const int m = blockIdx.x, lane = threadIdx.x;
  for (int c = 0; c < n_ctx; c++) {
    const size_t rc = (size_t)ctx[(size_t)m * n_ctx + c];
    const T* x = syn0 + rc * D;
    for (int o = 0; o < n_out; o++) {
      const size_t ro = (size_t)out[(size_t)m * n_out + o];
      const T* y = syn1 + ro * D;
      T part = (T)0;
      for (int d = lane; d < D; d += 64) part += x[d] * y[d];
      const T f = wave_sum(part);
      // d/df softplus(-f) = sigmoid(f) - 1 ; d/df softplus(f) = sigmoid(f)
      const T coef = scale * (sigmoid(f) - (o == 0 ? (T)1 : (T)0));
      for (int d = lane; d < D; d += 64) {
        atomicAdd(&g0[rc * D + d], coef * y[d]);
        atomicAdd(&g1[ro * D + d], coef * x[d]);
      }
    }
  }
}

template <typename R>
hipError_t launch_train_t(const TrainArgs& a, hipStream_t st, int precision) {
  const size_t lds = train_lds_bytes(a.Lmax, a.K, a.W, a.D, precision);
  const int np = (a.D / 2 + 63) / 64;
  dim3 grid(a.workers), block(64);
  switch (np) {
#define DGER_CASE(X)                                                      \
  case X:                                                                 \
    hipLaunchKernelGGL((k_train<R, X>), grid, block, lds, st, a);         \
    break;
    DGER_CASE(1) DGER_CASE(2) DGER_CASE(3) DGER_CASE(4)
    DGER_CASE(5) DGER_CASE(6) DGER_CASE(7) DGER_CASE(8)
#undef DGER_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

size_t train_lds_bytes(int Lmax, int K, int W, int D, int precision) {
  (void)precision;
  size_t off[7];
  return lds_layout(Lmax, K, W, D, off);
}

hipError_t launch_init_net(float* wih, uint64_t n, int D, hipStream_t s) {
  const u64 waves = (n + 64ull * kInitChunk - 1) / (64ull * kInitChunk);
  const u64 blocks = (waves + 3) / 4;
  hipLaunchKernelGGL(k_init_net, dim3((unsigned)blocks), dim3(256), 0, s, wih, (u64)n, D);
  return hipGetLastError();
}

hipError_t launch_table_fill(int32_t* table, uint64_t T, const int64_t* starts, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_table_fill, dim3(4096), dim3(256), 0, s, table, (u64)T, starts, n);
  return hipGetLastError();
}

hipError_t launch_f32_to_bf16(const float* src, uint16_t* dst, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(4096), dim3(256), 0, s, src, dst, (u64)n);
  return hipGetLastError();
}

hipError_t launch_bf16_to_f32(const uint16_t* src, float* dst, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_bf16_to_f32, dim3(4096), dim3(256), 0, s, src, dst, (u64)n);
  return hipGetLastError();
}

hipError_t launch_train(const TrainArgs& a, int precision, hipStream_t s) {
  if (precision == DGER_PRECISION_BF16) return launch_train_t<uint16_t>(a, s, precision);
  return launch_train_t<float>(a, s, precision);
}

hipError_t launch_step(int D, int os, int nin, int fin, const int32_t* meta, float alpha,
                       const float* exp_table, const float* O, const float* I, float* corr,
                       float* md, float* gin, hipStream_t s) {
  const int K = os - 2, W = (nin + 3) / 4 > 0 ? (nin + 3) / 4 : 1;
  size_t off[7];
  const size_t lds = lds_layout(1, K, W, D, off);
  const int np = (D / 2 + 63) / 64;
  switch (np) {
#define DGER_CASE(X)                                                                        \
  case X:                                                                                   \
    hipLaunchKernelGGL((k_step<X>), dim3(1), dim3(64), lds, s, D, os, nin, fin, meta, alpha, \
                       exp_table, O, I, corr, md, gin);                                     \
    break;
    DGER_CASE(1) DGER_CASE(2) DGER_CASE(3) DGER_CASE(4)
    DGER_CASE(5) DGER_CASE(6) DGER_CASE(7) DGER_CASE(8)
#undef DGER_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_loss_fwd(const T* syn0, const T* syn1, int D, const int32_t* ctx, const int32_t* out,
                           int n_mb, int n_ctx, int n_out, T* loss, hipStream_t s) {
  hipLaunchKernelGGL(k_loss_fwd<T>, dim3(n_mb), dim3(64), 0, s, syn0, syn1, D, ctx, out, n_ctx, n_out, loss);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_loss_bwd(const T* syn0, const T* syn1, int D, const int32_t* ctx, const int32_t* out,
                           int n_mb, int n_ctx, int n_out, T scale, T* g0, T* g1, hipStream_t s) {
  hipLaunchKernelGGL(k_loss_bwd<T>, dim3(n_mb), dim3(64), 0, s, syn0, syn1, D, ctx, out, n_ctx, n_out,
                     scale, g0, g1);
  return hipGetLastError();
}

template hipError_t launch_loss_fwd<float>(const float*, const float*, int, const int32_t*, const int32_t*,
                                           int, int, int, float*, hipStream_t);
template hipError_t launch_loss_fwd<double>(const double*, const double*, int, const int32_t*,
                                            const int32_t*, int, int, int, double*, hipStream_t);
template hipError_t launch_loss_bwd<float>(const float*, const float*, int, const int32_t*, const int32_t*,
                                           int, int, int, float, float*, float*, hipStream_t);
template hipError_t launch_loss_bwd<double>(const double*, const double*, int, const int32_t*,
                                            const int32_t*, int, int, int, double, double*, double*,
                                            hipStream_t);

}  // namespace dger
