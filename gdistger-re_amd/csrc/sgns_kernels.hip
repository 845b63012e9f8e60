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

#include <type_traits>

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

// Row storage traits.  A lane works on pairs of adjacent elements.  Model
// rows are loaded and stored at agent scope: workers on different XCDs race
// on the same rows (Hogwild), and plain write-back stores could leave
// different versions of a row in different XCDs' L2 caches after the pass.
template <typename R>
struct Row;

template <>
struct Row<float> {
  // Raw element pair as loaded; widen() converts it to two floats.
  __device__ static u64 load_raw(const void* base, size_t row, int D, int pr) {
    const u64* p = reinterpret_cast<const u64*>(static_cast<const float*>(base) + row * (size_t)D) + pr;
    return __hip_atomic_load(const_cast<u64*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static float2 widen(u64 v) {
    return make_float2(__uint_as_float((unsigned)v), __uint_as_float((unsigned)(v >> 32)));
  }
  __device__ static float2 load(const void* base, size_t row, int D, int pr) {
    return widen(load_raw(base, row, D, pr));
  }
  __device__ static void store(void* base, size_t row, int D, int pr, float2 v) {
    u64* p = reinterpret_cast<u64*>(static_cast<float*>(base) + row * (size_t)D) + pr;
    __hip_atomic_store(p, (u64)__float_as_uint(v.x) | ((u64)__float_as_uint(v.y) << 32), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static void atomic_add(void* base, size_t row, int D, int pr, float2 v) {
    float* p = static_cast<float*>(base) + row * (size_t)D + 2 * pr;
    unsafeAtomicAdd(p, v.x);
    unsafeAtomicAdd(p + 1, v.y);
  }
  // Whole-row store / target update (every lane of the wavefront calls them;
  // lanes past the row end store nothing).
  template <int NP>
  __device__ static void store_row(void* base, size_t row, int D, int lane, const float2 (&v)[NP], unsigned) {
#pragma unroll
    for (int u = 0; u < NP; u++)
      if (lane + 64 * u < (D >> 1)) store(base, row, D, lane + 64 * u, v[u]);
  }
  template <int NP>
  __device__ static void add_row(void* base, size_t row, int D, int lane, const float2 (&)[NP],
                                 const float2 (&d)[NP], unsigned) {
#pragma unroll
    for (int u = 0; u < NP; u++)
      if (lane + 64 * u < (D >> 1)) atomic_add(base, row, D, lane + 64 * u, d[u]);
  }
};

template <>
struct Row<uint16_t> {
  __device__ static u64 load_raw(const void* base, size_t row, int D, int pr) {
    const unsigned* p =
        reinterpret_cast<const unsigned*>(static_cast<const uint16_t*>(base) + row * (size_t)D) + pr;
    return __hip_atomic_load(const_cast<unsigned*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static float2 widen(u64 v) {
    return make_float2(bf16_to_f32((unsigned)v & 0xFFFFu), bf16_to_f32((unsigned)v >> 16));
  }
  __device__ static float2 load(const void* base, size_t row, int D, int pr) {
    return widen(load_raw(base, row, D, pr));
  }
  __device__ static void store(void* base, size_t row, int D, int pr, float2 v) {
    unsigned* p = reinterpret_cast<unsigned*>(static_cast<uint16_t*>(base) + row * (size_t)D) + pr;
    __hip_atomic_store(p, f32_to_bf16(v.x) | (f32_to_bf16(v.y) << 16), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static void atomic_add(void* base, size_t row, int D, int pr, float2 v) {
    typedef short __attribute__((ext_vector_type(2))) short2v;
    short2v* p = reinterpret_cast<short2v*>(static_cast<uint16_t*>(base) + row * (size_t)D) + pr;
    short2v x;
    x.x = (short)f32_to_bf16(v.x);
    x.y = (short)f32_to_bf16(v.y);
    (void)__builtin_amdgcn_flat_atomic_fadd_v2bf16(p, x);
  }
  // Whole-row store / target update (every lane of the wavefront calls them;
  // lanes past the row end store nothing).
  template <int NP>
  __device__ static void store_row(void* base, size_t row, int D, int lane, const float2 (&v)[NP], unsigned) {
#pragma unroll
    for (int u = 0; u < NP; u++)
      if (lane + 64 * u < (D >> 1)) store(base, row, D, lane + 64 * u, v[u]);
  }
  template <int NP>
  __device__ static void add_row(void* base, size_t row, int D, int lane, const float2 (&)[NP],
                                 const float2 (&d)[NP], unsigned) {
#pragma unroll
    for (int u = 0; u < NP; u++)
      if (lane + 64 * u < (D >> 1)) atomic_add(base, row, D, lane + 64 * u, d[u]);
  }
};

// FP8 rows (base is a device Fp8Mat): e4m3 codes scaled per row, FP32 master
// rows for the highest-count ranks.  A row store takes the row's largest
// magnitude over the wavefront, scales it to 448 (the largest e4m3 value)
// and rounds every element stochastically (unbiased, so that small updates
// are not lost to the 3-bit mantissa).  A target update of a coded row is a
// read-modify-write of the whole row from the worker's staged copy (the
// reference adds into the shared row without atomics, dsgl.hpp:1352-1356);
// master rows take FP32 atomic adds.
__device__ __forceinline__ unsigned hash32(unsigned x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}

template <>
struct Row<uint8_t> {
  __device__ static const Fp8Mat& mat(const void* base) { return *static_cast<const Fp8Mat*>(base); }
  __device__ static float2 load(const void* base, size_t row, int D, int pr) {
    const Fp8Mat& m = mat(base);
    if (row < (size_t)m.n_master) return Row<float>::load(m.master, row, D, pr);
    const uint8_t* r = m.rows + row * fp8_row_stride(D);
    const unsigned short q = __hip_atomic_load(const_cast<unsigned short*>(reinterpret_cast<const unsigned short*>(r) + pr),
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const float sc = __hip_atomic_load(const_cast<float*>(reinterpret_cast<const float*>(r + D)), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
    return make_float2(__builtin_amdgcn_cvt_f32_fp8((int)q, 0) * sc, __builtin_amdgcn_cvt_f32_fp8((int)q, 1) * sc);
  }
  // Decoded at load: the prefetch registers hold FP32 pairs.
  __device__ static u64 load_raw(const void* base, size_t row, int D, int pr) {
    const float2 v = load(base, row, D, pr);
    return (u64)__float_as_uint(v.x) | ((u64)__float_as_uint(v.y) << 32);
  }
  __device__ static float2 widen(u64 v) { return Row<float>::widen(v); }
  template <int NP>
  __device__ static void store_row(void* base, size_t row, int D, int lane, const float2 (&v)[NP], unsigned seed) {
    const Fp8Mat& m = mat(base);
    const int DP = D >> 1;
    if (row < (size_t)m.n_master) {
#pragma unroll
      for (int u = 0; u < NP; u++)
        if (lane + 64 * u < DP) Row<float>::store(m.master, row, D, lane + 64 * u, v[u]);
      return;
    }
    float mx = 0.f;
#pragma unroll
    for (int u = 0; u < NP; u++)
      if (lane + 64 * u < DP) mx = fmaxf(mx, fmaxf(fabsf(v[u].x), fabsf(v[u].y)));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    const float inv = mx > 0.f ? 448.f / mx : 0.f;
    uint8_t* r = m.rows + row * fp8_row_stride(D);
#pragma unroll
    for (int u = 0; u < NP; u++) {
      const int pr = lane + 64 * u;
      if (pr < DP) {
        const unsigned h = hash32(seed ^ hash32((unsigned)row * 1024u + (unsigned)pr));
        const float q0 = fminf(fmaxf(v[u].x * inv, -448.f), 448.f);
        const float q1 = fminf(fmaxf(v[u].y * inv, -448.f), 448.f);
        int w = __builtin_amdgcn_cvt_sr_fp8_f32(q0, (int)h, 0, 0);
        w = __builtin_amdgcn_cvt_sr_fp8_f32(q1, (int)hash32(h), w, 1);
        __hip_atomic_store(reinterpret_cast<unsigned short*>(r) + pr, (unsigned short)(w & 0xFFFF), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (lane == 0)
      __hip_atomic_store(reinterpret_cast<float*>(r + D), mx > 0.f ? mx / 448.f : 0.f, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
  template <int NP>
  __device__ static void add_row(void* base, size_t row, int D, int lane, const float2 (&old)[NP],
                                 const float2 (&d)[NP], unsigned seed) {
    const Fp8Mat& m = mat(base);
    if (row < (size_t)m.n_master) {
#pragma unroll
      for (int u = 0; u < NP; u++)
        if (lane + 64 * u < (D >> 1)) Row<float>::atomic_add(m.master, row, D, lane + 64 * u, d[u]);
      return;
    }
    float2 v[NP];
#pragma unroll
    for (int u = 0; u < NP; u++) v[u] = make_float2(old[u].x + d[u].x, old[u].y + d[u].y);
    store_row<NP>(base, row, D, lane, v, seed);
  }
  // Element-wise forms are not used for FP8 rows: the host admits the FP8
  // form only with shapes that take the whole-row paths.
  __device__ static void store(void*, size_t, int, int, float2) {}
  __device__ static void atomic_add(void*, size_t, int, int, float2) {}
};

// Table-quantised sigmoid gradient, dsgl.hpp:1249-1270.  The index is
// clamped to the table only so that an undefined input can never read out of
// bounds; for every finite f in [-6, 6] it is (int)((f + 6) * 83) <= 996.
// The table is read unconditionally and the saturated cases selected after,
// so that several lookups can be in flight at once.
__device__ __forceinline__ float sig_grad(float f, int label, float alpha, const float* exp_table) {
  int idx = (int)((f + 6.0f) * 83.0f);
  idx = min(max(idx, 0), kExpTable - 1);
  const float e = exp_table[idx];
  const float lab = (float)label;
  const float mid = (lab - e) * alpha;
  const float hi = (lab - 1.0f) * alpha;
  const float lo = lab * alpha;
  return f > 6.0f ? hi : (f < -6.0f ? lo : mid);
}

constexpr int kOC = 8;  // output rows per register chunk

// LDS carve-up of one worker (all offsets 16-byte aligned).
struct Lds {
  int* tok;     // 2 x Lmax ranks (sentence q at tok + q*Lmax)
  int* slot;    // 2 x Lmax: first position of the same rank in the sentence
  int* nxt;     // 2 x Lmax: next position of the same rank (kNone if none)
  int* prv;     // 2 x Lmax: previous position of the same rank (-1 if none)
  int* pent;    // 2 x Lmax: ring entry holding the row of each window position
  int* negw;    // K*Lmax negative draws
  int* negs;    // K*Lmax: first draw of the same rank
  unsigned* nflag;  // bit per negative-buffer row: the row holds its snapshot
  unsigned* nbloom; // kBloomWords: bit hash(rank) set for every rank of the negative batch
  int* cpos;    // 4W context positions
  int* csent;   // 4W context sentences
  int* ckey;    // 4W ring rows of the contexts (q*RR + entry)
  int* plans;   // two output plans (current step, next step), 3*KO8 ints each:
                // ranks [0, KO8), multiplicities [KO8, 2*KO8), negative-buffer rows [2*KO8, 3*KO8)
  float* ring;  // 2*RR rows x RST: the window rows of both sentences
  float* O;     // KO8 x OST staged output rows
  float* Ct;    // 4W x CST gradient coefficients, transposed: Ct[j][i]
};

constexpr int kNone = 1 << 30;
// Element pairs per lane for row length D, rounded up to 1, 2, 4 or 8 (the
// instantiated widths; lanes past the row end are masked).
__host__ __device__ inline int np_class(int D) {
  const int np = (D / 2 + 63) / 64;
  return np <= 1 ? 1 : np <= 2 ? 2 : np <= 4 ? 4 : 8;
}
constexpr int kMaxWpb = 4;  // workers per workgroup: one wavefront per SIMD, the full register file each
__host__ __device__ inline size_t exp_lds_bytes() { return ((size_t)kExpTable * 4 + 15) & ~(size_t)15; }

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }
__host__ __device__ inline int ko8(int K) { return ((K + 2) + kOC - 1) & ~(kOC - 1); }
__host__ __device__ inline int o_stride(int D) { return (D + 3) & ~3; }
__host__ __device__ inline int ring_stride(int D) { return ((D + 3) & ~3) + 4; }
__host__ __device__ inline int c_stride(int K) { return ko8(K); }
__host__ __device__ inline int nflag_words(int K, int Lmax) { return (K * Lmax + 31) / 32; }
constexpr int kBloomWords = 64;
__host__ __device__ inline unsigned bloom_bit(int rank) { return ((unsigned)rank * 0x9E3779B1u) >> 21; }
// MFMA path (at most kOC output rows): coefficient tile Cji[j][i],
// 32*ceil(4W/32) context rows of kOC output columns.
__host__ __device__ inline int cmf_stride(int K) { (void)K; return kOC; }
__host__ __device__ inline int cmf_rows(int W) { return ((4 * W + 31) / 32) * 32; }
__host__ __device__ inline size_t cmat_floats(int K, int W) {
  const size_t fma = (size_t)(4 * W) * c_stride(K), mf = (size_t)cmf_rows(W) * cmf_stride(K);
  return fma > mf ? fma : mf;
}

__host__ __device__ inline size_t lds_layout(int Lmax, int K, int W, int D, size_t* off) {
  const int KO = ko8(K), MC = 4 * W, RR = 2 * W + 1;
  size_t o = 0;
  off[0] = o; o = align16(o + 10 * (size_t)Lmax * 4);      // tok, slot, nxt, prv, pent (x2)
  off[1] = o; o = align16(o + (2 * (size_t)K * Lmax + nflag_words(K, Lmax) + kBloomWords) * 4);  // negw, negs, nflag, nbloom
  off[2] = o; o = align16(o + 3 * (size_t)MC * 4);         // cpos, csent, ckey
  off[3] = o; o = align16(o + 6 * (size_t)KO * 4);         // two output plans
  off[4] = o; o = align16(o + 2 * (size_t)RR * ring_stride(D) * 4);
  off[5] = o; o = align16(o + (size_t)KO * o_stride(D) * 4);
  off[6] = o; o = align16(o + cmat_floats(K, W) * 4);
  return o;
}

__device__ inline Lds carve(unsigned char* smem, int Lmax, int K, int W, int D) {
  size_t off[7];
  lds_layout(Lmax, K, W, D, off);
  const int KO = ko8(K), MC = 4 * W;
  Lds s;
  int* b0 = reinterpret_cast<int*>(smem + off[0]);
  s.tok = b0; s.slot = b0 + 2 * Lmax; s.nxt = b0 + 4 * Lmax; s.prv = b0 + 6 * Lmax; s.pent = b0 + 8 * Lmax;
  int* b1 = reinterpret_cast<int*>(smem + off[1]);
  s.negw = b1; s.negs = b1 + K * Lmax;
  s.nflag = reinterpret_cast<unsigned*>(b1 + 2 * K * Lmax);
  s.nbloom = s.nflag + nflag_words(K, Lmax);
  int* b2 = reinterpret_cast<int*>(smem + off[2]);
  s.cpos = b2; s.csent = b2 + MC; s.ckey = b2 + 2 * MC;
  int* b3 = reinterpret_cast<int*>(smem + off[3]);
  s.plans = b3;
  s.ring = reinterpret_cast<float*>(smem + off[4]);
  s.O = reinterpret_cast<float*>(smem + off[5]);
  s.Ct = reinterpret_cast<float*>(smem + off[6]);
  return s;
}

__device__ __forceinline__ int* plan_idx(const Lds& s, int K, int pl) { return s.plans + pl * 3 * ko8(K); }
__device__ __forceinline__ int* plan_meta(const Lds& s, int K, int pl) { return s.plans + pl * 3 * ko8(K) + ko8(K); }
__device__ __forceinline__ int* plan_slot(const Lds& s, int K, int pl) { return s.plans + pl * 3 * ko8(K) + 2 * ko8(K); }

// first[p] = smallest q <= p with v[q] == v[p], for p < n (values in LDS).
// The inner loop has a wave-uniform trip count and broadcast reads.
__device__ inline void first_occurrence(const int* v, int* first, int n, int lane) {
  for (int c0 = 0; c0 < n; c0 += 64) {
    const int p = c0 + lane;
    const int x = (p < n) ? v[p] : 0;
    const int lim = min(n, c0 + 64);
    int f = p;
#pragma unroll 8
    for (int q = 0; q < lim; q++) {
      const int y = v[q];
      f = (q < f && y == x) ? q : f;
    }
    if (p < n) first[p] = f;
  }
}

// first[p] as above through an LDS hash table of 2^cap_log2 slots: the
// value of each slot in tab[0, cap), the smallest position holding it in
// tab[cap, 2 cap) (atomic min).  Values are non-negative; n must stay below
// 3/4 of the slot count so that probes end quickly.
__device__ inline void first_occurrence_hash(const int* v, int* first, int n, int lane, unsigned* tab, int cap_log2) {
  const unsigned cap = 1u << cap_log2, empty = 0xFFFFFFFFu;
  unsigned* const keys = tab;
  unsigned* const pos = tab + cap;
  for (unsigned x = lane; x < cap; x += 64) {
    keys[x] = empty;
    pos[x] = empty;
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  for (int c0 = 0; c0 < n; c0 += 64) {
    const int p = c0 + lane;
    if (p < n) {
      const unsigned key = (unsigned)v[p];
      unsigned h = (key * 0x9E3779B1u) >> (32 - cap_log2);
      for (;;) {
        const unsigned old = atomicCAS(&keys[h], empty, key);
        if (old == empty || old == key) {
          atomicMin(&pos[h], (unsigned)p);
          break;
        }
        h = (h + 1) & (cap - 1);
      }
      first[p] = (int)h;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  for (int p = lane; p < n; p += 64) first[p] = (int)pos[first[p]];
}

// For each position p < n: first, previous and next position holding the
// same value (kNone / -1 where there is none).
__device__ inline void occurrences(const int* v, int* first, int* prev, int* next, int n, int lane) {
  for (int c0 = 0; c0 < n; c0 += 64) {
    const int p = c0 + lane;
    const int x = (p < n) ? v[p] : 0;
    int f = p, pv = -1, nx = kNone;
#pragma unroll 8
    for (int q = 0; q < n; q++) {
      const bool eq = v[q] == x;
      f = (eq && q < f) ? q : f;
      pv = (eq && q < p) ? q : pv;
      nx = (eq && q > p && nx == kNone) ? q : nx;
    }
    if (p < n) { first[p] = f; prev[p] = pv; next[p] = nx; }
  }
}

// GEMM1 + sigmoid (dsgl.hpp:1241-1272): lanes over contexts j (row
// rows + ckey[j]*RST); for each chunk of 8 output rows, f[i] = sum_k
// O[i][k]*I[j][k] in sequential k with one FMA per term.  Ct[j][i] = g*meta[i]
// for i < os, 0 for padding rows.
__device__ inline void step_gemm1(const Lds& s, const float* rows, int RST, const int* ometa, int D, int K,
                                  int os, int nin, int fin, float alpha, const float* exp_table, int lane) {
  const int OST = o_stride(D), CST = c_stride(K);
  const int D4 = D >> 2;
  for (int j = lane; j < nin; j += 64) {
    const float* x = rows + (size_t)s.ckey[j] * RST;
    for (int i0 = 0; i0 < os; i0 += kOC) {
      float acc[kOC];
#pragma unroll
      for (int ii = 0; ii < kOC; ii++) acc[ii] = 0.f;
      // operands of k4+1 are read while k4 is multiplied
      float4 xv = D4 > 0 ? *reinterpret_cast<const float4*>(x) : make_float4(0.f, 0.f, 0.f, 0.f);
      float4 ov[kOC];
#pragma unroll
      for (int ii = 0; ii < kOC; ii++)
        ov[ii] = D4 > 0 ? *reinterpret_cast<const float4*>(s.O + (i0 + ii) * OST) : make_float4(0.f, 0.f, 0.f, 0.f);
      for (int k4 = 0; k4 < D4; k4++) {
        const int kn = (k4 + 1 < D4) ? k4 + 1 : k4;
        const float4 xn = *reinterpret_cast<const float4*>(x + 4 * kn);
        float4 on[kOC];
#pragma unroll
        for (int ii = 0; ii < kOC; ii++) on[ii] = *reinterpret_cast<const float4*>(s.O + (i0 + ii) * OST + 4 * kn);
#pragma unroll
        for (int ii = 0; ii < kOC; ii++) {
          acc[ii] = __builtin_fmaf(ov[ii].x, xv.x, acc[ii]);
          acc[ii] = __builtin_fmaf(ov[ii].y, xv.y, acc[ii]);
          acc[ii] = __builtin_fmaf(ov[ii].z, xv.z, acc[ii]);
          acc[ii] = __builtin_fmaf(ov[ii].w, xv.w, acc[ii]);
        }
        xv = xn;
#pragma unroll
        for (int ii = 0; ii < kOC; ii++) ov[ii] = on[ii];
      }
      for (int k = 4 * D4; k < D; k++) {
        const float xk = x[k];
#pragma unroll
        for (int ii = 0; ii < kOC; ii++) acc[ii] = __builtin_fmaf(s.O[(i0 + ii) * OST + k], xk, acc[ii]);
      }
#pragma unroll
      for (int ii = 0; ii < kOC; ii++) {
        const int i = i0 + ii;
        const int label = (j < fin) ? (i == 0 ? 1 : 0) : (i == os - 1 ? 1 : 0);
        const float g = sig_grad(acc[ii], label, alpha, exp_table) * (float)ometa[i];
        s.Ct[(size_t)j * CST + i] = (i < os) ? g : 0.f;
      }
    }
  }
}

// GEMM2 (dsgl.hpp:1290): md[i] = sum_j C[i][j] * I[j][:] in sequential j;
// lanes over element pairs, 8 output rows per chunk; sink(i, pr, md) for i < os.
template <int NP, typename Sink>
__device__ inline void step_gemm2(const Lds& s, const float* rows, int RST, int D, int K, int os, int nin,
                                  int lane, Sink&& sink) {
  const int CST = c_stride(K), DP = D >> 1;
  for (int i0 = 0; i0 < os; i0 += kOC) {
    float2 acc[kOC][NP];
#pragma unroll
    for (int ii = 0; ii < kOC; ii++)
#pragma unroll
      for (int u = 0; u < NP; u++) acc[ii][u] = make_float2(0.f, 0.f);
    for (int j0 = 0; j0 < nin; j0 += 8) {
      float4 cb[8][2];
      float2 xb[8][NP];
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const int j = min(j0 + b, nin - 1);
        cb[b][0] = *reinterpret_cast<const float4*>(s.Ct + (size_t)j * CST + i0);
        cb[b][1] = *reinterpret_cast<const float4*>(s.Ct + (size_t)j * CST + i0 + 4);
        const float* xr = rows + (size_t)s.ckey[j] * RST;
#pragma unroll
        for (int u = 0; u < NP; u++) {
          const int pr = lane + 64 * u;
          xb[b][u] = (pr < DP) ? *reinterpret_cast<const float2*>(xr + 2 * pr) : make_float2(0.f, 0.f);
        }
      }
#pragma unroll
      for (int b = 0; b < 8; b++) {
        if (j0 + b < nin) {
          const float c[kOC] = {cb[b][0].x, cb[b][0].y, cb[b][0].z, cb[b][0].w,
                                cb[b][1].x, cb[b][1].y, cb[b][1].z, cb[b][1].w};
#pragma unroll
          for (int u = 0; u < NP; u++) {
#pragma unroll
            for (int ii = 0; ii < kOC; ii++) {
              acc[ii][u].x = __builtin_fmaf(c[ii], xb[b][u].x, acc[ii][u].x);
              acc[ii][u].y = __builtin_fmaf(c[ii], xb[b][u].y, acc[ii][u].y);
            }
          }
        }
      }
    }
#pragma unroll
    for (int ii = 0; ii < kOC; ii++) {
      const int i = i0 + ii;
      if (i < os) {
#pragma unroll
        for (int u = 0; u < NP; u++) {
          const int pr = lane + 64 * u;
          if (pr < DP) sink(i, pr, acc[ii][u]);
        }
      }
    }
  }
}

// GEMM3 (dsgl.hpp:1309) for contexts j0..j0+7: g[b] = sum_i C[i][j0+b] *
// O[i][:] in sequential i, for this lane's element pairs.  o0 holds output
// rows 0..7; rows beyond 8 are read from LDS.
template <int NP>
__device__ inline void step_gemm3_batch(const Lds& s, int D, int K, int os, int nin, int j0, int lane,
                                        const float2 (&o0)[kOC][NP], float2 (&g)[8][NP]) {
  const int OST = o_stride(D), CST = c_stride(K), DP = D >> 1;
  float4 cb[8][2];
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const int j = min(j0 + b, nin - 1);
    cb[b][0] = *reinterpret_cast<const float4*>(s.Ct + (size_t)j * CST);
    cb[b][1] = *reinterpret_cast<const float4*>(s.Ct + (size_t)j * CST + 4);
  }
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const int j = min(j0 + b, nin - 1);
#pragma unroll
    for (int u = 0; u < NP; u++) g[b][u] = make_float2(0.f, 0.f);
    for (int i0 = 0; i0 < os; i0 += kOC) {
      float4 c0 = cb[b][0], c1 = cb[b][1];
      if (i0 > 0) {
        c0 = *reinterpret_cast<const float4*>(s.Ct + (size_t)j * CST + i0);
        c1 = *reinterpret_cast<const float4*>(s.Ct + (size_t)j * CST + i0 + 4);
      }
      const float c[kOC] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
      for (int ii = 0; ii < kOC; ii++) {
        const bool on = i0 + ii < os;
#pragma unroll
        for (int u = 0; u < NP; u++) {
          const int pr = lane + 64 * u;
          float2 o = o0[ii][u];
          if (i0 > 0)
            o = (pr < DP) ? *reinterpret_cast<const float2*>(s.O + (i0 + ii) * OST + 2 * pr)
                          : make_float2(0.f, 0.f);
          const float gx = __builtin_fmaf(c[ii], o.x, g[b][u].x);
          const float gy = __builtin_fmaf(c[ii], o.y, g[b][u].y);
          g[b][u].x = on ? gx : g[b][u].x;
          g[b][u].y = on ? gy : g[b][u].y;
        }
      }
    }
  }
}

// Rows i >= os are zero: the staged-output area doubles as scratch between
// steps, and GEMM3 multiplies every row (by a zero coefficient past os).
template <int NP>
__device__ inline void load_o0(const Lds& s, int D, int os, int lane, float2 (&o0)[kOC][NP]) {
  const int OST = o_stride(D), DP = D >> 1;
#pragma unroll
  for (int ii = 0; ii < kOC; ii++)
#pragma unroll
    for (int u = 0; u < NP; u++) {
      const int pr = lane + 64 * u;
      o0[ii][u] = (ii < os && pr < DP) ? *reinterpret_cast<const float2*>(s.O + ii * OST + 2 * pr)
                                       : make_float2(0.f, 0.f);
    }
}

// ---- BF16 form: the three contractions on MFMA --------------------------------
// v_mfma_f32_16x16x32_bf16, lane l: A[l%16][8*(l/16)+e], B[8*(l/16)+e][l%16],
// D[4*(l/16)+r][l%16].  Operands are the FP32 working rows rounded to bf16;
// accumulation is FP32.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 cvt8(float4 a, float4 b) {
  bf16x8 r;
  r[0] = (__bf16)a.x; r[1] = (__bf16)a.y; r[2] = (__bf16)a.z; r[3] = (__bf16)a.w;
  r[4] = (__bf16)b.x; r[5] = (__bf16)b.y; r[6] = (__bf16)b.z; r[7] = (__bf16)b.w;
  return r;
}

__device__ __forceinline__ bool mfma_ok(int D, int K, int W) {
  return (D % 32) == 0 && D <= 256 && K + 2 <= 16 && W <= 16;
}

// GEMM1 + sigmoid: F[j][i] = sum_k I[j][k] O[i][k]; Cji[j][i] = g * meta[i]
// (0 outside j < nin, i < os), for at most kOC output rows.  Two 16-row
// context tiles per pass; columns i >= os take whatever the staged rows hold
// and are discarded.  Writes 32*ceil(nin/32) context rows.
template <int NP>
__device__ inline void mfma_gemm1(const Lds& s, const float* rows, int RST, const int* ometa, int D, int os,
                                  int nin, int fin, float alpha, const float* exp_table, int lane) {
  const int m = lane & 15, g = lane >> 4, OST = o_stride(D), CIS = cmf_stride(0);
  const int mi = min(m, kOC - 1);
  const float* orow = static_cast<const float*>(__builtin_assume_aligned(s.O + (size_t)mi * OST, 16));
  const float meta = (float)ometa[mi];
  const int nk = D >> 5;
  for (int jt = 0; jt < 2 * ((nin + 31) >> 5); jt += 2) {
    const float* xa = static_cast<const float*>(
        __builtin_assume_aligned(rows + (size_t)s.ckey[min(16 * jt + m, nin - 1)] * RST, 16));
    const float* xb = static_cast<const float*>(
        __builtin_assume_aligned(rows + (size_t)s.ckey[min(16 * jt + 16 + m, nin - 1)] * RST, 16));
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 4 * NP; ks++) {
      if (ks < nk) {
        const int k0 = 32 * ks + 8 * g;
        const float4 a0 = *reinterpret_cast<const float4*>(xa + k0);
        const float4 a1 = *reinterpret_cast<const float4*>(xa + k0 + 4);
        const float4 b0 = *reinterpret_cast<const float4*>(xb + k0);
        const float4 b1 = *reinterpret_cast<const float4*>(xb + k0 + 4);
        const float4 o0 = *reinterpret_cast<const float4*>(orow + k0);
        const float4 o1 = *reinterpret_cast<const float4*>(orow + k0 + 4);
        const bf16x8 ob = cvt8(o0, o1);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cvt8(a0, a1), ob, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cvt8(b0, b1), ob, acc1, 0, 0, 0);
      }
    }
    float c[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int j = 16 * jt + 16 * (r >> 2) + 4 * g + (r & 3);
      const int label = (j < fin) ? (m == 0 ? 1 : 0) : (m == os - 1 ? 1 : 0);
      c[r] = sig_grad((r < 4) ? acc0[r & 3] : acc1[r & 3], label, alpha, exp_table) * meta;
    }
    if (m < kOC) {
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const int j = 16 * jt + 16 * (r >> 2) + 4 * g + (r & 3);
        s.Ct[(size_t)j * CIS + m] = (j < nin && m < os) ? c[r] : 0.f;
      }
    }
  }
}

// ---- FP32 contractions with wave-uniform operands from registers ----------------
// For steps with at most 64 contexts and 8 output rows: output rows and the
// gradient coefficients are broadcast with v_readlane instead of LDS reads,
// which leaves LDS bandwidth to the context rows.  Same sequential FMA order
// as step_gemm1/2/3.
__device__ __forceinline__ float rdl(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// GEMM1 + sigmoid: lane j < nin computes c[i] = g(f[i][j]) * meta[i] for the
// 8 output rows, reading output rows as LDS broadcasts (operands of k4+1 are
// read while k4 is multiplied).
__device__ inline void fast_gemm1(const Lds& s, const float* ring, int RST, const int* ometa, int D, int os,
                                  int nin, int fin, float alpha, const float* exp_table, int lane, int myk) {
  const int OST = o_stride(D), D4 = D >> 2;
  const float* x = ring + (size_t)myk * RST;
  float acc[kOC];
#pragma unroll
  for (int ii = 0; ii < kOC; ii++) acc[ii] = 0.f;
  if (D4 > 0) {
    float4 xv = *reinterpret_cast<const float4*>(x);
    float4 ov[kOC];
#pragma unroll
    for (int ii = 0; ii < kOC; ii++) ov[ii] = *reinterpret_cast<const float4*>(s.O + ii * OST);
    for (int k4 = 0; k4 < D4; k4++) {
      const int kn = (k4 + 1 < D4) ? k4 + 1 : k4;
      const float4 xn = *reinterpret_cast<const float4*>(x + 4 * kn);
      float4 on[kOC];
#pragma unroll
      for (int ii = 0; ii < kOC; ii++) on[ii] = *reinterpret_cast<const float4*>(s.O + ii * OST + 4 * kn);
#pragma unroll
      for (int ii = 0; ii < kOC; ii++) {
        acc[ii] = __builtin_fmaf(ov[ii].x, xv.x, acc[ii]);
        acc[ii] = __builtin_fmaf(ov[ii].y, xv.y, acc[ii]);
        acc[ii] = __builtin_fmaf(ov[ii].z, xv.z, acc[ii]);
        acc[ii] = __builtin_fmaf(ov[ii].w, xv.w, acc[ii]);
      }
      xv = xn;
#pragma unroll
      for (int ii = 0; ii < kOC; ii++) ov[ii] = on[ii];
    }
  }
  for (int k = 4 * D4; k < D; k++) {
    const float xk = x[k];
#pragma unroll
    for (int ii = 0; ii < kOC; ii++) acc[ii] = __builtin_fmaf(s.O[ii * OST + k], xk, acc[ii]);
  }
  float c[kOC];
#pragma unroll
  for (int ii = 0; ii < kOC; ii++) {
    const int label = (lane < fin) ? (ii == 0 ? 1 : 0) : (ii == os - 1 ? 1 : 0);
    const float g = sig_grad(acc[ii], label, alpha, exp_table) * (float)ometa[ii];
    c[ii] = (ii < os) ? g : 0.f;
  }
  if (lane < nin) {
    float* ct = s.Ct + (size_t)lane * kOC;
    *reinterpret_cast<float4*>(ct) = make_float4(c[0], c[1], c[2], c[3]);
    *reinterpret_cast<float4*>(ct + 4) = make_float4(c[4], c[5], c[6], c[7]);
  }
}

// GEMM2 (md[i] = sum_j c[i][j] I[j] in sequential j), and with `fuse` (no
// repeated contexts) GEMM3 in the same pass: I[j] += sum_i c[i][j] O[i] in
// sequential i, written back to the ring right after the row is read.
// Coefficients come from Ct[j][0..kOC) (0 for i >= os, so those terms add
// zero) as LDS broadcasts.  The row and coefficients of context j+1 are read
// before row j is written (distinct rows).  Lanes past the row end compute on
// whatever LDS holds there and store nothing.
template <int NP>
__device__ inline void fast_gemm23(float* ring, int RST, const float* Ct, int D, int nin, int lane, int myk,
                                   const float2 (&o0)[kOC][NP], float2 (&md)[kOC][NP], bool fuse) {
  const int DP = D >> 1;
#pragma unroll
  for (int ii = 0; ii < kOC; ii++)
#pragma unroll
    for (int u = 0; u < NP; u++) md[ii][u] = make_float2(0.f, 0.f);
  const float* xr0 = ring + (size_t)__builtin_amdgcn_readlane(myk, 0) * RST;
  float2 xv[NP];
#pragma unroll
  for (int u = 0; u < NP; u++) xv[u] = *reinterpret_cast<const float2*>(xr0 + 2 * (lane + 64 * u));
  float4 ca = *reinterpret_cast<const float4*>(Ct), cb = *reinterpret_cast<const float4*>(Ct + 4);
  for (int j = 0; j < nin; j++) {
    const int jn = min(j + 1, nin - 1);
    float* xr = ring + (size_t)__builtin_amdgcn_readlane(myk, j) * RST;
    const float* xn_r = ring + (size_t)__builtin_amdgcn_readlane(myk, jn) * RST;
    float2 xn[NP];
#pragma unroll
    for (int u = 0; u < NP; u++) xn[u] = *reinterpret_cast<const float2*>(xn_r + 2 * (lane + 64 * u));
    const float4 na = *reinterpret_cast<const float4*>(Ct + (size_t)jn * kOC);
    const float4 nb = *reinterpret_cast<const float4*>(Ct + (size_t)jn * kOC + 4);
    const float cj[kOC] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
#pragma unroll
    for (int u = 0; u < NP; u++) {
      const int pr = lane + 64 * u;
#pragma unroll
      for (int ii = 0; ii < kOC; ii++) {
        md[ii][u].x = __builtin_fmaf(cj[ii], xv[u].x, md[ii][u].x);
        md[ii][u].y = __builtin_fmaf(cj[ii], xv[u].y, md[ii][u].y);
      }
      if (fuse) {
        float gx = 0.f, gy = 0.f;
#pragma unroll
        for (int ii = 0; ii < kOC; ii++) {
          gx = __builtin_fmaf(cj[ii], o0[ii][u].x, gx);
          gy = __builtin_fmaf(cj[ii], o0[ii][u].y, gy);
        }
        if (pr < DP) *reinterpret_cast<float2*>(xr + 2 * pr) = make_float2(xv[u].x + gx, xv[u].y + gy);
      }
    }
#pragma unroll
    for (int u = 0; u < NP; u++) xv[u] = xn[u];
    ca = na;
    cb = nb;
  }
}

// GEMM3 alone, for steps with repeated contexts: updates added into the ring
// row of each context in context order.  The row of context j+1 is read
// before row j is written; when both are the same row the value just
// computed is taken instead.
template <int NP>
__device__ inline void fast_gemm3(float* ring, int RST, const float* Ct, int D, int nin, int lane, int myk,
                                  const float2 (&o0)[kOC][NP]) {
  const int DP = D >> 1;
  const float* xr0 = ring + (size_t)__builtin_amdgcn_readlane(myk, 0) * RST;
  float2 xv[NP];
#pragma unroll
  for (int u = 0; u < NP; u++) xv[u] = *reinterpret_cast<const float2*>(xr0 + 2 * (lane + 64 * u));
  float4 ca = *reinterpret_cast<const float4*>(Ct), cb = *reinterpret_cast<const float4*>(Ct + 4);
  for (int j = 0; j < nin; j++) {
    const int jn = min(j + 1, nin - 1);
    const int kj = __builtin_amdgcn_readlane(myk, j);
    const int kn = __builtin_amdgcn_readlane(myk, jn);
    float* xr = ring + (size_t)kj * RST;
    const float* xn_r = ring + (size_t)kn * RST;
    float2 xn[NP];
#pragma unroll
    for (int u = 0; u < NP; u++) xn[u] = *reinterpret_cast<const float2*>(xn_r + 2 * (lane + 64 * u));
    const float4 na = *reinterpret_cast<const float4*>(Ct + (size_t)jn * kOC);
    const float4 nb = *reinterpret_cast<const float4*>(Ct + (size_t)jn * kOC + 4);
    const float cj[kOC] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
#pragma unroll
    for (int u = 0; u < NP; u++) {
      const int pr = lane + 64 * u;
      float gx = 0.f, gy = 0.f;
#pragma unroll
      for (int ii = 0; ii < kOC; ii++) {
        gx = __builtin_fmaf(cj[ii], o0[ii][u].x, gx);
        gy = __builtin_fmaf(cj[ii], o0[ii][u].y, gy);
      }
      const float2 nv = make_float2(xv[u].x + gx, xv[u].y + gy);
      if (pr < DP) *reinterpret_cast<float2*>(xr + 2 * pr) = nv;
      if (kn == kj) xn[u] = nv;
    }
#pragma unroll
    for (int u = 0; u < NP; u++) xv[u] = xn[u];
    ca = na;
    cb = nb;
  }
}

__device__ __forceinline__ float2 ld2(const float* p) { return *reinterpret_cast<const float2*>(p); }
__device__ __forceinline__ void st2(float* p, float2 v) { *reinterpret_cast<float2*>(p) = v; }

// Orders the accesses of one wavefront around a point of the step.  A
// workgroup holds several independent workers (one per wavefront), so there
// is no workgroup barrier inside the loop; the lanes of a wavefront observe
// each other's LDS and memory accesses in program order, so this only keeps
// the compiler from moving accesses across the point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// One worker per wavefront, a.wpb workers per workgroup sharing the sigmoid
// table in LDS; each worker's own LDS area follows the table.
template <typename R, int NP>
__global__ __launch_bounds__(64 * kMaxWpb) void k_train(TrainArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wk = blockIdx.x * a.wpb + wave;
  const int D = a.D, W = a.W, K = a.K, Lmax = a.Lmax;
  float* const expt = reinterpret_cast<float*>(smem);
  for (int x = threadIdx.x; x < kExpTable; x += blockDim.x) expt[x] = a.exp_table[x];
  __syncthreads();
  if (wk >= a.workers) return;
  const int DP = D >> 1;
  const int MC = 4 * W, RR = 2 * W + 1, RST = ring_stride(D), OST = o_stride(D), KO = ko8(K);
  size_t off_unused[7];
  const Lds s = carve(smem + exp_lds_bytes() + (size_t)wave * lds_layout(Lmax, K, W, D, off_unused), Lmax, K, W, D);
  // Rows with loads in flight per write-back batch, sized to keep the
  // allocator from spilling (measured per row type at NP = 1: 4 FP32 rows,
  // 16 bf16 / fp8 rows).
  constexpr int B0 = std::is_same<R, float>::value ? 4 : 16;
  constexpr int B = (B0 / NP) > 2 ? (B0 / NP) : 2;
  const bool can_plan = (K + 2 <= kOC);              // next-step outputs fit one register chunk
  int tab_log2 = 0;  // largest power-of-two word count in the staged-output + coefficient areas
  while ((2ull << tab_log2) <= (unsigned long long)KO * OST + cmat_floats(K, W)) tab_log2++;
  // The BF16-MFMA and FP8 forms run GEMM1 on MFMA where the shapes allow.
  const bool use_mfma = !std::is_same<R, float>::value && a.mfma && mfma_ok(D, K, W);
  // Output rows beyond the current step's count are read (and their results
  // discarded) by the 8-row GEMM chunks: start them finite.
  for (int x = lane; x < KO * OST; x += 64) s.O[x] = 0.f;
  wave_sync();

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

  // Working buffers (dsgl.hpp:844-857) in scratch memory: sentence q rows at
  // buf + q*Lmax*D indexed by the first position of each rank (rows of ranks
  // inside the current window live in the LDS ring instead); negative rows
  // at nbuf indexed by the first draw of each rank.
  float* const buf = a.fscratch + (size_t)wk * worker_scratch_floats(Lmax, K, D);
  float* const nbuf = buf + 2 * (size_t)Lmax * D;
  auto ring_row = [&](int q, int e) { return s.ring + (size_t)(q * RR + e) * RST; };

  const float a0 = a.alpha0;
  float alpha = a0;
  int iters_left = a.iter;
  u64 wc = 0, last_wc = 0, cur = start;
  int len[2] = {0, 0}, pos[2] = {0, 0};
  int done_sent = 0, refresh = 1, neg_count = 0, nbn = 1;
  u64 steps = 0, raw = 0, sentences = 0, n_ctx = 0, n_outctx = 0;
  const int stride = record_stride(W, K);
  int32_t* const rec = a.rec ? a.rec + (size_t)wk * (size_t)a.rec_cap * stride : nullptr;
  int plan = 0;           // which output plan the step uses
  bool planned = false;   // the step's outputs were planned and prefetched by the previous step
  int pre_os = 0;
  // Prefetched output rows of the next step, as loaded: FP32 pairs from the
  // negative buffer where bit i of pre_nb is set, model-row pairs otherwise.
  u64 pre_o[kOC][NP];
  unsigned pre_nb = 0;

  // Random bits for stochastic rounding of FP8 row stores (unused by the
  // other forms): worker, step and a per-site salt.
  auto seed_of = [&](unsigned salt) {
    return hash32((unsigned)wk * 0x9E3779B9u + (unsigned)steps * 0x85EBCA6Bu + salt);
  };
  // Raw element pair pr of an output row: negative-buffer row sl (FP32) when
  // from_nb, else Woh row rk.  Both 4-byte loads are issued unconditionally
  // from a selected address (no branch: a branch around loads makes the
  // compiler wait for each one at the join); for bf16 rows the second load
  // repeats the first.  FP8 rows are decoded at load instead.
  auto out_pair = [&](bool from_nb, int sl, int rk, int pr) -> u64 {
    if constexpr (std::is_same<R, uint8_t>::value) {
      return from_nb ? *reinterpret_cast<const u64*>(nbuf + (size_t)sl * D + 2 * pr)
                     : Row<R>::load_raw(a.woh, (size_t)rk, D, pr);
    } else {
      const char* nrow = reinterpret_cast<const char*>(nbuf + (size_t)sl * D) + 8 * pr;
      const char* wrow = static_cast<const char*>(a.woh) + (size_t)rk * D * sizeof(R) + 2 * sizeof(R) * pr;
      const unsigned* p1 = reinterpret_cast<const unsigned*>(from_nb ? nrow : wrow);
      const unsigned* p2 = reinterpret_cast<const unsigned*>(from_nb ? nrow + 4 : (sizeof(R) == 4 ? wrow + 4 : wrow));
      const unsigned x1 = __hip_atomic_load(const_cast<unsigned*>(p1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned x2 = __hip_atomic_load(const_cast<unsigned*>(p2), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return ((u64)x2 << 32) | x1;
    }
  };
  // Copies model rows src_rows[ids[k]] for the k < n with first[k] == k to
  // ring entry k of sentence q (k <= W) or to row k of bq, B rows per batch.
  auto gather = [&](const void* src_rows, const int* ids, const int* first, int n, int q, float* bq) {
    for (int k0 = 0; k0 < n; k0 += B) {
      float2 v[B][NP];
#pragma unroll
      for (int b = 0; b < B; b++) {
        const int k = k0 + b;
        const bool on = k < n && first[k] == k;
        const size_t row = on ? (size_t)ids[k] : 0;
#pragma unroll
        for (int u = 0; u < NP; u++) {
          const int pr = lane + 64 * u;
          v[b][u] = Row<R>::load(src_rows, row, D, min(pr, DP - 1));  // row 0 when off: unused
        }
      }
      // LDS and global destinations in separate loops, so that no store
      // goes through a generic pointer (those also count against LDS waits).
#pragma unroll
      for (int b = 0; b < B; b++) {
        const int k = k0 + b;
        if (k < n && k <= W && first[k] == k) {
          float* d = ring_row(q, k);
#pragma unroll
          for (int u = 0; u < NP; u++) {
            const int pr = lane + 64 * u;
            if (pr < DP) st2(d + 2 * pr, v[b][u]);
          }
        }
      }
#pragma unroll
      for (int b = 0; b < B; b++) {
        const int k = k0 + b;
        if (k < n && k > W && first[k] == k) {
          float* d = bq + (size_t)k * D;
#pragma unroll
          for (int u = 0; u < NP; u++) {
            const int pr = lane + 64 * u;
            if (pr < DP) st2(d + 2 * pr, v[b][u]);
          }
        }
      }
    }
  };
  // Write-backs (the memcpys of dsgl.hpp:1377-1407).  compact() lists the k
  // in [c0, c1) with on(k) into lst (the staged-output area, free between
  // steps) and returns the count; copy_back() writes the working rows
  // src + lst[i]*D over model rows rank(lst[i]), B rows per batch, issuing
  // the loads of batch i+1 before the stores of batch i.
  int* const lst = reinterpret_cast<int*>(s.O);
  const int lst_cap = KO * OST;
  auto compact = [&](int c0, int c1, auto&& on) -> int {
    int cnt = 0;
    for (int k0 = c0; k0 < c1; k0 += 64) {
      const int k = k0 + lane;
      const bool x = k < c1 && on(k);
      const u64 m = __ballot(x);
      if (x) lst[cnt + __popcll(m & lanemask_lt(lane))] = k;
      cnt += __popcll(m);
    }
    wave_sync();
    return cnt;
  };
  auto copy_back = [&](void* dst_rows, const float* src, int cnt, auto&& rank) {
    float2 v[B][NP];
    auto fetch = [&](int i0) {
#pragma unroll
      for (int bb = 0; bb < B; bb++) {
        const int i = i0 + bb;
        const int k = (i < cnt) ? lst[i] : 0;
#pragma unroll
        for (int u = 0; u < NP; u++) {
          const int pr = lane + 64 * u;
          v[bb][u] = ld2(src + (size_t)k * D + 2 * min(pr, DP - 1));  // row 0 when i >= cnt: unused
        }
      }
    };
    fetch(0);
    for (int i0 = 0; i0 < cnt; i0 += B) {
      float2 w[B][NP];
      int rk[B];
#pragma unroll
      for (int bb = 0; bb < B; bb++) {
        rk[bb] = (i0 + bb < cnt) ? rank(lst[i0 + bb]) : 0;
#pragma unroll
        for (int u = 0; u < NP; u++) w[bb][u] = v[bb][u];
      }
      if (i0 + B < cnt) fetch(i0 + B);
#pragma unroll
      for (int bb = 0; bb < B; bb++)
        if (i0 + bb < cnt) Row<R>::template store_row<NP>(dst_rows, (size_t)rk[bb], D, lane, w[bb], seed_of(0x51u));
    }
    wave_sync();
  };
  // Output list (dsgl.hpp:1136-1158) of a step with targets ta, tb and
  // negative group g into plan pl: the target, the distinct negatives of the
  // group in first-draw order with their multiplicities (a negative equal to
  // the target is dropped), then tb.  Lane k holds draw k.  Returns the count.
  auto make_outputs = [&](int pl, int ta, int tb, int g) -> int {
    const bool live = lane < K;
    const int v = live ? s.negw[g * K + lane] : -1;
    const int vslot = live ? s.negs[g * K + lane] : 0;
    int first = lane, mult = 0;
    for (int k2 = 0; k2 < K; k2++) {
      const int w2 = __builtin_amdgcn_readlane(v, k2);
      mult += (w2 == v) ? 1 : 0;
      first = (w2 == v && k2 < first) ? k2 : first;
    }
    const bool keep = live && first == lane && v != ta;
    const u64 km = __ballot(keep);
    const int o = 1 + __popcll(km & lanemask_lt(lane));
    if (keep) { plan_idx(s, K, pl)[o] = v; plan_meta(s, K, pl)[o] = mult; plan_slot(s, K, pl)[o] = vslot; }
    const int n = 2 + __popcll(km);
    if (lane == 0) { plan_idx(s, K, pl)[0] = ta; plan_meta(s, K, pl)[0] = 1; plan_idx(s, K, pl)[n - 1] = tb; plan_meta(s, K, pl)[n - 1] = 1; }
    return n;
  };

  auto nsnap = [&](int sl) -> bool { return (s.nflag[sl >> 5] >> (sl & 31)) & 1u; };
  auto nmark = [&](int sl) {
    if (lane == 0) s.nflag[sl >> 5] |= 1u << (sl & 31);
  };

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
#pragma unroll
    for (int q = 0; q < 2; q++) {
      int* const tq = s.tok + q * Lmax;
      int* const sq = s.slot + q * Lmax;
      for (;;) {
        if (len[q] != 0) break;
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
          const u64 km = __ballot(keep);
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
          if (keep) tq[L + __popcll(km & lanemask_lt(lane))] = r;
          L += nk;
          if (a.sample_on && draws > 0) rng = readlane64(jA, draws - 1) * rng + readlane64(jC, draws - 1);
          cur += (u64)consumed;
          wc += (u64)consumed;
          if (stop) break;
        }
        pos[q] = 0;
        len[q] = L > 1 ? L : 0;
        if (L <= 1) continue;
        wave_sync();
        occurrences(tq, sq, s.prv + q * Lmax, s.nxt + q * Lmax, L, lane);
        wave_sync();
        // Working rows of the sentence's distinct ranks: a rank first seen
        // inside the first window [0, W] goes straight to ring entry = that
        // position, the others to the scratch buffer.
        float* const bq = buf + (size_t)q * Lmax * D;
        gather(a.wih, tq, sq, L, q, bq);
        for (int p = lane; p <= min(W, L - 1); p += 64) s.pent[q * Lmax + p] = sq[p];
        planned = false;
      }
    }

    // End of the slice: in-flight sentences are dropped (dsgl.hpp:1031-1050).
    if (cur >= end) {
      if (lane == 0) atomicAdd(&a.counters[kWordCountActual], wc - last_wc);
      raw += wc;
      if (--iters_left == 0) break;
      wc = 0; cur = start; last_wc = 0;
      len[0] = 0; len[1] = 0;
      planned = false;
      continue;
    }

    // Shared negative batch (dsgl.hpp:1054-1088).
    if (refresh) {
      nbn = max(len[0], len[1]);
      const int NN = nbn * K;
      const u64 A64 = readlane64(jA, 63), C64 = readlane64(jC, 63);
      const u64 r0 = rng;
      u64 cs = r0;  // state before the first draw of the current 64-draw chunk
      for (int base = 0; base < NN; base += 64 * 8) {
        u64 sr[8];
        int v[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
          sr[c] = jA * cs + jC;
          cs = A64 * cs + C64;
        }
#pragma unroll
        for (int c = 0; c < 8; c++) {
          const int n = base + 64 * c + lane;
          v[c] = (n < NN) ? a.table[(sr[c] >> 16) % a.table_size] : 1;
        }
#pragma unroll
        for (int c = 0; c < 8; c++) {
          const int n = base + 64 * c + lane;
          if (n < NN) s.negw[n] = v[c] ? v[c] : (int)(sr[c] % (u64)(a.V - 1) + 1);
        }
      }
      {
        u64 f = r0;
        int left = NN;
        while (left >= 64) { f = A64 * f + C64; left -= 64; }
        if (left > 0) f = readlane64(jA, left - 1) * f + readlane64(jC, left - 1);
        rng = f;
      }
      wave_sync();
      // The staged-output and coefficient areas (contiguous, free here) hold
      // the hash table.
      if (tab_log2 >= 2 && 4 * NN <= 3 * (1 << (tab_log2 - 1)))
        first_occurrence_hash(s.negw, s.negs, NN, lane, reinterpret_cast<unsigned*>(s.O), tab_log2 - 1);
      else
        first_occurrence(s.negw, s.negs, NN, lane);
      for (int w = lane; w < kBloomWords; w += 64) s.nbloom[w] = 0u;
      wave_sync();
      for (int n = lane; n < NN; n += 64) {
        const unsigned h = bloom_bit(s.negw[n]);
        atomicOr(&s.nbloom[h >> 5], 1u << (h & 31));
      }
      wave_sync();
      // The negative buffer holds a snapshot of each row taken at the refresh
      // (dsgl.hpp:1081-1084).  Rows are snapshotted lazily: at first use as a
      // negative (read from Woh then), or, if this worker is about to update
      // the row as a target first, from that step's staged value.  Rows never
      // snapshotted are unchanged by this worker and are not written back.
      for (int w = lane; w < nflag_words(K, Lmax); w += 64) s.nflag[w] = 0u;
      refresh = 0;
      planned = false;
    }

    // ---- one step (dsgl.hpp:1089-1368) ----
    const int p0 = pos[0], p1 = pos[1];
    const int t1 = s.tok[p0];
    const int t2 = s.tok[Lmax + p1];
    rng = rng * kLcgMul + kLcgAdd;
    const int b = (int)(rng % (u64)W);
    const int hw = W - b;
    const int lo0 = max(0, p0 - hw), hi0 = min(len[0] - 1, p0 + hw);
    const int lo1 = max(0, p1 - hw), hi1 = min(len[1] - 1, p1 + hw);
    const int n0 = (p0 - lo0) + (hi0 - p0);
    const int n1 = (p1 - lo1) + (hi1 - p1);
    const int nin = n0 + n1, fin = n0;
    for (int j = lane; j < nin; j += 64) {
      const int q = (j < n0) ? 0 : 1;
      const int jj = q ? j - n0 : j;
      const int lo = q ? lo1 : lo0, pp = q ? p1 : p0;
      const int c = (jj < pp - lo) ? lo + jj : pp + 1 + (jj - (pp - lo));
      s.cpos[j] = c;
      s.csent[j] = q;
      s.ckey[j] = q * RR + s.pent[q * Lmax + c];
    }
    neg_count++;
    const int grp = neg_count % nbn;
    int os;
    if (planned) {
      os = pre_os;
    } else {
      os = make_outputs(plan, t1, t2, grp);
    }
    // Rows entering the windows at this step's transition, read now so the
    // load overlaps the step (their scratch rows are final: the rank is not
    // in the window).
    float2 pin[2][NP];
    bool pin_on[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int pq = q ? p1 : p0, Lq = len[q];
      const int y = pq + W + 1, x = pq - W;
      pin_on[q] = (pq + 1 < Lq) && (y < Lq) && (s.prv[q * Lmax + min(y, Lq - 1)] < max(x, 0));
      const float* src = buf + ((size_t)q * Lmax + (size_t)s.slot[q * Lmax + min(y, Lq - 1)]) * D;
#pragma unroll
      for (int u = 0; u < NP; u++) {
        const int pr = lane + 64 * u;
        pin[q][u] = ld2(src + 2 * min(pr, DP - 1));  // always a valid scratch row; used when pin_on
      }
    }
    wave_sync();
    // Repeated contexts share a ring row; the reduction path below keeps the
    // reference's order for them.  Each context writes its index to a tag
    // for its ring row (coefficient area, free until GEMM1) and reads it
    // back: another index there means another context shares the row.
    int* const tag = reinterpret_cast<int*>(s.Ct);
    for (int j = lane; j < nin; j += 64) tag[s.ckey[j]] = j;
    wave_sync();
    bool has_dup = false;
    for (int j0 = 0; j0 < nin; j0 += 64) {
      const int j = j0 + lane;
      has_dup = has_dup || (__ballot(j < nin && tag[s.ckey[min(j, nin - 1)]] != j) != 0);
    }
    wave_sync();
    // Output rows: prefetched by the previous step, or loaded now.
    if (planned) {
#pragma unroll
      for (int i = 0; i < kOC; i++) {
        if (i < os) {
#pragma unroll
          for (int u = 0; u < NP; u++) {
            const int pr = lane + 64 * u;
            if (pr < DP)
              st2(s.O + i * OST + 2 * pr, ((pre_nb >> i) & 1u) ? Row<float>::widen(pre_o[i][u]) : Row<R>::widen(pre_o[i][u]));
          }
        }
      }
    } else {
      for (int r0 = 0; r0 < os; r0 += kOC) {
        u64 v[kOC][NP];
        bool nb[kOC];
#pragma unroll
        for (int bb = 0; bb < kOC; bb++) {
          const int r = r0 + bb;
          nb[bb] = false;
          if (r < os) {
            const bool tgt = (r == 0) || (r == os - 1);
            const int sl = tgt ? 0 : plan_slot(s, K, plan)[r];
            const int rk = (r == 0) ? t1 : (r == os - 1) ? t2 : plan_idx(s, K, plan)[r];
            nb[bb] = !tgt && nsnap(sl);
#pragma unroll
            for (int u = 0; u < NP; u++) v[bb][u] = out_pair(nb[bb], sl, rk, min(lane + 64 * u, DP - 1));
          }
        }
#pragma unroll
        for (int bb = 0; bb < kOC; bb++) {
          const int r = r0 + bb;
          if (r < os) {
#pragma unroll
            for (int u = 0; u < NP; u++) {
              const int pr = lane + 64 * u;
              if (pr < DP)
                st2(s.O + r * OST + 2 * pr, nb[bb] ? Row<float>::widen(v[bb][u]) : Row<R>::widen(v[bb][u]));
            }
          }
        }
      }
    }
    if (!planned && lane > 0 && lane < os - 1) {
      const int sl = plan_slot(s, K, plan)[lane];
      atomicOr(&s.nflag[sl >> 5], 1u << (sl & 31));
    }
    // Targets that are pending negatives take their snapshot now, before this
    // step's update of the row.
    const unsigned hb1 = bloom_bit(t1), hb2 = bloom_bit(t2);
    const bool maybe1 = (s.nbloom[hb1 >> 5] >> (hb1 & 31)) & 1u;
    const bool maybe2 = (s.nbloom[hb2 >> 5] >> (hb2 & 31)) & 1u;
#pragma unroll
    for (int tq = 0; tq < 2; tq++) {
      const int r = tq ? t2 : t1;
      int sl = -1;
      for (int c0 = 0; c0 < nbn * K && sl < 0 && (tq ? maybe2 : maybe1); c0 += 64) {
        const int n = c0 + lane;
        const u64 hit = __ballot(n < nbn * K && s.negw[n] == r);
        if (hit) sl = s.negs[c0 + __ffsll((long long)hit) - 1];
      }
      if (sl >= 0 && !nsnap(sl)) {
        const float* src = s.O + (tq ? os - 1 : 0) * OST;
#pragma unroll
        for (int u = 0; u < NP; u++) {
          const int pr = lane + 64 * u;
          if (pr < DP) st2(nbuf + (size_t)sl * D + 2 * pr, ld2(src + 2 * pr));
        }
        nmark(sl);
      }
    }
    wave_sync();

    if (rec && (int64_t)steps < a.rec_cap && lane == 0) {
      int32_t* r = rec + (size_t)steps * stride;
      for (int z = 0; z < stride; z++) r[z] = -1;
      r[0] = t1; r[1] = t2; r[2] = b; r[3] = fin; r[4] = nin; r[5] = os;
      r[6] = __float_as_int(alpha);
      r[7] = grp;
      for (int j = 0; j < nin; j++) r[8 + j] = s.tok[s.csent[j] * Lmax + s.cpos[j]];
      for (int i = 0; i < os; i++) {
        r[8 + MC + i] = plan_idx(s, K, plan)[i];
        r[8 + MC + K + 2 + i] = plan_meta(s, K, plan)[i];
      }
    }

    // Contractions: MFMA (BF16 form), register-broadcast FMA (small steps),
    // or the LDS tile path (large windows / many negatives).
    const bool fast = nin <= 64 && os <= kOC;
    const int myk = s.ckey[min(lane, nin - 1)];
    float2 o0r[kOC][NP], mdr[kOC][NP];
    if (fast) {
      // GEMM1 leaves the coefficients in Ct[j][0..kOC).
      if (use_mfma)
        mfma_gemm1<NP>(s, s.ring, RST, plan_meta(s, K, plan), D, os, nin, fin, alpha, expt, lane);
      else
        fast_gemm1(s, s.ring, RST, plan_meta(s, K, plan), D, os, nin, fin, alpha, expt, lane, myk);
      wave_sync();
      load_o0<NP>(s, D, os, lane, o0r);
      fast_gemm23<NP>(s.ring, RST, s.Ct, D, nin, lane, myk, o0r, mdr, !has_dup);
    } else {
      step_gemm1(s, s.ring, RST, plan_meta(s, K, plan), D, K, os, nin, fin, alpha, expt, lane);
    }
    wave_sync();

    // GEMM2 and the output-row scatter: target rows straight into Woh
    // (row 0 first, then row os-1, as dsgl.hpp:1352-1356), negatives into the
    // negative buffer (dsgl.hpp:1358-1366) from their staged values.
    if (fast) {
      const int* oslot = plan_slot(s, K, plan);
      // Target rows (row 0, then row os-1; when both targets are one rank the
      // second update starts from the first's result).
      float2 last_o[NP], last_md[NP];
#pragma unroll
      for (int ii = 0; ii < kOC; ii++)
#pragma unroll
        for (int u = 0; u < NP; u++)
          if (ii == os - 1) { last_o[u] = o0r[ii][u]; last_md[u] = mdr[ii][u]; }
      Row<R>::template add_row<NP>(a.woh, (size_t)t1, D, lane, o0r[0], mdr[0], seed_of(0x71u));
      if (t2 == t1)
#pragma unroll
        for (int u = 0; u < NP; u++) last_o[u] = make_float2(o0r[0][u].x + mdr[0][u].x, o0r[0][u].y + mdr[0][u].y);
      Row<R>::template add_row<NP>(a.woh, (size_t)t2, D, lane, last_o, last_md, seed_of(0x72u));
#pragma unroll
      for (int ii = 1; ii < kOC; ii++) {
        if (ii < os - 1) {
#pragma unroll
          for (int u = 0; u < NP; u++) {
            const int pr = lane + 64 * u;
            if (pr < DP)
              st2(nbuf + (size_t)oslot[ii] * D + 2 * pr,
                  make_float2(o0r[ii][u].x + mdr[ii][u].x, o0r[ii][u].y + mdr[ii][u].y));
          }
        }
      }
    } else
    step_gemm2<NP>(s, s.ring, RST, D, K, os, nin, lane, [&](int i, int pr, float2 md) {
      if (i == 0) Row<R>::atomic_add(a.woh, (size_t)t1, D, pr, md);
      else if (i == os - 1) Row<R>::atomic_add(a.woh, (size_t)t2, D, pr, md);
      else {
        const float2 o = ld2(s.O + i * OST + 2 * pr);
        st2(nbuf + (size_t)plan_slot(s, K, plan)[i] * D + 2 * pr, make_float2(o.x + md.x, o.y + md.y));
      }
    });

    // Plan the next step's outputs and read its rows now (issued after this
    // step's updates of the same rows, so they see them), when neither
    // sentence ends here and no refresh can intervene.
    const bool cont = can_plan && (p0 + 1 < len[0]) && (p1 + 1 < len[1]);
    if (cont) {
      const int np = plan ^ 1;
      const int t1n = s.tok[p0 + 1], t2n = s.tok[Lmax + p1 + 1];
      const int osn = make_outputs(np, t1n, t2n, (neg_count + 1) % nbn);
      wave_sync();
      // Lane i < osn resolves row i's source, then every row is read.
      const bool neg_row = lane > 0 && lane < osn - 1;
      const int lsl = neg_row ? plan_slot(s, K, np)[lane] : 0;
      const int lrk = (lane == 0) ? t1n : (lane == osn - 1) ? t2n : (neg_row ? plan_idx(s, K, np)[lane] : 0);
      pre_nb = (unsigned)__ballot(neg_row && nsnap(lsl));
      // Every row and lane loads (rows past osn re-read the first target,
      // lanes past the row end its last pair): no branch around the loads,
      // so none of them is waited for here.
#pragma unroll
      for (int i = 0; i < kOC; i++) {
        const int sl = __builtin_amdgcn_readlane(lsl, i);
        const int rk = (i < osn) ? __builtin_amdgcn_readlane(lrk, i) : t1n;
        const bool from_nb = (pre_nb >> i) & 1u;
#pragma unroll
        for (int u = 0; u < NP; u++) pre_o[i][u] = out_pair(from_nb, sl, rk, min(lane + 64 * u, DP - 1));
      }
      pre_os = osn;
      if (neg_row) atomicOr(&s.nflag[lsl >> 5], 1u << (lsl & 31));
    }

    // GEMM3 and the context scatter (dsgl.hpp:1314-1333): each context's
    // update is added into its ring row in context order, so repeated
    // contexts accumulate as in the reference's sentence buffer.
    if (fast) {
      if (has_dup) fast_gemm3<NP>(s.ring, RST, s.Ct, D, nin, lane, myk, o0r);
    } else {
      float2 o0[kOC][NP];
      load_o0<NP>(s, D, os, lane, o0);
      for (int j0 = 0; j0 < nin; j0 += 8) {
        float2 g[8][NP];
        step_gemm3_batch<NP>(s, D, K, os, nin, j0, lane, o0, g);
        if (!has_dup) {
          float2 v[8][NP];
          float* rp[8];
#pragma unroll
          for (int bb = 0; bb < 8; bb++) {
            const int j = min(j0 + bb, nin - 1);
            rp[bb] = s.ring + (size_t)s.ckey[j] * RST;
#pragma unroll
            for (int u = 0; u < NP; u++) {
              const int pr = lane + 64 * u;
              v[bb][u] = (pr < DP) ? ld2(rp[bb] + 2 * pr) : make_float2(0.f, 0.f);
            }
          }
#pragma unroll
          for (int bb = 0; bb < 8; bb++) {
            if (j0 + bb < nin) {
#pragma unroll
              for (int u = 0; u < NP; u++) {
                const int pr = lane + 64 * u;
                if (pr < DP) st2(rp[bb] + 2 * pr, make_float2(v[bb][u].x + g[bb][u].x, v[bb][u].y + g[bb][u].y));
              }
            }
          }
        } else {
#pragma unroll
          for (int bb = 0; bb < 8; bb++) {
            const int j = j0 + bb;
            if (j < nin) {
              float* rr = s.ring + (size_t)s.ckey[j] * RST;
#pragma unroll
              for (int u = 0; u < NP; u++) {
                const int pr = lane + 64 * u;
                if (pr < DP) {
                  const float2 x = ld2(rr + 2 * pr);
                  st2(rr + 2 * pr, make_float2(x.x + g[bb][u].x, x.y + g[bb][u].y));
                }
              }
            }
          }
        }
      }
    }
    steps++;
    n_ctx += (u64)nin;
    n_outctx += (u64)(os * nin);
    wave_sync();

    // Transition: advance both sentences; a finished sentence writes its
    // buffer back over Wih (dsgl.hpp:1377-1395), a continuing one slides its
    // window by one position.
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int p = pos[q];
      const int Lq = len[q];
      int* const tq = s.tok + q * Lmax;
      int* const sq = s.slot + q * Lmax;
      int* const pv = s.prv + q * Lmax;
      int* const nx = s.nxt + q * Lmax;
      int* const pe = s.pent + q * Lmax;
      float* const bq = buf + (size_t)q * Lmax * D;
      pos[q] = p + 1;
      if (p + 1 >= Lq) {
        // Every distinct rank's working row overwrites its Wih row: ranks
        // still in the window from their ring entry, the others from the
        // scratch buffer (nx of a rank's first position marks the former; the
        // links are not needed any more).
        const int wlo = max(0, p - W);
        for (int w = wlo + lane; w < Lq; w += 64)
          if (pv[w] < wlo) nx[sq[w]] = -2;
        wave_sync();
        for (int w = wlo; w < Lq; w++) {
          if (pv[w] >= wlo) continue;
          const float* src = ring_row(q, pe[w]);
          float2 v[NP];
#pragma unroll
          for (int u = 0; u < NP; u++) v[u] = ld2(src + 2 * (lane + 64 * u));
          Row<R>::template store_row<NP>(a.wih, (size_t)tq[w], D, lane, v, seed_of(0x33u));
        }
        for (int c0 = 0; c0 < Lq; c0 += lst_cap) {
          const int cnt = compact(c0, min(Lq, c0 + lst_cap), [&](int k) { return sq[k] == k && nx[k] != -2; });
          copy_back(a.wih, bq, cnt, [&](int k) { return tq[k]; });
        }
        len[q] = 0;
        done_sent++;
        sentences++;
        continue;
      }
      const int x = p - W, y = p + W + 1;
      if (x >= 0) {
        // x is the first window position of its rank, so its row is in
        // entry x mod RR; it moves to the rank's next window position, or
        // back to the scratch buffer if the rank leaves the window.
        const int e = x % RR;
        const int z = nx[x];
        if (z <= y) {
          const int ez = z % RR;
          if (ez != e) {
#pragma unroll
            for (int u = 0; u < NP; u++) {
              const int pr = lane + 64 * u;
              if (pr < DP) st2(ring_row(q, ez) + 2 * pr, ld2(ring_row(q, e) + 2 * pr));
            }
          }
          const int sl = sq[x];
          for (int w = x + 1 + lane; w <= min(y, Lq - 1); w += 64)
            if (sq[w] == sl) pe[w] = ez;
        } else {
          float* dst = bq + (size_t)sq[x] * D;
#pragma unroll
          for (int u = 0; u < NP; u++) {
            const int pr = lane + 64 * u;
            if (pr < DP) st2(dst + 2 * pr, ld2(ring_row(q, e) + 2 * pr));
          }
        }
      }
      wave_sync();
      if (y < Lq) {
        const int pvy = pv[y];
        if (pvy >= max(x + 1, 0)) {
          if (lane == 0) pe[y] = pe[pvy];
        } else if (!(x >= 0 && pvy == x)) {
          const int ey = y % RR;
#pragma unroll
          for (int u = 0; u < NP; u++) {
            const int pr = lane + 64 * u;
            if (pr < DP) st2(ring_row(q, ey) + 2 * pr, pin[q][u]);
          }
          if (lane == 0) pe[y] = ey;
        }
      }
    }
    if (done_sent >= 2) {
      // Rows that took a snapshot only.
      const int NN = nbn * K;
      for (int c0 = 0; c0 < NN; c0 += lst_cap) {
        const int cnt = compact(c0, min(NN, c0 + lst_cap), [&](int k) { return s.negs[k] == k && nsnap(k); });
        copy_back(a.woh, nbuf, cnt, [&](int k) { return s.negw[k]; });
      }
      refresh = 1;
      done_sent = 0;
    }
    planned = cont;
    if (cont) plan ^= 1;
    wave_sync();
  }

  if (lane == 0) {
    a.worker_steps[wk] = steps;
    a.worker_alpha[wk] = alpha;
    atomicAdd(&a.counters[kSteps], steps);
    atomicAdd(&a.counters[kRawWords], raw);
    atomicAdd(&a.counters[kSentences], sentences);
    atomicAdd(&a.counters[kContexts], n_ctx);
    atomicAdd(&a.counters[kOutCtx], n_outctx);
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

// FP32 rows -> FP8 matrix, one wavefront per row (any D up to 64 * 2 * 8):
// master rows copied, coded rows scaled to 448 and rounded to nearest even.
__global__ __launch_bounds__(256) void k_fp8_quantize(const float* src, const Fp8Mat* m, uint32_t V, int D) {
  const u64 row = ((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= V) return;
  const float* x = src + row * (u64)D;
  if (row < (u64)m->n_master) {
    for (int k = lane; k < D; k += 64) m->master[row * (u64)D + k] = x[k];
    return;
  }
  float mx = 0.f;
  for (int k = lane; k < D; k += 64) mx = fmaxf(mx, fabsf(x[k]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  const float inv = mx > 0.f ? 448.f / mx : 0.f;
  uint8_t* r = m->rows + row * fp8_row_stride(D);
  for (int k = lane; k < D; k += 64) {
    const float q = fminf(fmaxf(x[k] * inv, -448.f), 448.f);
    r[k] = (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(q, 0.f, 0, false) & 0xFF);
  }
  if (lane == 0) *reinterpret_cast<float*>(r + D) = mx > 0.f ? mx / 448.f : 0.f;
}

__global__ __launch_bounds__(256) void k_fp8_dequantize(const Fp8Mat* m, float* dst, uint32_t V, int D) {
  const u64 row = ((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= V) return;
  float* y = dst + row * (u64)D;
  if (row < (u64)m->n_master) {
    for (int k = lane; k < D; k += 64)
      y[k] = __hip_atomic_load(m->master + row * (u64)D + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const uint8_t* r = m->rows + row * fp8_row_stride(D);
  const float sc = __hip_atomic_load(const_cast<float*>(reinterpret_cast<const float*>(r + D)), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
  for (int k = lane; k < D; k += 64) y[k] = __builtin_amdgcn_cvt_f32_fp8((int)r[k], 0) * sc;
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
  const int OST = o_stride(D), RST = ring_stride(D), CST = c_stride(K), DP = D >> 1;
  const Lds s = carve(smem, 1, K, W, D);
  for (int i = lane; i < ko8(K); i += 64) plan_meta(s, K, 0)[i] = i < os ? meta[i] : 0;
  for (int i = 0; i < ko8(K); i++)
    for (int d = lane; d < OST; d += 64) s.O[i * OST + d] = (i < os && d < D) ? O[(size_t)i * D + d] : 0.f;
  for (int j = 0; j < nin; j++)
    for (int d = lane; d < D; d += 64) s.ring[(size_t)j * RST + d] = I[(size_t)j * D + d];
  for (int j = lane; j < nin; j += 64) s.ckey[j] = j;
  __syncthreads();
  step_gemm1(s, s.ring, RST, plan_meta(s, K, 0), D, K, os, nin, fin, alpha, exp_table, lane);
  __syncthreads();
  for (int i = 0; i < os; i++)
    for (int j = lane; j < nin; j += 64) corr[(size_t)i * nin + j] = s.Ct[(size_t)j * CST + i];
  step_gemm2<NP>(s, s.ring, RST, D, K, os, nin, lane,
                 [&](int i, int pr, float2 v) { st2(md + (size_t)i * D + 2 * pr, v); });
  float2 o0[kOC][NP];
  load_o0<NP>(s, D, os, lane, o0);
  for (int j0 = 0; j0 < nin; j0 += 8) {
    float2 g[8][NP];
    step_gemm3_batch<NP>(s, D, K, os, nin, j0, lane, o0, g);
#pragma unroll
    for (int bb = 0; bb < 8; bb++) {
      if (j0 + bb < nin) {
#pragma unroll
        for (int u = 0; u < NP; u++) {
          const int pr = lane + 64 * u;
          if (pr < DP) st2(gin + (size_t)(j0 + bb) * D + 2 * pr, g[bb][u]);
        }
      }
    }
  }
}

// GEMM1 + sigmoid of the BF16-MFMA step (mfma_gemm1, the trainer's code for
// steps of at most kOC outputs) on given rows: corr[i][j] = the coefficient
// of output i and context j.
template <int NP>
__global__ __launch_bounds__(64) void k_step_mfma_gemm1(int D, int os, int nin, int fin, const int32_t* meta,
                                                        float alpha, const float* exp_table, const float* O,
                                                        const float* I, float* corr) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x;
  const int K = os - 2, W = (nin + 3) / 4 > 0 ? (nin + 3) / 4 : 1;
  const int OST = o_stride(D), RST = ring_stride(D);
  const Lds s = carve(smem, 1, K, W, D);
  for (int i = lane; i < ko8(K); i += 64) plan_meta(s, K, 0)[i] = i < os ? meta[i] : 0;
  for (int i = 0; i < ko8(K); i++)
    for (int d = lane; d < OST; d += 64) s.O[i * OST + d] = (i < os && d < D) ? O[(size_t)i * D + d] : 0.f;
  for (int j = 0; j < nin; j++)
    for (int d = lane; d < D; d += 64) s.ring[(size_t)j * RST + d] = I[(size_t)j * D + d];
  for (int j = lane; j < nin; j += 64) s.ckey[j] = j;
  __syncthreads();
  mfma_gemm1<NP>(s, s.ring, RST, plan_meta(s, K, 0), D, os, nin, fin, alpha, exp_table, lane);
  __syncthreads();
  for (int i = 0; i < os; i++)
    for (int j = lane; j < nin; j += 64) corr[(size_t)i * nin + j] = s.Ct[(size_t)j * cmf_stride(K) + i];
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

// Lets the trainer instance for a's shape take its workgroup's LDS (above the
// 64 KB default); done before a launch, outside any stream capture.
template <typename R>
hipError_t prepare_train_t(const TrainArgs& a, int precision) {
  if (a.wpb < 1 || a.wpb > kMaxWpb) return hipErrorInvalidValue;
  const int lds = (int)(exp_lds_bytes() + (size_t)a.wpb * train_lds_bytes(a.Lmax, a.K, a.W, a.D, precision));
  switch (np_class(a.D)) {
#define DGER_CASE(X) \
  case X:            \
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_train<R, X>), hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    DGER_CASE(1) DGER_CASE(2) DGER_CASE(4) DGER_CASE(8)
#undef DGER_CASE
    default: return hipErrorInvalidValue;
  }
}

__global__ __launch_bounds__(64) void k_zero_u64(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

template <typename R>
hipError_t launch_train_t(const TrainArgs& a, hipStream_t st, int precision) {
  if (a.wpb < 1 || a.wpb > kMaxWpb) return hipErrorInvalidValue;
  const size_t lds = exp_lds_bytes() + (size_t)a.wpb * train_lds_bytes(a.Lmax, a.K, a.W, a.D, precision);
  const int np = np_class(a.D);
  dim3 grid((a.workers + a.wpb - 1) / a.wpb), block(64 * a.wpb);
  switch (np) {
#define DGER_CASE(X)                                                      \
  case X:                                                                 \
    hipLaunchKernelGGL((k_train<R, X>), grid, block, lds, st, a);         \
    break;
    DGER_CASE(1) DGER_CASE(2) DGER_CASE(4) DGER_CASE(8)
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

int train_workers_per_block(int Lmax, int K, int W, int D, int precision, size_t lds_per_block) {
  const size_t per = train_lds_bytes(Lmax, K, W, D, precision);
  if (lds_per_block < exp_lds_bytes() + per) return 0;
  const size_t n = (lds_per_block - exp_lds_bytes()) / per;
  return (int)(n < (size_t)kMaxWpb ? n : (size_t)kMaxWpb);
}

size_t train_block_lds_bytes(int Lmax, int K, int W, int D, int precision, int wpb) {
  return exp_lds_bytes() + (size_t)wpb * train_lds_bytes(Lmax, K, W, D, precision);
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

hipError_t launch_fp8_quantize(const float* src, const Fp8Mat* m, uint32_t V, int D, hipStream_t s) {
  const unsigned blocks = (unsigned)(((uint64_t)V + 3) / 4);
  hipLaunchKernelGGL(k_fp8_quantize, dim3(blocks), dim3(256), 0, s, src, m, V, D);
  return hipGetLastError();
}

hipError_t launch_fp8_dequantize(const Fp8Mat* m, float* dst, uint32_t V, int D, hipStream_t s) {
  const unsigned blocks = (unsigned)(((uint64_t)V + 3) / 4);
  hipLaunchKernelGGL(k_fp8_dequantize, dim3(blocks), dim3(256), 0, s, m, dst, V, D);
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

hipError_t prepare_train(const TrainArgs& a, int precision) {
  if (precision == DGER_PRECISION_FP8) return prepare_train_t<uint8_t>(a, precision);
  if (precision == DGER_PRECISION_BF16 || precision == DGER_PRECISION_BF16_MFMA)
    return prepare_train_t<uint16_t>(a, precision);
  return prepare_train_t<float>(a, precision);
}

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s) {
  if (n < 0 || n > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_train(const TrainArgs& a, int precision, hipStream_t s) {
  if (precision == DGER_PRECISION_FP8) return launch_train_t<uint8_t>(a, s, precision);
  if (precision == DGER_PRECISION_BF16 || precision == DGER_PRECISION_BF16_MFMA)
    return launch_train_t<uint16_t>(a, s, precision);
  return launch_train_t<float>(a, s, precision);
}

hipError_t launch_step_mfma_gemm1(int D, int os, int nin, int fin, const int32_t* meta, float alpha,
                                  const float* exp_table, const float* O, const float* I, float* corr,
                                  hipStream_t s) {
  if (D % 32 != 0 || D < 32 || D > 256 || os < 2 || os > kOC || nin < 1 || nin > 64 || fin < 0 || fin > nin)
    return hipErrorInvalidValue;
  const int K = os - 2, W = (nin + 3) / 4;
  size_t off[7];
  const size_t lds = lds_layout(1, K, W, D, off);
  if (np_class(D) == 1)
    hipLaunchKernelGGL((k_step_mfma_gemm1<1>), dim3(1), dim3(64), lds, s, D, os, nin, fin, meta, alpha, exp_table,
                       O, I, corr);
  else
    hipLaunchKernelGGL((k_step_mfma_gemm1<2>), dim3(1), dim3(64), lds, s, D, os, nin, fin, meta, alpha, exp_table,
                       O, I, corr);
  return hipGetLastError();
}

hipError_t launch_step(int D, int os, int nin, int fin, const int32_t* meta, float alpha,
                       const float* exp_table, const float* O, const float* I, float* corr,
                       float* md, float* gin, hipStream_t s) {
  const int K = os - 2, W = (nin + 3) / 4 > 0 ? (nin + 3) / 4 : 1;
  size_t off[7];
  const size_t lds = lds_layout(1, K, W, D, off);
  switch (np_class(D)) {
#define DGER_CASE(X)                                                                        \
  case X:                                                                                   \
    hipLaunchKernelGGL((k_step<X>), dim3(1), dim3(64), lds, s, D, os, nin, fin, meta, alpha, \
                       exp_table, O, I, corr, md, gin);                                     \
    break;
    DGER_CASE(1) DGER_CASE(2) DGER_CASE(4) DGER_CASE(8)
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
