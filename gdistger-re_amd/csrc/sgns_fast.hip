// gfx950 window-slot trainer of the BF16-MFMA form (DistGER DSGL,
// include/dsgl.hpp:786-1408).
//
// One 64-lane wavefront per Hogwild worker and row length 128, as in
// sgns_kernels.hip: worker k trains the corpus slice and LCG stream of the
// reference's compute thread k+1, so the index stream (sentences,
// subsampling, windows, negative groups, output lists, learning rate) is the
// reference's.  Rows of length 256 take two wavefronts per worker (one
// workgroup), each holding one 128-element half of every row: both run the
// same index stream, each reads, updates and writes back its half, and GEMM1's
// partial dot products of the two halves are added through LDS (two workgroup
// barriers per step).  What differs from the reference is where the rows live
// and how a step is computed:
//
//   * Window slots.  The FP32 working row of every rank inside a sentence's
//     window [p-W, p+W] sits in registers ("slots", 24 per sentence).  A row
//     is read from Wih when its rank enters the window and written back to
//     Wih (bf16: round to nearest, or stochastic rounding with -rounding sr)
//     when its last occurrence leaves it or the sentence ends.  The reference keeps the whole sentence in an FP32
//     buffer and writes it back at the sentence end (dsgl.hpp:959-962,
//     :1377-1395); here a rank that leaves the window and comes back later
//     in the sentence is read again from Wih.
//   * Slot copy.  A bf16 copy of every slot row sits in LDS (GEMM1 reads the
//     slots across lanes from it).  In the instance without masters and
//     stochastic rounding it is refreshed after each step's update and the
//     rows entering the windows are written into it as they arrive, so the
//     write-backs to Wih (round to nearest) read their rows from it, four
//     rows per store instruction, instead of from the slots' registers.
//   * Direct output updates.  Target and negative rows take their updates
//     straight in Woh (bf16 atomic adds).  The reference adds target updates
//     to Woh directly too (dsgl.hpp:1352-1356) but buffers negatives per
//     sentence pair and writes the buffer back over Woh (dsgl.hpp:1081-1084,
//     :1358-1366, :1396-1404).
//   * Repeated contexts of a step share their slot; their coefficients are
//     equal (same row, same label), so the slot takes one coefficient scaled
//     by the repeat count instead of the reference's sequential adds
//     (dsgl.hpp:1314-1333).
//   * The three contractions of a step (dsgl.hpp:1241, :1290, :1309) run on
//     MFMA with bf16 operands and FP32 accumulation, over all 48 slots (a
//     slot that is no context of the step has coefficient 0):
//       GEMM1  F^T = O I^T         v_mfma_f32_16x16x32_bf16, 3 x 4 per step
//       GEMM2  md  = C^T I         v_mfma_f32_16x16x32_bf16, 8 x 2
//       GEMM3  I  += C O           v_mfma_f32_16x16x16_bf16, 3 x 8, the
//                                  slots' registers as the accumulator.
//
// Register layout (D = 128 elements per wavefront, lane = 16 g + m):
//   slot s = 16 rt + 4 g' + r lives in the lanes of group g' = (s >> 2) & 3;
//   lane (g', m) holds elements 8m .. 8m+7 of it in R[rt][t][r], t = 0..7.
// This is the output layout of GEMM3 (16x16x16: D[4g+r][n] for rows = slots,
// columns = elements 8n + t of tile t) and the B operand layout of GEMM2 (the
// K index 8g + e of block kb is slot 16 (2kb + e/4) + 4g + e%4).  GEMM1 needs
// the slots across lanes instead; it reads a bf16 copy of the slots that each
// step writes to LDS.
//
// Output rows (the step's target, distinct negatives, second target; at most
// 16) are read from Woh in two layouts: for GEMM1's A operand lane (g, m)
// reads row m, elements 32 ks + 8 g .. +8 (ks = 0..3); for GEMM3's B operand
// lane (g, m) reads rows 4g .. 4g+3, elements 8m .. 8m+7.  The next step's
// rows are read right after this step's updates (so they include them).

#include <mutex>
#include <type_traits>
#include <hip/hip_runtime.h>

#include "dger_sgns.h"
#include "sgns_internal.h"

namespace dger {
namespace {

typedef unsigned long long u64;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short short4v __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

constexpr int D = kFastDim;           // row elements per wavefront (128)
constexpr int NT = D / 16;            // element tiles of GEMM2 / GEMM3 (8)
constexpr int KB1 = D / 32;           // K blocks of GEMM1 (4)
constexpr int kSlots = 48;            // 3 MFMA tiles of 16 slots
constexpr int kSentSlots = 24;        // slots of one sentence
constexpr int kOutRows = 16;          // output rows per step (MFMA tile)
constexpr int RB = 2 * D;             // bytes of a bf16 row in LDS
constexpr int CBS = 144;              // bytes per output row of the transposed coefficients
constexpr int MDS = 4 * D + 16;       // bytes per output row of the staged Woh updates (FP32)
constexpr int kMdOff = kOutRows * CBS; // staged Woh updates: in rc, after the coefficients
constexpr int kStageOff = 4096;        // FP32 master output rows of the next step, in rc (past a new
                                       // sentence's first-window rows, which the loader stages at its start)
constexpr int kNone = 1 << 30;
constexpr short kNone16 = 0x7FFF;     // "no next occurrence" in the 16-bit position arrays
constexpr int kExpOne = kExpTable, kExpZero = kExpTable + 1;  // sigmoid 1 and 0 after the table

// Byte offset of 16-byte chunk c of row r in an LDS bf16 matrix: the chunks
// of a row are permuted by r so that sixteen lanes reading the same chunk of
// sixteen consecutive rows hit different banks.
__device__ __forceinline__ unsigned swz(int r, int c) { return (unsigned)r * RB + (((unsigned)(c ^ r) & 15u) << 4); }
constexpr unsigned kRsrcWord3 = 0x00020000u;  // raw buffer descriptor word 3 for gfx9
constexpr int kSc1 = 16;                       // cache-policy bit: coherent across XCDs (agent scope)

static_assert(D % 32 == 0 && NT == 8, "the slot layout is written for D = 128");

__device__ __forceinline__ u64 readlane64(u64 v, int l) {
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)v, l);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), l);
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 lanemask_lt(int lane) { return (1ull << lane) - 1ull; }

__device__ __forceinline__ unsigned hash32(unsigned x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}

// Two floats to a bf16 pair, round to nearest even.
__device__ __forceinline__ unsigned pack_rne(float a, float b) {
  bf16x2 v;
  v[0] = (__bf16)a;
  v[1] = (__bf16)b;
  return __builtin_bit_cast(unsigned, v);
}
// Stochastic rounding of finite x to bf16 with 16 random bits.
__device__ __forceinline__ void unpack8(const u32x4& v, float (&e)[8]) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    e[2 * k] = __uint_as_float(v[k] << 16);
    e[2 * k + 1] = __uint_as_float(v[k] & 0xFFFF0000u);
  }
}

// Table-quantised sigmoid gradient (dsgl.hpp:1249-1270); the index is
// clamped only so that an undefined input cannot read out of bounds.
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

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// 16-byte row-segment loads and stores at agent scope through a raw buffer
// descriptor (the matrices stay below 4 GiB; checked on the host).
struct Mat {
  __amdgpu_buffer_rsrc_t rs;
  __device__ Mat(const void* base, unsigned bytes)
      : rs(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, (int)kRsrcWord3)) {}
  __device__ __forceinline__ u32x4 load(unsigned off) const {
    return __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, kSc1);
  }
  __device__ __forceinline__ void store(unsigned off, const u32x4& v) const {
    __builtin_amdgcn_raw_buffer_store_b128(v, rs, (int)off, 0, kSc1);
  }
  __device__ __forceinline__ u32x2 load2(unsigned off) const {
    return __builtin_amdgcn_raw_buffer_load_b64(rs, (int)off, 0, kSc1);
  }
  __device__ __forceinline__ unsigned load1(unsigned off) const {
    return __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, kSc1);
  }
  __device__ __forceinline__ void store2(unsigned off, const u32x2& v) const {
    __builtin_amdgcn_raw_buffer_store_b64(v, rs, (int)off, 0, kSc1);
  }
  __device__ __forceinline__ void store1(unsigned off, unsigned v) const {
    __builtin_amdgcn_raw_buffer_store_b32(v, rs, (int)off, 0, kSc1);
  }
  // 4 bytes per lane straight into LDS at dst + 4 * lane (dst wave-uniform);
  // the data is there after wait_vm_all().
  __device__ __forceinline__ void load_lds(unsigned off, unsigned char* dst) const {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)dst, 4, off, 0, 0, kSc1);
  }
};

// Waits for every vector memory access of the wave (s_waitcnt vmcnt(0)),
// then orders the LDS accesses after it.
__device__ __forceinline__ void wait_vm_all() {
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// For each position p < n: the previous position holding the same rank when
// it lies at most `reach` positions away (-1 otherwise); returns, in lane
// p < nfirst, the first position of p's rank.  The trainer only asks whether
// a rank recurs within one window span (2 window + 1 positions), so nearer
// occurrences are all it needs: each lane compares its rank with the `reach`
// positions before it (LDS reads that do not wait for each other).  The
// first positions (p < nfirst <= reach) follow the previous-occurrence chain.
__device__ inline int occurrences_near(const int* v, short* prev, int n, int nfirst, int reach, int lane) {
  for (int c0 = 0; c0 < n; c0 += 64) {
    const int p = c0 + lane;
    const int x = (p < n) ? v[p] : -1;
    int pv = -1;
    // farthest first, four reads at a time, so that the nearest match is the
    // last one kept and the reads of a group are all in flight (no read
    // depends on a comparison)
    for (int d0 = reach; d0 >= 1; d0 -= 4) {
      int y[4];
#pragma unroll
      for (int j = 0; j < 4; j++) y[j] = v[min(max(p - (d0 - j), 0), n - 1)];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int d = d0 - j, q = p - d;
        pv = ((y[j] == x) & (q >= 0) & (d >= 1)) ? q : pv;
      }
    }
    if (p < n) prev[p] = (short)pv;
  }
  wave_sync();
  int f = lane;
  for (int it = 0; it < nfirst; it++) {
    const int q = (lane < nfirst) ? prev[f] : -1;
    if (__ballot(q >= 0) == 0) break;
    f = (q >= 0) ? q : f;
  }
  return f;
}

// ---- remainders ----------------------------------------------------------------
// Remainders by divisors that stay fixed for many uses, without a division
// each time.  Div32: M = floor((2^32 - 1) / d), so that for every 32-bit x,
// mulhi(x, M) is floor(x / d) or one less (M d > 2^32 - 1 - d) and one
// conditional subtraction finishes the remainder.
struct Div32 {
  unsigned d, M;
};
__device__ __forceinline__ Div32 make_div32(unsigned d) { return Div32{d, 0xFFFFFFFFu / d}; }
__device__ __forceinline__ unsigned mod32(unsigned x, const Div32& v) {
  const unsigned r = x - __umulhi(x, v.M) * v.d;
  return r >= v.d ? r - v.d : r;
}
// x mod d for a 64-bit x and d < 2^16: long division by 16-bit digits.
__device__ __forceinline__ unsigned mod64_small(u64 x, const Div32& v) {
  unsigned r = mod32((unsigned)(x >> 32), v);
  r = mod32((r << 16) | (unsigned)((x >> 16) & 0xFFFFu), v);
  return mod32((r << 16) | (unsigned)(x & 0xFFFFu), v);
}
// x mod d from inv = 1 / d rounded to double, for x < 2^63 and x / d < 2^51:
// x * inv is within 3 * 2^-53 * x / d < 1 of x / d, so its floor is the
// quotient or one off, and one correction step finishes the remainder.
__device__ __forceinline__ u64 mod_recip(u64 x, unsigned d, double inv) {
  const u64 q = (u64)__builtin_floor((double)x * inv);
  long long r = (long long)(x - q * (u64)d);
  r += (r < 0) ? (long long)d : 0;
  r -= (r >= (long long)d) ? (long long)d : 0;
  return (u64)r;
}
// x mod d for any 64-bit x and d < 2^31: the high word first, then the
// 64-bit value below d * 2^32 that remains.
__device__ __forceinline__ unsigned mod64_recip(u64 x, unsigned d, double inv) {
  const u64 h = mod_recip(x >> 32, d, inv);
  return (unsigned)mod_recip((h << 32) | (x & 0xFFFFFFFFull), d, inv);
}

// ---- the slots' registers -------------------------------------------------------
typedef float f32x32 __attribute__((ext_vector_type(32)));
// Two register layouts of the slots, the same values: SlotsV keeps each tile
// as one 32-float vector, so that a slot (wave-uniform index) is read and
// written by indexed register moves without branches; SlotsQ keeps the
// MFMA quads apart and reaches a slot through a switch.  The instance
// without masters and stochastic rounding uses SlotsV; the full instance,
// whose other live registers leave no room for the tile vectors' indexed
// writes, uses SlotsQ.
struct SlotsV {
  f32x32 T[3];  // T[rt][4 t + r]: element 8m + t of slot 16 rt + 4 g + r
};
struct SlotsQ {
  f32x4 R[3][NT];  // R[rt][t][r]: element 8m + t of slot 16 rt + 4 g + r
};
__device__ __forceinline__ float sget(const SlotsV& S, int rt, int t, int r) { return S.T[rt][4 * t + r]; }
__device__ __forceinline__ float sget(const SlotsQ& S, int rt, int t, int r) { return S.R[rt][t][r]; }
__device__ __forceinline__ f32x4 sget4(const SlotsV& S, int rt, int t) {
  return (f32x4){S.T[rt][4 * t], S.T[rt][4 * t + 1], S.T[rt][4 * t + 2], S.T[rt][4 * t + 3]};
}
__device__ __forceinline__ f32x4 sget4(const SlotsQ& S, int rt, int t) { return S.R[rt][t]; }
__device__ __forceinline__ void sset4(SlotsV& S, int rt, int t, const f32x4& c) {
  S.T[rt][4 * t] = c[0]; S.T[rt][4 * t + 1] = c[1]; S.T[rt][4 * t + 2] = c[2]; S.T[rt][4 * t + 3] = c[3];
}
__device__ __forceinline__ void sset4(SlotsQ& S, int rt, int t, const f32x4& c) { S.R[rt][t] = c; }
__device__ __forceinline__ void sset(SlotsV& S, int rt, int t, int r, float x) { S.T[rt][4 * t + r] = x; }
__device__ __forceinline__ void szero(SlotsV& S) {
#pragma unroll
  for (int rt = 0; rt < 3; rt++) S.T[rt] = (f32x32)(0.f);
}
__device__ __forceinline__ void szero(SlotsQ& S) {
#pragma unroll
  for (int rt = 0; rt < 3; rt++)
#pragma unroll
    for (int t = 0; t < NT; t++) S.R[rt][t] = (f32x4){0.f, 0.f, 0.f, 0.f};
}

// Slot r (wave-uniform) of one tile vector takes e in the lanes where own
// holds (indexed moves under the lane mask).
__device__ __forceinline__ f32x32 tile_put(f32x32 v, int r, const float (&e)[8], bool own) {
  if (own) {
#pragma unroll
    for (int t = 0; t < NT; t++) v[4 * t + r] = e[t];
  }
  return v;
}

// Slot s (wave-uniform) takes the row whose elements 8m .. 8m+7 this lane
// holds in e; only the owner group's lanes change their registers.
__device__ __forceinline__ void slot_put_f(SlotsV& S, int s, const float (&e)[8], int g) {
  const bool own = g == ((s >> 2) & 3);
  const int rt = s >> 4, r = s & 3;
  S.T[0] = tile_put(S.T[0], r, e, own && rt == 0);
  S.T[1] = tile_put(S.T[1], r, e, own && rt == 1);
  S.T[2] = tile_put(S.T[2], r, e, own && rt == 2);
}
__device__ __forceinline__ void slot_put_f(SlotsQ& S, int s, const float (&e)[8], int g) {
  const bool own = g == ((s >> 2) & 3);
  switch (((s >> 4) << 2) | (s & 3)) {
#define DGER_PUT(RT, RR)                                             \
  case RT * 4 + RR:                                                  \
    _Pragma("unroll") for (int t = 0; t < NT; t++) S.R[RT][t][RR] = own ? e[t] : S.R[RT][t][RR]; \
    break;
    DGER_PUT(0, 0) DGER_PUT(0, 1) DGER_PUT(0, 2) DGER_PUT(0, 3)
    DGER_PUT(1, 0) DGER_PUT(1, 1) DGER_PUT(1, 2) DGER_PUT(1, 3)
    DGER_PUT(2, 0) DGER_PUT(2, 1) DGER_PUT(2, 2) DGER_PUT(2, 3)
#undef DGER_PUT
    default: break;
  }
}

// The same from a bf16 row (elements 8m .. 8m+7 in v).
template <class SL>
__device__ __forceinline__ void slot_put(SL& S, int s, const u32x4& v, int g) {
  float e[8];
  unpack8(v, e);
  slot_put_f(S, s, e, g);
}

// Elements 8m .. 8m+7 of slot s held by this lane (meaningful in the owner group).
__device__ __forceinline__ void slot_get(const SlotsV& S, int s, float (&e)[8]) {
  const int rt = s >> 4, r = s & 3;
  const f32x32 t0 = S.T[0], t1 = S.T[1], t2 = S.T[2];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const float a = t0[4 * t + r], b = t1[4 * t + r], c = t2[4 * t + r];
    e[t] = rt == 0 ? a : (rt == 1 ? b : c);
  }
}
__device__ __forceinline__ void slot_get(const SlotsQ& S, int s, float (&e)[8]) {
  switch (((s >> 4) << 2) | (s & 3)) {
#define DGER_GET(RT, RR)                                                     \
  case RT * 4 + RR:                                                          \
    _Pragma("unroll") for (int t = 0; t < NT; t++) e[t] = S.R[RT][t][RR];     \
    break;
    DGER_GET(0, 0) DGER_GET(0, 1) DGER_GET(0, 2) DGER_GET(0, 3)
    DGER_GET(1, 0) DGER_GET(1, 1) DGER_GET(1, 2) DGER_GET(1, 3)
    DGER_GET(2, 0) DGER_GET(2, 1) DGER_GET(2, 2) DGER_GET(2, 3)
#undef DGER_GET
    default:
      _Pragma("unroll") for (int t = 0; t < NT; t++) e[t] = 0.f;
      break;
  }
}

// Eight floats to bf16 with stochastic rounding (random bits from seed):
// v_cvt_sr_bf16_f32 adds the high 16 bits of its random operand to the
// float's bits and truncates (NaN gives the canonical NaN, denormals zero),
// so each pair takes the two halves of one xorshift state.
__device__ __forceinline__ u32x4 pack8_sr(const float (&e)[8], unsigned seed) {
  u32x4 w;
  unsigned h = hash32(seed);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    bf16x2 v = {};
    v = __builtin_amdgcn_cvt_sr_bf16_f32(v, e[2 * k], h, false);
    v = __builtin_amdgcn_cvt_sr_bf16_f32(v, e[2 * k + 1], h << 16, true);
    w[k] = __builtin_bit_cast(unsigned, v);
    h ^= h << 13; h ^= h >> 17; h ^= h << 5;
  }
  return w;
}

// Eight floats to bf16, round to nearest even.
__device__ __forceinline__ u32x4 pack8_rne(const float (&e)[8]) {
  u32x4 w;
#pragma unroll
  for (int k = 0; k < 4; k++) w[k] = pack_rne(e[2 * k], e[2 * k + 1]);
  return w;
}

// ---- FP8 rows ----------------------------------------------------------------
// A coded row is kH D e4m3 codes (element j at byte j) and its FP32 scale
// right after them, at a row stride of fp8_row_stride(kH D) bytes: element
// value = code * scale, scale = (largest magnitude of the row) / 448.
template <int kH>
__device__ __forceinline__ unsigned row8(int rk, int h) {
  return (unsigned)rk * (unsigned)fp8_row_stride(kH * D) + (unsigned)h * D;
}
template <int kH>
__device__ __forceinline__ unsigned scale8(int rk) {
  return (unsigned)rk * (unsigned)fp8_row_stride(kH * D) + (unsigned)(kH * D);
}
// Eight codes (bytes of c) times the row scale.
__device__ __forceinline__ void fp8_unpack8(const u32x2& c, float sc, float (&e)[8]) {
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const auto lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)c[k], false);
    const auto hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)c[k], true);
    e[4 * k] = lo[0] * sc;
    e[4 * k + 1] = lo[1] * sc;
    e[4 * k + 2] = hi[0] * sc;
    e[4 * k + 3] = hi[1] * sc;
  }
}
// Eight floats to codes for the scale 1 / inv (inv = 448 / row maximum, 0
// for an all-zero row), each rounded stochastically with its own random bits.
__device__ __forceinline__ u32x2 fp8_pack8_sr(const float (&e)[8], float inv, unsigned seed) {
  u32x2 w;
  unsigned h = hash32(seed);
  auto next = [&]() {
    const int r = (int)h;
    h ^= h << 13; h ^= h >> 17; h ^= h << 5;
    return r;
  };
  auto q = [&](float x) { return fminf(fmaxf(x * inv, -448.f), 448.f); };
#pragma unroll
  for (int k = 0; k < 2; k++) {
    int x = 0;
    x = __builtin_amdgcn_cvt_sr_fp8_f32(q(e[4 * k]), next(), x, 0);
    x = __builtin_amdgcn_cvt_sr_fp8_f32(q(e[4 * k + 1]), next(), x, 1);
    x = __builtin_amdgcn_cvt_sr_fp8_f32(q(e[4 * k + 2]), next(), x, 2);
    x = __builtin_amdgcn_cvt_sr_fp8_f32(q(e[4 * k + 3]), next(), x, 3);
    w[k] = (unsigned)x;
  }
  return w;
}
// Largest value over the 16 lanes of each lane group (one slot row or one
// output row of the MFMA layouts).
__device__ __forceinline__ float max16(float x) {
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) x = fmaxf(x, __shfl_xor(x, o, 64));
  return x;
}
__device__ __forceinline__ float absmax8(const float (&e)[8]) {
  float x = 0.f;
#pragma unroll
  for (int j = 0; j < 8; j++) x = fmaxf(x, fabsf(e[j]));
  return x;
}

// ---- one step's contractions ----------------------------------------------------
// LDS of the step: rc (bf16 slot rows, swizzled), cb (coefficients
// transposed for GEMM2: [output][owner group][16 slots in K order], CBS
// bytes per output).  cb overlays the start of rc: GEMM1 reads rc before the
// coefficients are written, and the next step's slot copy comes after GEMM2
// has read them.  Entries 12..15 of each group hold stale finite bf16 values
// that GEMM2 multiplies by zero.
struct StepLds {
  unsigned char* rc;
  unsigned char* cb;
};

// bf16 copy of every slot row into rc (GEMM1 reads the slots across lanes).
template <class SL>
__device__ __forceinline__ void ring_copy(const SL& S, unsigned char* rc, int g, int m) {
#pragma unroll
  for (int rt = 0; rt < 3; rt++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      u32x4 v;
#pragma unroll
      for (int k = 0; k < 4; k++) v[k] = pack_rne(sget(S, rt, 2 * k, r), sget(S, rt, 2 * k + 1, r));
      *reinterpret_cast<u32x4*>(rc + swz(16 * rt + 4 * g + r, m)) = v;
    }
}

// GEMM1 + sigmoid + coefficients.  ost: the step's output rows in bf16
// (GEMM1's A operand is row m, elements 32 ks + 8 g .. +8).
// On return C[rt][r] = coefficient of slot 16 rt + m and
// output 4 g + r (0 unless the slot is a context and the output exists), cb
// holds the coefficients for GEMM2.  meta4[r] = multiplicity of output 4g+r
// (0 past the step's outputs); os = number of outputs; cnt (lane s < 48) =
// the step's contexts in slot s.
// kH = 2 (rows of two halves, one per wavefront of the workgroup): the
// products over this half are written to xo (3 KB at rc + kMdOff, free
// until GEMM2's staged updates), the other half's are read from xp, and both
// wavefronts add them in the same order (own + other is exact-commutative),
// so they get the same F.  The second barrier keeps xo until the other
// wavefront has read it.
template <int kH = 1>
__device__ __forceinline__ void gemm1_coef(const unsigned char* ost, const StepLds& L, const int (&meta4)[4], int os,
                                           float alpha, const float* expt, int g, int m, int cnt, float (&C)[3][4],
                                           unsigned char* xo = nullptr, const unsigned char* xp = nullptr) {
  f32x4 f[3];
#pragma unroll
  for (int rt = 0; rt < 3; rt++) f[rt] = (f32x4){0.f, 0.f, 0.f, 0.f};
  // Every operand is read from LDS before the first MFMA (left to itself
  // the scheduler keeps only one or two reads ahead of the MFMAs, so that
  // each waits out an LDS round trip).
  bf16x8 av[KB1], bv[KB1][3];
#pragma unroll
  for (int ks = 0; ks < KB1; ks++) {
    av[ks] = *reinterpret_cast<const bf16x8*>(ost + swz(m, 4 * ks + g));
#pragma unroll
    for (int rt = 0; rt < 3; rt++) bv[ks][rt] = *reinterpret_cast<const bf16x8*>(L.rc + swz(16 * rt + m, 4 * ks + g));
  }
  __builtin_amdgcn_sched_group_barrier(0x100, 4 * KB1, 0);  // the 16 LDS reads
#pragma unroll
  for (int ks = 0; ks < KB1; ks++)
#pragma unroll
    for (int rt = 0; rt < 3; rt++) f[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[ks], bv[ks][rt], f[rt], 0, 0, 0);
  __builtin_amdgcn_sched_group_barrier(0x008, 3 * KB1, 0);  // then the 12 MFMAs
  if constexpr (kH > 1) {
    const int lane = 16 * g + m;
#pragma unroll
    for (int rt = 0; rt < 3; rt++) *reinterpret_cast<f32x4*>(xo + rt * 1024 + 16 * lane) = f[rt];
    __syncthreads();
#pragma unroll
    for (int rt = 0; rt < 3; rt++) f[rt] = f[rt] + *reinterpret_cast<const f32x4*>(xp + rt * 1024 + 16 * lane);
    __syncthreads();
  }
  // f[rt][r] = F[slot 16 rt + m][output 4 g + r].  The sigmoid is one
  // table read: the bin of x, or the entries 1 (x > 6) and 0 (x < -6) after
  // the table, so that no branch is needed.  All reads are issued before the
  // coefficients are stored (the stores could alias them).
  float tab[3][4], mu[3], me[4];
#pragma unroll
  for (int r = 0; r < 4; r++) me[r] = (float)meta4[r];
#pragma unroll
  for (int rt = 0; rt < 3; rt++) {
    mu[rt] = (float)__builtin_amdgcn_ds_bpermute(4 * (16 * rt + m), cnt);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const float x = f[rt][r];
      int idx = (int)((x + 6.0f) * 83.0f);
      idx = min(max(idx, 0), kExpTable - 1);
      idx = (x > 6.0f) ? kExpOne : idx;
      idx = (x < -6.0f) ? kExpZero : idx;
      tab[rt][r] = expt[idx];
    }
  }
#pragma unroll
  for (int rt = 0; rt < 3; rt++) {
    const bool second = 16 * rt + m >= kSentSlots;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = 4 * g + r;
      const float lab = (second ? (i == os - 1) : (i == 0)) ? 1.f : 0.f;
      // (label - sigmoid) * alpha * count: a count of 0 (no context in the
      // slot, or an output past os) gives exactly 0; the count (a product of
      // small integers) is exact in float
      C[rt][r] = (lab - tab[rt][r]) * (alpha * (me[r] * mu[rt]));
    }
  }
#pragma unroll
  for (int rt = 0; rt < 3; rt++)
#pragma unroll
    for (int r = 0; r < 4; r++)
      *reinterpret_cast<__bf16*>(L.cb + (4 * g + r) * CBS + (m >> 2) * 32 + (4 * rt + (m & 3)) * 2) = (__bf16)C[rt][r];
}

// GEMM2 for the element tiles t0 .. t0 + NT/2 - 1: md[t - t0][r] = the
// update of output row 4g + r, element 8m + t, from the coefficients in cb
// (a0, a1: GEMM2's A operand, read once per step) and the slots' values
// before this step's update.
template <class SL>
__device__ __forceinline__ void gemm2_half(const SL& S, const bf16x8& a0, const bf16x8& a1, int t0,
                                           f32x4 (&md)[NT / 2]) {
#pragma unroll
  for (int t = t0; t < t0 + NT / 2; t++) {
    u32x4 b0, b1;
    b0[0] = pack_rne(sget(S, 0, t, 0), sget(S, 0, t, 1));
    b0[1] = pack_rne(sget(S, 0, t, 2), sget(S, 0, t, 3));
    b0[2] = pack_rne(sget(S, 1, t, 0), sget(S, 1, t, 1));
    b0[3] = pack_rne(sget(S, 1, t, 2), sget(S, 1, t, 3));
    b1[0] = pack_rne(sget(S, 2, t, 0), sget(S, 2, t, 1));
    b1[1] = pack_rne(sget(S, 2, t, 2), sget(S, 2, t, 3));
    b1[2] = 0u;
    b1[3] = 0u;
    f32x4 c = (f32x4){0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, __builtin_bit_cast(bf16x8, b0), c, 0, 0, 0);
    md[t - t0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, __builtin_bit_cast(bf16x8, b1), c, 0, 0, 0);
  }
}

// GEMM3: every slot row += sum over outputs of C[slot][i] O[i] (zero for
// slots that are no context).  The B operand comes from the output rows in
// ost: rows 4g .. 4g+3, elements 8m .. 8m+7.
template <class SL>
__device__ __forceinline__ void gemm3(SL& S, const float (&C)[3][4], const unsigned char* ost, int g, int m) {
  u32x4 o3[4];
#pragma unroll
  for (int e = 0; e < 4; e++) o3[e] = *reinterpret_cast<const u32x4*>(ost + swz(4 * g + e, m));
  short4v a[3];
#pragma unroll
  for (int rt = 0; rt < 3; rt++) {
    const unsigned lo = pack_rne(C[rt][0], C[rt][1]), hi = pack_rne(C[rt][2], C[rt][3]);
    a[rt] = __builtin_bit_cast(short4v, (u64)lo | ((u64)hi << 32));
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const unsigned sel = (t & 1) ? 0x07060302u : 0x05040100u;
    const unsigned lo = __builtin_amdgcn_perm(o3[1][t >> 1], o3[0][t >> 1], sel);
    const unsigned hi = __builtin_amdgcn_perm(o3[3][t >> 1], o3[2][t >> 1], sel);
    const short4v b = __builtin_bit_cast(short4v, (u64)lo | ((u64)hi << 32));
#pragma unroll
    for (int rt = 0; rt < 3; rt++) sset4(S, rt, t, __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[rt], b, sget4(S, rt, t), 0, 0, 0));
  }
}

// Per-worker totals in LDS: contexts, output-context pairs, sentences, raw
// words; kStatXfer passes the learning rate's word total between the halves
// of a row.
enum { kStatCtx = 0, kStatOutCtx, kStatSentences, kStatRaw, kStatXfer, kStatCount };
// One lane adds v to a per-worker total (an LDS atomic without return).
__device__ __forceinline__ void stat_add(u64* stats, int k, u64 v) {
  __hip_atomic_fetch_add(stats + k, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// LDS carve-up of one worker of the window-slot trainer (16-byte aligned).
struct FastLds {
  unsigned char* rc;  // kSlots x RB: bf16 copy of the slots (swizzled); also the
                      // staging area of a sentence's first rows, and (after GEMM1
                      // has read it) the coefficients cb
  unsigned char* ost; // kOutRows x RB: the step's output rows (bf16, swizzled)
  unsigned char* cb;  // kOutRows x CBS, inside rc
  int* srank;         // kSlots: rank held by each slot
  int* plans;         // two plans: ranks [0, 16), multiplicities [16, 32)
  int* tok;           // 2 x Lmax ranks
  short* prv;         // 2 x Lmax: previous occurrence of each position's rank
  u64* stats;         // kStatCount per-worker totals (kept here, not in scalar registers)
};

__host__ __device__ inline size_t al16(size_t x) { return (x + 15) & ~(size_t)15; }

static_assert(kOutRows * CBS <= kSlots * RB, "the coefficients live inside the slot copy");
static_assert(kMdOff + kOutRows * MDS <= kSlots * RB, "the staged updates live inside the slot copy");
static_assert(kStageOff >= kSentSlots / 2 * RB && kStageOff + kOutRows * 4 * D <= kSlots * RB,
              "the staged master rows live inside the slot copy, past the first-window rows");
static_assert(2 * 4 * D <= kOutRows * RB, "two entering master rows fit in the output rows' space");
static_assert(kMdOff + 3 * 1024 <= kSlots * RB, "GEMM1's half-row products live inside the slot copy");
static_assert(kMdOff + kSlots * 4 <= kStageOff, "the FP8 write-backs' maxima sit before the staged master rows");

// One wavefront's LDS bytes.
__host__ __device__ inline size_t fast_layout(int Lmax, int K, size_t* off) {
  (void)K;
  size_t o = 0;
  off[0] = o; o = al16(o + (size_t)(kSlots + kOutRows) * RB);
  off[1] = o;
  off[2] = o; o = al16(o + (size_t)kSlots * 4);
  off[3] = o; o = al16(o + 2 * 2 * (size_t)kOutRows * 4);
  off[4] = o; o = al16(o + 2 * (size_t)Lmax * 4);
  off[5] = o; o = al16(o + 2 * (size_t)Lmax * 2);
  off[1] = o; o = al16(o + kStatCount * 8);
  return o;
}

__device__ inline FastLds fast_carve(unsigned char* base, int Lmax, int K) {
  size_t off[6];
  fast_layout(Lmax, K, off);
  FastLds s;
  s.rc = base + off[0];
  s.ost = base + off[0] + (size_t)kSlots * RB;
  s.cb = s.rc;
  s.srank = reinterpret_cast<int*>(base + off[2]);
  s.plans = reinterpret_cast<int*>(base + off[3]);
  s.tok = reinterpret_cast<int*>(base + off[4]);
  s.prv = reinterpret_cast<short*>(base + off[5]);
  s.stats = reinterpret_cast<u64*>(base + off[1]);
  return s;
}

// The sigmoid table in LDS: the kExpTable entries, then 1 and 0 (the values
// above and below the table's range, at kExpOne and kExpZero).
__host__ __device__ inline size_t exp_bytes() { return al16((size_t)(kExpTable + 2) * 4); }
__device__ inline void load_exp_table(float* expt, const float* src, int tid, int nthreads) {
  for (int x = tid; x < kExpTable; x += nthreads) expt[x] = src[x];
  if (tid == 0) { expt[kExpOne] = 1.f; expt[kExpZero] = 0.f; }
}

// Byte offset of half h of row rk in a matrix of kH halves per row: bf16
// rows (rowb) and FP32 master rows (rowm).
template <int kH>
__device__ __forceinline__ unsigned rowb(int rk, int h) { return (unsigned)rk * (kH * D * 2) + (unsigned)h * (D * 2); }
template <int kH>
__device__ __forceinline__ unsigned rowm(int rk, int h) { return (unsigned)rk * (kH * D * 4) + (unsigned)h * (D * 4); }

// Output rows of a plan for GEMM3's B operand (rows 4g .. 4g+3, elements
// 8m .. 8m+7 of this wavefront's half h); rows past os read row 0 of the plan
// (their results are not used).  A row with an FP32 master (rank < nm) is
// copied instead into the LDS staging area (row i at stage + i * 4D bytes,
// straight from memory), and out_rows_from_stage fills it in once the loads
// have been waited for.
template <int kH>
__device__ __forceinline__ void load_output_l3(const Mat& woh, const Mat& om, int nm, const int* plan, int os, int h,
                                               int g, int m, int lane, u32x4 (&o3)[4], unsigned char* stage) {
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int i = 4 * g + e;
    const int r3 = plan[i < os ? i : 0];
    if (nm == 0 || r3 >= nm) o3[e] = woh.load(rowb<kH>(r3, h) + 16 * m);
  }
  if (nm > 0) {
    // the master rows among the outputs (usually a few), lane i: output i
    const int pr = plan[lane < kOutRows ? lane : 0];
    for (u64 mm = __ballot(lane < os && pr < nm); mm; mm &= mm - 1) {
      const int i = (int)__builtin_ctzll(mm);
      const int rk = __builtin_amdgcn_readlane(pr, i);
      om.load_lds(rowm<kH>(rk, h) + 4 * lane, stage + i * (D * 4));
      om.load_lds(rowm<kH>(rk, h) + 256 + 4 * lane, stage + i * (D * 4) + 256);
    }
  }
}
__device__ __forceinline__ void out_rows_from_stage(u32x4 (&o3)[4], int nm, const int* plan, int os, int g, int m,
                                                    const unsigned char* stage) {
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int i = 4 * g + e;
    const int src = i < os ? i : 0;
    if (plan[src] < nm) {
      const f32x4 lo = *reinterpret_cast<const f32x4*>(stage + src * (D * 4) + 32 * m);
      const f32x4 hi = *reinterpret_cast<const f32x4*>(stage + src * (D * 4) + 32 * m + 16);
      o3[e] = (u32x4){pack_rne(lo[0], lo[1]), pack_rne(lo[2], lo[3]), pack_rne(hi[0], hi[1]), pack_rne(hi[2], hi[3])};
    }
  }
}
// Elements 8m .. 8m+7 of half h of FP32 master row rk.
template <int kH>
__device__ __forceinline__ void master_get(const Mat& M, int rk, int h, int m, float (&e)[8]) {
  const u32x4 lo = M.load(rowm<kH>(rk, h) + 32 * m), hi = M.load(rowm<kH>(rk, h) + 32 * m + 16);
#pragma unroll
  for (int k = 0; k < 4; k++) { e[k] = __uint_as_float(lo[k]); e[4 + k] = __uint_as_float(hi[k]); }
}
template <int kH>
__device__ __forceinline__ void master_put(const Mat& M, int rk, int h, int m, const float (&e)[8]) {
  u32x4 lo, hi;
#pragma unroll
  for (int k = 0; k < 4; k++) { lo[k] = __float_as_uint(e[k]); hi[k] = __float_as_uint(e[4 + k]); }
  M.store(rowm<kH>(rk, h) + 32 * m, lo);
  M.store(rowm<kH>(rk, h) + 32 * m + 16, hi);
}

// A window row (half h) goes back to Wih: its FP32 master (if it has one)
// exactly, or the bf16 row with stochastic (or nearest) rounding.
template <int kH>
__device__ __forceinline__ void write_back(const Mat& wih, const Mat& wm, bool master, int rk, int h, int m,
                                           const float (&e)[8], int round_nearest, unsigned seed) {
  if (master)
    master_put<kH>(wm, rk, h, m, e);
  else
    wih.store(rowb<kH>(rk, h) + 16 * m, round_nearest ? pack8_rne(e) : pack8_sr(e, seed));
}

// kFull: the instance with FP32 master rows, stochastic rounding of the Woh
// updates (a.n_master, a.round_nearest) and minibatches smaller than a step
// (a.minibatch) available; the other instance has none of them and none of
// their code.  Both share negative groups over 2^a.share_shift steps.
// WPB wavefronts per workgroup; kH halves per row (kH = 2: the workgroup is
// one worker, its two wavefronts hold halves 0 and 1 of every row).  kFp8:
// the FP8 form, in the full instance: Wih rows are e4m3 codes with a per-row
// FP32 scale, decoded into the slots when read and quantized (stochastic
// rounding, the row's largest magnitude over both halves) when written back;
// Woh rows stay bf16 and take stochastically rounded atomic adds, and the top
// ranks of both have FP32 masters.  (Coded Woh rows would take a whole-row
// read-modify-write at every step they appear in, so that every update would
// requantize the row and concurrent workers would overwrite each other's.)
template <int WPB, bool kFull, int kH, bool kFp8>
__global__ __launch_bounds__(64 * WPB, 2) void k_train_fast(TrainArgs a) {
  static_assert(kH == 1 || WPB == kH, "a row split into halves takes one workgroup per worker");
  static_assert(!kFp8 || kFull, "the FP8 form runs in the full instance");
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, m = lane & 15;
  // The worker index is wave-uniform; readfirstlane lets the compiler keep it
  // (and the slice, LCG state and LDS bases derived from it) in scalar registers.
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int h = kH > 1 ? wave % kH : 0;  // this wavefront's half of every row
  const int wk = blockIdx.x * a.wpb + wave / kH;
  const int W = a.W, K = a.K, Lmax = a.Lmax;
  float* const expt = reinterpret_cast<float*>(smem);
  load_exp_table(expt, a.exp_table, threadIdx.x, blockDim.x);
  __syncthreads();
  if (wk >= a.workers) return;
  size_t off_unused[6];
  const size_t wave_lds = fast_layout(Lmax, K, off_unused);
  const FastLds s = fast_carve(smem + exp_bytes() + (size_t)wave * wave_lds, Lmax, K);
  // kH = 2: the other half's GEMM1 products and half 0's word total
  const FastLds s_other = fast_carve(smem + exp_bytes() + (size_t)(wave ^ (kH - 1)) * wave_lds, Lmax, K);
  u64* const xfer = (h == 0 ? s.stats : s_other.stats) + kStatXfer;
  // (the FP8 form codes Wih only; Woh stays bf16)
  const size_t wih_bytes = kFp8 ? (size_t)a.V * fp8_row_stride(kH * D) : (size_t)a.V * kH * D * 2;
  const Mat wih(a.wih, (unsigned)wih_bytes), woh(a.woh, (unsigned)((size_t)a.V * kH * D * 2));
  // FP32 master rows of the hottest ranks (rank < nmaster).
  const int nmaster = kFull ? a.n_master : 0;
  const int round_nearest = kFull ? a.round_nearest : 1;
  auto is_master = [&](int rk) { return kFull && rk < nmaster; };
  // Contexts per minibatch (0: a step's contexts at once) and negative
  // sharing: steps n with equal n >> hs draw from one group.
  const int mb = kFull ? a.minibatch : 0;
  const int hs = a.share_shift;
  // Neither masters nor stochastic rounding of the Woh updates: the bf16
  // paths below take their shorter forms.
  const bool plain = !kFull || (nmaster == 0 && round_nearest);
  const Mat wm(a.wmaster, (unsigned)((size_t)nmaster * kH * D * 4)),
      om(a.omaster, (unsigned)((size_t)nmaster * kH * D * 4));
  // The shared negative batch of the worker: K x Lmax draws in device memory
  // (one copy per wavefront, so that the halves share no device memory).
  int* const negw = reinterpret_cast<int*>(a.fscratch) + (size_t)(wk * kH + h) * fast_scratch_ints(Lmax, K);

  // Slot copies and coefficients start at zero (unused slots and the K
  // padding of GEMM2 must multiply as finite zeros).
  for (int x = lane; x < ((kSlots + kOutRows) * RB) / 16; x += 64) reinterpret_cast<u32x4*>(s.rc)[x] = (u32x4){0u, 0u, 0u, 0u};
  if (lane < kSlots) s.srank[lane] = 0;
  if (lane < kStatCount) s.stats[lane] = 0;
  if (lane < 2 * 2 * kOutRows) s.plans[lane] = 0;
  typename std::conditional<kFull, SlotsQ, SlotsV>::type S;
  szero(S);
  wave_sync();

  // Corpus slice and seed of compute thread wk+1 (dsgl.hpp:791, :811-821).
  const u64 N = a.n_tokens;
  const u64 nparts = (u64)a.workers;
  const u64 per = N / nparts;
  const u64 start = per * (u64)wk;
  u64 cnt_tok = per;
  if (N % nparts > 0 && wk == a.workers - 1) cnt_tok = N - (nparts - 1) * per;
  const u64 end = start + cnt_tok;
  u64 rng = (u64)wk;

  // Lane l holds the (l+1)-step LCG map s -> jA*s + jC.
  u64 jA = 1, jC = 0;
  for (int i = 0; i <= lane; i++) { jA = jA * kLcgMul; jC = jC * kLcgMul + kLcgAdd; }

  const float a0 = a.alpha0;
  float alpha = a0;
  int iters_left = a.iter;
  u64 wc = 0, last_wc = 0, cur = start;
  int len[2] = {0, 0}, pos[2] = {0, 0};
  unsigned freem[2] = {0xFFFFFFu, 0xFFFFFFu};  // free slots of each sentence (bit k: slot 24q + k)
  int done_sent = 0, refresh = 1, neg_count = 0, nbn = 1;
  const Div32 divW = make_div32((unsigned)W);
  // Plain instance: the window draws (one per step, consecutive LCG draws
  // between sentence loads) are taken 64 at a time, lane-parallel: lane l
  // holds the state after l + 1 draws from rng (wsv) and its draw (wbv);
  // wd of them are used.  Before anything else reads rng, put_back()
  // advances rng past the used ones.
  u64 wsv = 0;
  int wbv = 0, wd = 0;
  bool wpre = false;
  auto put_back = [&]() {
    if (wpre) {
      if (wd > 0) rng = readlane64(wsv, wd - 1);
      wpre = false;
    }
  };
  Div32 divN = make_div32(1u);
  u64 steps = 0;
  u64 n_ctx = 0, n_outctx = 0;  // contexts and output-context pairs trained
  // Lane s < 48 holds in bit j (j = 0 .. 2W) whether window position
  // p - W + j of slot s's sentence holds slot s's rank (p = the sentence's
  // current position).
  unsigned pm = 0;
  const int stride = record_stride(W, K);
  int32_t* const rec = (a.rec && h == 0) ? a.rec + (size_t)wk * (size_t)a.rec_cap * stride : nullptr;
  int plan = 0;
  bool planned = false;
  int pre_os = 0;
  u32x4 o3[4];              // output rows of the next step (GEMM3 layout), on their way to ost
  u32x4 in_row[2];          // bf16 rows entering the windows at a transition
  int pend[2] = {-1, -1};   // their slots, filled at the next step's start
  bool pend_hot[2] = {false, false};  // the row comes from an FP32 master (staged in ost)
  int in_rk[2] = {0, 0};    // ranks of those rows
  bool staged = false;      // FP32 master rows are on their way into LDS
  unsigned char* const ostage = s.rc + kStageOff;  // FP32 master output rows of the next step
  unsigned salt = hash32((unsigned)wk * 0x9E3779B9u + 0x6A09E667u + (unsigned)h * 0x3C6EF372u);
  int j0 = 0;      // first context of the step's current minibatch
  int b_step = 0;  // the step's window shrink, drawn at its first minibatch

  // Learning-rate decay, sentence loads, slice end, refresh: as the reference
  // loop (dsgl.hpp:886-1088).
  for (;;) {
    // Lane-derived values are recomputed in every iteration from the lane
    // index, counted from the full execution mask in volatile asm so that it
    // is neither hoisted out of the loop nor kept live across it.
    int lane_o;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane_o));
    const int lane = lane_o, g = lane_o >> 4, m = lane_o & 15;
    if (wc - last_wc > 10000) {
      const u64 diff = wc - last_wc;
      u64 old = 0;
      if (lane == 0 && h == 0) old = atomicAdd(&a.counters[kWordCountActual], diff);
      u64 total = readlane64(old, 0) + diff;
      if constexpr (kH > 1) {
        // both halves run this at the same point of the same index stream;
        // half 0 passes its total to the other
        if (lane == 0 && h == 0) *xfer = total;
        __syncthreads();
        total = *xfer;
        __syncthreads();
      }
      last_wc = wc;
      alpha = a0 * (1 - (float)total / a.alpha_den);
      if (alpha < a0 * 0.0001f) alpha = a0 * 0.0001f;
    }

#pragma unroll
    for (int q = 0; q < 2; q++) {
      int* const tq = s.tok + q * Lmax;
      for (;;) {
        if (len[q] != 0) break;
        if (cur >= end) break;
        if constexpr (!kFull) put_back();
        int L = 0;
        for (;;) {
          const u64 idx = cur + (u64)lane;
          const int id = (idx < end) ? a.corpus[idx] : -1;
          const bool term = (id == -1) || (idx + 1 >= end);
          const u64 tm = __ballot(term);
          const int tl = tm ? (__ffsll((long long)tm) - 1) : 64;
          const bool regular = lane < tl;
          int rk = 0;
          bool keep = regular;
          if (regular) {
            rk = a.id_rank[id];
            if (a.sample_on) {
              const u64 sr = jA * rng + jC;
              keep = !(a.keep_thr[rk] < (float)(unsigned)(sr & 0xFFFFu) / 65536.f);
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
          if (keep) tq[L + __popcll(km & lanemask_lt(lane))] = rk;
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
        // Slots of the first window [0, W]: the k-th distinct rank takes slot 24q + k.
        const int nw = min(W, L - 1) + 1;
        // lane p < nw: first position of p's rank
        const int first = occurrences_near(tq, s.prv + q * Lmax, L, nw, 2 * W + 1, lane);
        wave_sync();
        const bool isnew = lane < nw && s.prv[q * Lmax + lane] < 0;
        const u64 newmask = __ballot(isnew);
        const int nnew = __popcll(newmask);
        int myslot = 0;
        if (lane < nw) myslot = kSentSlots * q + __popcll(newmask & lanemask_lt(first));
        wave_sync();
        pm = (lane >= kSentSlots * q && lane < kSentSlots * (q + 1)) ? 0u : pm;
        for (int p = 0; p < nw; p++) {
          const int sl = __builtin_amdgcn_readlane(myslot, p);
          pm |= (lane == sl) ? (1u << (p + W)) : 0u;
        }
        if (isnew) s.srank[myslot] = tq[lane];
        freem[q] = 0xFFFFFFu & ~((1u << nnew) - 1u);
        wave_sync();
        // Read the new slots' rows into LDS (the bf16 slot copy is rewritten
        // at the next step, so it serves as the staging area), then move
        // them into their slots.
        // (the plain instance stages them in ost, which holds nothing
        // between steps, and also writes them into the slot copy)
        unsigned char* const stg = kFull ? s.rc : s.ost;
        const int krk = (lane < nnew) ? s.srank[kSentSlots * q + lane] : 0;  // rank of new row `lane`
        for (int k = 0; k < nnew; k++) {
          const int rk = __builtin_amdgcn_readlane(krk, k);
          // FP8: codes in the first D bytes, the scale after them
          wih.load_lds(kFp8 ? (lane < D / 4 ? row8<kH>(rk, h) + 4 * lane : scale8<kH>(rk)) : rowb<kH>(rk, h) + 4 * lane,
                       stg + k * (D * 2));
        }
        wait_vm_all();
        if constexpr (!kFull) {
          // All new rows at once: each lane fills those of its slot
          // registers that are new slots of this sentence (fixed registers,
          // a select per element), and the slot copy takes the rows four
          // per store.
#pragma unroll
          for (int rt = 0; rt < 3; rt++) {
            if ((q == 0 && rt == 2) || (q == 1 && rt == 0)) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
              const int k = 16 * rt + 4 * g + r - kSentSlots * q;
              const bool mine = k >= 0 && k < nnew;
              const u32x4 v = *reinterpret_cast<const u32x4*>(stg + (mine ? k : 0) * (D * 2) + 16 * m);
              float e[8];
              unpack8(v, e);
#pragma unroll
              for (int t = 0; t < NT; t++) sset(S, rt, t, r, mine ? e[t] : sget(S, rt, t, r));
            }
          }
          for (int k = g; k < nnew; k += 4)
            *reinterpret_cast<u32x4*>(s.rc + swz(kSentSlots * q + k, m)) =
                *reinterpret_cast<const u32x4*>(stg + k * (D * 2) + 16 * m);
        } else
        for (int k = 0; k < nnew; k++) {
          const int rk = __builtin_amdgcn_readfirstlane(s.srank[kSentSlots * q + k]);
          if (is_master(rk)) {
            float e[8];
            master_get<kH>(wm, rk, h, m, e);
            slot_put_f(S, kSentSlots * q + k, e, g);
          } else if constexpr (kFp8) {
            float e[8];
            fp8_unpack8(*reinterpret_cast<const u32x2*>(stg + k * (D * 2) + 8 * m),
                        *reinterpret_cast<const float*>(stg + k * (D * 2) + D), e);
            slot_put_f(S, kSentSlots * q + k, e, g);
          } else {
            slot_put(S, kSentSlots * q + k, *reinterpret_cast<const u32x4*>(stg + k * (D * 2) + 16 * m), g);
          }
        }
        wave_sync();
        planned = false;
      }
    }

    // End of the slice: in-flight sentences are dropped (dsgl.hpp:1031-1050).
    if (cur >= end) {
      if (lane == 0 && h == 0) atomicAdd(&a.counters[kWordCountActual], wc - last_wc);
      if (lane == 0) stat_add(s.stats, kStatRaw, wc);
      if (--iters_left == 0) break;
      wc = 0; cur = start; last_wc = 0;
      len[0] = 0; len[1] = 0;
      freem[0] = 0xFFFFFFu; freem[1] = 0xFFFFFFu;
      pend[0] = -1; pend[1] = -1;
      planned = false;
      continue;
    }

    // Shared negative batch (dsgl.hpp:1054-1080): K * max(L1, L2) draws, or
    // K per 2^hs steps of the longer sentence.
    if (refresh) {
      if constexpr (!kFull) put_back();
      nbn = (max(len[0], len[1]) + (1 << hs) - 1) >> hs;
      divN = make_div32((unsigned)nbn);
      const int NN = nbn * K;
      const unsigned ts = (unsigned)a.table_size, vm1 = (unsigned)(a.V - 1);
      const double inv_ts = 1.0 / (double)ts, inv_vm1 = 1.0 / (double)vm1;
      const u64 A64 = readlane64(jA, 63), C64 = readlane64(jC, 63);
      const u64 r0 = rng;
      u64 cs = r0;
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
          v[c] = (n < NN) ? a.table[mod_recip(sr[c] >> 16, ts, inv_ts)] : 1;
        }
#pragma unroll
        for (int c = 0; c < 8; c++) {
          const int n = base + 64 * c + lane;
          if (n < NN) negw[n] = v[c] ? v[c] : (int)(mod64_recip(sr[c], vm1, inv_vm1) + 1);
        }
      }
      {
        u64 f = r0;
        int left = NN;
        while (left >= 64) { f = A64 * f + C64; left -= 64; }
        if (left > 0) f = readlane64(jA, left - 1) * f + readlane64(jC, left - 1);
        rng = f;
      }
      refresh = 0;
      planned = false;
      wave_sync();
    }

    // ---- one step (dsgl.hpp:1089-1368) ----
    const int p0 = pos[0], p1 = pos[1];
    // The step's scalar inputs in two gathers (one wait), then read lane by
    // lane: gr = ranks at p, p+1, p-W, p+W+1 (lane 2k+q: sentence q, k-th of
    // these), gs = previous occurrence of p+W+1 (lane q).  Positions are
    // clamped into the sentence; the code below reads only those that exist.
    int gr, gs;
    {
      const int q = lane & 1, k = (lane >> 1) & 3;
      const int pq = q ? p1 : p0, last_pos = max(len[q], 1) - 1;
      const int at = k == 0 ? pq : (k == 1 ? pq + 1 : (k == 2 ? pq - W : pq + W + 1));
      gr = s.tok[q * Lmax + min(max(at, 0), last_pos)];
      gs = s.prv[q * Lmax + min(pq + W + 1, last_pos)];
    }
    const int t1 = __builtin_amdgcn_readlane(gr, 0);
    const int t2 = __builtin_amdgcn_readlane(gr, 1);
    int b;
    if constexpr (!kFull) {
      if (!wpre || wd == 64) {
        if (wpre) rng = readlane64(wsv, 63);
        wsv = jA * rng + jC;
        wbv = (int)mod64_small(wsv, divW);
        wd = 0;
        wpre = true;
      }
      b = __builtin_amdgcn_readlane(wbv, wd);
      wd++;
    } else if (j0 == 0) {
      rng = rng * kLcgMul + kLcgAdd;
      b = (int)mod64_small(rng, divW);
      b_step = b;
    } else {
      b = b_step;
    }
    const int hw = W - b;
    const int lo0 = max(0, p0 - hw), hi0 = min(len[0] - 1, p0 + hw);
    const int lo1 = max(0, p1 - hw), hi1 = min(len[1] - 1, p1 + hw);
    const int n0 = (p0 - lo0) + (hi0 - p0);
    const int n1 = (p1 - lo1) + (hi1 - p1);
    const int nin = n0 + n1, fin = n0;
    // This minibatch: contexts [jb, je) of the step's list.
    const int jb = (kFull && mb > 0) ? j0 : 0;
    const int je = (kFull && mb > 0) ? min(nin, j0 + mb) : nin;
    const bool last = je >= nin;

    // Positions entering the windows at this step's transition.
    int in_slot_src[2];  // position whose slot the entering position shares, or -1 (new row)
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int pq = q ? p1 : p0, Lq = len[q];
      const int x = pq - W, y = pq + W + 1;
      in_slot_src[q] = -2;  // nothing enters
      if (pq + 1 < Lq && y < Lq) {
        const int pv = __builtin_amdgcn_readlane(gs, q);
        in_slot_src[q] = (pv >= 0 && pv >= x) ? pv : -1;
      }
    }

    // Context multiplicity per slot: the slot's window positions within the
    // step's window, minus the centre (positions outside the sentence are
    // never set).  A minibatch takes contexts [jb, je) of the step's list
    // (sentence 0's window positions in order, then sentence 1's): for each
    // sentence a contiguous run of its set positions, so one mask each.
    int my_ctx_rank = -1;
    int cnt;  // lane s < 48: the step's (minibatch's) contexts in slot s
    {
      const unsigned ctxm = ((((2u << (2 * hw)) - 1u) << (W - hw)) & ~(1u << W));
      unsigned m0 = ctxm, m1 = ctxm;
      if (kFull && mb > 0) {
        // contexts [a, b) of one sentence's list: the j-th is window bit
        // W - left + j below the centre, W + 1 + j - left above it
        auto run_mask = [&](int a_, int b_, int left) -> unsigned {
          if (a_ >= b_) return 0u;
          const int lo_bit = a_ < left ? W - left + a_ : W + 1 + (a_ - left);
          const int hi_bit = (b_ - 1) < left ? W - left + (b_ - 1) : W + 1 + (b_ - 1 - left);
          return ((2u << hi_bit) - 1u) & ~((1u << lo_bit) - 1u) & ~(1u << W);
        };
        m0 = run_mask(max(jb, 0), min(je, n0), p0 - lo0);
        m1 = run_mask(max(jb - n0, 0), min(je - n0, n1), p1 - lo1);
      }
      cnt = __builtin_popcount(pm & (lane < kSentSlots ? m0 : m1));  // 0 in lanes >= 48 (pm = 0 there)
      if (rec && lane < nin) {
        const int q = (lane < n0) ? 0 : 1;
        const int jj = q ? lane - n0 : lane;
        const int lo = q ? lo1 : lo0, pp = q ? p1 : p0;
        const int c = (jj < pp - lo) ? lo + jj : pp + 1 + (jj - (pp - lo));
        my_ctx_rank = s.tok[q * Lmax + c];
      }
    }

    // Output list (dsgl.hpp:1136-1158): target, distinct negatives of the
    // group with multiplicities (a negative equal to the target is dropped),
    // second target; rows past the count have multiplicity 0.
    // v: negative k2 of the group in lane k2 < K.
    if (!kFull || j0 == 0) neg_count++;
    const int grp = (int)mod32((unsigned)neg_count >> hs, divN);
    // The next step's group, read now for the planning after GEMM2.
    // (every lane loads, from a clamped index, so that the load is not
    // waited for here; lanes >= K are masked at the use)
    const int vnext = negw[(int)mod32((unsigned)(neg_count + 1) >> hs, divN) * K + min(lane, max(K - 1, 0))];
    auto make_outputs = [&](int pl, int ta, int tb, int v) -> int {
      int* const pr = s.plans + pl * 2 * kOutRows;
      const bool live = lane < K;
      int first = lane, mult = 0;
      for (int k2 = 0; k2 < K; k2++) {
        const int w2 = __builtin_amdgcn_readlane(v, k2);
        mult += (w2 == v) ? 1 : 0;
        first = (w2 == v && k2 < first) ? k2 : first;
      }
      const bool keep = live && first == lane && v != ta;
      const u64 km = __ballot(keep);
      const int o = 1 + __popcll(km & lanemask_lt(lane));
      const int n = 2 + __popcll(km);
      if (lane < kOutRows) pr[kOutRows + lane] = 0;
      wave_sync();
      if (keep) { pr[o] = v; pr[kOutRows + o] = mult; }
      if (lane == 0) { pr[0] = ta; pr[kOutRows] = 1; pr[n - 1] = tb; pr[kOutRows + n - 1] = 1; }
      wave_sync();
      return __builtin_amdgcn_readfirstlane(n);  // the same in every lane
    };
    int os;
    if (planned) {
      os = pre_os;
    } else {
      os = make_outputs(plan, t1, t2, (lane < K) ? negw[grp * K + lane] : -1);
      load_output_l3<kH>(woh, om, nmaster, s.plans + plan * 2 * kOutRows, os, h, g, m, lane, o3, ostage);
      if (nmaster > 0) {
        wait_vm_all();
        out_rows_from_stage(o3, nmaster, s.plans + plan * 2 * kOutRows, os, g, m, ostage);
      }
    }
    const int* const cplan = s.plans + plan * 2 * kOutRows;

    if (rec && (int64_t)steps < a.rec_cap && (!kFull || j0 == 0)) {
      int32_t* r = rec + (size_t)steps * stride;
      for (int z = lane; z < stride; z += 64) r[z] = -1;
      wave_sync();
      if (lane == 0) {
        r[0] = t1; r[1] = t2; r[2] = b; r[3] = fin; r[4] = nin; r[5] = os;
        r[6] = __float_as_int(alpha);
        r[7] = grp;
      }
      if (lane < nin) r[8 + lane] = my_ctx_rank;
      if (lane < os) { r[8 + 4 * W + lane] = cplan[lane]; r[8 + 4 * W + K + 2 + lane] = cplan[kOutRows + lane]; }
    }

    // Contractions.
    // Rows that entered the windows at the last transition take their slots
    // now (their loads, issued during that step, are waited for here
    // together with this step's output rows).
    if (staged) {
      wait_vm_all();
      staged = false;
    }
#pragma unroll
    for (int q = 0; q < 2; q++)
      if (pend[q] >= 0) {
        if (pend_hot[q]) {
          const f32x4 lo = *reinterpret_cast<const f32x4*>(s.ost + q * (D * 4) + 32 * m);
          const f32x4 hi = *reinterpret_cast<const f32x4*>(s.ost + q * (D * 4) + 32 * m + 16);
          const float e[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
          slot_put_f(S, pend[q], e, g);
        } else if constexpr (kFp8) {
          float e[8];
          fp8_unpack8((u32x2){in_row[q][0], in_row[q][1]}, __uint_as_float(in_row[q][2]), e);
          slot_put_f(S, pend[q], e, g);
        } else {
          slot_put(S, pend[q], in_row[q], g);
          if (!kFull && g == 0) *reinterpret_cast<u32x4*>(s.rc + swz(pend[q], m)) = in_row[q];
        }
        pend[q] = -1;
      }
    if (planned && nmaster > 0) out_rows_from_stage(o3, nmaster, cplan, os, g, m, ostage);
    // The full instance copies the slots here; the plain one keeps the copy
    // made at the end of the last step (plus the rows put since).
    if (kFull) ring_copy(S, s.rc, g, m);
    // the output rows, bf16, for GEMM1's A operand and GEMM3's B operand
#pragma unroll
    for (int e = 0; e < 4; e++) *reinterpret_cast<u32x4*>(s.ost + swz(4 * g + e, m)) = o3[e];
    int meta4[4];
    {
      const int4 mm = *reinterpret_cast<const int4*>(cplan + kOutRows + 4 * g);
      meta4[0] = mm.x; meta4[1] = mm.y; meta4[2] = mm.z; meta4[3] = mm.w;
    }
    wave_sync();
    const StepLds L{s.rc, s.cb};
    float C[3][4];
    gemm1_coef<kH>(s.ost, L, meta4, os, alpha, expt, g, m, cnt, C, s.rc + kMdOff, s_other.rc + kMdOff);
    wave_sync();
    {
      // md in FP32, staged row by row in LDS (rc is free until the next
      // step's slot copy); then Woh row of output i += md row i, one row at a
      // time: an FP32 master row takes FP32 atomic adds (lanes l and 64 + l),
      // a bf16 row bf16 atomic adds of element pairs (lane l: elements 2l,
      // 2l+1), each instruction covering whole 64-byte segments.
      unsigned char* const mdl = s.rc + kMdOff;
      const bf16x8 c0 = *reinterpret_cast<const bf16x8*>(s.cb + m * CBS + g * 32);
      const bf16x8 c1 = *reinterpret_cast<const bf16x8*>(s.cb + m * CBS + g * 32 + 16);
#pragma unroll
      for (int h = 0; h < 2; h++) {
        f32x4 md[NT / 2];
        gemm2_half(S, c0, c1, h * (NT / 2), md);
        if (plain) {
#pragma unroll
          for (int r = 0; r < 4; r++) {
            u32x2 w;
            w[0] = pack_rne(md[0][r], md[1][r]);
            w[1] = pack_rne(md[2][r], md[3][r]);
            *reinterpret_cast<u32x2*>(mdl + (4 * g + r) * MDS + 16 * m + 8 * h) = w;
          }
        } else {
#pragma unroll
          for (int r = 0; r < 4; r++)
            *reinterpret_cast<f32x4*>(mdl + (4 * g + r) * MDS + 32 * m + 16 * h) =
                (f32x4){md[0][r], md[1][r], md[2][r], md[3][r]};
        }
        // keep the halves apart (their registers would not fit together)
        __builtin_amdgcn_sched_barrier(0);
      }
      const int orank = (lane < kOutRows) ? cplan[lane] : 0;
      // random bits of this step's stochastic roundings (rotated per row)
      const unsigned rseed = plain ? 0u : hash32(salt ^ 0x85EBCA6Bu ^ ((unsigned)lane * 0x9E3779B9u));
      wave_sync();
      typedef short __attribute__((ext_vector_type(2))) short2v;
      __attribute__((address_space(1))) short2v* const wb = (__attribute__((address_space(1))) short2v*)(a.woh);
      __attribute__((address_space(1))) float* const ob = (__attribute__((address_space(1))) float*)(a.omaster);
      if (plain) {
        // no masters, nearest rounding: bf16 pairs, all rows read first
        // (at most 8 output rows unless negative > 6: half the reads and tests)
        auto add_rows = [&](auto nrows) {
          constexpr int NR = decltype(nrows)::value;
          unsigned v[NR];
          // this lane's column pointer; a row's 64-bit offset is added per row
          // (one scalar shift and one vector add, no 32-bit wrap above 2^26 ranks)
          __attribute__((address_space(1))) short2v* const lp = wb + (h * (D / 2) + lane);
#pragma unroll
          for (int i = 0; i < NR; i++) v[i] = *reinterpret_cast<const unsigned*>(mdl + i * MDS + 4 * lane);
#pragma unroll
          for (int i = 0; i < NR; i++)
            if (i < os) {
              const size_t rk = (unsigned)__builtin_amdgcn_readlane(orank, i);
              (void)__builtin_amdgcn_global_atomic_fadd_v2bf16(lp + rk * (kH * D / 2), __builtin_bit_cast(short2v, v[i]));
            }
        };
        if (os <= 8)
          add_rows(std::integral_constant<int, 8>());
        else
          add_rows(std::integral_constant<int, kOutRows>());
      } else
      for (int i = 0; i < os; i++) {
        const int rk = __builtin_amdgcn_readlane(orank, i);
        if (is_master(rk)) {
          const float v0 = *reinterpret_cast<const float*>(mdl + i * MDS + 4 * lane);
          const float v1 = *reinterpret_cast<const float*>(mdl + i * MDS + 256 + 4 * lane);
          (void)__builtin_amdgcn_global_atomic_fadd_f32(ob + (size_t)rk * (kH * D) + h * D + lane, v0);
          (void)__builtin_amdgcn_global_atomic_fadd_f32(ob + (size_t)rk * (kH * D) + h * D + 64 + lane, v1);
        } else {
          const float2 mv = *reinterpret_cast<const float2*>(mdl + i * MDS + 8 * lane);
          unsigned u;
          if (round_nearest) {
            u = pack_rne(mv.x, mv.y);
          } else {
            // Stochastic rounding of (row value read for this step + md):
            // the atomic adds the exact bf16 difference t - old, so that a
            // row nobody else changed meanwhile ends at t, an unbiased
            // rounding of old + md.
            const unsigned oldp = *reinterpret_cast<const unsigned*>(s.ost + swz(i, lane >> 2) + 4 * (lane & 3));
            const float o0 = __uint_as_float(oldp << 16), o1 = __uint_as_float(oldp & 0xFFFF0000u);
            const unsigned rb = __builtin_amdgcn_alignbit(rseed, rseed, (unsigned)(7 * i) & 31u);
            bf16x2 t = {};
            t = __builtin_amdgcn_cvt_sr_bf16_f32(t, o0 + mv.x, rb, false);
            t = __builtin_amdgcn_cvt_sr_bf16_f32(t, o1 + mv.y, rb << 16, true);
            const unsigned tb = __builtin_bit_cast(unsigned, t);
            u = pack_rne(__uint_as_float(tb << 16) - o0, __uint_as_float(tb & 0xFFFF0000u) - o1);
          }
          (void)__builtin_amdgcn_global_atomic_fadd_v2bf16(wb + (unsigned)rk * (kH * D / 2) + h * (D / 2) + lane,
                                                           __builtin_bit_cast(short2v, u));
        }
      }
    }

    // Wih rows entering the windows at the transition below; they take their
    // slots at the start of the next step (the loads are issued after the
    // atomics so that the contractions above do not hold them in registers).
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int y = (q ? p1 : p0) + W + 1;
      const int rk = (last && in_slot_src[q] == -1) ? __builtin_amdgcn_readlane(gr, 6 + q) : -1;
      pend_hot[q] = rk >= 0 && is_master(rk);  // staged into ost after GEMM3 below
      in_rk[q] = rk;
      if (!pend_hot[q]) {
        if constexpr (kFp8) {
          const u32x2 c = wih.load2(row8<kH>(rk >= 0 ? rk : 0, h) + 8 * m);
          in_row[q] = (u32x4){c[0], c[1], wih.load1(scale8<kH>(rk >= 0 ? rk : 0)), 0u};
        } else {
          in_row[q] = wih.load(rowb<kH>(rk >= 0 ? rk : 0, h) + 16 * m);
        }
      }
    }
    // Plan the next step's outputs and read its rows (after this step's
    // updates of the same rows), when neither sentence ends here.
    const bool cont = last && (p0 + 1 < len[0]) && (p1 + 1 < len[1]);
    int osn = 0;
    if (cont) {
      const int np = plan ^ 1;
      osn = make_outputs(np, __builtin_amdgcn_readlane(gr, 2), __builtin_amdgcn_readlane(gr, 3), (lane < K) ? vnext : -1);
      load_output_l3<kH>(woh, om, nmaster, s.plans + np * 2 * kOutRows, osn, h, g, m, lane, o3, ostage);
      staged = nmaster > 0;
    } else {
      // the next step plans afresh: the rows are dead (and need no register
      // across this step)
#pragma unroll
      for (int e = 0; e < 4; e++) o3[e] = (u32x4){0u, 0u, 0u, 0u};
    }
    gemm3(S, C, s.ost, g, m);
    if (!kFull) {
      // bf16 copy of the updated slots: the write-backs below and the next
      // step's GEMM1 read it (GEMM2's coefficients and the staged updates in
      // rc have been read by now; LDS accesses of a wave run in order)
      wave_sync();
      ring_copy(S, s.rc, g, m);
      wave_sync();
    }
    // entering rows with FP32 masters: into ost, which GEMM3 no longer needs
#pragma unroll
    for (int q = 0; q < 2; q++)
      if (pend_hot[q]) {
        wm.load_lds(rowm<kH>(in_rk[q], h) + 4 * lane, s.ost + q * (D * 4));
        wm.load_lds(rowm<kH>(in_rk[q], h) + 256 + 4 * lane, s.ost + q * (D * 4) + 256);
        staged = true;
      }

    // (wave-uniform totals in registers: no lane mask, no LDS)
    n_ctx += (u64)(je - jb);
    n_outctx += (u64)(os * (je - jb));
    if (kFull && !last) {
      // The next minibatch of this step: its output rows are read again
      // (after the updates above); the windows stay.
      j0 = je;
      planned = false;
      salt = hash32(salt + 0x9E3779B9u);
      wave_sync();
      continue;
    }
    j0 = 0;
    steps++;

    // Transition: a finished sentence writes its live slots back; a
    // continuing one slides its window by one position.
    const unsigned pm_old = pm;
    pm >>= 1;  // both windows move on by one position
    // FP8: the slots to write back are collected here and quantized after the
    // loop (a row's scale needs its largest magnitude over both halves); a
    // slot freed below keeps its registers until the next step's start, and
    // its rank in srank until the write-backs are done.
    u64 wbm = 0;                              // bit s: slot s goes back to Wih
    int ent_sl[2] = {-1, -1}, ent_rk[2] = {0, 0};  // entering rows' slots and ranks, recorded after
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int p = pos[q];
      const int Lq = len[q];
      pos[q] = p + 1;
      if (p + 1 >= Lq) {
        if (!kFull) {
          // from the slot copy, four rows per store (group g takes the g-th)
          for (unsigned live = ~freem[q] & 0xFFFFFFu; live;) {
            int sl4 = -1;
#pragma unroll
            for (int k = 0; k < 4; k++) {
              const int b = live ? __builtin_ctz(live) : -1;
              sl4 = (g == k && b >= 0) ? kSentSlots * q + b : sl4;
              live &= live - 1;
            }
            if (sl4 >= 0) {
              const int rk = s.srank[sl4];
              wih.store(rowb<kH>(rk, h) + 16 * m, *reinterpret_cast<const u32x4*>(s.rc + swz(sl4, m)));
            }
          }
        } else if (kFp8) {
          wbm |= (u64)(~freem[q] & 0xFFFFFFu) << (kSentSlots * q);
        } else
        for (unsigned live = ~freem[q] & 0xFFFFFFu; live; live &= live - 1) {
          const int sl = kSentSlots * q + __builtin_ctz(live);
          float e[8];
          slot_get(S, sl, e);
          const int rk = __builtin_amdgcn_readfirstlane(s.srank[sl]);
          if (g == ((sl >> 2) & 3))
            write_back<kH>(wih, wm, is_master(rk), rk, h, m, e, round_nearest,
                           salt ^ (unsigned)(steps * 64 + lane) ^ (unsigned)(sl << 24));
        }
        freem[q] = 0xFFFFFFu;
        len[q] = 0;
        pm = (lane >= kSentSlots * q && lane < kSentSlots * (q + 1)) ? 0u : pm;
        done_sent++;
        if (lane == 0) stat_add(s.stats, kStatSentences, 1);
        continue;
      }
      // x's slot: the one holding the position that just left (bit 0); a
      // shared entering position takes its previous occurrence's slot; x's
      // rank leaves if its slot holds no position any more.
      const int x = p - W;
      const bool inq = lane >= kSentSlots * q && lane < kSentSlots * (q + 1);
      int sx = 0;
      if (x >= 0) sx = (int)__builtin_ctzll(__ballot(inq && (pm_old & 1u)));
      if (in_slot_src[q] >= 0) {
        const int j = in_slot_src[q] - x - 1;  // bit of that occurrence after the slide
        const int sy = j < 0 ? sx : (int)__builtin_ctzll(__ballot(inq && ((pm >> j) & 1u)));
        pm |= (lane == sy) ? (1u << (2 * W)) : 0u;
      }
      if (x >= 0 && __builtin_amdgcn_readlane(pm, sx) == 0) {
        const int rkx = __builtin_amdgcn_readlane(gr, 4 + q);
        if (!kFull) {
          if (g == q) wih.store(rowb<kH>(rkx, h) + 16 * m, *reinterpret_cast<const u32x4*>(s.rc + swz(sx, m)));
        } else if (kFp8) {
          wbm |= 1ull << sx;
        } else {
          float e[8];
          slot_get(S, sx, e);
          if (g == ((sx >> 2) & 3))
            write_back<kH>(wih, wm, is_master(rkx), rkx, h, m, e, round_nearest,
                           salt ^ (unsigned)(steps * 64 + lane) ^ (unsigned)(sx << 24));
        }
        freem[q] |= 1u << (sx - kSentSlots * q);
      }
      if (in_slot_src[q] == -1) {
        const int k = __ffs((int)freem[q]) - 1;
        freem[q] &= ~(1u << k);
        const int sl = kSentSlots * q + k;
        pend[q] = sl;
        pm |= (lane == sl) ? (1u << (2 * W)) : 0u;
        const int rky = __builtin_amdgcn_readlane(gr, 6 + q);
        if (kFp8) {
          ent_sl[q] = sl;
          ent_rk[q] = rky;
        } else if (lane == 0) {
          s.srank[sl] = rky;
        }
      }
    }
    if constexpr (kFp8) {
      // Slot 16 rt + 4 g + r sits at fixed registers of lane group g, so the
      // write-backs go by register position: each of the 12 positions that
      // holds a slot to write back serves four slots at once, one per lane
      // group.  Two halves pass their largest magnitudes through xw (one
      // barrier per such position; a master row takes its FP32 values).  xw
      // is the slot copy past the coefficients, which the full instance
      // rewrites only at the next step's start (the staged master output rows
      // start past it); a last barrier keeps it until the other half has
      // read it.
      float* const xw = reinterpret_cast<float*>(s.rc + kMdOff);
      const float* const xw_other = reinterpret_cast<const float*>(s_other.rc + kMdOff);
      if (wbm != 0) {
#pragma unroll
        for (int rt = 0; rt < 3; rt++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            if ((wbm & (0x1111ull << (16 * rt + r))) == 0) continue;
            const int sl = 16 * rt + 4 * g + r;  // this lane group's slot
            const bool mine = (wbm >> sl) & 1u;
            float e[8];
#pragma unroll
            for (int t = 0; t < NT; t++) e[t] = sget(S, rt, t, r);
            const int rk = s.srank[sl];
            const bool hot = is_master(rk);
            float mx = max16(absmax8(e));
            if constexpr (kH > 1) {
              if (m == 0) xw[sl] = mx;
              __syncthreads();
              mx = fmaxf(mx, xw_other[sl]);
            }
            if (mine) {
              if (hot) {
                master_put<kH>(wm, rk, h, m, e);
              } else {
                const float inv = mx > 0.f ? 448.f / mx : 0.f;
                wih.store2(row8<kH>(rk, h) + 8 * m,
                           fp8_pack8_sr(e, inv, salt ^ (unsigned)(steps * 64 + lane) ^ (unsigned)(sl << 24)));
                if (m == 0) wih.store1(scale8<kH>(rk), __float_as_uint(mx > 0.f ? mx / 448.f : 0.f));
              }
            }
          }
        if constexpr (kH > 1) __syncthreads();
      }
#pragma unroll
      for (int q = 0; q < 2; q++)
        if (ent_sl[q] >= 0 && lane == 0) s.srank[ent_sl[q]] = ent_rk[q];
    }
    if (done_sent >= 2) {
      refresh = 1;
      done_sent = 0;
    }
    planned = cont;
    if (cont) { plan = __builtin_amdgcn_readfirstlane(plan ^ 1); pre_os = osn; }
    salt = hash32(salt + 0x9E3779B9u);
    wave_sync();
  }

  if (lane == 0 && h == 0) {
    a.worker_steps[wk] = steps;
    a.worker_alpha[wk] = alpha;
    atomicAdd(&a.counters[kSteps], steps);
    atomicAdd(&a.counters[kRawWords], s.stats[kStatRaw]);
    atomicAdd(&a.counters[kSentences], s.stats[kStatSentences]);
    atomicAdd(&a.counters[kContexts], n_ctx);
    atomicAdd(&a.counters[kOutCtx], n_outctx);
  }
}

// One step of the window-slot form on given rows (test-facing): O (os x kH D
// float), slot rows I (48 x kH D float, slots 0..23 first sentence), context
// count per slot, output multiplicities.  Writes the coefficients (os x 48),
// md (os x kH D) and the updated slot rows (48 x kH D), each through the same
// device code as the trainer; wavefront h of the workgroup takes half h of
// every row.
template <int kH>
__global__ __launch_bounds__(64 * kH) void k_step_fast(int os, const int32_t* meta, const int32_t* mult, float alpha,
                                                       const float* exp_table, const float* O, const float* I,
                                                       float* corr, float* md_out, float* I_out) {
  constexpr int kWaveLds = (kSlots + kOutRows) * RB + 2 * kSlots * 4 + 4096 + 64;
  __shared__ __align__(16) unsigned char lds_all[kH * kWaveLds];
  const int lane = threadIdx.x & 63, g = lane >> 4, m = lane & 15;
  const int h = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int DS = kH * D;  // row length of O, I, md, I_out
  unsigned char* const lds = lds_all + h * kWaveLds;
  unsigned char* rc = lds;
  unsigned char* ost = lds + kSlots * RB;
  unsigned char* cb = rc;
  int* ml = reinterpret_cast<int*>(ost + kOutRows * RB);
  int* plan = ml + kSlots;
  float* expt = reinterpret_cast<float*>(plan + kSlots);
  load_exp_table(expt, exp_table, lane, 64);
  if (lane < kSlots) ml[lane] = mult[lane];
  if (lane < kOutRows) { plan[lane] = lane; plan[kOutRows + lane] = lane < os ? meta[lane] : 0; }
  SlotsV S;
#pragma unroll
  for (int rt = 0; rt < 3; rt++)
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) S.T[rt][4 * (t) + (r)] = I[(size_t)(16 * rt + 4 * g + r) * DS + h * D + 8 * m + t];
  // Output rows as bf16 in LDS (rows past os repeat row 0).
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int i = 4 * g + e;
    const float* src = O + (size_t)(i < os ? i : 0) * DS + h * D + 8 * m;
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = pack_rne(src[2 * k], src[2 * k + 1]);
    *reinterpret_cast<u32x4*>(ost + swz(i, m)) = o;
  }
  __syncthreads();
  ring_copy(S, rc, g, m);
  int meta4[4];
#pragma unroll
  for (int r = 0; r < 4; r++) meta4[r] = plan[kOutRows + 4 * g + r];
  __syncthreads();
  const StepLds L{rc, cb};
  const int cnt = (lane < kSlots) ? ml[lane] : 0;
  float C[3][4];
  gemm1_coef<kH>(ost, L, meta4, os, alpha, expt, g, m, cnt, C, rc + kMdOff,
                 lds_all + (h ^ (kH - 1)) * kWaveLds + kMdOff);
  if (h == 0)
#pragma unroll
    for (int rt = 0; rt < 3; rt++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (4 * g + r < os) corr[(size_t)(4 * g + r) * kSlots + 16 * rt + m] = C[rt][r];
  __syncthreads();
  const bf16x8 c0 = *reinterpret_cast<const bf16x8*>(cb + m * CBS + g * 32);
  const bf16x8 c1 = *reinterpret_cast<const bf16x8*>(cb + m * CBS + g * 32 + 16);
#pragma unroll
  for (int hh = 0; hh < 2; hh++) {
    f32x4 md[NT / 2];
    gemm2_half(S, c0, c1, hh * (NT / 2), md);
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (4 * g + r < os)
#pragma unroll
        for (int t = 0; t < NT / 2; t++)
          md_out[(size_t)(4 * g + r) * DS + h * D + 8 * m + hh * (NT / 2) + t] = md[t][r];
  }
  gemm3(S, C, ost, g, m);
#pragma unroll
  for (int rt = 0; rt < 3; rt++)
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        I_out[(size_t)(16 * rt + 4 * g + r) * DS + h * D + 8 * m + t] = S.T[rt][4 * (t) + (r)];
}

}  // namespace

bool fast_supported(int D_, int K, int W, uint32_t V, int64_t table_size) {
  const int nh = fast_halves(D_);
  return nh > 0 && K + 2 <= kOutRows && 2 * W + 1 <= kSentSlots && W >= 1 && V >= 2 &&
         (uint64_t)V * (uint64_t)D_ * 2 < (1ull << 31) && table_size >= 1 && table_size < (1ll << 31);
}

size_t fast_lds_bytes(int Lmax, int K) {
  size_t off[6];
  return fast_layout(Lmax, K, off);
}

size_t fast_block_lds_bytes(int Lmax, int K, int wpb, int D_) {
  return exp_bytes() + (size_t)wpb * (size_t)fast_halves(D_) * fast_lds_bytes(Lmax, K);
}

int fast_workers_per_block(int Lmax, int K, size_t lds_per_block, int D_) {
  const int nh = fast_halves(D_);
  if (nh < 1) return 0;
  const size_t per = (size_t)nh * fast_lds_bytes(Lmax, K);
  if (lds_per_block < exp_bytes() + per) return 0;
  if (nh > 1) return 1;  // the workgroup's barriers pair the halves of one worker
  const size_t n = (lds_per_block - exp_bytes()) / per;
  return (int)(n < (size_t)kFastMaxWpb ? n : (size_t)kFastMaxWpb);
}

namespace {
// The trainer instance for a run: masters, stochastic rounding or
// minibatches take the full instance; halves per row from D.
template <bool kFull, int kH, bool kFp8 = false>
const void* train_fast_fn() {
  return reinterpret_cast<const void*>(&k_train_fast<kH == 1 ? kFastMaxWpb : kH, kFull, kH, kFp8>);
}
bool fast_full(const TrainArgs& a) { return a.fp8 || a.n_master > 0 || !a.round_nearest || a.minibatch > 0; }
const void* pick_train_fast(const TrainArgs& a) {
  const bool full = fast_full(a);
  if (fast_halves(a.D) == 2) {
    if (a.fp8) return train_fast_fn<true, 2, true>();
    return full ? train_fast_fn<true, 2>() : train_fast_fn<false, 2>();
  }
  if (a.fp8) return train_fast_fn<true, 1, true>();
  return full ? train_fast_fn<true, 1>() : train_fast_fn<false, 1>();
}
bool fast_args_ok(const TrainArgs& a) {
  const int nh = fast_halves(a.D);
  return nh > 0 && a.wpb >= 1 && a.wpb <= (nh == 1 ? kFastMaxWpb : 1);
}
}  // namespace

hipError_t prepare_train_fast(const TrainArgs& a) {
  if (!fast_args_ok(a)) return hipErrorInvalidValue;
  const int lds = (int)fast_block_lds_bytes(a.Lmax, a.K, a.wpb, a.D);
  return hipFuncSetAttribute(pick_train_fast(a), hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

hipError_t launch_train_fast(const TrainArgs& a, hipStream_t st) {
  if (!fast_args_ok(a)) return hipErrorInvalidValue;
  const size_t lds = fast_block_lds_bytes(a.Lmax, a.K, a.wpb, a.D);
  const dim3 grid((a.workers + a.wpb - 1) / a.wpb), block(64 * a.wpb * fast_halves(a.D));
  TrainArgs args = a;
  void* params[] = {&args};
  return hipLaunchKernel(pick_train_fast(a), grid, block, params, lds, st);
}

hipError_t launch_step_fast(int D_, int os, const int32_t* meta, const int32_t* mult, float alpha,
                            const float* exp_table, const float* O, const float* I, float* corr, float* md,
                            float* I_out, hipStream_t st) {
  if (fast_halves(D_) == 2)
    hipLaunchKernelGGL(k_step_fast<2>, dim3(1), dim3(128), 0, st, os, meta, mult, alpha, exp_table, O, I, corr, md,
                       I_out);
  else if (fast_halves(D_) == 1)
    hipLaunchKernelGGL(k_step_fast<1>, dim3(1), dim3(64), 0, st, os, meta, mult, alpha, exp_table, O, I, corr, md,
                       I_out);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// ---- scratch of the allocation-free entry point ------------------------------------
// One file-scope device array: each GPU the library runs on holds its own copy
// (kNoAllocScratchBytes) until the process ends.  In host code its name is a
// stand-in, not a device address: the address comes from hipGetSymbolAddress,
// taken once per device with that device current and kept here.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, void** ptr) {
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes the device current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  return hipSuccess;
}

}  // namespace dger
