// Internal interface between the host layer (sgns_host.cpp) and the gfx950
// kernels (sgns_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <string>

// Walks kept in HBM by the corpus producer (walk.hip): device pointer,
// token count (separators included), longest walk, vertex count, device.
struct dger_walk_corpus {
  int32_t* d = nullptr;
  uint64_t n_tokens = 0;
  uint32_t max_walk = 0;
  uint32_t v_num = 0;
  int device = 0;
};

namespace dger {

constexpr unsigned long long kLcgMul = 25214903917ULL;  // dsgl.hpp:453
constexpr unsigned long long kLcgAdd = 11ULL;
constexpr int kMaxSentence = 1000;                     // dsgl.hpp:23
constexpr int kExpTable = 1001;                        // dsgl.hpp:20, :1557

// Counters kept in device memory for one training run.
enum Counter : int {
  kWordCountActual = 0,  // word_count_actual (dsgl.hpp:535), shared by all workers
  kSteps = 1,
  kRawWords = 2,
  kSentences = 3,
  kContexts = 4,   // sum over steps of the context count
  kOutCtx = 5,     // sum over steps of (output rows x contexts): MACs per contraction / D
  kNumCounters = 6,
};

struct TrainArgs {
  const int32_t* corpus;     // n_tokens vertex ids, -1 after each walk
  uint64_t n_tokens;
  const int32_t* id_rank;    // V: vertex id -> vocabulary rank
  const float* keep_thr;     // V: subsampling threshold per rank
  const int32_t* table;      // noise table over ranks
  uint64_t table_size;
  const float* exp_table;    // 1001-entry sigmoid table
  void* wih;                 // V x D (float or bf16 rows)
  void* woh;                 // V x D
  float* fscratch;           // per worker: 2 sentence buffers + negative buffer
  unsigned long long* counters;
  int32_t* rec;              // optional index-stream records (device)
  int64_t rec_cap;
  unsigned long long* worker_steps;  // per worker
  float* worker_alpha;       // per worker, final learning rate
  int32_t V, D, W, K, iter, workers, Lmax;
  int32_t wpb;  // workers (wavefronts) per workgroup
  int32_t sample_on;
  int32_t mfma;              // BF16 rows: run the contractions on MFMA
  float alpha0;
  float alpha_den;           // (float)(iter * train_words + 1)
  // Window-slot trainer: FP32 master rows of the ranks below n_master (Wih,
  // Woh; n_master x D floats each), and the rounding of bf16 row updates
  // (0: stochastic, so that small updates survive in expectation; 1: nearest).
  float* wmaster;
  float* omaster;
  int32_t n_master;
  int32_t round_nearest;
  // Window-slot trainer batching: 2^share_shift consecutive steps draw from
  // one negative group (negative sharing 2 << share_shift); minibatch > 0
  // splits a step's contexts into minibatches of at most that many.
  int32_t share_shift;
  int32_t minibatch;
  // Window-slot trainer, FP8 form: wih / woh are coded rows (fp8_row_stride
  // bytes each) instead of bf16 rows.
  int32_t fp8;
};

// Floats of per-worker scratch: two sentence buffers (Lmax rows each) and the
// negative buffer (K * Lmax rows).
#if defined(__HIPCC__)
#define DGER_HD __host__ __device__
#else
#define DGER_HD
#endif

DGER_HD inline size_t worker_scratch_floats(int Lmax, int K, int D) {
  return (size_t)(2 + K) * (size_t)Lmax * (size_t)D;
}

size_t train_lds_bytes(int Lmax, int K, int W, int D, int precision);
// Workers per workgroup that fit lds_per_block bytes (0 if not even one), at
// most 4; and the workgroup's LDS bytes for wpb workers.
int train_workers_per_block(int Lmax, int K, int W, int D, int precision, size_t lds_per_block);
size_t train_block_lds_bytes(int Lmax, int K, int W, int D, int precision, int wpb);
DGER_HD inline int record_stride(int W, int K) { return 8 + 4 * W + 2 * (K + 2); }

// Window-slot trainer of the BF16-MFMA form (sgns_fast.hip): bf16 rows of
// length kFastDim per wavefront (one wavefront per worker at D = 128, two at
// D = 256, each holding one half of every row), window slots in registers,
// contractions on MFMA.
constexpr int kFastDim = 128;
constexpr int kFastMaxWpb = 8;  // workers (wavefronts) per workgroup at D = 128: two per SIMD
// Wavefronts per worker (halves of a row) for row length D; 0: not supported.
DGER_HD inline int fast_halves(int D) { return D == kFastDim ? 1 : (D == 2 * kFastDim ? 2 : 0); }
bool fast_supported(int D, int K, int W, uint32_t V, int64_t table_size);
size_t fast_lds_bytes(int Lmax, int K);  // one wavefront's LDS
// Device memory per wavefront of the window-slot trainer, in 4-byte words:
// its copy of the worker's shared negative batch (K x Lmax ranks).
DGER_HD inline size_t fast_scratch_ints(int Lmax, int K) { return (size_t)(K > 0 ? K : 1) * (size_t)Lmax; }
// The same per worker (all its wavefronts).
DGER_HD inline size_t fast_worker_scratch_ints(int Lmax, int K, int D) {
  return (size_t)fast_halves(D) * fast_scratch_ints(Lmax, K);
}
// LDS of a workgroup of wpb workers; workers per workgroup that fit
// lds_per_block (0 if not even one; always 1 at D = 256).
size_t fast_block_lds_bytes(int Lmax, int K, int wpb, int D);
int fast_workers_per_block(int Lmax, int K, size_t lds_per_block, int D);
hipError_t prepare_train_fast(const TrainArgs& a);
// Device scratch of the allocation-free trainer entry point: a static device
// array of kNoAllocScratchBytes per GPU.  noalloc_scratch returns this
// device's address of it (looked up once per device; `device` must be current).
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
hipError_t noalloc_scratch(int device, void** ptr);
hipError_t launch_train_fast(const TrainArgs& a, hipStream_t s);
// One step of the window-slot form on host-staged rows of length D (test-facing).
hipError_t launch_step_fast(int D, int os, const int32_t* meta, const int32_t* mult, float alpha,
                            const float* exp_table, const float* O, const float* I, float* corr, float* md,
                            float* I_out, hipStream_t s);

hipError_t launch_init_net(float* wih, uint64_t n, int D, hipStream_t s);
hipError_t launch_table_fill(int32_t* table, uint64_t table_size, const int64_t* starts, int n_starts,
                             hipStream_t s);
hipError_t prepare_train(const TrainArgs& a, int precision);
hipError_t launch_train(const TrainArgs& a, int precision, hipStream_t s);
// FP8 form: a model matrix is V rows of D e4m3 codes followed by the row's
// FP32 scale (row stride D + 4 rounded up to 16 bytes, so that every row's
// codes start 16-byte aligned), plus FP32 master rows for ranks
// < n_master that take the place of the coded rows.  The trainer gets a
// device pointer to this descriptor instead of the row array.
struct Fp8Mat {
  uint8_t* rows;
  float* master;
  int32_t n_master;
  int32_t D;
};
DGER_HD inline size_t fp8_row_stride(int D) { return ((size_t)D + 4 + 15) & ~(size_t)15; }
// FP32 rows (rank order) <-> an FP8 matrix; quantize rounds to nearest even.
hipError_t launch_fp8_quantize(const float* src, const Fp8Mat* m, uint32_t V, int D, hipStream_t s);
hipError_t launch_fp8_dequantize(const Fp8Mat* m, float* dst, uint32_t V, int D, hipStream_t s);

// Records msg as the calling thread's last error and returns code (the
// dger_last_error() of every entry point in the library).
int host_fail(int code, const std::string& msg);

// Link-prediction evaluation (link_eval.hip).
hipError_t launch_link_scores(const void* wih, bool bf16_rows, const int32_t* id_rank, uint32_t V, int D,
                              const int32_t* pairs, uint64_t n, float* scores, hipStream_t s);
size_t link_auc_scratch_bytes(uint64_t n);
hipError_t launch_link_auc(const float* scores, const int8_t* labels, uint64_t n, void* scratch,
                           unsigned long long* counts3, uint64_t* npos, uint64_t* nneg,
                           unsigned long long* twice_u, hipStream_t s);
// Zeroes n <= 64 counters with a kernel (capturable in a graph).
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s);
hipError_t launch_f32_to_bf16(const float* src, uint16_t* dst, uint64_t n, hipStream_t s);
hipError_t launch_bf16_to_f32(const uint16_t* src, float* dst, uint64_t n, hipStream_t s);
// GEMM1 + sigmoid of the BF16-MFMA form on given rows (test-facing): D a
// multiple of 32 up to 256, 2 <= os <= 8, 1 <= nin <= 64.
hipError_t launch_step_mfma_gemm1(int D, int os, int nin, int fin, const int32_t* meta, float alpha,
                                  const float* exp_table, const float* O, const float* I, float* corr,
                                  hipStream_t s);
hipError_t launch_step(int D, int os, int nin, int fin, const int32_t* meta, float alpha,
                       const float* exp_table, const float* O, const float* I, float* corr,
                       float* md, float* gin, hipStream_t s);
template <typename T>
hipError_t launch_loss_fwd(const T* syn0, const T* syn1, int D, const int32_t* ctx,
                           const int32_t* out, int n_mb, int n_ctx, int n_out, T* loss,
                           hipStream_t s);
template <typename T>
hipError_t launch_loss_bwd(const T* syn0, const T* syn1, int D, const int32_t* ctx,
                           const int32_t* out, int n_mb, int n_ctx, int n_out, T scale, T* g0,
                           T* g1, hipStream_t s);

}  // namespace dger
