/*
 * dger_sgns.h — C ABI of the MI355X (gfx950) skip-gram negative-sampling
 * trainer that stands in for DistGER's DSGL trainer.
 *
 * Reference interface replaced:
 *   int dsgl(int argc, char **argv, vector<int> *vertex_cn,
 *            vector<vertex_id_t> *new_sort, WalkEngine<real_t,uint32_t> *graph)
 *   (include/dsgl.hpp:1464), which reads graph->local_corpus (vertex ids with
 *   -1 after every walk, include/path.hpp:91-96,110) and vertex_cn (corpus
 *   count per vertex id, include/path.hpp:95).
 *
 * Conventions: plain pointers and sizes, no C++ or HIP types.  Every entry
 * point returns DGER_OK (0) or a negative DGER_E_* code and never exits the
 * process; dger_last_error() describes the last failure of the calling
 * thread.  The caller owns every host buffer it passes; the library owns its
 * device buffers (freed by dger_ctx_destroy, or before dger_sgns_train and
 * dger_sgns_main return).  Embedding matrices exchanged with the host are
 * V x size float32, rows in vocabulary-rank order (rank 0 = most frequent
 * vertex), exactly the layout of the reference's Wih / Woh.
 */
#ifndef DGER_SGNS_H
#define DGER_SGNS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGER_OK 0
#define DGER_E_ARG (-1)         /* invalid option or argument */
#define DGER_E_CORPUS (-2)      /* corpus holds an id outside [-1, v_num) */
#define DGER_E_VOCAB (-3)       /* vocabulary too small for the call */
#define DGER_E_HIP (-4)         /* HIP runtime failure */
#define DGER_E_ALLOC (-5)       /* device or host allocation failed */
#define DGER_E_IO (-6)          /* output file could not be written */
#define DGER_E_STATE (-7)       /* call made in the wrong order */
#define DGER_E_UNSUPPORTED (-8) /* option value outside what the kernels support */

#define DGER_PRECISION_FP32 0   /* FP32 rows, FP32 FMA arithmetic */
#define DGER_PRECISION_BF16 1   /* BF16 rows in HBM, FP32 working buffers, FP32 FMA contractions */
#define DGER_PRECISION_BF16_MFMA 2 /* BF16 rows, contractions on MFMA (bf16 operands, FP32 accumulate).
                                   At size 128 or 256, window <= 11, negative <= 14 it runs the
                                   window-slot trainer (at 256 two wavefronts per worker, one half
                                   of every row each): window rows in registers, read from Wih on entering the
                                   window and written back (round to nearest, or stochastic
                                   rounding with -rounding sr) on leaving it,
                                   target and negative rows updated in Woh directly, all three
                                   contractions on MFMA; the index stream is the reference's.  Other
                                   shapes keep the reference's buffers with GEMM1 on MFMA. */
#define DGER_PRECISION_FP8 3    /* FP8 e4m3 rows with a per-row FP32 scale, FP32 master rows for the
                                   top 1% ranks by count, stochastic rounding on every row store;
                                   contractions as BF16_MFMA.  Needs size % 4 == 0, window <= 16,
                                   negative <= 6.  At size 128 or 256 it runs on the window-slot
                                   trainer, where Wih (the embeddings) is coded and Woh stays bf16
                                   with stochastically rounded atomic adds (a coded Woh row would be
                                   requantized at every step it is used); elsewhere both are coded. */

/* Options.  The first block mirrors the flags dsgl() parses
 * (include/dsgl.hpp:1496-1535) with the same defaults (dsgl.hpp:67-76);
 * the second block holds this library's GPU options. */
typedef struct dger_sgns_opts {
  int32_t size;            /* -size        hidden size D (100) */
  int32_t window;          /* -window      (5) */
  int32_t negative;        /* -negative    (5) */
  int32_t batch_size;      /* -batch-size  (11); the reference replaces it by 2*window+1 */
  float sample;            /* -sample      (1e-3) */
  float alpha;             /* -alpha       (0.1) */
  int32_t min_count;       /* -min-count   (5); parsed, unused, as in the reference */
  int32_t iter;            /* -iter        (5) */
  int32_t threads;         /* -threads     (12); splits the unigram power sum */
  int32_t debug;           /* -debug       (2) */
  int32_t binary;          /* -binary      (0) */
  int32_t disk;            /* -disk        parsed, no effect */
  float sync_period;       /* -sync-period (0.1); no effect on one GPU */
  int32_t min_sync_words;  /* -min-sync-words (1024); no effect on one GPU */
  int32_t full_sync_times; /* -full-sync-times (0); no effect on one GPU */
  int32_t message_size;    /* -message-size (1024); no effect on one GPU */
  char eoutput[512];       /* -eoutput     embedding file ("" = none) */
  char save_vocab[512];    /* -save-vocab  parsed, unused, as in the reference */
  char read_vocab[512];    /* -read-vocab  parsed, unused, as in the reference */
  char save_emb[512];      /* -save-emb    parsed, unused, as in the reference */
  /* GPU options */
  int32_t workers;         /* -workers     concurrent Hogwild workers; 0 = auto.
                              Worker k trains the corpus slice and LCG seed of the
                              reference's compute thread k+1 under -threads workers+1. */
  int32_t precision;       /* -precision   fp32 (0), bf16 (1), bf16-mfma (2), fp8 (3) */
  int64_t table_size;      /* -table-size  noise-table slots (1e8, dsgl.hpp:82) */
  int32_t device;          /* -device      HIP device ordinal (0) */
  int32_t graph;           /* -graph       1: the trainer launch (counter reset + kernel) is captured
                              once in a HIP graph and replayed by later passes (1) */
  int32_t autotune;        /* -autotune    with -workers 0, the first pass times candidates on a
                              corpus prefix and keeps the fastest: 1 worker counts and
                              workgroup shapes; 2 (window-slot trainer) also -share 2/4/8 and
                              the minibatch (one per step, or -batch-size 2*window+1 honoured),
                              choices that change the training and that the caller gates on
                              embedding quality (0) */
  int32_t rounding;        /* -rounding    bf16 row stores and Woh updates of the window-slot
                              trainer: sr (0) stochastic rounding, so that updates below half
                              a bf16 step survive in expectation (multi-epoch or small-alpha
                              runs, where nearest rounding stalls); rne (1) nearest (1) */
  int32_t masters;         /* -masters     1: the window-slot trainer keeps FP32 master rows
                              for the top 1 % ranks by count in Wih and Woh, read and
                              updated in FP32 (0) */
  int32_t batch_parity;    /* -batch-parity 1: -batch-size is replaced by 2*window+1 and a step's
                              contexts form one minibatch, as in the reference (dsgl.hpp:832,
                              :1163-1167); 0: -batch-size is honoured as the minibatch size: a
                              step's contexts (first window, then second, in order) are trained
                              in minibatches of at most batch_size contexts, each updating the
                              output and window rows before the next is formed (window-slot
                              trainer only) (1) */
  int32_t share;           /* -share       target positions sharing one negative group: 2, the
                              reference's two sentences per step (dsgl.hpp:1136-1158), or 4, 8,
                              16 (window-slot trainer), where share/2 consecutive steps draw from
                              one group and a refresh draws ceil(max(L1, L2) / (share/2)) groups
                              (2) */
} dger_sgns_opts;

typedef struct dger_sgns_stats {
  uint64_t steps;          /* minibatch steps; each trains two target positions */
  uint64_t trained_words;  /* post-subsampling target positions trained (2 * steps) */
  uint64_t raw_words;      /* tokens consumed incl. separators and discards (dsgl.hpp:913) */
  uint64_t sentences;      /* sentences trained to their end */
  float final_alpha;       /* learning rate after the last update of worker 0 */
  int32_t workers;         /* workers used */
  int32_t vocab_size;
  int32_t max_sentence;    /* longest sentence slot the kernels were sized for */
  double kernel_ms;        /* HIP-event time of the trainer kernel launches */
  double words_per_sec;    /* trained_words / kernel time */
  uint64_t contexts;       /* sum over steps of the contexts of both windows */
  uint64_t out_ctx;        /* sum over steps of (output rows x contexts); each of the three
                              contractions is D * out_ctx multiply-adds */
  int32_t share;           /* negative sharing the run used */
  int32_t minibatch;       /* contexts per minibatch the run used; 0 = a step's contexts at once */
} dger_sgns_stats;

/* Fills *out with the reference defaults (dsgl.hpp:67-76) and GPU defaults. */
int dger_sgns_default_opts(dger_sgns_opts *out);

/* Parses argv like dsgl() (dsgl.hpp:1429-1441, 1496-1535): exact-match flags,
 * the value in the next argument, unknown flags ignored.  Starts from the
 * defaults.  Returns DGER_E_ARG if a flag is the last argument. */
int dger_sgns_parse_args(int argc, char **argv, dger_sgns_opts *out);

/* Trains on a host corpus (vertex ids, -1 after every walk) with counts
 * vertex_cn[v_num]; blocking.  emb_out (optional, V x size float32, rank
 * order) receives the input embeddings Wih; rank_id_out (optional, V int32)
 * receives the vertex id of each rank.  Replaces Train_SGNS_MPI
 * (dsgl.hpp:522-1427) plus preparatory_work (dsgl.hpp:511-519). */
int dger_sgns_train(const dger_sgns_opts *opts, const int32_t *corpus, uint64_t n_tokens,
                    const int32_t *vertex_cn, uint32_t v_num, float *emb_out,
                    int32_t *rank_id_out, dger_sgns_stats *stats);
/* dger_sgns_train with its device work -- uploads, trainer launches, model
 * copies -- ordered on the caller's HIP stream (a hipStream_t passed as a
 * pointer; NULL = the library's own stream), so that it follows work the
 * caller queued there without a host wait.  Still blocking: the stream is
 * synchronised before the call returns.  The stream must belong to
 * opts->device (else DGER_E_ARG). */
int dger_sgns_train_stream(const dger_sgns_opts *opts, const int32_t *corpus, uint64_t n_tokens,
                           const int32_t *vertex_cn, uint32_t v_num, float *emb_out,
                           int32_t *rank_id_out, dger_sgns_stats *stats, void *stream);

/* The dsgl() drop-in (dsgl.hpp:1464-1584): parse argv, train, and when
 * -eoutput is given write the embedding file in my_saveModel's format
 * (dsgl.hpp:1443-1461).  Like the reference's assert (dsgl.hpp:370) it
 * rejects vocabularies of 1000 vertices or fewer (DGER_E_VOCAB). */
int dger_sgns_main(int argc, char **argv, const int32_t *corpus, uint64_t n_tokens,
                   const int32_t *vertex_cn, uint32_t v_num);

/* Thread-local description of the last failure ("" if none). */
const char *dger_last_error(void);

/* ---- Context API: device-resident corpus and model across calls ---------- */
typedef struct dger_ctx dger_ctx;

/* Builds the vocabulary (dsgl.hpp:314-382), the sigmoid table
 * (dsgl.hpp:1557-1562), the noise table (dsgl.hpp:121-147) and the initial
 * model (dsgl.hpp:433-457) on the device. */
int dger_ctx_create(const dger_sgns_opts *opts, const int32_t *vertex_cn, uint32_t v_num,
                    dger_ctx **out);
/* Orders the context's later device work on the caller's HIP stream (a
 * hipStream_t passed as a pointer; NULL = back to the context's own stream).
 * Work already queued on the previous stream is waited for first.  The
 * stream must belong to the context's device (else DGER_E_ARG) and outlive
 * its use by the context. */
int dger_ctx_set_stream(dger_ctx *ctx, void *stream);
/* Copies the corpus to device memory (validated on the host first). */
int dger_ctx_set_corpus(dger_ctx *ctx, const int32_t *corpus, uint64_t n_tokens);

/* Hands walks left in HBM by dger_walk_run_device (include/dger_walk.h) to
 * the trainer without a host round trip (a device-to-device copy; the
 * handle stays the caller's).  Replaces the reference's text dump and
 * reload of local_corpus (include/path.hpp:67-119).  The walks must be on
 * the context's device and use ids below its v_num. */
struct dger_walk_corpus;
int dger_ctx_set_corpus_walks(dger_ctx *ctx, const struct dger_walk_corpus *walks);
/* Re-initialises Wih (LCG, seed 1) and zeroes Woh. */
int dger_ctx_reset_model(dger_ctx *ctx);
/* One training run (all -iter epochs) with `workers` workers (0 = opts).
 * rec (optional, host) receives rec_cap index-stream records per worker,
 * each dger_record_stride() int32 wide; worker_steps (optional, host,
 * `workers` entries) receives each worker's step count. */
int dger_ctx_train(dger_ctx *ctx, int32_t workers, int32_t *rec, int64_t rec_cap,
                   uint64_t *worker_steps, dger_sgns_stats *stats);
/* dger_ctx_train without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate, such as code under HIP graph
 * capture, tight serving loops or harnesses that forbid hipMalloc.  Runs the
 * window-slot trainer (DGER_PRECISION_BF16_MFMA and DGER_PRECISION_FP8
 * contexts it supports; other contexts: DGER_E_UNSUPPORTED) with its per-run
 * scratch -- counters, per-worker step counts and rates, per-wavefront
 * negative batches -- in a static device array of 16 MiB per GPU (held until
 * the process ends), so that 4 * (negative * max_sentence * size / 128 + 3)
 * bytes per worker plus 2 KiB must fit:
 * larger requests fail with DGER_E_ARG and name the worker count that fits.
 * No autotuning, records or graph replay; stats as dger_ctx_train.  Every call
 * on one GPU uses that GPU's array: calls that overlap on one GPU race. */
int dger_ctx_train_noalloc(dger_ctx *ctx, int32_t workers, dger_sgns_stats *stats);
/* Copies Wih / Woh (V x size float32, rank order) to the host; either may be NULL. */
int dger_ctx_get_model(dger_ctx *ctx, float *wih, float *woh);
/* Replaces Wih / Woh from the host; either may be NULL. */
int dger_ctx_set_model(dger_ctx *ctx, const float *wih, const float *woh);
/* Vocabulary: rank -> vertex id (V int32) and the corpus word total. */
int dger_ctx_get_vocab(dger_ctx *ctx, int32_t *rank_id, uint64_t *train_words);
/* Copies the noise table (table_size int32) to the host. */
int dger_ctx_get_table(dger_ctx *ctx, int32_t *table);
/* Writes Wih in my_saveModel's format (dsgl.hpp:1443-1461), through
 * dger_emb_write. */
int dger_ctx_save(dger_ctx *ctx, const char *path, int32_t binary);
/* Writes v embedding rows (rows: v x d float32; ids[r]: the vertex id
 * printed before row r) in my_saveModel's format (dsgl.hpp:1443-1461): a
 * "v d" header line, then per row "id " followed by d values as "%f " text
 * (the bytes printf writes, formatted without printf by several threads) or,
 * with binary = 1, as raw float32, and a newline.  Host only (no device). */
int dger_emb_write(const char *path, uint32_t v, int32_t d, const int32_t *ids, const float *rows,
                   int32_t binary);
void dger_ctx_destroy(dger_ctx *ctx);

/* Width in int32 of one index-stream record:
 * [t1, t2, b, n_ctx_first, n_ctx, n_out, alpha bits, negative group,
 *  contexts[4*window], outputs[negative+2], multiplicities[negative+2]]. */
int32_t dger_record_stride(int32_t window, int32_t negative);

/* ---- Test-facing kernels --------------------------------------------------- */
/* One minibatch step (dsgl.hpp:1241-1310) on host-staged rows, run by the same
 * device code as the trainer: O (n_out x D), I (n_ctx x D), meta (n_out),
 * contexts [0, n_ctx_first) label row 0, the rest label row n_out-1.
 * Outputs corr (n_out x n_ctx), md (n_out x D), gin (n_ctx x D). */
int dger_sgns_step(int32_t D, int32_t n_out, int32_t n_ctx, int32_t n_ctx_first,
                   const int32_t *meta, float alpha, const float *O, const float *I,
                   float *corr, float *md, float *gin);

/* GEMM1 and the sigmoid gradient (dsgl.hpp:1241-1272) of the BF16-MFMA form's
 * general trainer -- the code its minibatches of at most 8 outputs and 64
 * contexts run (sizes other than the window-slot trainer's) -- on
 * host-staged rows: bf16 operands, FP32 accumulation.  D a multiple of 32 up
 * to 256, n_out in [2, 8], n_ctx in [1, 64]; contexts [0, n_ctx_first) label
 * row 0, the rest row n_out-1.  corr (n_out x n_ctx) receives the
 * coefficients meta * (label - sigmoid) * alpha. */
int dger_sgns_step_mfma_gemm1(int32_t D, int32_t n_out, int32_t n_ctx, int32_t n_ctx_first,
                              const int32_t *meta, float alpha, const float *O, const float *I,
                              float *corr);

/* One step of the BF16-MFMA form's window-slot trainer on host-staged rows
 * (the same device code as that trainer; see DGER_PRECISION_BF16_MFMA).
 * Row length 128.  O (n_out x 128, n_out in [2, 16]) holds the step's output
 * rows, meta (n_out) their multiplicities; I (48 x 128) holds the 48 window
 * slots: slots 0..23 belong to the first sentence (their contexts label
 * output 0 positive), 24..47 to the second (output n_out-1 positive); mult
 * (48) is the number of the step's contexts in each slot (0 = none).  The
 * contractions take bf16 operands and accumulate in FP32.  Outputs: corr
 * (n_out x 48) the coefficients mult * meta * (label - sigmoid) * alpha, md
 * (n_out x 128) the output-row updates, I_out (48 x 128) the slot rows after
 * the update.  Each step of dsgl.hpp:1241-1333 with repeated contexts merged
 * into one slot. */
int dger_sgns_step_fast(int32_t n_out, const int32_t *meta, const int32_t *mult, float alpha,
                        const float *O, const float *I, float *corr, float *md, float *I_out);

/* The same step at row length `size` (128 or 256; 128 x becomes size x in
 * every shape above).  At 256 the step runs as the trainer does at that
 * size: two wavefronts, each holding one 128-element half of every row, that
 * add their halves' GEMM1 products through LDS.  Each step of
 * dsgl.hpp:1241-1333 with repeated contexts merged into one slot. */
int dger_sgns_step_fast_d(int32_t size, int32_t n_out, const int32_t *meta, const int32_t *mult, float alpha,
                          const float *O, const float *I, float *corr, float *md, float *I_out);

/* SGNS logistic loss and its adjoint over minibatches of shared negatives,
 * on DEVICE pointers.  Minibatch m pairs contexts ctx[m*n_ctx .. +n_ctx) (rows
 * of syn0) with outputs out[m*n_out .. +n_out) (rows of syn1neg), whose first
 * entry is the positive target and the rest negatives:
 *   loss = sum_m sum_c [ softplus(-s1[o0].s0[c]) + sum_{k>=1} softplus(s1[ok].s0[c]) ]
 * bwd accumulates dloss/dsyn0 and dloss/dsyn1neg (times `scale`) into
 * grad0 / grad1 (zero them first).  No reference counterpart exists as a
 * separate function; the update of dsgl.hpp:1241-1368 is -alpha times this
 * gradient with a table sigmoid. */
int dger_loss_fwd_f32(const float *syn0, const float *syn1, int32_t D, const int32_t *ctx,
                      const int32_t *out, int32_t n_mb, int32_t n_ctx, int32_t n_out,
                      float *loss_out /* device, 1 float */);
int dger_loss_bwd_f32(const float *syn0, const float *syn1, int32_t D, const int32_t *ctx,
                      const int32_t *out, int32_t n_mb, int32_t n_ctx, int32_t n_out,
                      float scale, float *grad0, float *grad1);
int dger_loss_fwd_f64(const double *syn0, const double *syn1, int32_t D, const int32_t *ctx,
                      const int32_t *out, int32_t n_mb, int32_t n_ctx, int32_t n_out,
                      double *loss_out);
int dger_loss_bwd_f64(const double *syn0, const double *syn1, int32_t D, const int32_t *ctx,
                      const int32_t *out, int32_t n_mb, int32_t n_ctx, int32_t n_out,
                      double scale, double *grad0, double *grad1);

/* Minimal device-memory helpers so FFI callers (ctypes tests) can stage the
 * device pointers the loss kernels take. */
int dger_dev_alloc(void **ptr, uint64_t bytes);
int dger_dev_free(void *ptr);
int dger_dev_upload(void *dst, const void *src, uint64_t bytes);
int dger_dev_download(void *dst, const void *src, uint64_t bytes);
int dger_dev_zero(void *ptr, uint64_t bytes);
/* Link prediction on the context's current model (north-star evaluation;
   held-out split as src/tools/gtrain.cpp:300-343).  pairs: n x 2 vertex ids.
   dger_ctx_link_scores: scores[i] = dot(Wih[u_i], Wih[v_i]) computed on the
   device.  Every id must lie in [0, v_num): a pair outside it fails the call
   with DGER_E_ARG before anything is uploaded.  dger_ctx_link_auc: ROC AUC of
   those scores against labels (1 = held-out edge, 0 = non-edge) by the
   Mann-Whitney statistic with ties counting one half, computed on the device;
   NaN when a class is empty. */
int dger_ctx_link_scores(dger_ctx *ctx, const int32_t *pairs, uint64_t n, float *scores);
int dger_ctx_link_auc(dger_ctx *ctx, const int32_t *pairs, const int8_t *labels, uint64_t n, double *auc);

/* Device-to-device copy (also the known-size reference for traffic counters). */
int dger_dev_copy(void *dst, const void *src, uint64_t bytes);
int dger_device_count(int32_t *count);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
#include <vector>
/* Header-only C++ shim with the reference's call shape (dsgl.hpp:1464) for a
 * caller that holds the corpus in a std::vector; new_sort is accepted and
 * unused, as it is in the reference's training path. */
inline int dsgl(int argc, char **argv, std::vector<int> *vertex_cn,
                std::vector<uint32_t> *new_sort, const std::vector<int> *local_corpus) {
  (void)new_sort;
  return dger_sgns_main(argc, argv, local_corpus->data(), (uint64_t)local_corpus->size(),
                        vertex_cn->data(), (uint32_t)vertex_cn->size());
}
#endif

#endif /* DGER_SGNS_H */
