/*
 * dger_walk.h — C ABI of the GPU corpus producer: DistGER's
 * information-oriented random walks on one MI355X, writing the corpus the
 * trainer of dger_sgns.h consumes.
 *
 * Reference interfaces replaced:
 *   WalkEngine::random_walk / internal_random_walk / internal_walk_epoch
 *   (include/walk.hpp:517, :609-700, :789-1043) for one partition: every
 *   vertex local, walker_num * init_round walkers, walker w starting at
 *   w % v_num (walk.hpp:501-509, :614), the path dump that builds
 *   local_corpus and vertex_cn (include/path.hpp:67-119), and the
 *   Correlation termination rule (walk.hpp:260-357, minLength = 20 at
 *   walk.hpp:404, R = 0.999 at walk.hpp:24).
 *   The per-edge common-neighbour counts the reference loads from a file
 *   (include/graph.hpp:826, include/storage.hpp:64-100) can be computed by
 *   dger_walk_common_neighbors.
 *
 * Conventions are those of dger_sgns.h: DGER_OK or a negative DGER_E_* code,
 * dger_last_error() for the message, caller-owned host buffers.  The graph
 * is CSR: rowptr[v_num + 1] (int64, rowptr[0] = 0), col[rowptr[v_num]]
 * (int32 neighbour ids), cnb[] (int32, common neighbours of the edge's two
 * ends, same order as col).
 */
#ifndef DGER_WALK_H
#define DGER_WALK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dger_walk_opts {
  uint64_t n_walkers;     /* walkers first_walker .. first_walker + n_walkers - 1
                             (reference: walker_num * init_round, walk.hpp:614) */
  uint64_t first_walker;  /* id of the first walker; walker w starts at w % v_num */
  uint32_t min_length;    /* steps before the correlation rule may end a walk (minLength) */
  uint32_t max_length;    /* most vertices per walk (no counterpart; walks past it are cut) */
  double r_threshold;     /* a walk ends once R^2 < r_threshold or R < 0 (R = 0.999) */
  uint64_t seed;          /* random stream: splitmix64 keyed by (seed, walker id, draw index) */
  int32_t device;         /* HIP device ordinal */
  uint32_t chunk_walkers; /* walkers per launch; 0 = as many as fit 8 GiB of paths */
} dger_walk_opts;

typedef struct dger_walk_stats {
  uint64_t walks;   /* walkers run */
  uint64_t tokens;  /* vertices written (separators excluded) */
  uint64_t moves;   /* neighbour draws that reached the acceptance test */
  uint64_t traces;  /* moves that stepped to the neighbour and back */
  uint64_t capped;  /* walks cut at max_length */
  double walk_ms;   /* device time of the walk, scan and corpus-packing kernels */
} dger_walk_stats;

/* Walks kept in device memory (HBM) for the trainer; see dger_walk_run_device. */
typedef struct dger_walk_corpus dger_walk_corpus;

void dger_walk_default_opts(dger_walk_opts* o);

/* cnb_out[e] = |N(u) ∩ N(col[e])| for edge e of row u.  Rows must be sorted
 * ascending without repeats (DGER_E_ARG otherwise). */
int dger_walk_common_neighbors(const int64_t* rowptr, const int32_t* col, uint32_t v_num, int32_t* cnb_out,
                               int32_t device);

/* Runs the walkers and writes the corpus (each walk's vertices, then -1;
 * walks in walker-id order) and vertex_cn[v_num] (occurrences of each vertex
 * in the corpus).  If corpus_cap is below the corpus length, returns
 * DGER_E_ARG with *corpus_len set to the length needed and the corpus
 * contents unspecified; n_walkers * (max_length + 1) always suffices. */
int dger_walk_run(const int64_t* rowptr, const int32_t* col, const int32_t* cnb, uint32_t v_num,
                  const dger_walk_opts* o, int32_t* corpus, uint64_t corpus_cap, uint64_t* corpus_len,
                  int32_t* vertex_cn, dger_walk_stats* st);

/* Same walks, left in HBM: *out receives a handle to the corpus on o->device
 * (the reference's local_corpus, path.hpp:96,110, without the text dump);
 * vertex_cn[v_num] is written to the host.  Hand it to the trainer with
 * dger_ctx_set_corpus_walks (include/dger_sgns.h), which copies it on the
 * device; free it with dger_walk_corpus_destroy. */
int dger_walk_run_device(const int64_t* rowptr, const int32_t* col, const int32_t* cnb, uint32_t v_num,
                         const dger_walk_opts* o, dger_walk_corpus** out, int32_t* vertex_cn, dger_walk_stats* st);

/* Token count (separators included) and the longest walk of a handle. */
int dger_walk_corpus_info(const dger_walk_corpus* h, uint64_t* n_tokens, uint32_t* max_walk);

/* Copies the corpus to the host (corpus_cap >= its token count). */
int dger_walk_corpus_download(const dger_walk_corpus* h, int32_t* corpus, uint64_t corpus_cap);

void dger_walk_corpus_destroy(dger_walk_corpus* h);

#ifdef __cplusplus
}
#endif

#endif /* DGER_WALK_H */
