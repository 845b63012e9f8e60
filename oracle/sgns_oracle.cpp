// SGNS ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of DistGER's DSGL skip-gram negative-sampling trainer
// (include/dsgl.hpp).  Only the test suite, the smoke check and the CPU
// baseline leg of bench.py load this library, and only as the checker or the
// timed CPU baseline.  The product (libdger_sgns.so) never links or calls it.
//
// Every function cites the reference lines it restates.  Where the reference
// leaves an order unspecified, this file fixes one and says so:
//   * the three per-step GEMMs (cblas_sgemm, dsgl.hpp:1241,1290,1309) sum in
//     plain sequential order with one fused multiply-add per term;
//   * the OpenMP reduction of sum(cn^0.75) (dsgl.hpp:125-132) combines the
//     static-schedule partial sums in thread order (deterministic for
//     -threads <= 2, where the reference is too).
//
// Build: oracle/Makefile (exact build: -ffp-contract=off, fmaf calls;
// "fast" build: -DORC_FAST -ffast-math -fopenmp, used only for CPU timing).
//
// Parity status of each function is listed in the header comment of
// oracle/oracle.py and in the tests under tests/test_oracle_*.py.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORC_FAST
#define ORC_FMA(a, b, c) ((a) * (b) + (c))
#else
#define ORC_FMA(a, b, c) std::fma((a), (b), (c))
#endif

namespace {

typedef unsigned long long u64;

// word2vec linear congruential generator, dsgl.hpp:453, :925, :1063, :1100.
const u64 kLcgMul = 25214903917ULL;
const u64 kLcgAdd = 11ULL;
inline u64 lcg_next(u64 s) { return s * kLcgMul + kLcgAdd; }

// dsgl.hpp:20-23
const int kExpTableSize = 1000;
const int kMaxExp = 6;
const int kMaxSentence = 1000;

// Table-quantised sigmoid gradient of dsgl.hpp:1249-1270 for one corrM entry.
// The index expression (f + 6) * 83 uses the integer quotient 1000/6/2 = 83.
template <typename T>
inline T sig_grad(T f, int label, T alpha, const float* exp_table, int* idx_out) {
  if (f > (T)kMaxExp) {
    if (idx_out) *idx_out = -1;
    return (T)(label - 1) * alpha;
  }
  if (f < (T)(-kMaxExp)) {
    if (idx_out) *idx_out = -2;
    return (T)label * alpha;
  }
  int idx = (int)((f + (T)kMaxExp) * (T)(kExpTableSize / kMaxExp / 2));
  if (idx_out) *idx_out = idx;
  return ((T)label - (T)exp_table[idx]) * alpha;
}

// One minibatch step (dsgl.hpp:1241-1310) on staged rows.
//   O    : os  x D output rows (target, unique negatives, target2)
//   I    : nin x D input (context) rows
//   corr : os  x nin  gradient coefficients (GEMM1 + sigmoid, dsgl.hpp:1241-1272)
//   md   : os  x D    = corr . I           (GEMM2, dsgl.hpp:1290)
//   gin  : nin x D    = corr^T . O         (GEMM3, dsgl.hpp:1309)
// Contexts j < fin belong to the first sentence: row 0 is their positive;
// contexts j >= fin belong to the second: row os-1 is their positive.
template <typename T>
void step_math(int D, int os, int nin, int fin, const int* meta, T alpha,
               const float* exp_table, const T* O, const T* I, T* corr, T* md,
               T* gin, int* sig_idx) {
  for (int i = 0; i < os; i++) {
    for (int j = 0; j < nin; j++) {
      T f = (T)0;
      const T* o = O + (size_t)i * D;
      const T* x = I + (size_t)j * D;
      for (int k = 0; k < D; k++) f = ORC_FMA(o[k], x[k], f);
      int label = (j < fin) ? (i == 0 ? 1 : 0) : (i == os - 1 ? 1 : 0);
      int* ip = sig_idx ? &sig_idx[i * nin + j] : nullptr;
      T g = sig_grad<T>(f, label, alpha, exp_table, ip);
      corr[i * nin + j] = g * (T)meta[i];
    }
  }
  for (int i = 0; i < os; i++) {
    for (int d = 0; d < D; d++) {
      T acc = (T)0;
      for (int j = 0; j < nin; j++) acc = ORC_FMA(corr[i * nin + j], I[(size_t)j * D + d], acc);
      md[(size_t)i * D + d] = acc;
    }
  }
  for (int j = 0; j < nin; j++) {
    for (int d = 0; d < D; d++) {
      T acc = (T)0;
      for (int i = 0; i < os; i++) acc = ORC_FMA(corr[i * nin + j], O[(size_t)i * D + d], acc);
      gin[(size_t)j * D + d] = acc;
    }
  }
}

struct Tables {
  int V = 0, D = 0;
  std::vector<int> rank_id;     // v_vocab[r].id
  std::vector<int> id_rank;     // hash2v_vocab
  const int* cn_by_id = nullptr;  // cn_vocab (indexed by vertex id)
  uint64_t train_words = 0;
  std::vector<float> exp_table; // 1001 entries
  std::vector<float> keep_thr;  // per rank: sqrtf(ratio) + ratio
  std::vector<int> table;       // unigram^0.75 noise table over ranks
  long long table_size = 0;
};

}  // namespace

extern "C" {

// ---------------------------------------------------------------- LCG ----
// n sequential steps of the LCG (dsgl.hpp:453).
uint64_t orc_lcg_steps(uint64_t s, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) s = lcg_next(s);
  return s;
}

// The same n steps by repeated squaring of the affine map s -> a*s + c
// (mod 2^64).  Used to check the sequential form and the GPU jump-ahead.
uint64_t orc_lcg_jump(uint64_t s, uint64_t n) {
  u64 a = kLcgMul, c = kLcgAdd;   // current power: s -> a*s + c
  u64 ra = 1, rc = 0;             // accumulated map
  while (n) {
    if (n & 1) { ra = a * ra; rc = a * rc + c; }
    c = a * c + c;
    a = a * a;
    n >>= 1;
  }
  return ra * s + rc;
}

// -------------------------------------------------------------- vocab ----
// my_LearnVocabFromTrainFile + my_SortVocab (dsgl.hpp:314-382, :246-252):
// vertices ranked by corpus count, descending, by std::sort with the same
// comparator on (cn, id) records built in id order; the unstable tie order is
// whatever libstdc++'s introsort produces for that input.
int orc_build_vocab(const int32_t* cn, int32_t V, int32_t* rank_id, int32_t* id_rank,
                    uint64_t* train_words) {
  struct Rec { int cn; uint32_t id; };
  if (V <= 0) return -1;
  std::vector<Rec> v((size_t)V);
  u64 tw = 0;
  for (int32_t i = 0; i < V; i++) { v[i].cn = cn[i]; v[i].id = (uint32_t)i; }
  for (int32_t i = 0; i < V; i++) tw += (u64)(long long)cn[i];
  std::sort(v.begin(), v.end(), [](Rec a, Rec b) { return a.cn > b.cn; });
  for (int32_t r = 0; r < V; r++) {
    rank_id[r] = (int32_t)v[r].id;
    id_rank[v[r].id] = r;
  }
  *train_words = tw;
  return 0;
}

// InitNet (dsgl.hpp:433-457): Wih filled in rank-major order from the LCG
// seeded with 1; Woh is zero (not produced here).
void orc_init_net(int32_t V, int32_t D, float* wih) {
  u64 s = 1;
  size_t n = (size_t)V * (size_t)D;
  for (size_t i = 0; i < n; i++) {
    s = lcg_next(s);
    wih[i] = (((float)(s & 0xFFFF) / 65536.f) - 0.5f) / (float)(size_t)D;
  }
}

// expTable (dsgl.hpp:1557-1562): float exp of a float argument, then x/(x+1).
void orc_exp_table(float* t) {
  for (int i = 0; i < kExpTableSize + 1; i++) {
    float e = std::exp((i / (float)kExpTableSize * 2 - 1) * kMaxExp);
    t[i] = e / (e + 1);
  }
}

// Subsampling threshold per rank (dsgl.hpp:921-926).  The reference indexes
// cn_vocab (a per-vertex-id array) with the RANK w; this keeps that.
void orc_keep_thresholds(const int32_t* cn_by_id, int32_t V, uint64_t train_words, float sample,
                         float* thr) {
  for (int32_t w = 0; w < V; w++) {
    float ratio = (sample * (float)train_words) / (float)cn_by_id[w];
    thr[w] = std::sqrt(ratio) + ratio;
  }
}

// my_InitUnigramTable (dsgl.hpp:121-147).  sum(cn^0.75) is a double sum
// split in `threads` static chunks (libgomp's schedule(static) split) whose
// partials are combined in thread order; the running CDF d1 is a float that
// advances by at most one rank per slot.
double orc_unigram_pow_sum(const int32_t* cn_rank, int32_t V, int32_t threads) {
  if (threads < 1) threads = 1;
  double total = 0.;
  int64_t n = V;
  int64_t q = n / threads, t = n % threads;
  for (int th = 0; th < threads; th++) {
    int64_t qq = q, tt = t;
    if (th < tt) { tt = 0; qq++; }
    int64_t s0 = qq * th + tt, e0 = s0 + qq;
    double part = 0.;
    for (int64_t i = s0; i < e0; i++) part += std::pow((double)cn_rank[i], (double)0.75f);
    total += part;
  }
  return total;
}

int orc_unigram_table(const int32_t* cn_rank, int32_t V, int64_t table_size, int32_t threads,
                      int32_t* table) {
  if (V <= 0 || table_size <= 0) return -1;
  double pw = orc_unigram_pow_sum(cn_rank, V, threads);
  int i = 0;
  float d1 = (float)(std::pow((double)cn_rank[i], (double)0.75f) / pw);
  for (int64_t a = 0; a < table_size; a++) {
    table[a] = i;
    if ((float)a / (float)table_size > d1) {
      i++;
      if (i >= V) i = V - 1;
      d1 = (float)((double)d1 + std::pow((double)cn_rank[i], (double)0.75f) / pw);
    }
  }
  return 0;
}

// ------------------------------------------------------------ trainer ----
typedef struct orc_opts {
  int32_t dim;          // -size
  int32_t window;       // -window
  int32_t negative;     // -negative
  int32_t iter;         // -iter
  int32_t threads;      // -threads (only the pow-sum split uses it)
  int32_t workers;      // compute threads = reference -threads minus one
  float sample;         // -sample
  float alpha;          // -alpha
  int64_t table_size;   // 1e8 in the reference
  int32_t init_model;   // 1: InitNet + zero Woh; 0: use the given matrices
  int32_t parallel;     // 0: workers run one after another; 1: OpenMP Hogwild
  int32_t semantics;    // 0: the reference's buffers; 1: window slots (the BF16-MFMA
                        //    window-slot trainer's update rules, see run_worker_slots)
  int32_t share_shift;  // window slots only: 2^share_shift consecutive steps draw from
                        //    one negative group (0 = the reference's one group per step)
  int32_t minibatch;    // window slots only: contexts per minibatch (0 = a step at once)
} orc_opts;

typedef struct orc_stats {
  uint64_t steps;         // minibatch steps (2 target positions each)
  uint64_t trained_words; // 2 * steps
  uint64_t raw_words;     // tokens consumed, separators and discards included
  uint64_t sentences;     // sentences trained to their end
  float final_alpha;
  int32_t vocab_size;
  double train_seconds;   // wall time of the worker loop only (as dsgl.hpp:1571-1573)
} orc_stats;

int32_t orc_record_stride(int32_t window, int32_t negative) {
  return 8 + 4 * window + 2 * (negative + 2);
}

}  // extern "C"

namespace {

struct Sentence {
  int len = 0, pos = 0;
  std::vector<int> tok;    // ranks (sen / sen2)
  std::vector<int> uniq;   // sen_words: distinct ranks in first-occurrence order
  std::vector<int> slot;   // per position: index into uniq
  std::vector<float> buf;  // inputBuffer: uniq.size() x D working copy of Wih rows
};

struct Shared {
  const orc_opts* o;
  const Tables* t;
  const int32_t* corpus;
  u64 n_tokens;
  float* wih;
  float* woh;
  std::atomic<u64> word_count_actual{0};
};

struct WorkerResult {
  u64 steps = 0, raw = 0, sentences = 0;
  float alpha = 0.f;
};

// One compute thread of Train_SGNS_MPI (dsgl.hpp:786-1408) with id = w + 1.
WorkerResult run_worker(Shared& sh, int w, int32_t* rec, int64_t rec_cap) {
  const orc_opts& o = *sh.o;
  const Tables& t = *sh.t;
  const int D = o.dim, W = o.window, K = o.negative;
  const int nparts = o.workers;
  const u64 N = sh.n_tokens;
  const int stride = orc_record_stride(W, K);
  WorkerResult res;

  // Corpus slice and seed (dsgl.hpp:791, :811-821).  One process: rank 0.
  u64 per = N / (u64)nparts;
  u64 start = per * (u64)w;
  u64 count = per;
  if (N % (u64)nparts > 0 && w == nparts - 1) count = N - (u64)(nparts - 1) * per;
  const u64 end = start + count;
  u64 rng = (u64)w;

  const float a0 = o.alpha;
  float alpha = a0;
  int iters_left = o.iter;
  u64 wc = 0, last_wc = 0, cur = start;

  Sentence s[2];
  for (auto& x : s) { x.tok.assign(kMaxSentence, 0); x.slot.assign(kMaxSentence, 0); }
  int done_sentences = 0;
  bool refresh = true;
  int neg_count = 0, nbn = 1;
  std::vector<int> neg_words, neg_uniq;
  std::vector<float> negbuf;

  const int max_in = 4 * W;
  std::vector<int> inputs(max_in), in_sent(max_in), in_slot(max_in);
  std::vector<int> oidx(K + 2), ometa(K + 2), oslot(K + 2);
  std::vector<float> O((size_t)(K + 2) * D), I((size_t)max_in * D);
  std::vector<float> corr((size_t)(K + 2) * max_in), md((size_t)(K + 2) * D), gin((size_t)max_in * D);

  // Loads the next sentence of length >= 2 into x (dsgl.hpp:902-965).
  auto load = [&](Sentence& x) {
    while (x.len == 0) {
      if (cur >= end || cur >= N) break;
      for (;;) {
        int id = sh.corpus[cur];
        wc++;
        cur++;
        if (id == -1 || cur >= end) break;
        int r = t.id_rank[id];
        if (o.sample > 0) {
          rng = lcg_next(rng);
          if (t.keep_thr[r] < (float)(rng & 0xFFFF) / 65536.f) continue;
        }
        x.tok[x.len++] = r;
        if (x.len >= kMaxSentence) break;
      }
      x.pos = 0;
      if (x.len == 1 || x.len == 0) { x.len = 0; continue; }
      x.uniq.clear();
      for (int p = 0; p < x.len; p++) {
        auto it = std::find(x.uniq.begin(), x.uniq.end(), x.tok[p]);
        if (it == x.uniq.end()) { x.slot[p] = (int)x.uniq.size(); x.uniq.push_back(x.tok[p]); }
        else x.slot[p] = (int)(it - x.uniq.begin());
      }
      x.buf.resize(x.uniq.size() * (size_t)D);
      for (size_t u = 0; u < x.uniq.size(); u++)
        std::memcpy(&x.buf[u * D], sh.wih + (size_t)x.uniq[u] * D, sizeof(float) * D);
    }
  };

  for (;;) {
    // Learning-rate decay in >10000-token chunks (dsgl.hpp:886-897).
    if (wc - last_wc > 10000) {
      u64 diff = wc - last_wc;
      u64 total = sh.word_count_actual.fetch_add(diff) + diff;
      last_wc = wc;
      alpha = a0 * (1 - (float)(total * 1ULL) / (float)((u64)o.iter * t.train_words + 1));
      if (alpha < a0 * 0.0001f) alpha = a0 * 0.0001f;
    }
    load(s[0]);
    load(s[1]);
    // End of the slice: in-flight sentences are dropped (dsgl.hpp:1031-1050).
    if (cur >= end || cur >= N) {
      sh.word_count_actual.fetch_add(wc - last_wc);
      res.raw += wc;
      if (--iters_left == 0) break;
      wc = 0; cur = start; last_wc = 0;
      s[0].len = 0; s[1].len = 0;
      continue;
    }
    // Shared negative batch for the sentence pair (dsgl.hpp:1054-1088).
    if (refresh) {
      neg_words.clear();
      neg_uniq.clear();
      nbn = std::max(s[0].len, s[1].len);
      for (int i = 0; i < nbn; i++) {
        for (int k = 0; k < K; k++) {
          rng = lcg_next(rng);
          int v = t.table[(rng >> 16) % (u64)t.table_size];
          if (!v) v = (int)(rng % (u64)(t.V - 1) + 1);
          neg_words.push_back(v);
          if (std::find(neg_uniq.begin(), neg_uniq.end(), v) == neg_uniq.end()) neg_uniq.push_back(v);
        }
      }
      negbuf.resize(neg_uniq.size() * (size_t)D);
      for (size_t u = 0; u < neg_uniq.size(); u++)
        std::memcpy(&negbuf[u * D], sh.woh + (size_t)neg_uniq[u] * D, sizeof(float) * D);
      refresh = false;
    }
    // One step: both sentences' current targets share one window draw and
    // one group of K negatives (dsgl.hpp:1089-1160).
    const int t1 = s[0].tok[s[0].pos];
    const int t2 = s[1].tok[s[1].pos];
    rng = lcg_next(rng);
    const int b = (int)(rng % (u64)W);
    int nin = 0, fin = 0;
    for (int q = 0; q < 2; q++) {
      const Sentence& x = s[q];
      for (int i = b; i < 2 * W + 1 - b; i++) {
        if (i == W) continue;
        int c = x.pos - W + i;
        if (c < 0) continue;
        if (c >= x.len) break;
        inputs[nin] = x.tok[c];
        in_sent[nin] = q;
        in_slot[nin] = x.slot[c];
        nin++;
      }
      if (q == 0) fin = nin;
    }
    neg_count++;
    const int grp = neg_count % nbn;
    int os = 1;
    oidx[0] = t1;
    ometa[0] = 1;
    for (int k = 0; k < K; k++) {
      int v = neg_words[(size_t)grp * K + k];
      int hit = -1;
      for (int i = 0; i < os; i++) if (oidx[i] == v) { hit = i; break; }
      if (hit < 0) { oidx[os] = v; ometa[os] = 1; os++; }
      else ometa[hit]++;
    }
    ometa[0] = 1;            // a negative equal to the target is dropped
    oidx[os] = t2;
    ometa[os] = 1;
    os++;

    // Stage rows (dsgl.hpp:1167-1211).
    for (int j = 0; j < nin; j++) {
      const Sentence& x = s[in_sent[j]];
      std::memcpy(&I[(size_t)j * D], &x.buf[(size_t)in_slot[j] * D], sizeof(float) * D);
    }
    std::memcpy(&O[0], sh.woh + (size_t)t1 * D, sizeof(float) * D);
    std::memcpy(&O[(size_t)(os - 1) * D], sh.woh + (size_t)t2 * D, sizeof(float) * D);
    for (int i = 1; i < os - 1; i++) {
      oslot[i] = (int)(std::find(neg_uniq.begin(), neg_uniq.end(), oidx[i]) - neg_uniq.begin());
      std::memcpy(&O[(size_t)i * D], &negbuf[(size_t)oslot[i] * D], sizeof(float) * D);
    }

    if (rec && (int64_t)res.steps < rec_cap) {
      int32_t* r = rec + (size_t)res.steps * stride;
      for (int z = 0; z < stride; z++) r[z] = -1;
      r[0] = t1; r[1] = t2; r[2] = b; r[3] = fin; r[4] = nin; r[5] = os;
      std::memcpy(&r[6], &alpha, 4);
      r[7] = grp;
      for (int j = 0; j < nin; j++) r[8 + j] = inputs[j];
      for (int i = 0; i < os; i++) { r[8 + 4 * W + i] = oidx[i]; r[8 + 4 * W + K + 2 + i] = ometa[i]; }
    }

    step_math<float>(D, os, nin, fin, ometa.data(), alpha, t.exp_table.data(), O.data(), I.data(),
                     corr.data(), md.data(), gin.data(), nullptr);

    // Scatter (dsgl.hpp:1314-1368): contexts into the sentence buffers in
    // order (duplicates accumulate), targets straight into Woh, negatives into
    // the shared negative buffer.
    for (int j = 0; j < nin; j++) {
      Sentence& x = s[in_sent[j]];
      float* dst = &x.buf[(size_t)in_slot[j] * D];
      const float* src = &gin[(size_t)j * D];
      for (int d = 0; d < D; d++) dst[d] += src[d];
    }
    {
      float* w1 = sh.woh + (size_t)t1 * D;
      float* w2 = sh.woh + (size_t)t2 * D;
      const float* m1 = &md[0];
      const float* m2 = &md[(size_t)(os - 1) * D];
      for (int d = 0; d < D; d++) { w1[d] += m1[d]; w2[d] += m2[d]; }
    }
    for (int i = 1; i < os - 1; i++) {
      float* dst = &negbuf[(size_t)oslot[i] * D];
      const float* src = &md[(size_t)i * D];
      for (int d = 0; d < D; d++) dst[d] += src[d];
    }
    res.steps++;

    // Write-backs (dsgl.hpp:1377-1407): finished sentences' buffers overwrite
    // their Wih rows; after two finished sentences the negative buffer
    // overwrites its Woh rows and a new negative batch is drawn.
    for (int q = 0; q < 2; q++) {
      Sentence& x = s[q];
      x.pos++;
    }
    for (int q = 0; q < 2; q++) {
      Sentence& x = s[q];
      if (x.pos >= x.len) {
        for (size_t u = 0; u < x.uniq.size(); u++)
          std::memcpy(sh.wih + (size_t)x.uniq[u] * D, &x.buf[u * D], sizeof(float) * D);
        x.len = 0;
        done_sentences++;
        res.sentences++;
      }
    }
    if (done_sentences >= 2) {
      for (size_t u = 0; u < neg_uniq.size(); u++)
        std::memcpy(sh.woh + (size_t)neg_uniq[u] * D, &negbuf[u * D], sizeof(float) * D);
      refresh = true;
      done_sentences = 0;
    }
  }
  res.alpha = alpha;
  return res;
}

// One compute thread with the window-slot update rules of the library's
// BF16-MFMA trainer (gdistger-re_amd/csrc/sgns_fast.hip), in FP32.  The index
// stream is the reference's (same loads, draws, windows and output lists as
// run_worker, dsgl.hpp:786-1160); the updates differ from dsgl.hpp:1314-1407:
//   * a rank's working row exists while one of its positions is inside the
//     window [p-W, p+W]; it is read from Wih when the rank enters (rows of
//     the first window at the sentence load, an entering row at the start of
//     the step whose transition brings it in) and written back to Wih when
//     its last in-window position leaves, or at the sentence end;
//   * target and negative rows are read from Woh at the step and updated
//     there directly (no negative buffer);
//   * repeated contexts share their row: one coefficient times their count.
// Two batching options change the training itself (the library's -share and
// -batch-parity 0; with share_shift = minibatch = 0 the stream is the
// reference's):
//   * share_shift h: a refresh draws ceil(max(L1, L2) / 2^h) groups of K and
//     step n uses group (n >> h) mod that count;
//   * minibatch m > 0: the step's context list (first window, then second,
//     in order) is split into runs of at most m; each run reads the output
//     rows from Woh again and updates them and the working rows before the
//     next run is formed.
WorkerResult run_worker_slots(Shared& sh, int w, int32_t* rec, int64_t rec_cap) {
  const orc_opts& o = *sh.o;
  const Tables& t = *sh.t;
  const int D = o.dim, W = o.window, K = o.negative;
  const int nparts = o.workers;
  const u64 N = sh.n_tokens;
  const int stride = orc_record_stride(W, K);
  WorkerResult res;
  u64 per = N / (u64)nparts;
  u64 start = per * (u64)w;
  u64 count = per;
  if (N % (u64)nparts > 0 && w == nparts - 1) count = N - (u64)(nparts - 1) * per;
  const u64 end = start + count;
  u64 rng = (u64)w;
  const float a0 = o.alpha;
  float alpha = a0;
  int iters_left = o.iter;
  u64 wc = 0, last_wc = 0, cur = start;

  struct Win {
    int len = 0, pos = 0;
    std::vector<int> tok;
    std::vector<int> rank;    // ranks with a row in the window
    std::vector<float> row;   // their rows (rank.size() x D)
    int find(int r) const {
      for (size_t i = 0; i < rank.size(); i++) if (rank[i] == r) return (int)i;
      return -1;
    }
  };
  Win s[2];
  for (auto& x : s) x.tok.assign(kMaxSentence, 0);
  int done_sentences = 0;
  bool refresh = true;
  int neg_count = 0, nbn = 1;
  std::vector<int> neg_words;
  std::vector<int> oidx(K + 2), ometa(K + 2);
  std::vector<float> O((size_t)(K + 2) * D), md((size_t)(K + 2) * D);

  auto in_window_before = [&](const Win& x, int y, int lo) {  // rank of y at a position in [lo, y)
    for (int c = std::max(lo, 0); c < y; c++) if (x.tok[c] == x.tok[y]) return true;
    return false;
  };
  auto write_back = [&](Win& x, int e) {
    std::memcpy(sh.wih + (size_t)x.rank[e] * D, &x.row[(size_t)e * D], sizeof(float) * D);
    x.rank.erase(x.rank.begin() + e);
    x.row.erase(x.row.begin() + (size_t)e * D, x.row.begin() + (size_t)(e + 1) * D);
  };
  auto add_row = [&](Win& x, int r, const float* src) {
    x.rank.push_back(r);
    x.row.insert(x.row.end(), src, src + D);
  };
  auto load = [&](Win& x) {
    while (x.len == 0) {
      if (cur >= end || cur >= N) break;
      for (;;) {
        int id = sh.corpus[cur];
        wc++;
        cur++;
        if (id == -1 || cur >= end) break;
        int r = t.id_rank[id];
        if (o.sample > 0) {
          rng = lcg_next(rng);
          if (t.keep_thr[r] < (float)(rng & 0xFFFF) / 65536.f) continue;
        }
        x.tok[x.len++] = r;
        if (x.len >= kMaxSentence) break;
      }
      x.pos = 0;
      if (x.len == 1 || x.len == 0) { x.len = 0; continue; }
      x.rank.clear();
      x.row.clear();
      for (int c = 0; c <= std::min(W, x.len - 1); c++)
        if (x.find(x.tok[c]) < 0) add_row(x, x.tok[c], sh.wih + (size_t)x.tok[c] * D);
    }
  };

  std::vector<float> in_row[2] = {std::vector<float>(D), std::vector<float>(D)};
  for (;;) {
    if (wc - last_wc > 10000) {
      u64 diff = wc - last_wc;
      u64 total = sh.word_count_actual.fetch_add(diff) + diff;
      last_wc = wc;
      alpha = a0 * (1 - (float)(total * 1ULL) / (float)((u64)o.iter * t.train_words + 1));
      if (alpha < a0 * 0.0001f) alpha = a0 * 0.0001f;
    }
    load(s[0]);
    load(s[1]);
    if (cur >= end || cur >= N) {
      sh.word_count_actual.fetch_add(wc - last_wc);
      res.raw += wc;
      if (--iters_left == 0) break;
      wc = 0; cur = start; last_wc = 0;
      s[0].len = 0; s[1].len = 0;
      continue;
    }
    if (refresh) {
      neg_words.clear();
      nbn = (std::max(s[0].len, s[1].len) + (1 << o.share_shift) - 1) >> o.share_shift;
      for (int i = 0; i < nbn; i++)
        for (int k = 0; k < K; k++) {
          rng = lcg_next(rng);
          int v = t.table[(rng >> 16) % (u64)t.table_size];
          if (!v) v = (int)(rng % (u64)(t.V - 1) + 1);
          neg_words.push_back(v);
        }
      refresh = false;
    }
    const int t1 = s[0].tok[s[0].pos];
    const int t2 = s[1].tok[s[1].pos];
    rng = lcg_next(rng);
    const int b = (int)(rng % (u64)W);
    // contexts of both windows in order: sentence and position
    int nin = 0, fin = 0;
    std::vector<int> inputs, in_q, in_c;
    for (int q = 0; q < 2; q++) {
      const Win& x = s[q];
      for (int i = b; i < 2 * W + 1 - b; i++) {
        if (i == W) continue;
        int c = x.pos - W + i;
        if (c < 0) continue;
        if (c >= x.len) break;
        inputs.push_back(x.tok[c]);
        in_q.push_back(q);
        in_c.push_back(c);
        nin++;
      }
      if (q == 0) fin = nin;
    }
    neg_count++;
    const int grp = (int)(((unsigned)neg_count >> o.share_shift) % (unsigned)nbn);
    int os = 1;
    oidx[0] = t1;
    ometa[0] = 1;
    for (int k = 0; k < K; k++) {
      int v = neg_words[(size_t)grp * K + k];
      int hit = -1;
      for (int i = 0; i < os; i++) if (oidx[i] == v) { hit = i; break; }
      if (hit < 0) { oidx[os] = v; ometa[os] = 1; os++; }
      else ometa[hit]++;
    }
    ometa[0] = 1;
    oidx[os] = t2;
    ometa[os] = 1;
    os++;
    // rows entering the windows at this step's transition, read now
    bool enters[2] = {false, false};
    for (int q = 0; q < 2; q++) {
      const Win& x = s[q];
      const int p = x.pos, y = p + W + 1;
      if (p + 1 < x.len && y < x.len && !in_window_before(x, y, p - W)) {
        enters[q] = true;
        std::memcpy(in_row[q].data(), sh.wih + (size_t)x.tok[y] * D, sizeof(float) * D);
      }
    }
    if (rec && (int64_t)res.steps < rec_cap) {
      int32_t* r = rec + (size_t)res.steps * stride;
      for (int z = 0; z < stride; z++) r[z] = -1;
      r[0] = t1; r[1] = t2; r[2] = b; r[3] = fin; r[4] = nin; r[5] = os;
      std::memcpy(&r[6], &alpha, 4);
      r[7] = grp;
      for (int j = 0; j < nin; j++) r[8 + j] = inputs[j];
      for (int i = 0; i < os; i++) { r[8 + 4 * W + i] = oidx[i]; r[8 + 4 * W + K + 2 + i] = ometa[i]; }
    }
    const int mbs = o.minibatch > 0 ? o.minibatch : std::max(nin, 1);
    for (int jb = 0; jb < std::max(nin, 1); jb += mbs) {
    const int je = std::min(nin, jb + mbs);
    // this minibatch's contexts as distinct rows with counts
    std::vector<int> ctx_q, ctx_e, ctx_mult;
    for (int j = jb; j < je; j++) {
      const int q = in_q[j];
      const int e = s[q].find(s[q].tok[in_c[j]]);
      size_t u = 0;
      while (u < ctx_e.size() && !(ctx_q[u] == q && ctx_e[u] == e)) u++;
      if (u == ctx_e.size()) { ctx_q.push_back(q); ctx_e.push_back(e); ctx_mult.push_back(0); }
      ctx_mult[u]++;
    }
    for (int i = 0; i < os; i++) std::memcpy(&O[(size_t)i * D], sh.woh + (size_t)oidx[i] * D, sizeof(float) * D);
    // coefficients over the distinct context rows (count times the table
    // gradient), GEMM2 on the rows before the update, then GEMM3
    const size_t nu = ctx_e.size();
    std::vector<float> C(nu * (size_t)os);
    for (size_t u = 0; u < nu; u++) {
      const float* x = &s[ctx_q[u]].row[(size_t)ctx_e[u] * D];
      for (int i = 0; i < os; i++) {
        float f = 0.f;
        for (int k = 0; k < D; k++) f = ORC_FMA(O[(size_t)i * D + k], x[k], f);
        const int label = ctx_q[u] == 0 ? (i == 0) : (i == os - 1);
        C[u * os + i] = sig_grad<float>(f, label, alpha, t.exp_table.data(), nullptr) * (float)(ometa[i] * ctx_mult[u]);
      }
    }
    for (int i = 0; i < os; i++)
      for (int d = 0; d < D; d++) {
        float acc = 0.f;
        for (size_t u = 0; u < nu; u++) acc = ORC_FMA(C[u * os + i], s[ctx_q[u]].row[(size_t)ctx_e[u] * D + d], acc);
        md[(size_t)i * D + d] = acc;
      }
    for (size_t u = 0; u < nu; u++) {
      float* x = &s[ctx_q[u]].row[(size_t)ctx_e[u] * D];
      for (int d = 0; d < D; d++) {
        float acc = 0.f;
        for (int i = 0; i < os; i++) acc = ORC_FMA(C[u * os + i], O[(size_t)i * D + d], acc);
        x[d] += acc;
      }
    }
    for (int i = 0; i < os; i++) {
      float* dst = sh.woh + (size_t)oidx[i] * D;
      for (int d = 0; d < D; d++) dst[d] += md[(size_t)i * D + d];
    }
    }
    res.steps++;
    // transition
    for (int q = 0; q < 2; q++) {
      Win& x = s[q];
      const int p = x.pos;
      x.pos = p + 1;
      if (p + 1 >= x.len) {
        while (!x.rank.empty()) write_back(x, 0);
        x.len = 0;
        done_sentences++;
        res.sentences++;
        continue;
      }
      const int xo = p - W, y = p + W + 1;
      if (xo >= 0) {
        bool stays = false;
        for (int c = xo + 1; c <= std::min(y, x.len - 1); c++) if (x.tok[c] == x.tok[xo]) stays = true;
        if (!stays) write_back(x, x.find(x.tok[xo]));
      }
      if (enters[q]) add_row(x, x.tok[y], in_row[q].data());
    }
    if (done_sentences >= 2) {
      refresh = true;
      done_sentences = 0;
    }
  }
  res.alpha = alpha;
  return res;
}

int prepare_tables(const orc_opts* o, const int32_t* cn_by_id, int32_t V, Tables* t) {
  if (V < 2 || o->dim <= 0 || o->window <= 0 || o->negative < 0) return -1;
  t->V = V;
  t->D = o->dim;
  t->rank_id.resize(V);
  t->id_rank.resize(V);
  orc_build_vocab(cn_by_id, V, t->rank_id.data(), t->id_rank.data(), &t->train_words);
  t->cn_by_id = cn_by_id;
  t->exp_table.resize(kExpTableSize + 1);
  orc_exp_table(t->exp_table.data());
  t->keep_thr.resize(V);
  orc_keep_thresholds(cn_by_id, V, t->train_words, o->sample, t->keep_thr.data());
  std::vector<int32_t> cn_rank(V);
  for (int r = 0; r < V; r++) cn_rank[r] = cn_by_id[t->rank_id[r]];
  t->table_size = o->table_size;
  t->table.resize((size_t)o->table_size);
  orc_unigram_table(cn_rank.data(), V, o->table_size, o->threads, t->table.data());
  return 0;
}

}  // namespace

extern "C" {

// Train on an in-memory corpus (ids with -1 separators) the way dsgl() does
// with -threads = workers + 1 on one process.  wih/woh are V x D, rank-major.
// rec: optional per-worker index-stream records (rec_cap steps per worker);
// worker_steps: optional per-worker step counts.
int orc_train(const orc_opts* o, const int32_t* corpus, uint64_t n_tokens, const int32_t* cn_by_id,
              int32_t V, float* wih, float* woh, int32_t* rec, int64_t rec_cap,
              uint64_t* worker_steps, orc_stats* stats) {
  if (o->workers < 1 || o->iter < 1) return -1;
  if (o->share_shift < 0 || o->share_shift > 3 || o->minibatch < 0) return -1;
  if (!o->semantics && (o->share_shift || o->minibatch)) return -1;  // the reference batches one way
  for (uint64_t i = 0; i < n_tokens; i++)
    if (corpus[i] < -1 || corpus[i] >= V) return -2;
  Tables t;
  if (prepare_tables(o, cn_by_id, V, &t)) return -1;
  if (o->init_model) {
    orc_init_net(V, o->dim, wih);
    std::memset(woh, 0, sizeof(float) * (size_t)V * o->dim);
  }
  Shared sh;
  sh.o = o;
  sh.t = &t;
  sh.corpus = corpus;
  sh.n_tokens = n_tokens;
  sh.wih = wih;
  sh.woh = woh;
  const int stride = orc_record_stride(o->window, o->negative);
  std::vector<WorkerResult> res(o->workers);
  const auto t0 = std::chrono::steady_clock::now();
  if (o->parallel) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 1) num_threads(o->workers)
#endif
    for (int w = 0; w < o->workers; w++)
      res[w] = (o->semantics ? run_worker_slots : run_worker)(sh, w, rec ? rec + (size_t)w * rec_cap * stride : nullptr,
                                                             rec_cap);
  } else {
    for (int w = 0; w < o->workers; w++)
      res[w] = (o->semantics ? run_worker_slots : run_worker)(sh, w, rec ? rec + (size_t)w * rec_cap * stride : nullptr,
                                                             rec_cap);
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->train_seconds = secs;
    for (int w = 0; w < o->workers; w++) {
      stats->steps += res[w].steps;
      stats->raw_words += res[w].raw;
      stats->sentences += res[w].sentences;
    }
    stats->trained_words = 2 * stats->steps;
    stats->final_alpha = res[o->workers - 1].alpha;
    stats->vocab_size = V;
  }
  if (worker_steps)
    for (int w = 0; w < o->workers; w++) worker_steps[w] = res[w].steps;
  return 0;
}

// One minibatch step in FP32 (the oracle's fixed sequential-FMA order).
void orc_step_f32(int32_t D, int32_t os, int32_t nin, int32_t fin, const int32_t* meta, float alpha,
                  const float* exp_table, const float* O, const float* I, float* corr, float* md,
                  float* gin, int32_t* sig_idx) {
  step_math<float>(D, os, nin, fin, meta, alpha, exp_table, O, I, corr, md, gin, sig_idx);
}

// The same step in FP64 (same table sigmoid, index taken from the FP64 dot).
void orc_step_f64(int32_t D, int32_t os, int32_t nin, int32_t fin, const int32_t* meta, double alpha,
                  const float* exp_table, const double* O, const double* I, double* corr,
                  double* md, double* gin, int32_t* sig_idx) {
  step_math<double>(D, os, nin, fin, meta, alpha, exp_table, O, I, corr, md, gin, sig_idx);
}

}  // extern "C"
