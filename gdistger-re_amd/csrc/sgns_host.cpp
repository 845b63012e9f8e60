// Host layer of the SGNS trainer: the C ABI of include/dger_sgns.h.
//
// Reference boundary replaced: dsgl() (include/dsgl.hpp:1464-1584) with its
// preparatory_work() (dsgl.hpp:511-519) and Train_SGNS_MPI()
// (dsgl.hpp:522-1427).  Host work here is the O(V) part the reference also
// does serially (vocabulary sort, sigmoid table, the per-rank boundaries of
// the noise table); every O(corpus) or O(V*D) pass runs on the GPU.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dger_sgns.h"
#include "sgns_internal.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define DGER_HIP(expr)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(DGER_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

// Device allocation that frees itself.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, n ? n : 1);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// dsgl.hpp:1429-1441: index of the exact-match flag, or -1.
int arg_pos(const char* flag, int argc, char** argv) {
  for (int a = 1; a < argc; a++)
    if (argv[a] && !std::strcmp(flag, argv[a])) return a;
  return -1;
}

// word2vec sigmoid table, dsgl.hpp:1557-1562 (float exp of a float argument).
void build_exp_table(float* t) {
  for (int i = 0; i < dger::kExpTable; i++) {
    float e = std::exp((i / (float)1000 * 2 - 1) * 6);
    t[i] = e / (e + 1);
  }
}

// sum(cn^0.75) in double over `threads` static chunks combined in order
// (dsgl.hpp:125-132; OpenMP schedule(static) split).
double pow_sum(const std::vector<int32_t>& cn_rank, int threads) {
  if (threads < 1) threads = 1;
  const int64_t n = (int64_t)cn_rank.size();
  const int64_t q = n / threads, t = n % threads;
  double total = 0.;
  for (int th = 0; th < threads; th++) {
    int64_t qq = q, tt = t;
    if (th < tt) { tt = 0; qq++; }
    const int64_t s0 = qq * th + tt, e0 = s0 + qq;
    double part = 0.;
    for (int64_t i = s0; i < e0; i++) part += std::pow((double)cn_rank[i], (double)0.75f);
    total += part;
  }
  return total;
}

// Rank boundaries of the lagged-CDF noise table (dsgl.hpp:134-146).  The
// reference advances at most one rank per slot: the k-th advance happens at
// the first slot a after the previous advance with (float)a/(float)T > d1,
// where d1 is the float running sum.  starts[k] = first slot holding rank k.
std::vector<int64_t> table_starts(const std::vector<int32_t>& cn_rank, int64_t T, int threads) {
  const int V = (int)cn_rank.size();
  const double pw = pow_sum(cn_rank, threads);
  const float Tf = (float)T;
  auto first_above = [&](float d) -> int64_t {
    int64_t lo = 0, hi = T;
    while (lo < hi) {
      const int64_t mid = lo + (hi - lo) / 2;
      if ((float)mid / Tf > d) hi = mid; else lo = mid + 1;
    }
    return lo;
  };
  std::vector<int64_t> starts;
  starts.reserve(V);
  starts.push_back(0);
  float d1 = (float)(std::pow((double)cn_rank[0], (double)0.75f) / pw);
  int64_t prev = -1;
  for (int k = 1; k <= V - 1; k++) {
    const int64_t a = std::max(prev + 1, first_above(d1));
    if (a >= T) break;
    starts.push_back(a + 1);
    prev = a;
    d1 = (float)((double)d1 + std::pow((double)cn_rank[k], (double)0.75f) / pw);
  }
  return starts;
}

size_t row_bytes(int precision) { return precision == DGER_PRECISION_FP32 ? 4 : 2; }

}  // namespace

int dger::host_fail(int code, const std::string& msg) { return fail(code, msg); }

struct dger_ctx {
  dger_sgns_opts o;
  hipStream_t stream = nullptr;      // where device work is queued
  hipStream_t own_stream = nullptr;  // the context's own stream
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t V = 0;
  int D = 0;
  uint64_t train_words = 0;
  std::vector<int32_t> cn, rank_id, id_rank;
  DevBuf id_rank_d, keep_d, table_d, exp_d, wih_d, woh_d, corpus_d, scratch_d, counters_d,
      wsteps_d, walpha_d, rec_d, tmp_d;
  uint64_t n_tokens = 0;
  int Lmax = 0;
  bool have_corpus = false;
  // Captured trainer launch, valid while its arguments are unchanged.
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  dger::TrainArgs graph_args{};
  int tuned_workers = 0;  // autotuner's choice (0 = not tuned)
  // Negative sharing and minibatch size of the runs (the options', or the
  // autotuner's choice under -autotune 2).
  int share = 2, minibatch = 0;
  DevBuf link_pairs_d, link_scores_d, link_labels_d, link_scratch_d, link_counts_d;
  // FP8 form: master rows of Wih / Woh and the two matrix descriptors.
  DevBuf wm_d, om_d, desc_d;
  int n_master = 0;
  // BF16-MFMA form on the window-slot trainer (sgns_fast.hip): set when the
  // shape is one it supports, else the form runs on the buffered trainer.
  bool fast = false;

  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }
  ~dger_ctx() {
    drop_graph();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (own_stream) (void)hipStreamDestroy(own_stream);
  }
};

extern "C" {

const char* dger_last_error(void) { return g_err.c_str(); }

int32_t dger_record_stride(int32_t window, int32_t negative) { return dger::record_stride(window, negative); }

int dger_sgns_default_opts(dger_sgns_opts* o) {
  if (!o) return fail(DGER_E_ARG, "null opts");
  std::memset(o, 0, sizeof(*o));
  // dsgl.hpp:67-76
  o->size = 100;
  o->window = 5;
  o->negative = 5;
  o->batch_size = 11;
  o->sample = 1e-3f;
  o->alpha = 0.1f;
  o->min_count = 5;
  o->iter = 5;
  o->threads = 12;
  o->debug = 2;
  o->binary = 0;
  o->sync_period = 0.1f;
  o->min_sync_words = 1024;
  o->full_sync_times = 0;
  o->message_size = 1024;
  o->workers = 0;
  o->precision = DGER_PRECISION_FP32;
  o->table_size = 100000000LL;
  o->device = 0;
  o->graph = 1;
  o->autotune = 0;
  o->rounding = 1;
  o->masters = 0;
  o->batch_parity = 1;
  o->share = 2;
  return DGER_OK;
}

int dger_sgns_parse_args(int argc, char** argv, dger_sgns_opts* o) {
  int rc = dger_sgns_default_opts(o);
  if (rc) return rc;
  int i;
  auto val = [&](const char* flag) -> const char* {
    i = arg_pos(flag, argc, argv);
    if (i < 0) return nullptr;
    if (i + 1 >= argc || !argv[i + 1]) return "";
    return argv[i + 1];
  };
  auto missing = [&](const char* flag) { return fail(DGER_E_ARG, std::string("missing value for ") + flag); };
  auto copy_str = [&](char* dst, const char* v) {
    std::strncpy(dst, v, 511);
    dst[511] = 0;
  };
  const char* v;
#define DGER_INT(flag, field)                        \
  if ((v = val(flag))) {                             \
    if (!*v) return missing(flag);                   \
    o->field = (int32_t)std::atoi(v);                \
  }
#define DGER_FLT(flag, field)                        \
  if ((v = val(flag))) {                             \
    if (!*v) return missing(flag);                   \
    o->field = (float)std::atof(v);                  \
  }
#define DGER_STR(flag, field)                        \
  if ((v = val(flag))) {                             \
    if (!*v) return missing(flag);                   \
    copy_str(o->field, v);                           \
  }
  DGER_INT("-size", size)
  DGER_STR("-save-vocab", save_vocab)
  DGER_STR("-read-vocab", read_vocab)
  DGER_INT("-debug", debug)
  DGER_INT("-binary", binary)
  DGER_FLT("-alpha", alpha)
  DGER_STR("-eoutput", eoutput)
  DGER_INT("-window", window)
  DGER_FLT("-sample", sample)
  DGER_INT("-negative", negative)
  DGER_INT("-threads", threads)
  DGER_INT("-iter", iter)
  DGER_INT("-min-count", min_count)
  DGER_INT("-batch-size", batch_size)
  if (arg_pos("-disk", argc, argv) > 0) o->disk = 1;
  DGER_FLT("-sync-period", sync_period)
  DGER_INT("-min-sync-words", min_sync_words)
  DGER_INT("-full-sync-times", full_sync_times)
  DGER_INT("-message-size", message_size)
  DGER_STR("-save-emb", save_emb)
  DGER_INT("-workers", workers)
  DGER_INT("-device", device)
  DGER_INT("-graph", graph)
  DGER_INT("-autotune", autotune)
  DGER_INT("-masters", masters)
  DGER_INT("-batch-parity", batch_parity)
  DGER_INT("-share", share)
  if ((v = val("-rounding"))) {
    if (!*v) return missing("-rounding");
    if (!std::strcmp(v, "sr") || !std::strcmp(v, "0")) o->rounding = 0;
    else if (!std::strcmp(v, "rne") || !std::strcmp(v, "1")) o->rounding = 1;
    else return fail(DGER_E_ARG, std::string("unknown -rounding ") + v);
  }
  if ((v = val("-precision"))) {
    if (!*v) return missing("-precision");
    if (!std::strcmp(v, "bf16") || !std::strcmp(v, "1")) o->precision = DGER_PRECISION_BF16;
    else if (!std::strcmp(v, "bf16-mfma") || !std::strcmp(v, "2")) o->precision = DGER_PRECISION_BF16_MFMA;
    else if (!std::strcmp(v, "fp32") || !std::strcmp(v, "0")) o->precision = DGER_PRECISION_FP32;
    else if (!std::strcmp(v, "fp8") || !std::strcmp(v, "3")) o->precision = DGER_PRECISION_FP8;
    else return fail(DGER_E_ARG, std::string("unknown -precision ") + v);
  }
  if ((v = val("-table-size"))) {
    if (!*v) return missing("-table-size");
    o->table_size = std::atoll(v);
  }
#undef DGER_INT
#undef DGER_FLT
#undef DGER_STR
  return DGER_OK;
}

static int check_opts(const dger_sgns_opts* o) {
  if (o->size < 2 || o->size > 1024 || (o->size & 1))
    return fail(DGER_E_UNSUPPORTED, "-size must be even and in [2, 1024]");
  if (o->window < 1 || o->window > 64) return fail(DGER_E_UNSUPPORTED, "-window must be in [1, 64]");
  if (o->negative < 0 || o->negative > 62) return fail(DGER_E_UNSUPPORTED, "-negative must be in [0, 62]");
  if (o->iter < 1) return fail(DGER_E_ARG, "-iter must be >= 1");
  if (o->table_size < 1) return fail(DGER_E_ARG, "-table-size must be >= 1");
  if (o->precision != DGER_PRECISION_FP32 && o->precision != DGER_PRECISION_BF16 &&
      o->precision != DGER_PRECISION_BF16_MFMA && o->precision != DGER_PRECISION_FP8)
    return fail(DGER_E_ARG, "unknown precision");
  // The FP8 form updates whole rows; its shapes must take the register path
  // (at most 64 contexts, 8 output rows) and rows must keep the scale aligned.
  if (o->precision == DGER_PRECISION_FP8 && (o->size % 4 != 0 || o->window > 16 || o->negative > 6))
    return fail(DGER_E_UNSUPPORTED, "-precision fp8 needs -size divisible by 4, -window <= 16, -negative <= 6");
  if (o->workers < 0) return fail(DGER_E_ARG, "-workers must be >= 0");
  if (o->share != 2 && o->share != 4 && o->share != 8 && o->share != 16)
    return fail(DGER_E_ARG, "-share must be 2, 4, 8 or 16");
  if (o->batch_parity != 0 && o->batch_parity != 1) return fail(DGER_E_ARG, "-batch-parity must be 0 or 1");
  if (!o->batch_parity && o->batch_size < 1) return fail(DGER_E_ARG, "-batch-size must be >= 1");
  if (o->autotune < 0 || o->autotune > 2) return fail(DGER_E_ARG, "-autotune must be 0, 1 or 2");
  return DGER_OK;
}

// Minibatch size a run trains with: 0 (a step's contexts at once) under the
// parity switch, where the reference replaces -batch-size by 2*window+1
// (dsgl.hpp:832) and never splits a step (dsgl.hpp:1163-1167); else
// -batch-size, or 0 when no step can exceed it (2 windows of 2*window).
static int run_minibatch(const dger_sgns_opts& o) {
  if (o.batch_parity) return 0;
  return o.batch_size >= 4 * o.window ? 0 : o.batch_size;
}

// Whether matrix m (0: Wih, 1: Woh) holds FP8 codes: both matrices in the
// general trainer's FP8 form, Wih only in the window-slot trainer's (its Woh
// rows are bf16).
static bool coded(const dger_ctx* c, int m) {
  return c->o.precision == DGER_PRECISION_FP8 && (m == 0 || !c->fast);
}

// The model in the context's storage form from FP32 rows (rank order) in
// tmp_d (Wih when m == 0, Woh when m == 1).
static int store_model_from_tmp(dger_ctx* c, int m) {
  const uint64_t n = (uint64_t)c->V * (uint64_t)c->D;
  DevBuf* dst = m ? &c->woh_d : &c->wih_d;
  if (coded(c, m)) {
    DGER_HIP(dger::launch_fp8_quantize(c->tmp_d.as<float>(), c->desc_d.as<dger::Fp8Mat>() + m, c->V, c->D, c->stream));
  } else if (c->o.precision != DGER_PRECISION_FP32) {
    DGER_HIP(dger::launch_f32_to_bf16(c->tmp_d.as<float>(), dst->as<uint16_t>(), n, c->stream));
    if (c->fast && c->n_master > 0)  // the FP32 master rows of the top ranks
      DGER_HIP(hipMemcpyAsync((m ? c->om_d : c->wm_d).p, c->tmp_d.p, (size_t)c->n_master * c->D * 4,
                              hipMemcpyDeviceToDevice, c->stream));
  } else {
    DGER_HIP(hipMemcpyAsync(dst->p, c->tmp_d.p, n * 4, hipMemcpyDeviceToDevice, c->stream));
  }
  return DGER_OK;
}

// FP32 rows (rank order) of Wih (m == 0) or Woh (m == 1) into tmp_d.
static int load_model_to_tmp(dger_ctx* c, int m) {
  const uint64_t n = (uint64_t)c->V * (uint64_t)c->D;
  DGER_HIP(c->tmp_d.ensure(n * 4));
  DevBuf* src = m ? &c->woh_d : &c->wih_d;
  if (coded(c, m)) {
    DGER_HIP(dger::launch_fp8_dequantize(c->desc_d.as<dger::Fp8Mat>() + m, c->tmp_d.as<float>(), c->V, c->D, c->stream));
  } else if (c->o.precision != DGER_PRECISION_FP32) {
    DGER_HIP(dger::launch_bf16_to_f32(src->as<uint16_t>(), c->tmp_d.as<float>(), n, c->stream));
    if (c->fast && c->n_master > 0)  // the top ranks' rows are their FP32 masters
      DGER_HIP(hipMemcpyAsync(c->tmp_d.p, (m ? c->om_d : c->wm_d).p, (size_t)c->n_master * c->D * 4,
                              hipMemcpyDeviceToDevice, c->stream));
  } else {
    DGER_HIP(hipMemcpyAsync(c->tmp_d.p, src->p, n * 4, hipMemcpyDeviceToDevice, c->stream));
  }
  return DGER_OK;
}

int dger_ctx_reset_model(dger_ctx* c) {
  if (!c) return fail(DGER_E_ARG, "null ctx");
  const uint64_t n = (uint64_t)c->V * (uint64_t)c->D;
  if (c->o.precision == DGER_PRECISION_FP8) {
    DGER_HIP(c->tmp_d.ensure(n * 4));
    DGER_HIP(dger::launch_init_net(c->tmp_d.as<float>(), n, c->D, c->stream));
    if (const int rc = store_model_from_tmp(c, 0); rc != DGER_OK) return rc;
    DGER_HIP(hipMemsetAsync(c->woh_d.p, 0, c->woh_d.bytes, c->stream));
    DGER_HIP(hipMemsetAsync(c->om_d.p, 0, c->om_d.bytes, c->stream));
    DGER_HIP(hipStreamSynchronize(c->stream));
    return DGER_OK;
  }
  if (c->o.precision != DGER_PRECISION_FP32) {
    DGER_HIP(c->tmp_d.ensure(n * 4));
    DGER_HIP(dger::launch_init_net(c->tmp_d.as<float>(), n, c->D, c->stream));
    DGER_HIP(dger::launch_f32_to_bf16(c->tmp_d.as<float>(), c->wih_d.as<uint16_t>(), n, c->stream));
    if (c->fast && c->n_master > 0) {
      DGER_HIP(hipMemcpyAsync(c->wm_d.p, c->tmp_d.p, (size_t)c->n_master * c->D * 4, hipMemcpyDeviceToDevice,
                              c->stream));
      DGER_HIP(hipMemsetAsync(c->om_d.p, 0, (size_t)c->n_master * c->D * 4, c->stream));
    }
  } else {
    DGER_HIP(dger::launch_init_net(c->wih_d.as<float>(), n, c->D, c->stream));
  }
  DGER_HIP(hipMemsetAsync(c->woh_d.p, 0, n * row_bytes(c->o.precision), c->stream));
  DGER_HIP(hipStreamSynchronize(c->stream));
  return DGER_OK;
}

int dger_ctx_create(const dger_sgns_opts* opts, const int32_t* vertex_cn, uint32_t v_num, dger_ctx** out) {
  if (!opts || !vertex_cn || !out) return fail(DGER_E_ARG, "null argument");
  *out = nullptr;
  int rc = check_opts(opts);
  if (rc) return rc;
  if (v_num < 2) return fail(DGER_E_VOCAB, "vocabulary needs at least 2 vertices");
  for (uint32_t i = 0; i < v_num; i++)
    if (vertex_cn[i] < 0) return fail(DGER_E_ARG, "negative vertex count");
  int ndev = 0;
  DGER_HIP(hipGetDeviceCount(&ndev));
  if (opts->device < 0 || opts->device >= ndev) return fail(DGER_E_ARG, "device ordinal out of range");
  DGER_HIP(hipSetDevice(opts->device));

  dger_ctx* c = new dger_ctx();
  c->o = *opts;
  c->V = v_num;
  c->D = opts->size;
  c->fast = (opts->precision == DGER_PRECISION_BF16_MFMA || opts->precision == DGER_PRECISION_FP8) &&
            dger::fast_supported(opts->size, opts->negative, opts->window, v_num, opts->table_size);
  auto bail = [&](int code) { delete c; return code; };
  if (opts->batch_parity) c->o.batch_size = 2 * opts->window + 1;  // dsgl.hpp:832
  c->share = opts->share;
  c->minibatch = run_minibatch(c->o);
  if ((c->share != 2 || c->minibatch != 0 || opts->autotune == 2) && !c->fast)
    return bail(fail(DGER_E_UNSUPPORTED,
                     "-share, -batch-parity 0 and -autotune 2 need the window-slot trainer "
                     "(-precision bf16-mfma or fp8, -size 128 or 256, -window <= 11, -negative <= 14)"));
  // The allocation-free entry point's scratch address on this device is
  // looked up once here (dger_ctx_train_noalloc reports a failure).
  if (c->fast) {
    void* sp = nullptr;
    (void)dger::noalloc_scratch(opts->device, &sp);
  }
  if (hipStreamCreate(&c->own_stream) != hipSuccess || hipEventCreate(&c->ev0) != hipSuccess ||
      hipEventCreate(&c->ev1) != hipSuccess)
    return bail(fail(DGER_E_HIP, "stream/event creation failed"));
  c->stream = c->own_stream;

  // Vocabulary (dsgl.hpp:314-382): ranks by count, std::sort with the
  // reference's comparator on (cn, id) records in id order.
  struct Rec { int cn; uint32_t id; };
  std::vector<Rec> v(v_num);
  c->cn.assign(vertex_cn, vertex_cn + v_num);
  c->train_words = 0;
  for (uint32_t i = 0; i < v_num; i++) {
    v[i].cn = vertex_cn[i];
    v[i].id = i;
    c->train_words += (uint64_t)(int64_t)vertex_cn[i];
  }
  std::sort(v.begin(), v.end(), [](Rec a, Rec b) { return a.cn > b.cn; });
  c->rank_id.resize(v_num);
  c->id_rank.resize(v_num);
  std::vector<int32_t> cn_rank(v_num);
  for (uint32_t r = 0; r < v_num; r++) {
    c->rank_id[r] = (int32_t)v[r].id;
    c->id_rank[v[r].id] = (int32_t)r;
    cn_rank[r] = v[r].cn;
  }
  // Subsampling thresholds per rank; cn_vocab is indexed by the rank as in
  // dsgl.hpp:923.
  std::vector<float> keep(v_num);
  for (uint32_t w = 0; w < v_num; w++) {
    const float ratio = (opts->sample * (float)c->train_words) / (float)vertex_cn[w];
    keep[w] = std::sqrt(ratio) + ratio;
  }
  float exp_table[dger::kExpTable];
  build_exp_table(exp_table);
  const std::vector<int64_t> starts = table_starts(cn_rank, opts->table_size, opts->threads);

  const uint64_t n = (uint64_t)v_num * (uint64_t)c->D;
  const size_t rb = row_bytes(opts->precision);
  if (c->id_rank_d.ensure(v_num * 4) || c->keep_d.ensure(v_num * 4) ||
      c->table_d.ensure((size_t)opts->table_size * 4) || c->exp_d.ensure(sizeof(exp_table)) ||
      c->wih_d.ensure(opts->precision == DGER_PRECISION_FP8 ? (size_t)v_num * dger::fp8_row_stride(c->D) : n * rb) ||
      c->woh_d.ensure(coded(c, 1) ? (size_t)v_num * dger::fp8_row_stride(c->D) : n * rb) ||
      c->counters_d.ensure(dger::kNumCounters * 8) || c->tmp_d.ensure(starts.size() * 8))
    return bail(fail(DGER_E_ALLOC, "device allocation failed"));
  if (hipMemcpyAsync(c->id_rank_d.p, c->id_rank.data(), v_num * 4, hipMemcpyHostToDevice, c->stream) ||
      hipMemcpyAsync(c->keep_d.p, keep.data(), v_num * 4, hipMemcpyHostToDevice, c->stream) ||
      hipMemcpyAsync(c->exp_d.p, exp_table, sizeof(exp_table), hipMemcpyHostToDevice, c->stream) ||
      hipMemcpyAsync(c->tmp_d.p, starts.data(), starts.size() * 8, hipMemcpyHostToDevice, c->stream))
    return bail(fail(DGER_E_HIP, "upload failed"));
  if (c->fast && opts->masters) {
    // Window-slot trainer: FP32 masters for the top 1 % ranks by count.
    c->n_master = (int)std::max<uint64_t>(1, ((uint64_t)v_num + 99) / 100);
    if (c->wm_d.ensure((size_t)c->n_master * c->D * 4) || c->om_d.ensure((size_t)c->n_master * c->D * 4))
      return bail(fail(DGER_E_ALLOC, "device allocation failed"));
  }
  if (opts->precision == DGER_PRECISION_FP8) {
    // FP32 masters for the top 1 % ranks by count (at least one row).
    c->n_master = (int)std::max<uint64_t>(1, ((uint64_t)v_num + 99) / 100);
    if (c->wm_d.ensure((size_t)c->n_master * c->D * 4) || c->om_d.ensure((size_t)c->n_master * c->D * 4) ||
        c->desc_d.ensure(2 * sizeof(dger::Fp8Mat)))
      return bail(fail(DGER_E_ALLOC, "device allocation failed"));
    dger::Fp8Mat d[2];
    d[0] = {c->wih_d.as<uint8_t>(), c->wm_d.as<float>(), c->n_master, c->D};
    d[1] = {c->woh_d.as<uint8_t>(), c->om_d.as<float>(), c->n_master, c->D};
    if (hipMemcpy(c->desc_d.p, d, sizeof(d), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(DGER_E_HIP, "upload failed"));
  }
  if (dger::launch_table_fill(c->table_d.as<int32_t>(), (uint64_t)opts->table_size, c->tmp_d.as<int64_t>(),
                              (int)starts.size(), c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(fail(DGER_E_HIP, "noise table fill failed"));
  rc = dger_ctx_reset_model(c);
  if (rc) return bail(rc);
  *out = c;
  return DGER_OK;
}

int dger_ctx_set_stream(dger_ctx* c, void* stream) {
  if (!c) return fail(DGER_E_ARG, "null context");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
  if (stream) {
    hipDevice_t dev = -1;
    DGER_HIP(hipStreamGetDevice(s, &dev));
    if ((int)dev != c->o.device) return fail(DGER_E_ARG, "stream belongs to another device");
  }
  DGER_HIP(hipSetDevice(c->o.device));
  DGER_HIP(hipStreamSynchronize(c->stream));
  c->stream = s;
  return DGER_OK;
}

int dger_ctx_set_corpus(dger_ctx* c, const int32_t* corpus, uint64_t n_tokens) {
  if (!c || (!corpus && n_tokens)) return fail(DGER_E_ARG, "null argument");
  int run = 0, lmax = 0;
  for (uint64_t i = 0; i < n_tokens; i++) {
    const int32_t id = corpus[i];
    if (id < -1 || id >= (int64_t)c->V) return fail(DGER_E_CORPUS, "corpus id out of range at " + std::to_string(i));
    if (id == -1) run = 0;
    else if (++run > lmax) lmax = run;
  }
  c->Lmax = std::max(2, std::min(lmax, dger::kMaxSentence));
  if (c->corpus_d.ensure((n_tokens ? n_tokens : 1) * 4)) return fail(DGER_E_ALLOC, "corpus allocation failed");
  if (n_tokens) DGER_HIP(hipMemcpyAsync(c->corpus_d.p, corpus, n_tokens * 4, hipMemcpyHostToDevice, c->stream));
  DGER_HIP(hipStreamSynchronize(c->stream));
  c->n_tokens = n_tokens;
  c->have_corpus = true;
  return DGER_OK;
}

int dger_ctx_set_corpus_walks(dger_ctx* c, const dger_walk_corpus* w) {
  if (!c || !w) return fail(DGER_E_ARG, "null argument");
  if (w->device != c->o.device) return fail(DGER_E_ARG, "walks are on another device");
  if (w->v_num > c->V) return fail(DGER_E_CORPUS, "walks use ids beyond the context's v_num");
  const uint64_t n_tokens = w->n_tokens;
  c->Lmax = std::max(2, std::min((int)std::min<uint32_t>(w->max_walk, 1u << 30), dger::kMaxSentence));
  if (c->corpus_d.ensure((n_tokens ? n_tokens : 1) * 4)) return fail(DGER_E_ALLOC, "corpus allocation failed");
  if (n_tokens) DGER_HIP(hipMemcpyAsync(c->corpus_d.p, w->d, n_tokens * 4, hipMemcpyDeviceToDevice, c->stream));
  DGER_HIP(hipStreamSynchronize(c->stream));
  c->n_tokens = n_tokens;
  c->have_corpus = true;
  return DGER_OK;
}

// Default worker count: as many workgroups per CU as its LDS holds, at most
// one wavefront per SIMD for the general trainer (it uses more than 256
// registers) and two for the window-slot trainer.
static int auto_workers(dger_ctx* c, size_t block_lds, int wpb) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->o.device) != hipSuccess) return 1024;
  const size_t lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                              : (size_t)160 * 1024;
  int per_cu = (int)std::max<size_t>(1, lds_cu / std::max<size_t>(block_lds, 1)) * wpb;
  per_cu = std::min(per_cu, c->fast ? dger::kFastMaxWpb : 4);
  int w = prop.multiProcessorCount * per_cu;
  // Keep slices long enough that dropped in-flight sentences stay rare.
  const uint64_t max_by_corpus = std::max<uint64_t>(1, c->n_tokens / 4096);
  if ((uint64_t)w > max_by_corpus) w = (int)max_by_corpus;
  return std::max(1, w);
}

static int ensure_worker_buffers(dger_ctx* c, int workers) {
  const size_t per = c->fast ? dger::fast_worker_scratch_ints(c->Lmax, c->o.negative, c->D)
                             : dger::worker_scratch_floats(c->Lmax, c->o.negative, c->D);
  if (c->scratch_d.ensure((size_t)workers * per * 4) || c->wsteps_d.ensure((size_t)workers * 8) ||
      c->walpha_d.ensure((size_t)workers * 4))
    return fail(DGER_E_ALLOC, "scratch allocation failed");
  return DGER_OK;
}

// Kernel arguments of one training pass over the first n_tokens of the corpus
// (value-initialised so that equal passes compare equal byte for byte).
static dger::TrainArgs make_args(dger_ctx* c, int workers, int wpb, int32_t* rec, int64_t rec_cap,
                                 uint64_t n_tokens) {
  dger::TrainArgs a{};
  a.corpus = c->corpus_d.as<int32_t>();
  a.n_tokens = n_tokens;
  a.id_rank = c->id_rank_d.as<int32_t>();
  a.keep_thr = c->keep_d.as<float>();
  a.table = c->table_d.as<int32_t>();
  a.table_size = (uint64_t)c->o.table_size;
  a.exp_table = c->exp_d.as<float>();
  const bool fp8 = c->o.precision == DGER_PRECISION_FP8;
  // (the general trainer reads FP8 matrices through their descriptors, the
  // window-slot trainer the coded rows and the masters directly)
  a.wih = (fp8 && !c->fast) ? (void*)c->desc_d.as<dger::Fp8Mat>() : c->wih_d.p;
  a.woh = (fp8 && !c->fast) ? (void*)(c->desc_d.as<dger::Fp8Mat>() + 1) : c->woh_d.p;
  a.fscratch = c->scratch_d.as<float>();
  a.counters = c->counters_d.as<unsigned long long>();
  a.rec = rec;
  a.rec_cap = rec_cap;
  a.worker_steps = c->wsteps_d.as<unsigned long long>();
  a.worker_alpha = c->walpha_d.as<float>();
  a.V = (int32_t)c->V;
  a.D = c->D;
  a.W = c->o.window;
  a.K = c->o.negative;
  a.iter = c->o.iter;
  a.workers = workers;
  a.Lmax = c->Lmax;
  a.wpb = wpb;
  a.sample_on = c->o.sample > 0 ? 1 : 0;
  a.mfma = (c->o.precision == DGER_PRECISION_BF16_MFMA || fp8) ? 1 : 0;
  a.alpha0 = c->o.alpha;
  a.alpha_den = (float)((uint64_t)c->o.iter * c->train_words + 1);
  if (c->fast) {
    a.wmaster = c->wm_d.as<float>();
    a.omaster = c->om_d.as<float>();
    a.n_master = c->n_master;
    a.round_nearest = c->o.rounding == 1 ? 1 : 0;
    a.share_shift = c->share == 16 ? 3 : c->share == 8 ? 2 : c->share == 4 ? 1 : 0;
    a.minibatch = c->minibatch;
    if (fp8) {  // coded rows, stochastic rounding on every row store
      a.fp8 = 1;
      a.round_nearest = 0;
    }
  }
  return a;
}

// One pass: counter reset (a fill kernel, so that the pair can be captured)
// and the trainer kernel, timed with events on the context's stream.  With
// use_graph the pair is captured once and the graph replayed while the
// arguments stay the same.
static hipError_t launch_pass_kernel(dger_ctx* c, const dger::TrainArgs& a) {
  return c->fast ? dger::launch_train_fast(a, c->stream) : dger::launch_train(a, c->o.precision, c->stream);
}

static int run_pass(dger_ctx* c, const dger::TrainArgs& a, bool use_graph, float* ms) {
  if (c->fast) {
    DGER_HIP(dger::prepare_train_fast(a));
  } else {
    DGER_HIP(dger::prepare_train(a, c->o.precision));
  }
  if (use_graph && (!c->graph_exec || std::memcmp(&c->graph_args, &a, sizeof(a)) != 0)) {
    c->drop_graph();
    DGER_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = dger::launch_zero_u64(a.counters, dger::kNumCounters, c->stream);
    if (e == hipSuccess) e = launch_pass_kernel(c, a);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      c->graph_exec = nullptr;
      return fail(DGER_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    }
    c->graph = g;
    c->graph_args = a;
  }
  DGER_HIP(hipEventRecord(c->ev0, c->stream));
  if (use_graph) {
    DGER_HIP(hipGraphLaunch(c->graph_exec, c->stream));
  } else {
    DGER_HIP(dger::launch_zero_u64(a.counters, dger::kNumCounters, c->stream));
    DGER_HIP(launch_pass_kernel(c, a));
  }
  DGER_HIP(hipEventRecord(c->ev1, c->stream));
  DGER_HIP(hipStreamSynchronize(c->stream));
  DGER_HIP(hipEventElapsedTime(ms, c->ev0, c->ev1));
  if (c->fast && c->n_master > 0) {
    // The bf16 rows of the master ranks (read by link evaluation and the
    // next pass only through their masters) follow their masters; coded
    // rows are read through their masters directly.
    const uint64_t nmd = (uint64_t)c->n_master * c->D;
    if (!coded(c, 0)) DGER_HIP(dger::launch_f32_to_bf16(c->wm_d.as<float>(), c->wih_d.as<uint16_t>(), nmd, c->stream));
    DGER_HIP(dger::launch_f32_to_bf16(c->om_d.as<float>(), c->woh_d.as<uint16_t>(), nmd, c->stream));
    DGER_HIP(hipStreamSynchronize(c->stream));
  }
  return DGER_OK;
}

// Autotuner: times one pass over a corpus prefix per candidate and keeps the
// fastest in trained words per second.  Level 1: one to four workers per CU
// (eight for the window-slot trainer at size 128, one workgroup per CU; at
// size 256 one worker per workgroup, up to the workgroups the CU's LDS
// holds).  Level 2 (the
// window-slot trainer) then also times, at that worker count, negative
// sharing 2, 4, 8 with and without -batch-size 2*window+1 minibatches; these
// change the training, so the caller gates the choice on embedding quality.
// The model is saved before and restored after.
static int autotune_run(dger_ctx* c, int cus, int per_cu_max, int wpb_max, int* out) {
  DevBuf* model[4] = {&c->wih_d, &c->woh_d, &c->wm_d, &c->om_d};
  DevBuf save[4];
  for (int m = 0; m < 4; m++) {
    if (!model[m]->p) continue;
    if (save[m].ensure(model[m]->bytes)) return fail(DGER_E_ALLOC, "autotune: model copy allocation failed");
    DGER_HIP(hipMemcpyAsync(save[m].p, model[m]->p, model[m]->bytes, hipMemcpyDeviceToDevice, c->stream));
  }
  const uint64_t sample = std::min<uint64_t>(c->n_tokens, 16ull << 20);
  auto time_pass = [&](int w, int wpb, double* rate) -> int {
    if (const int rc = ensure_worker_buffers(c, w); rc != DGER_OK) return rc;
    const dger::TrainArgs a = make_args(c, w, wpb, nullptr, 0, sample);
    float ms = 0.f;
    if (const int rc = run_pass(c, a, false, &ms); rc != DGER_OK) return rc;
    unsigned long long steps = 0;
    DGER_HIP(hipMemcpy(&steps, c->counters_d.as<unsigned long long>() + dger::kSteps, 8, hipMemcpyDeviceToHost));
    *rate = ms > 0 ? 2.0 * (double)steps / ms : 0.0;
    return DGER_OK;
  };
  double best = -1.0;
  int best_w = 0, best_wpb = 1;
  for (int per_cu = 1; per_cu <= per_cu_max; per_cu++) {
    const int w = std::max<int>(1, std::min<uint64_t>((uint64_t)cus * per_cu, std::max<uint64_t>(1, sample / 4096)));
    const int wpb = std::min(std::min(per_cu, wpb_max), w);
    double rate = 0.0;
    if (const int rc = time_pass(w, wpb, &rate); rc != DGER_OK) return rc;
    if (c->o.debug > 2) fprintf(stderr, "dger: autotune %d workers: %.1f M words/s\n", w, rate * 1e-3);
    if (rate > best) {
      best = rate;
      best_w = w;
      best_wpb = wpb;
    }
  }
  if (c->o.autotune >= 2 && c->fast) {
    const int share0 = c->share, mb0 = c->minibatch;  // timed above
    int best_share = share0, best_mb = mb0;
    for (int share = 2; share <= 8; share *= 2)
      for (const int mb : {0, 2 * c->o.window + 1}) {
        if (share == share0 && mb == mb0) continue;
        c->share = share;
        c->minibatch = mb;
        double rate = 0.0;
        if (const int rc = time_pass(best_w, best_wpb, &rate); rc != DGER_OK) return rc;
        if (c->o.debug > 2)
          fprintf(stderr, "dger: autotune share %d minibatch %d: %.1f M words/s\n", share, mb, rate * 1e-3);
        if (rate > best) {
          best = rate;
          best_share = share;
          best_mb = mb;
        }
      }
    c->share = best_share;
    c->minibatch = best_mb;
  }
  for (int m = 0; m < 4; m++)
    if (model[m]->p)
      DGER_HIP(hipMemcpyAsync(model[m]->p, save[m].p, model[m]->bytes, hipMemcpyDeviceToDevice, c->stream));
  DGER_HIP(hipStreamSynchronize(c->stream));
  *out = best_w;
  return DGER_OK;
}

// Worker count and workers per workgroup of a training run: `workers` if
// positive, else the options' count, the tuned count, the autotuner (when
// allowed) or the default.
static int pick_workers(dger_ctx* c, int32_t workers, bool allow_autotune, int* out_workers, int* out_wpb) {
  hipDeviceProp_t prop;
  DGER_HIP(hipGetDeviceProperties(&prop, c->o.device));
  // Workgroups may take more LDS than the default per-block figure by opting in.
  const size_t lds_block_max = std::max(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin);
  auto block_lds = [&](int n) {
    return c->fast ? dger::fast_block_lds_bytes(c->Lmax, c->o.negative, n, c->D)
                   : dger::train_block_lds_bytes(c->Lmax, c->o.negative, c->o.window, c->D, c->o.precision, n);
  };
  int wpb = c->fast ? dger::fast_workers_per_block(c->Lmax, c->o.negative, lds_block_max, c->D)
                    : dger::train_workers_per_block(c->Lmax, c->o.negative, c->o.window, c->D, c->o.precision,
                                                    lds_block_max);
  if (wpb < 1)
    return fail(DGER_E_UNSUPPORTED,
                "sentence length / window / negative need " +
                    std::to_string(block_lds(1)) +
                    " bytes of LDS, more than the device allows");
  const int wpb_max = wpb;
  if (workers <= 0) workers = c->o.workers;
  if (workers <= 0 && c->tuned_workers > 0) workers = c->tuned_workers;
  if (workers <= 0 && c->o.autotune && allow_autotune) {
    const size_t lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                                : (size_t)160 * 1024;
    const int per_cu_max =
        (c->fast && dger::fast_halves(c->D) > 1)
            ? (int)std::max<size_t>(1, lds_cu / block_lds(1))
            : std::min(c->fast ? dger::kFastMaxWpb : 4, wpb_max);
    const int rc = autotune_run(c, prop.multiProcessorCount, per_cu_max, wpb_max, &workers);
    if (rc != DGER_OK) return rc;
    c->tuned_workers = workers;
  }
  if (workers <= 0)
    workers = auto_workers(c, block_lds(wpb), wpb);
  wpb = std::min(wpb, workers);
  if (c->o.debug > 2)
    fprintf(stderr, "dger: %d workers, %d per workgroup, LDS %zu bytes per workgroup (limit %zu, opt-in %zu)\n",
            workers, wpb,
            block_lds(wpb),
            (size_t)prop.sharedMemPerBlock, (size_t)prop.sharedMemPerBlockOptin);
  *out_workers = workers;
  *out_wpb = wpb;
  return DGER_OK;
}

// Run statistics from the device counters and the first worker's final rate.
static int read_stats(dger_ctx* c, const dger::TrainArgs& a, float ms, dger_sgns_stats* stats) {
  unsigned long long cnt[dger::kNumCounters];
  DGER_HIP(hipMemcpy(cnt, a.counters, sizeof(cnt), hipMemcpyDeviceToHost));
  float alpha0 = 0.f;
  DGER_HIP(hipMemcpy(&alpha0, a.worker_alpha, 4, hipMemcpyDeviceToHost));
  std::memset(stats, 0, sizeof(*stats));
  stats->steps = cnt[dger::kSteps];
  stats->trained_words = 2 * cnt[dger::kSteps];
  stats->raw_words = cnt[dger::kRawWords];
  stats->sentences = cnt[dger::kSentences];
  stats->contexts = cnt[dger::kContexts];
  stats->out_ctx = cnt[dger::kOutCtx];
  stats->final_alpha = alpha0;
  stats->workers = a.workers;
  stats->vocab_size = (int32_t)c->V;
  stats->max_sentence = c->Lmax;
  stats->share = c->share;
  stats->minibatch = c->minibatch;
  stats->kernel_ms = ms;
  stats->words_per_sec = ms > 0 ? (double)stats->trained_words / (ms * 1e-3) : 0.0;
  return DGER_OK;
}

int dger_ctx_train(dger_ctx* c, int32_t workers, int32_t* rec, int64_t rec_cap, uint64_t* worker_steps,
                   dger_sgns_stats* stats) {
  if (!c) return fail(DGER_E_ARG, "null ctx");
  if (!c->have_corpus) return fail(DGER_E_STATE, "dger_ctx_set_corpus must come first");
  if (rec && rec_cap <= 0) return fail(DGER_E_ARG, "rec_cap must be positive");
  DGER_HIP(hipSetDevice(c->o.device));
  int wpb = 0;
  if (const int rc = pick_workers(c, workers, !rec, &workers, &wpb); rc != DGER_OK) return rc;
  const int stride = dger::record_stride(c->o.window, c->o.negative);
  if (const int rc = ensure_worker_buffers(c, workers); rc != DGER_OK) return rc;
  if (rec && c->rec_d.ensure((size_t)workers * (size_t)rec_cap * stride * 4))
    return fail(DGER_E_ALLOC, "record allocation failed");

  const dger::TrainArgs a = make_args(c, workers, wpb, rec ? c->rec_d.as<int32_t>() : nullptr, rec ? rec_cap : 0,
                                      c->n_tokens);
  float ms = 0.f;
  if (const int rc = run_pass(c, a, c->o.graph != 0 && !rec, &ms); rc != DGER_OK) return rc;

  if (worker_steps)
    DGER_HIP(hipMemcpy(worker_steps, c->wsteps_d.p, (size_t)workers * 8, hipMemcpyDeviceToHost));
  if (rec)
    DGER_HIP(hipMemcpy(rec, c->rec_d.p, (size_t)workers * (size_t)rec_cap * stride * 4, hipMemcpyDeviceToHost));
  if (stats)
    if (const int rc = read_stats(c, a, ms, stats); rc != DGER_OK) return rc;
  return DGER_OK;
}

static size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

// Offsets in the allocation-free entry point's scratch: counters, per-worker
// steps, per-worker final rates, per-worker negative batches; returns the
// bytes needed.
static size_t noalloc_layout(dger_ctx* c, int workers, size_t* off) {
  off[0] = 0;
  off[1] = al256((size_t)dger::kNumCounters * 8);
  off[2] = off[1] + al256((size_t)workers * 8);
  off[3] = off[2] + al256((size_t)workers * 4);
  return off[3] + (size_t)workers * dger::fast_worker_scratch_ints(c->Lmax, c->o.negative, c->D) * 4;
}

int dger_ctx_train_noalloc(dger_ctx* c, int32_t workers, dger_sgns_stats* stats) {
  if (!c) return fail(DGER_E_ARG, "null ctx");
  if (!c->have_corpus) return fail(DGER_E_STATE, "dger_ctx_set_corpus must come first");
  if (!c->fast)
    return fail(DGER_E_UNSUPPORTED,
                "dger_ctx_train_noalloc runs the window-slot trainer only (DGER_PRECISION_BF16_MFMA or "
                "DGER_PRECISION_FP8, size 128 or 256, window <= 11, negative <= 14); use dger_ctx_train");
  DGER_HIP(hipSetDevice(c->o.device));
  void* sp = nullptr;
  DGER_HIP(dger::noalloc_scratch(c->o.device, &sp));
  int wpb = 0;
  if (const int rc = pick_workers(c, workers, false, &workers, &wpb); rc != DGER_OK) return rc;
  size_t off[4];
  const size_t need = noalloc_layout(c, workers, off);
  if (need > dger::kNoAllocScratchBytes) {
    int fit = 0;
    for (int w = 1; w <= workers; w *= 2)
      if (noalloc_layout(c, w, off) <= dger::kNoAllocScratchBytes) fit = w;
    return fail(DGER_E_ARG, std::to_string(workers) + " workers need " + std::to_string(need) +
                                " bytes of scratch; dger_ctx_train_noalloc holds " +
                                std::to_string(dger::kNoAllocScratchBytes) + " (at most " + std::to_string(fit) +
                                " workers fit at this sentence length and negative count)");
  }
  unsigned char* const base = static_cast<unsigned char*>(sp);
  dger::TrainArgs a = make_args(c, workers, wpb, nullptr, 0, c->n_tokens);
  a.counters = reinterpret_cast<unsigned long long*>(base + off[0]);
  a.worker_steps = reinterpret_cast<unsigned long long*>(base + off[1]);
  a.worker_alpha = reinterpret_cast<float*>(base + off[2]);
  a.fscratch = reinterpret_cast<float*>(base + off[3]);
  float ms = 0.f;
  if (const int rc = run_pass(c, a, false, &ms); rc != DGER_OK) return rc;
  if (stats)
    if (const int rc = read_stats(c, a, ms, stats); rc != DGER_OK) return rc;
  return DGER_OK;
}

int dger_ctx_get_model(dger_ctx* c, float* wih, float* woh) {
  if (!c) return fail(DGER_E_ARG, "null ctx");
  DGER_HIP(hipSetDevice(c->o.device));
  const uint64_t n = (uint64_t)c->V * (uint64_t)c->D;
  float* outs[2] = {wih, woh};
  DevBuf* srcs[2] = {&c->wih_d, &c->woh_d};
  for (int m = 0; m < 2; m++) {
    if (!outs[m]) continue;
    if (c->o.precision != DGER_PRECISION_FP32) {
      if (const int rc = load_model_to_tmp(c, m); rc != DGER_OK) return rc;
      DGER_HIP(hipMemcpyAsync(outs[m], c->tmp_d.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    } else {
      DGER_HIP(hipMemcpyAsync(outs[m], srcs[m]->p, n * 4, hipMemcpyDeviceToHost, c->stream));
    }
    DGER_HIP(hipStreamSynchronize(c->stream));
  }
  return DGER_OK;
}

int dger_ctx_set_model(dger_ctx* c, const float* wih, const float* woh) {
  if (!c) return fail(DGER_E_ARG, "null ctx");
  DGER_HIP(hipSetDevice(c->o.device));
  const uint64_t n = (uint64_t)c->V * (uint64_t)c->D;
  const float* ins[2] = {wih, woh};
  DevBuf* dsts[2] = {&c->wih_d, &c->woh_d};
  for (int m = 0; m < 2; m++) {
    if (!ins[m]) continue;
    if (c->o.precision != DGER_PRECISION_FP32) {
      DGER_HIP(c->tmp_d.ensure(n * 4));
      DGER_HIP(hipMemcpyAsync(c->tmp_d.p, ins[m], n * 4, hipMemcpyHostToDevice, c->stream));
      if (const int rc = store_model_from_tmp(c, m); rc != DGER_OK) return rc;
    } else {
      DGER_HIP(hipMemcpyAsync(dsts[m]->p, ins[m], n * 4, hipMemcpyHostToDevice, c->stream));
    }
    DGER_HIP(hipStreamSynchronize(c->stream));
  }
  return DGER_OK;
}

// Uploads n vertex-id pairs and scores them on the device into c->link_scores_d.
static int link_scores_device(dger_ctx* c, const int32_t* pairs, uint64_t n) {
  if (!c) return fail(DGER_E_ARG, "null ctx");
  if (!pairs && n) return fail(DGER_E_ARG, "null pairs");
  for (uint64_t i = 0; i < 2 * n; i++)
    if (pairs[i] < 0 || (uint32_t)pairs[i] >= c->V)
      return fail(DGER_E_ARG, "pair " + std::to_string(i / 2) + " holds a vertex id outside [0, v_num)");
  DGER_HIP(hipSetDevice(c->o.device));
  if (c->link_pairs_d.ensure(n * 8) || c->link_scores_d.ensure(n * 4))
    return fail(DGER_E_ALLOC, "link evaluation buffers");
  DGER_HIP(hipMemcpyAsync(c->link_pairs_d.p, pairs, n * 8, hipMemcpyHostToDevice, c->stream));
  const void* rows = c->wih_d.p;
  bool bf16_rows = c->o.precision != DGER_PRECISION_FP32;
  if (c->o.precision == DGER_PRECISION_FP8) {  // scored from the decoded FP32 rows
    if (const int rc = load_model_to_tmp(c, 0); rc != DGER_OK) return rc;
    rows = c->tmp_d.p;
    bf16_rows = false;
  }
  DGER_HIP(dger::launch_link_scores(rows, bf16_rows, c->id_rank_d.as<int32_t>(),
                                    c->V, c->D, c->link_pairs_d.as<int32_t>(), n, c->link_scores_d.as<float>(),
                                    c->stream));
  return DGER_OK;
}

int dger_ctx_link_scores(dger_ctx* c, const int32_t* pairs, uint64_t n, float* scores) {
  if (!scores && n) return fail(DGER_E_ARG, "null scores");
  if (const int rc = link_scores_device(c, pairs, n); rc != DGER_OK) return rc;
  DGER_HIP(hipMemcpyAsync(scores, c->link_scores_d.p, n * 4, hipMemcpyDeviceToHost, c->stream));
  DGER_HIP(hipStreamSynchronize(c->stream));
  return DGER_OK;
}

int dger_ctx_link_auc(dger_ctx* c, const int32_t* pairs, const int8_t* labels, uint64_t n, double* auc) {
  if (!auc || (!labels && n)) return fail(DGER_E_ARG, "null pointer");
  if (n > (uint64_t)INT32_MAX) return fail(DGER_E_UNSUPPORTED, "at most 2^31-1 pairs");
  if (const int rc = link_scores_device(c, pairs, n); rc != DGER_OK) return rc;
  if (c->link_labels_d.ensure(n ? n : 1) || c->link_scratch_d.ensure(dger::link_auc_scratch_bytes(n)) ||
      c->link_counts_d.ensure(3 * 8))
    return fail(DGER_E_ALLOC, "link evaluation buffers");
  DGER_HIP(hipMemcpyAsync(c->link_labels_d.p, labels, n, hipMemcpyHostToDevice, c->stream));
  uint64_t npos = 0, nneg = 0;
  unsigned long long twice_u = 0;
  DGER_HIP(dger::launch_link_auc(c->link_scores_d.as<float>(), c->link_labels_d.as<int8_t>(), n,
                                 c->link_scratch_d.p, c->link_counts_d.as<unsigned long long>(), &npos, &nneg,
                                 &twice_u, c->stream));
  *auc = (npos && nneg) ? 0.5 * (double)twice_u / ((double)npos * (double)nneg) : std::nan("");
  return DGER_OK;
}

int dger_ctx_get_vocab(dger_ctx* c, int32_t* rank_id, uint64_t* train_words) {
  if (!c) return fail(DGER_E_ARG, "null ctx");
  if (rank_id) std::memcpy(rank_id, c->rank_id.data(), c->rank_id.size() * 4);
  if (train_words) *train_words = c->train_words;
  return DGER_OK;
}

int dger_ctx_get_table(dger_ctx* c, int32_t* table) {
  if (!c || !table) return fail(DGER_E_ARG, "null argument");
  DGER_HIP(hipSetDevice(c->o.device));
  DGER_HIP(hipMemcpy(table, c->table_d.p, (size_t)c->o.table_size * 4, hipMemcpyDeviceToHost));
  return DGER_OK;
}

// my_saveModel (dsgl.hpp:1443-1461): "V D\n", then per rank the vertex id and
// D values ("%f " each, or raw floats with -binary), one line per rank.
int dger_ctx_save(dger_ctx* c, const char* path, int32_t binary) {
  if (!c || !path) return fail(DGER_E_ARG, "null argument");
  std::vector<float> wih((size_t)c->V * c->D);
  const int rc = dger_ctx_get_model(c, wih.data(), nullptr);
  if (rc) return rc;
  return dger_emb_write(path, c->V, c->D, c->rank_id.data(), wih.data(), binary);
}

void dger_ctx_destroy(dger_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->o.device);
  delete c;
}

int dger_sgns_train(const dger_sgns_opts* opts, const int32_t* corpus, uint64_t n_tokens,
                    const int32_t* vertex_cn, uint32_t v_num, float* emb_out, int32_t* rank_id_out,
                    dger_sgns_stats* stats) {
  return dger_sgns_train_stream(opts, corpus, n_tokens, vertex_cn, v_num, emb_out, rank_id_out, stats, nullptr);
}

int dger_sgns_train_stream(const dger_sgns_opts* opts, const int32_t* corpus, uint64_t n_tokens,
                           const int32_t* vertex_cn, uint32_t v_num, float* emb_out, int32_t* rank_id_out,
                           dger_sgns_stats* stats, void* stream) {
  dger_ctx* c = nullptr;
  int rc = dger_ctx_create(opts, vertex_cn, v_num, &c);
  if (rc) return rc;
  rc = dger_ctx_set_stream(c, stream);
  if (!rc) rc = dger_ctx_set_corpus(c, corpus, n_tokens);
  if (!rc) rc = dger_ctx_train(c, 0, nullptr, 0, nullptr, stats);
  if (!rc && emb_out) rc = dger_ctx_get_model(c, emb_out, nullptr);
  if (!rc && rank_id_out) rc = dger_ctx_get_vocab(c, rank_id_out, nullptr);
  dger_ctx_destroy(c);
  return rc;
}

int dger_sgns_main(int argc, char** argv, const int32_t* corpus, uint64_t n_tokens,
                   const int32_t* vertex_cn, uint32_t v_num) {
  dger_sgns_opts o;
  int rc = dger_sgns_parse_args(argc, argv, &o);
  if (rc) return rc;
  if (v_num <= 1000) return fail(DGER_E_VOCAB, "dsgl() needs more than 1000 vertices (dsgl.hpp:370)");
  dger_ctx* c = nullptr;
  rc = dger_ctx_create(&o, vertex_cn, v_num, &c);
  if (rc) return rc;
  dger_sgns_stats st;
  rc = dger_ctx_set_corpus(c, corpus, n_tokens);
  if (!rc) rc = dger_ctx_train(c, 0, nullptr, 0, nullptr, &st);
  if (!rc && o.debug > 0)
    std::printf("> [Train SGNS TIME:] %lf  trained words/s: %.1f  workers: %d\n", st.kernel_ms * 1e-3,
                st.words_per_sec, st.workers);
  if (!rc && o.eoutput[0]) rc = dger_ctx_save(c, o.eoutput, o.binary);
  dger_ctx_destroy(c);
  return rc;
}

// ---- test-facing kernels ------------------------------------------------------
int dger_sgns_step(int32_t D, int32_t n_out, int32_t n_ctx, int32_t n_ctx_first, const int32_t* meta,
                   float alpha, const float* O, const float* I, float* corr, float* md, float* gin) {
  if (D < 2 || D > 1024 || (D & 1) || n_out < 2 || n_ctx < 1 || n_ctx_first < 0 || n_ctx_first > n_ctx)
    return fail(DGER_E_ARG, "bad step shape");
  float exp_table[dger::kExpTable];
  build_exp_table(exp_table);
  DevBuf dm, de, dO, dI, dc, dmd, dg;
  if (dm.ensure(n_out * 4) || de.ensure(sizeof(exp_table)) || dO.ensure((size_t)n_out * D * 4) ||
      dI.ensure((size_t)n_ctx * D * 4) || dc.ensure((size_t)n_out * n_ctx * 4) ||
      dmd.ensure((size_t)n_out * D * 4) || dg.ensure((size_t)n_ctx * D * 4))
    return fail(DGER_E_ALLOC, "device allocation failed");
  DGER_HIP(hipMemcpy(dm.p, meta, n_out * 4, hipMemcpyHostToDevice));
  DGER_HIP(hipMemcpy(de.p, exp_table, sizeof(exp_table), hipMemcpyHostToDevice));
  DGER_HIP(hipMemcpy(dO.p, O, (size_t)n_out * D * 4, hipMemcpyHostToDevice));
  DGER_HIP(hipMemcpy(dI.p, I, (size_t)n_ctx * D * 4, hipMemcpyHostToDevice));
  DGER_HIP(dger::launch_step(D, n_out, n_ctx, n_ctx_first, dm.as<int32_t>(), alpha, de.as<float>(),
                             dO.as<float>(), dI.as<float>(), dc.as<float>(), dmd.as<float>(), dg.as<float>(),
                             nullptr));
  DGER_HIP(hipDeviceSynchronize());
  DGER_HIP(hipMemcpy(corr, dc.p, (size_t)n_out * n_ctx * 4, hipMemcpyDeviceToHost));
  DGER_HIP(hipMemcpy(md, dmd.p, (size_t)n_out * D * 4, hipMemcpyDeviceToHost));
  DGER_HIP(hipMemcpy(gin, dg.p, (size_t)n_ctx * D * 4, hipMemcpyDeviceToHost));
  return DGER_OK;
}

int dger_sgns_step_mfma_gemm1(int32_t D, int32_t n_out, int32_t n_ctx, int32_t n_ctx_first, const int32_t* meta,
                              float alpha, const float* O, const float* I, float* corr) {
  if (D < 32 || D > 256 || D % 32 || n_out < 2 || n_out > 8 || n_ctx < 1 || n_ctx > 64 || n_ctx_first < 0 ||
      n_ctx_first > n_ctx || !meta || !O || !I || !corr)
    return fail(DGER_E_ARG, "bad MFMA step shape (size a multiple of 32 up to 256, 2..8 outputs, 1..64 contexts)");
  float exp_table[dger::kExpTable];
  build_exp_table(exp_table);
  DevBuf dm, de, dO, dI, dc;
  if (dm.ensure(n_out * 4) || de.ensure(sizeof(exp_table)) || dO.ensure((size_t)n_out * D * 4) ||
      dI.ensure((size_t)n_ctx * D * 4) || dc.ensure((size_t)n_out * n_ctx * 4))
    return fail(DGER_E_ALLOC, "device allocation failed");
  DGER_HIP(hipMemcpy(dm.p, meta, n_out * 4, hipMemcpyHostToDevice));
  DGER_HIP(hipMemcpy(de.p, exp_table, sizeof(exp_table), hipMemcpyHostToDevice));
  DGER_HIP(hipMemcpy(dO.p, O, (size_t)n_out * D * 4, hipMemcpyHostToDevice));
  DGER_HIP(hipMemcpy(dI.p, I, (size_t)n_ctx * D * 4, hipMemcpyHostToDevice));
  DGER_HIP(dger::launch_step_mfma_gemm1(D, n_out, n_ctx, n_ctx_first, dm.as<int32_t>(), alpha, de.as<float>(),
                                        dO.as<float>(), dI.as<float>(), dc.as<float>(), nullptr));
  DGER_HIP(hipDeviceSynchronize());
  DGER_HIP(hipMemcpy(corr, dc.p, (size_t)n_out * n_ctx * 4, hipMemcpyDeviceToHost));
  return DGER_OK;
}

int dger_sgns_step_fast(int32_t n_out, const int32_t* meta, const int32_t* mult, float alpha, const float* O,
                        const float* I, float* corr, float* md, float* I_out) {
  return dger_sgns_step_fast_d(dger::kFastDim, n_out, meta, mult, alpha, O, I, corr, md, I_out);
}

int dger_sgns_step_fast_d(int32_t size, int32_t n_out, const int32_t* meta, const int32_t* mult, float alpha,
                          const float* O, const float* I, float* corr, float* md, float* I_out) {
  const int D = size, S = 48;
  if (dger::fast_halves(D) < 1) return fail(DGER_E_ARG, "window-slot step: size must be 128 or 256");
  if (n_out < 2 || n_out > 16 || !meta || !mult || !O || !I || !corr || !md || !I_out)
    return fail(DGER_E_ARG, "bad window-slot step arguments");
  float exp_table[dger::kExpTable];
  build_exp_table(exp_table);
  DevBuf dm, dmu, de, dO, dI, dc, dmd, dI2;
  if (dm.ensure(n_out * 4) || dmu.ensure(S * 4) || de.ensure(sizeof(exp_table)) || dO.ensure((size_t)n_out * D * 4) ||
      dI.ensure((size_t)S * D * 4) || dc.ensure((size_t)n_out * S * 4) || dmd.ensure((size_t)n_out * D * 4) ||
      dI2.ensure((size_t)S * D * 4))
    return fail(DGER_E_ALLOC, "device allocation failed");
  DGER_HIP(hipMemcpy(dm.p, meta, n_out * 4, hipMemcpyHostToDevice));
  DGER_HIP(hipMemcpy(dmu.p, mult, S * 4, hipMemcpyHostToDevice));
  DGER_HIP(hipMemcpy(de.p, exp_table, sizeof(exp_table), hipMemcpyHostToDevice));
  DGER_HIP(hipMemcpy(dO.p, O, (size_t)n_out * D * 4, hipMemcpyHostToDevice));
  DGER_HIP(hipMemcpy(dI.p, I, (size_t)S * D * 4, hipMemcpyHostToDevice));
  DGER_HIP(dger::launch_step_fast(D, n_out, dm.as<int32_t>(), dmu.as<int32_t>(), alpha, de.as<float>(), dO.as<float>(),
                                  dI.as<float>(), dc.as<float>(), dmd.as<float>(), dI2.as<float>(), nullptr));
  DGER_HIP(hipDeviceSynchronize());
  DGER_HIP(hipMemcpy(corr, dc.p, (size_t)n_out * S * 4, hipMemcpyDeviceToHost));
  DGER_HIP(hipMemcpy(md, dmd.p, (size_t)n_out * D * 4, hipMemcpyDeviceToHost));
  DGER_HIP(hipMemcpy(I_out, dI2.p, (size_t)S * D * 4, hipMemcpyDeviceToHost));
  return DGER_OK;
}

#define DGER_LOSS_CHECK()                                                                   \
  if (!syn0 || !syn1 || !ctx || !out || D < 1 || n_mb < 1 || n_ctx < 1 || n_out < 1)        \
    return fail(DGER_E_ARG, "bad loss arguments");

int dger_loss_fwd_f32(const float* syn0, const float* syn1, int32_t D, const int32_t* ctx, const int32_t* out,
                      int32_t n_mb, int32_t n_ctx, int32_t n_out, float* loss_out) {
  DGER_LOSS_CHECK()
  DGER_HIP(dger::launch_loss_fwd<float>(syn0, syn1, D, ctx, out, n_mb, n_ctx, n_out, loss_out, nullptr));
  DGER_HIP(hipDeviceSynchronize());
  return DGER_OK;
}

int dger_loss_bwd_f32(const float* syn0, const float* syn1, int32_t D, const int32_t* ctx, const int32_t* out,
                      int32_t n_mb, int32_t n_ctx, int32_t n_out, float scale, float* g0, float* g1) {
  DGER_LOSS_CHECK()
  DGER_HIP(dger::launch_loss_bwd<float>(syn0, syn1, D, ctx, out, n_mb, n_ctx, n_out, scale, g0, g1, nullptr));
  DGER_HIP(hipDeviceSynchronize());
  return DGER_OK;
}

int dger_loss_fwd_f64(const double* syn0, const double* syn1, int32_t D, const int32_t* ctx,
                      const int32_t* out, int32_t n_mb, int32_t n_ctx, int32_t n_out, double* loss_out) {
  DGER_LOSS_CHECK()
  DGER_HIP(dger::launch_loss_fwd<double>(syn0, syn1, D, ctx, out, n_mb, n_ctx, n_out, loss_out, nullptr));
  DGER_HIP(hipDeviceSynchronize());
  return DGER_OK;
}

int dger_loss_bwd_f64(const double* syn0, const double* syn1, int32_t D, const int32_t* ctx,
                      const int32_t* out, int32_t n_mb, int32_t n_ctx, int32_t n_out, double scale, double* g0,
                      double* g1) {
  DGER_LOSS_CHECK()
  DGER_HIP(dger::launch_loss_bwd<double>(syn0, syn1, D, ctx, out, n_mb, n_ctx, n_out, scale, g0, g1, nullptr));
  DGER_HIP(hipDeviceSynchronize());
  return DGER_OK;
}

int dger_dev_alloc(void** ptr, uint64_t bytes) {
  if (!ptr) return fail(DGER_E_ARG, "null pointer");
  DGER_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return DGER_OK;
}
int dger_dev_free(void* ptr) {
  DGER_HIP(hipFree(ptr));
  return DGER_OK;
}
int dger_dev_upload(void* dst, const void* src, uint64_t bytes) {
  DGER_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return DGER_OK;
}
int dger_dev_download(void* dst, const void* src, uint64_t bytes) {
  DGER_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return DGER_OK;
}
int dger_dev_zero(void* ptr, uint64_t bytes) {
  DGER_HIP(hipMemset(ptr, 0, bytes));
  return DGER_OK;
}
int dger_dev_copy(void* dst, const void* src, uint64_t bytes) {
  if ((!dst || !src) && bytes) return fail(DGER_E_ARG, "null pointer");
  DGER_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return DGER_OK;
}
int dger_device_count(int32_t* count) {
  if (!count) return fail(DGER_E_ARG, "null pointer");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return DGER_OK;
}

}  // extern "C"
