// Information-oriented random walks on the GPU: the corpus producer that
// feeds the trainer (packed int32 vertex ids, -1 after every walk, plus the
// per-vertex counts), written straight into HBM.
//
// Reference behaviour followed:
//   * walker w starts at vertex w % v_num (include/walk.hpp:501-509), its
//     statistics start at H=0, S=1, E[H]=0, E[S]=1, E[HS]=0, E[HH]=0,
//     E[SS]=1, N=1, R=1 and the start vertex counts as visited once
//     (walk.hpp:851-858);
//   * one move (walk.hpp:884-944): draw a neighbour uniformly; stop at a
//     vertex without neighbours or at a degree-2 vertex whose first
//     neighbour is itself; redraw a self loop on the first move; accept with
//     p_norm = tanh((1 / (deg_src - common)) * max(deg_src, deg_dst) /
//     min(deg_src, deg_dst)) and, when p_norm <= p for a uniform p in
//     [0, 1], step to the neighbour and straight back ("trace");
//   * every visit updates the entropy / length correlation
//     (walk.hpp:286-357, Correlation::iterate with step_k = 1) with the
//     number of earlier visits of that vertex in the same walk; after more
//     than min_length steps the walk ends as soon as R^2 < r_threshold or
//     R < 0 (walk.hpp:867-873, 955-961, 981-987);
//   * the corpus lists each walk's vertices followed by -1, and vertex_cn
//     counts every listed vertex (include/path.hpp:91-110).
// One process, one GPU: every vertex is local, so the cross-partition
// message path of the reference (emit / traceFlag) never runs.
//
// Differences kept on purpose: the reference draws from a per-thread
// mt19937 and the process-wide rand(); here every draw is a splitmix64
// output keyed by (seed, walker id, draw index) (Steele, Lea and Flood,
// "Fast splittable pseudorandom number generators", 2014), so a walk does not
// depend on which thread runs it.  The high 32 bits pick the neighbour
// (multiply-shift), the low 31 bits give p = u / (2^31 - 1).  A walk is cut
// at max_length vertices, and a first move that keeps drawing self loops
// gives up after 64 draws; the reference has neither limit.
//
// Layout in HBM: per chunk of walkers the paths are stored step-major,
// path[i * n + t] for walker t, so lanes of a wavefront that are at the same
// step write neighbouring words.  The earlier-visit count of a vertex is a
// scan of the walker's own path (walks are tens of vertices), the first 32
// steps from a per-lane LDS mirror.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <utility>

#include "dger_sgns.h"
#include "dger_walk.h"
#include "sgns_internal.h"

namespace dger {
namespace {

typedef unsigned long long u64;

__device__ __forceinline__ u64 walk_draw(u64 seed, u64 walker, u64 ctr) {
  u64 z = seed + walker * 0x9E3779B97F4A7C15ull + (ctr + 1) * 0xD1B54A32D192ED03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Entropy / length statistics of one walk (walk.hpp:260-357).
struct Corr {
  double H, S, EH, ES, EHS, EHH, ESS, N, R;

  // One visit of a vertex seen fi times before in this walk, step_k = 1.
  __device__ void iterate(double fi) {
    const double Snp1 = S + 1.0;
    // pow(Sn/Sn+1, Sn) * (1/Sn+1)^1 and pow(fj/fi, fi) * fj^1 (walk.hpp:286-298)
    const double a = pow(S / Snp1, S) * (1.0 / Snp1);
    double t;
    if (fi == 0.0) {
      t = log2(a);
    } else {
      const double fj = fi + 1.0;
      t = log2(a * (pow(fj / fi, fi) * fj));
    }
    const double Hn1 = (S * H - t) / (S + 1.0);
    const double Np1 = N + 1.0;
    const double EH1 = EH + (Hn1 - EH) / Np1;
    const double ES1 = ES + (Snp1 - ES) / Np1;
    const double EHS1 = EHS + (Hn1 * Snp1 - EHS) / Np1;
    const double EHH1 = EHH + (Hn1 * Hn1 - EHH) / Np1;
    const double ESS1 = ESS + (Snp1 * Snp1 - ESS) / Np1;
    const double DH = EHH1 - EH1 * EH1;
    const double DS = ESS1 - ES1 * ES1;
    const double e = EHS1 - EH1 * ES1;
    R = e == 0.0 ? 0.0 : e / sqrt(DH * DS);
    H = Hn1; S = Snp1; EH = EH1; ES = ES1; EHS = EHS1; EHH = EHH1; ESS = ESS1; N = Np1;
  }
};

constexpr uint32_t kLdsSteps = 32;   // 32 KB of LDS per 256-lane workgroup

// One thread per walker.  path is step-major over the chunk's n walkers;
// len1[t] = walk length + 1 (the walk's share of the corpus, separator
// included).  vertex_cn gets one count per vertex written: visits of the
// current vertex (the start and every trace-return) gather in a register
// and are added when the walker moves on or stops, other visits are added
// at once; the adds return nothing, so they overlap the walk's loads.
__global__ __launch_bounds__(256) void k_walk(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                              const int32_t* __restrict__ cnb, uint32_t V, u64 first, u64 n,
                                              uint32_t min_length, uint32_t max_length, double r_thr, u64 seed,
                                              int32_t* __restrict__ path, u64* __restrict__ len1,
                                              int32_t* __restrict__ vertex_cn, u64* __restrict__ counters) {
  // each lane's column holds its walk's first kLdsSteps vertices
  __shared__ int32_t lpath[kLdsSteps][256];
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  u64 tokens = 0, moves = 0, traces = 0, capped = 0;
  if (t < n) {
    const u64 w = first + t;
    uint32_t cur = (uint32_t)(w % V);
    path[t] = (int32_t)cur;
    lpath[0][threadIdx.x] = (int32_t)cur;
    uint32_t len = 1, step = 0;
    u64 ctr = 0;
    int self_draws = 0;
    int32_t cur_visits = 1;   // visits of cur not yet added to vertex_cn
    Corr c{0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0, 1.0};
    // earlier visits of v: the first kLdsSteps steps from the LDS mirror,
    // the rest (long walks) from the path in HBM
    auto seen = [&](uint32_t v) {
      uint32_t k = 0;
      const uint32_t m = len < kLdsSteps ? len : kLdsSteps;
      for (uint32_t i = 0; i < m; ++i) k += lpath[i][threadIdx.x] == (int32_t)v;
      for (uint32_t i = kLdsSteps; i < len; ++i) k += path[(u64)i * n + t] == (int32_t)v;
      return (double)k;
    };
    auto put = [&](uint32_t x) {
      path[(u64)len * n + t] = (int32_t)x;
      if (len < kLdsSteps) lpath[len][threadIdx.x] = (int32_t)x;
    };
    auto ends = [&]() { return step > min_length && (c.R * c.R < r_thr || c.R < 0.0); };
    while (true) {
      const int64_t b = rowptr[cur];
      const int64_t deg = rowptr[cur + 1] - b;
      if (deg == 0) break;
      const u64 r = walk_draw(seed, w, ctr++);
      const int64_t e = b + (int64_t)(((r >> 32) * (u64)deg) >> 32);
      const uint32_t cand = (uint32_t)col[e];
      if (deg == 2 && (uint32_t)col[b] == cur) break;
      if (step == 0 && cand == cur) {
        if (++self_draws >= 64) break;
        continue;
      }
      const int sdeg = (int)deg;
      const int ddeg = (int)(rowptr[cand + 1] - rowptr[cand]);
      const double p = (double)(uint32_t)(r & 0x7FFFFFFFu) / 2147483647.0;
      // ((1 / (s - c)) * max) / min, evaluated left to right as in walk.hpp:925
      const double p_accept = (1.0 / (double)(sdeg - cnb[e])) * (double)(sdeg > ddeg ? sdeg : ddeg) /
                              (double)(sdeg < ddeg ? sdeg : ddeg);
      const bool trace = tanh(p_accept) <= p;
      ++moves;
      traces += trace;
      if (len >= max_length) { capped = 1; break; }
      const double f1 = seen(cand);
      put(cand);
      if (cand == cur) ++cur_visits;
      else atomicAdd(&vertex_cn[cand], 1);
      ++len; ++step;
      c.iterate(f1);
      if (ends()) break;
      if (trace) {
        if (len >= max_length) { capped = 1; break; }
        const double f2 = seen(cur);
        put(cur);
        ++cur_visits;
        ++len; ++step;
        c.iterate(f2);
        if (ends()) break;
      } else if (cand != cur) {
        atomicAdd(&vertex_cn[cur], cur_visits);   // cand's visit is already counted
        cur = cand;
        cur_visits = 0;
      }
    }
    if (cur_visits) atomicAdd(&vertex_cn[cur], cur_visits);
    len1[t] = (u64)len + 1;
    tokens = len;
  }
  // wavefront sums, one atomic per wavefront and counter
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    tokens += __shfl_xor(tokens, off, 64);
    moves += __shfl_xor(moves, off, 64);
    traces += __shfl_xor(traces, off, 64);
    capped += __shfl_xor(capped, off, 64);
  }
  if ((threadIdx.x & 63) == 0 && tokens) {
    atomicAdd(&counters[0], tokens);
    atomicAdd(&counters[1], moves);
    atomicAdd(&counters[2], traces);
    atomicAdd(&counters[3], capped);
  }
}

// Walk t's vertices go to corpus[off[t] ...], then -1; off is the inclusive
// sum of len + 1, so walk t starts at off[t - 1].  max_walk keeps the
// longest walk (one atomic per wavefront).
__global__ __launch_bounds__(256) void k_walk_pack(const int32_t* __restrict__ path, const u64* __restrict__ len1,
                                                   const u64* __restrict__ off_incl, u64 n, int32_t* __restrict__ corpus,
                                                   u64* __restrict__ max_walk) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  u64 longest = 0;
  if (t < n) {
    const u64 o = t == 0 ? 0 : off_incl[t - 1];
    const uint32_t len = (uint32_t)(len1[t] - 1);
    for (uint32_t i = 0; i < len; ++i) corpus[o + i] = path[(u64)i * n + t];
    corpus[o + len] = -1;
    longest = len;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const u64 x = __shfl_xor(longest, off, 64);
    longest = x > longest ? x : longest;
  }
  if ((threadIdx.x & 63) == 0 && longest) atomicMax(max_walk, longest);
}

// |N(u) ∩ N(v)| for every directed edge (u, v = col[e]); rows sorted
// ascending without repeats.  Each element of the shorter row is searched
// in the longer one.  An edge whose shorter row has more than kLaneRow
// entries (hub to hub) is appended to `heavy` and left to
// k_common_neighbors_heavy, so no lane carries a long serial search.
constexpr int64_t kLaneRow = 32;

__device__ __forceinline__ int32_t row_has(const int32_t* __restrict__ col, int64_t lo, int64_t hi, int32_t x) {
  const int64_t end = hi;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (col[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < end && col[lo] == x;
}

__global__ __launch_bounds__(256) void k_common_neighbors(const int64_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ col,
                                                          const int32_t* __restrict__ src, u64 nnz,
                                                          int32_t* __restrict__ cnb, u64* __restrict__ heavy,
                                                          u64* __restrict__ n_heavy) {
  const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const int32_t u = src[e], v = col[e];
  int64_t a0 = rowptr[u], a1 = rowptr[u + 1], b0 = rowptr[v], b1 = rowptr[v + 1];
  if (a1 - a0 > b1 - b0) {
    int64_t x = a0; a0 = b0; b0 = x;
    x = a1; a1 = b1; b1 = x;
  }
  if (a1 - a0 > kLaneRow) {
    heavy[atomicAdd(n_heavy, 1ull)] = e;
    return;
  }
  int32_t k = 0;
  for (int64_t i = a0; i < a1; ++i) k += row_has(col, b0, b1, col[i]);
  cnb[e] = k;
}

// One wavefront per heavy edge: lanes split the shorter row.
__global__ __launch_bounds__(256) void k_common_neighbors_heavy(const int64_t* __restrict__ rowptr,
                                                                const int32_t* __restrict__ col,
                                                                const int32_t* __restrict__ src,
                                                                const u64* __restrict__ heavy, u64 n_heavy,
                                                                int32_t* __restrict__ cnb) {
  const u64 h = ((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (h >= n_heavy) return;   // whole wavefronts leave together
  const u64 e = heavy[h];
  const int32_t u = src[e], v = col[e];
  int64_t a0 = rowptr[u], a1 = rowptr[u + 1], b0 = rowptr[v], b1 = rowptr[v + 1];
  if (a1 - a0 > b1 - b0) {
    int64_t x = a0; a0 = b0; b0 = x;
    x = a1; a1 = b1; b1 = x;
  }
  int32_t k = 0;
  for (int64_t i = a0 + lane; i < a1; i += 64) k += row_has(col, b0, b1, col[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) k += __shfl_xor(k, off, 64);
  if (lane == 0) cnb[e] = k;
}

__global__ void k_row_ids(const int64_t* __restrict__ rowptr, uint32_t V, int32_t* __restrict__ src) {
  const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= V) return;
  for (int64_t i = rowptr[u]; i < rowptr[u + 1]; ++i) src[i] = (int32_t)u;
}

struct Dev {
  void* p = nullptr;
  ~Dev() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 1); }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

#define WALK_HIP(expr)                                                                       \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return host_fail(DGER_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Host-side graph checks shared by both entry points.
int check_graph(const int64_t* rowptr, const int32_t* col, uint32_t V, bool sorted) {
  if (!rowptr || V == 0) return host_fail(DGER_E_ARG, "null graph or v_num == 0");
  if (rowptr[0] != 0) return host_fail(DGER_E_ARG, "rowptr[0] must be 0");
  for (uint32_t u = 0; u < V; ++u) {
    if (rowptr[u + 1] < rowptr[u]) return host_fail(DGER_E_ARG, "rowptr must be non-decreasing");
    if (rowptr[u + 1] - rowptr[u] > 0x7FFFFFFF) return host_fail(DGER_E_UNSUPPORTED, "degree above 2^31 - 1");
  }
  const int64_t nnz = rowptr[V];
  if (nnz > 0 && !col) return host_fail(DGER_E_ARG, "null col");
  for (int64_t i = 0; i < nnz; ++i)
    if (col[i] < 0 || (uint32_t)col[i] >= V) return host_fail(DGER_E_ARG, "neighbour id outside [0, v_num)");
  if (sorted)
    for (uint32_t u = 0; u < V; ++u)
      for (int64_t i = rowptr[u] + 1; i < rowptr[u + 1]; ++i)
        if (col[i] <= col[i - 1]) return host_fail(DGER_E_ARG, "rows must be sorted ascending without repeats");
  return DGER_OK;
}

}  // namespace
}  // namespace dger

using namespace dger;

extern "C" {

void dger_walk_default_opts(dger_walk_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->n_walkers = 0;
  o->first_walker = 0;
  o->min_length = 20;
  o->max_length = 1024;
  o->r_threshold = 0.999;
  o->seed = 1;
  o->device = 0;
  o->chunk_walkers = 0;
}

int dger_walk_common_neighbors(const int64_t* rowptr, const int32_t* col, uint32_t v_num, int32_t* cnb_out,
                               int32_t device) {
  if (int rc = check_graph(rowptr, col, v_num, true)) return rc;
  const uint64_t nnz = (uint64_t)rowptr[v_num];
  if (nnz && !cnb_out) return host_fail(DGER_E_ARG, "null cnb_out");
  if (nnz == 0) return DGER_OK;
  WALK_HIP(hipSetDevice(device));
  Dev d_row, d_col, d_src, d_cnb, d_heavy, d_nh;
  WALK_HIP(d_row.alloc((v_num + 1) * sizeof(int64_t)));
  WALK_HIP(d_col.alloc(nnz * sizeof(int32_t)));
  WALK_HIP(d_src.alloc(nnz * sizeof(int32_t)));
  WALK_HIP(d_cnb.alloc(nnz * sizeof(int32_t)));
  WALK_HIP(d_heavy.alloc(nnz * sizeof(u64)));
  WALK_HIP(d_nh.alloc(sizeof(u64)));
  WALK_HIP(hipMemcpy(d_row.p, rowptr, (v_num + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  WALK_HIP(hipMemcpy(d_col.p, col, nnz * sizeof(int32_t), hipMemcpyHostToDevice));
  WALK_HIP(hipMemset(d_nh.p, 0, sizeof(u64)));
  hipLaunchKernelGGL(k_row_ids, dim3((v_num + 255) / 256), dim3(256), 0, 0, d_row.as<int64_t>(), v_num,
                     d_src.as<int32_t>());
  WALK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_common_neighbors, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, 0,
                     d_row.as<int64_t>(), d_col.as<int32_t>(), d_src.as<int32_t>(), (u64)nnz, d_cnb.as<int32_t>(),
                     d_heavy.as<u64>(), d_nh.as<u64>());
  WALK_HIP(hipGetLastError());
  u64 n_heavy = 0;
  WALK_HIP(hipMemcpy(&n_heavy, d_nh.p, sizeof(u64), hipMemcpyDeviceToHost));
  if (n_heavy) {
    // four wavefronts (heavy edges) per 256-lane workgroup
    hipLaunchKernelGGL(k_common_neighbors_heavy, dim3((unsigned)((n_heavy + 3) / 4)), dim3(256), 0, 0,
                       d_row.as<int64_t>(), d_col.as<int32_t>(), d_src.as<int32_t>(), d_heavy.as<u64>(), n_heavy,
                       d_cnb.as<int32_t>());
    WALK_HIP(hipGetLastError());
  }
  WALK_HIP(hipMemcpy(cnb_out, d_cnb.p, nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
  return DGER_OK;
}

int dger_walk_run_device(const int64_t* rowptr, const int32_t* col, const int32_t* cnb, uint32_t v_num,
                         const dger_walk_opts* o, dger_walk_corpus** out, int32_t* vertex_cn, dger_walk_stats* st) {
  if (!o || !out || !vertex_cn) return host_fail(DGER_E_ARG, "null argument");
  *out = nullptr;
  if (int rc = check_graph(rowptr, col, v_num, false)) return rc;
  const uint64_t nnz = (uint64_t)rowptr[v_num];
  if (nnz && !cnb) return host_fail(DGER_E_ARG, "null cnb");
  if (o->min_length < 1) return host_fail(DGER_E_ARG, "min_length must be >= 1");
  if (o->max_length < 2 || o->max_length > (1u << 20)) return host_fail(DGER_E_UNSUPPORTED, "max_length must be in [2, 2^20]");
  if (!(o->r_threshold >= 0.0 && o->r_threshold <= 1.0)) return host_fail(DGER_E_ARG, "r_threshold must be in [0, 1]");
  int ndev = 0;
  WALK_HIP(hipGetDeviceCount(&ndev));
  if (o->device < 0 || o->device >= ndev) return host_fail(DGER_E_ARG, "device ordinal out of range");
  WALK_HIP(hipSetDevice(o->device));
  const uint64_t total = o->n_walkers;
  // Paths of one chunk take chunk * max_length * 4 bytes; default about 8 GiB
  // (a small share of the 288 GB of HBM; fewer, larger launches).
  uint64_t chunk = o->chunk_walkers ? o->chunk_walkers : ((uint64_t)1 << 33) / (4ull * o->max_length);
  if (chunk > total) chunk = total;
  if (chunk == 0) chunk = 1;
  if (chunk > (1ull << 30)) return host_fail(DGER_E_UNSUPPORTED, "chunk_walkers must be <= 2^30");
  Dev d_row, d_col, d_cnb, d_path, d_len, d_off, d_cnt, d_tmp, d_vcn;
  WALK_HIP(d_row.alloc((v_num + 1) * sizeof(int64_t)));
  WALK_HIP(d_col.alloc(nnz * sizeof(int32_t)));
  WALK_HIP(d_cnb.alloc(nnz * sizeof(int32_t)));
  WALK_HIP(d_path.alloc(chunk * o->max_length * sizeof(int32_t)));
  WALK_HIP(d_len.alloc(chunk * sizeof(u64)));
  WALK_HIP(d_off.alloc(chunk * sizeof(u64)));
  WALK_HIP(d_vcn.alloc(v_num * sizeof(int32_t)));
  WALK_HIP(d_cnt.alloc(5 * sizeof(u64)));
  // The corpus grows in HBM: capacity starts at room for walks of
  // 2 * min_length + 4 vertices per walker and doubles (device-to-device
  // copy) when a chunk does not fit.
  Dev d_corp;
  const uint64_t guess = std::min<uint64_t>(o->max_length, 2ull * o->min_length + 4);
  uint64_t corp_cap = std::min<uint64_t>(total, (uint64_t)1 << 31) * (guess + 1);
  if (corp_cap == 0) corp_cap = 1;
  WALK_HIP(d_corp.alloc(corp_cap * sizeof(int32_t)));
  size_t tmp_bytes = 0;
  WALK_HIP(hipMemcpy(d_row.p, rowptr, (v_num + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  if (nnz) {
    WALK_HIP(hipMemcpy(d_col.p, col, nnz * sizeof(int32_t), hipMemcpyHostToDevice));
    WALK_HIP(hipMemcpy(d_cnb.p, cnb, nnz * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  WALK_HIP(hipMemset(d_vcn.p, 0, v_num * sizeof(int32_t)));
  WALK_HIP(hipMemset(d_cnt.p, 0, 5 * sizeof(u64)));
  hipEvent_t ev0, ev1;
  WALK_HIP(hipEventCreate(&ev0));
  WALK_HIP(hipEventCreate(&ev1));
  float walk_ms = 0.f;
  uint64_t out_len = 0;
  int rc = DGER_OK;
  for (uint64_t done = 0; done < total && rc == DGER_OK; done += chunk) {
    const uint64_t n = total - done < chunk ? total - done : chunk;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipError_t e = hipEventRecord(ev0, 0);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_walk, dim3(blocks), dim3(256), 0, 0, d_row.as<int64_t>(), d_col.as<int32_t>(),
                         d_cnb.as<int32_t>(), v_num, (u64)(o->first_walker + done), (u64)n, o->min_length,
                         o->max_length, o->r_threshold, (u64)o->seed, d_path.as<int32_t>(), d_len.as<u64>(),
                         d_vcn.as<int32_t>(), d_cnt.as<u64>());
      e = hipGetLastError();
    }
    // inclusive sum of len + 1 (one separator per walk)
    size_t need = 0;
    if (e == hipSuccess) e = hipcub::DeviceScan::InclusiveSum(nullptr, need, d_len.as<u64>(), d_off.as<u64>(), (int)n, 0);
    if (e == hipSuccess && need > tmp_bytes) {
      if (d_tmp.p) (void)hipFree(d_tmp.p);
      d_tmp.p = nullptr;
      e = d_tmp.alloc(need);
      tmp_bytes = need;
    }
    if (e == hipSuccess)
      e = hipcub::DeviceScan::InclusiveSum(d_tmp.p, tmp_bytes, d_len.as<u64>(), d_off.as<u64>(), (int)n, 0);
    u64 chunk_len = 0;
    if (e == hipSuccess) e = hipMemcpy(&chunk_len, d_off.as<u64>() + (n - 1), sizeof(u64), hipMemcpyDeviceToHost);
    if (e == hipSuccess && out_len + chunk_len > corp_cap) {
      uint64_t cap2 = corp_cap * 2;
      while (cap2 < out_len + chunk_len) cap2 *= 2;
      Dev bigger;
      e = bigger.alloc(cap2 * sizeof(int32_t));
      if (e == hipSuccess) e = hipMemcpy(bigger.p, d_corp.p, out_len * sizeof(int32_t), hipMemcpyDeviceToDevice);
      if (e == hipSuccess) {
        std::swap(bigger.p, d_corp.p);
        corp_cap = cap2;
      }
    }
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_walk_pack, dim3(blocks), dim3(256), 0, 0, d_path.as<int32_t>(), d_len.as<u64>(),
                         d_off.as<u64>(), (u64)n, d_corp.as<int32_t>() + out_len, d_cnt.as<u64>() + 4);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ev1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
    if (e != hipSuccess) rc = host_fail(DGER_E_HIP, std::string("walk chunk: ") + hipGetErrorString(e));
    walk_ms += ms;
    out_len += chunk_len;
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  if (rc != DGER_OK) return rc;
  WALK_HIP(hipMemcpy(vertex_cn, d_vcn.p, v_num * sizeof(int32_t), hipMemcpyDeviceToHost));
  u64 cnt[5] = {0, 0, 0, 0, 0};
  WALK_HIP(hipMemcpy(cnt, d_cnt.p, sizeof(cnt), hipMemcpyDeviceToHost));
  if (st) {
    st->walks = total;
    st->tokens = cnt[0];
    st->moves = cnt[1];
    st->traces = cnt[2];
    st->capped = cnt[3];
    st->walk_ms = walk_ms;
  }
  dger_walk_corpus* h = new (std::nothrow) dger_walk_corpus;
  if (!h) return host_fail(DGER_E_ALLOC, "corpus handle allocation failed");
  h->d = d_corp.as<int32_t>();
  d_corp.p = nullptr;
  h->n_tokens = out_len;
  h->max_walk = (uint32_t)cnt[4];
  h->v_num = v_num;
  h->device = o->device;
  *out = h;
  return DGER_OK;
}

int dger_walk_corpus_info(const dger_walk_corpus* h, uint64_t* n_tokens, uint32_t* max_walk) {
  if (!h) return host_fail(DGER_E_ARG, "null corpus handle");
  if (n_tokens) *n_tokens = h->n_tokens;
  if (max_walk) *max_walk = h->max_walk;
  return DGER_OK;
}

int dger_walk_corpus_download(const dger_walk_corpus* h, int32_t* corpus, uint64_t corpus_cap) {
  if (!h) return host_fail(DGER_E_ARG, "null corpus handle");
  if (corpus_cap < h->n_tokens) return host_fail(DGER_E_ARG, "corpus_cap below the corpus length");
  if (h->n_tokens && !corpus) return host_fail(DGER_E_ARG, "null corpus");
  WALK_HIP(hipSetDevice(h->device));
  if (h->n_tokens) WALK_HIP(hipMemcpy(corpus, h->d, h->n_tokens * sizeof(int32_t), hipMemcpyDeviceToHost));
  return DGER_OK;
}

void dger_walk_corpus_destroy(dger_walk_corpus* h) {
  if (!h) return;
  if (h->d && hipSetDevice(h->device) == hipSuccess) (void)hipFree(h->d);
  delete h;
}

int dger_walk_run(const int64_t* rowptr, const int32_t* col, const int32_t* cnb, uint32_t v_num,
                  const dger_walk_opts* o, int32_t* corpus, uint64_t corpus_cap, uint64_t* corpus_len,
                  int32_t* vertex_cn, dger_walk_stats* st) {
  if (!corpus_len) return host_fail(DGER_E_ARG, "null argument");
  dger_walk_corpus* h = nullptr;
  if (int rc = dger_walk_run_device(rowptr, col, cnb, v_num, o, &h, vertex_cn, st)) return rc;
  *corpus_len = h->n_tokens;
  int rc = DGER_OK;
  if (corpus_cap < h->n_tokens)
    rc = host_fail(DGER_E_ARG, "corpus_cap too small; *corpus_len holds the length needed");
  else
    rc = dger_walk_corpus_download(h, corpus, corpus_cap);
  dger_walk_corpus_destroy(h);
  return rc;
}

}  // extern "C"
