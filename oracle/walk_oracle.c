/*
 * Walk oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C, one-walker-at-a-time restatement of DistGER's information-oriented
 * random walk for one partition (every vertex local), used to check the GPU
 * corpus producer.  Only tests/ may load it, as the checker.
 *
 * Follows include/walk.hpp:
 *   orc_corr_iterate  Correlation::iterate with step_k = 1 (walk.hpp:286-357):
 *                     entropy recursion Hn (iterateHn, :300-308), running
 *                     means (iterateEX, :313-315), R = cov / sqrt(var var).
 *   orc_walk_run      internal_random_walk + internal_walk_epoch
 *                     (walk.hpp:609-700, :789-1043): walker w starts at
 *                     w % V (:501-509), statistics start at
 *                     (0,1,0,1,0,0,1, N=1, R=1) (:851-858), moves, traces
 *                     and the termination rule (:884-991); the path dump
 *                     (include/path.hpp:91-110).
 * Random draws replace the reference's mt19937 / rand() with splitmix64
 * keyed by (seed, walker, draw index): high 32 bits pick the neighbour by
 * multiply-shift, the low 31 bits give p = u / (2^31 - 1).  A walk is cut at
 * max_length vertices and a first move gives up after 64 self-loop draws.
 *
 * Pins (tests/test_oracle_walk.py): Hn equals the Shannon entropy (bits) of
 * the visit counts of every prefix and R the Pearson correlation of the
 * (H, S) sequence, both recomputed from scratch with numpy; the trace rate
 * equals 1 - tanh(p_accept) on graphs where p_accept is one constant
 * (complete graph: 1; complete bipartite K(a,a): 1/a); every walk ends at
 * the first step past min_length where R^2 < r_threshold or R < 0 (checked
 * against the numpy R of the printed path); consecutive path vertices are
 * adjacent; vertex_cn equals the corpus histogram.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double H, S, EH, ES, EHS, EHH, ESS, N, R;
} orc_corr;

static uint64_t draw(uint64_t seed, uint64_t walker, uint64_t ctr) {
  uint64_t z = seed + walker * 0x9E3779B97F4A7C15ull + (ctr + 1) * 0xD1B54A32D192ED03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void orc_corr_init(orc_corr* c) {
  c->H = 0.0; c->S = 1.0; c->EH = 0.0; c->ES = 1.0; c->EHS = 0.0;
  c->EHH = 0.0; c->ESS = 1.0; c->N = 1.0; c->R = 1.0;
}

void orc_corr_iterate(orc_corr* c, double fi) {
  double Snp1 = c->S + 1.0;
  double a = pow(c->S / Snp1, c->S) * (1.0 / Snp1);
  double t;
  if (fi == 0.0) {
    t = log2(a);
  } else {
    double fj = fi + 1.0;
    t = log2(a * (pow(fj / fi, fi) * fj));
  }
  double Hn1 = (c->S * c->H - t) / (c->S + 1.0);
  double Np1 = c->N + 1.0;
  double EH1 = c->EH + (Hn1 - c->EH) / Np1;
  double ES1 = c->ES + (Snp1 - c->ES) / Np1;
  double EHS1 = c->EHS + (Hn1 * Snp1 - c->EHS) / Np1;
  double EHH1 = c->EHH + (Hn1 * Hn1 - c->EHH) / Np1;
  double ESS1 = c->ESS + (Snp1 * Snp1 - c->ESS) / Np1;
  double DH = EHH1 - EH1 * EH1;
  double DS = ESS1 - ES1 * ES1;
  double e = EHS1 - EH1 * ES1;
  c->R = e == 0.0 ? 0.0 : e / sqrt(DH * DS);
  c->H = Hn1; c->S = Snp1; c->EH = EH1; c->ES = ES1; c->EHS = EHS1; c->EHH = EHH1; c->ESS = ESS1; c->N = Np1;
}

/* Runs the correlation over a given visit sequence (path[0] is the start,
 * counted once with no update): H[i], R[i] after visit i (H[0] = 0, R[0] = 1). */
void orc_corr_trace(const int32_t* path, int64_t len, double* H, double* R) {
  orc_corr c;
  orc_corr_init(&c);
  if (len <= 0) return;
  H[0] = c.H;
  R[0] = c.R;
  for (int64_t i = 1; i < len; ++i) {
    int64_t k = 0;
    for (int64_t j = 0; j < i; ++j) k += path[j] == path[i];
    orc_corr_iterate(&c, (double)k);
    H[i] = c.H;
    R[i] = c.R;
  }
}

/* stats: walks, tokens, moves, traces, capped.  Returns the corpus length, or
 * -1 if it would exceed corpus_cap. */
int64_t orc_walk_run(const int64_t* rowptr, const int32_t* col, const int32_t* cnb, uint32_t V, uint64_t first,
                     uint64_t n_walkers, uint32_t min_length, uint32_t max_length, double r_thr, uint64_t seed,
                     int32_t* corpus, int64_t corpus_cap, int32_t* vertex_cn, uint64_t* stats) {
  int32_t* path = (int32_t*)malloc(sizeof(int32_t) * (max_length ? max_length : 1));
  int64_t out = 0;
  memset(vertex_cn, 0, sizeof(int32_t) * V);
  memset(stats, 0, sizeof(uint64_t) * 5);
  for (uint64_t t = 0; t < n_walkers; ++t) {
    uint64_t w = first + t;
    uint32_t cur = (uint32_t)(w % V);
    uint32_t len = 1, step = 0;
    uint64_t ctr = 0;
    int self_draws = 0;
    orc_corr c;
    orc_corr_init(&c);
    path[0] = (int32_t)cur;
    for (;;) {
      int64_t b = rowptr[cur], deg = rowptr[cur + 1] - b;
      if (deg == 0) break;
      uint64_t r = draw(seed, w, ctr++);
      int64_t e = b + (int64_t)(((r >> 32) * (uint64_t)deg) >> 32);
      uint32_t cand = (uint32_t)col[e];
      if (deg == 2 && (uint32_t)col[b] == cur) break;
      if (step == 0 && cand == cur) {
        if (++self_draws >= 64) break;
        continue;
      }
      int sdeg = (int)deg;
      int ddeg = (int)(rowptr[cand + 1] - rowptr[cand]);
      double p = (double)(uint32_t)(r & 0x7FFFFFFFu) / 2147483647.0;
      double p_accept = (1.0 / (double)(sdeg - cnb[e])) * (double)(sdeg > ddeg ? sdeg : ddeg) /
                        (double)(sdeg < ddeg ? sdeg : ddeg);
      int trace = tanh(p_accept) <= p;
      stats[2] += 1;
      stats[3] += (uint64_t)trace;
      if (len >= max_length) { stats[4] += 1; break; }
      int64_t k = 0;
      for (uint32_t j = 0; j < len; ++j) k += path[j] == (int32_t)cand;
      path[len++] = (int32_t)cand;
      ++step;
      orc_corr_iterate(&c, (double)k);
      if (step > min_length && (c.R * c.R < r_thr || c.R < 0.0)) break;
      if (trace) {
        if (len >= max_length) { stats[4] += 1; break; }
        k = 0;
        for (uint32_t j = 0; j < len; ++j) k += path[j] == (int32_t)cur;
        path[len++] = (int32_t)cur;
        ++step;
        orc_corr_iterate(&c, (double)k);
        if (step > min_length && (c.R * c.R < r_thr || c.R < 0.0)) break;
      } else {
        cur = cand;
      }
    }
    if (out + len + 1 > corpus_cap) { free(path); return -1; }
    for (uint32_t j = 0; j < len; ++j) {
      corpus[out++] = path[j];
      vertex_cn[path[j]] += 1;
    }
    corpus[out++] = -1;
    stats[0] += 1;
    stats[1] += len;
  }
  free(path);
  return out;
}
