// Embedding file writer in my_saveModel's format (include/dsgl.hpp:1443-1461):
// a "V D\n" header, then per row the vertex id as "%u ", the D values either
// as "%f " text or as raw float32 bytes (-binary 1), and "\n".
//
// The text path produces the bytes printf("%f ") would, without printf: a
// float32 x = m * 2^e (m < 2^24) times 10^6 is m * 10^6 * 2^e, an integer
// (e >= 0) or a dyadic fraction (e < 0) whose rounding to an integer, half
// to even, is what glibc prints for the exact binary value in the default
// rounding mode.  Values of magnitude 2^43 and above, infinities and NaNs go
// to snprintf.  Rows are formatted by several threads into their own
// buffers and written in row order.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "dger_sgns.h"
#include "sgns_internal.h"

namespace dger {

// Decimal digits of v at p; returns the count.
inline int put_u64(uint64_t v, char* p) {
  char t[20];
  int n = 0;
  do {
    t[n++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  for (int i = 0; i < n; i++) p[i] = t[n - 1 - i];
  return n;
}

// "%f" of x at p (at most 64 bytes); returns the length.
int format_f6(float x, char* p) {
  uint32_t bits;
  std::memcpy(&bits, &x, 4);
  const uint32_t ex = (bits >> 23) & 0xFFu, man = bits & 0x7FFFFFu;
  const uint64_t m = ex ? (uint64_t)(man | 0x800000u) : (uint64_t)man;
  const int e = ex ? (int)ex - 150 : -149;  // |x| = m * 2^e
  if (ex == 0xFFu || e > 19) return std::snprintf(p, 64, "%f", (double)x);
  const uint64_t N = m * 1000000u;  // < 2^44
  uint64_t q;
  if (e >= 0) {
    q = N << e;
  } else {
    const int s = -e;
    if (s >= 63) {
      q = 0;  // N < 2^44 is below half a unit
    } else {
      q = N >> s;
      const uint64_t r = N & ((1ull << s) - 1), half = 1ull << (s - 1);
      if (r > half || (r == half && (q & 1))) q++;
    }
  }
  int n = 0;
  if (bits >> 31) p[n++] = '-';
  n += put_u64(q / 1000000u, p + n);
  p[n++] = '.';
  uint32_t f = (uint32_t)(q % 1000000u);
  for (int i = 5; i >= 0; i--) {
    p[n + i] = (char)('0' + f % 10);
    f /= 10;
  }
  return n + 6;
}

}  // namespace dger

extern "C" int dger_emb_write(const char* path, uint32_t v, int32_t d, const int32_t* ids, const float* rows,
                              int32_t binary) {
  if (!path || (v && (!ids || !rows)) || d < 1) return dger::host_fail(DGER_E_ARG, "bad embedding write arguments");
  FILE* fo = std::fopen(path, "wb");
  if (!fo) return dger::host_fail(DGER_E_IO, std::string("cannot open ") + path);
  bool ok = std::fprintf(fo, "%ld %ld\n", (long)v, (long)d) > 0;
  // Batches of rows: each thread formats its block into its own buffer, then
  // the blocks are written in order.
  const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint32_t block = binary ? 4096 : 1024;
  std::vector<std::string> buf(nthreads);
  auto format_block = [&](uint32_t r0, uint32_t r1, std::string& out) {
    const size_t row_max = 12 + (binary ? (size_t)d * 4 : (size_t)d * 66) + 1;
    out.resize((size_t)(r1 - r0) * row_max);
    char* p = &out[0];
    for (uint32_t r = r0; r < r1; r++) {
      p += dger::put_u64((uint32_t)ids[r], p);
      *p++ = ' ';
      const float* x = rows + (size_t)r * d;
      if (binary) {
        std::memcpy(p, x, (size_t)d * 4);
        p += (size_t)d * 4;
      } else {
        for (int b = 0; b < d; b++) {
          p += dger::format_f6(x[b], p);
          *p++ = ' ';
        }
      }
      *p++ = '\n';
    }
    out.resize((size_t)(p - &out[0]));
  };
  for (uint64_t base = 0; ok && base < v; base += (uint64_t)block * nthreads) {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
      const uint64_t r0 = std::min<uint64_t>(v, base + (uint64_t)t * block);
      const uint64_t r1 = std::min<uint64_t>(v, r0 + block);
      if (r0 >= r1) {
        buf[t].clear();
        continue;
      }
      if (t == nthreads - 1 || nthreads == 1)
        format_block((uint32_t)r0, (uint32_t)r1, buf[t]);
      else
        th.emplace_back(format_block, (uint32_t)r0, (uint32_t)r1, std::ref(buf[t]));
    }
    for (auto& x : th) x.join();
    for (int t = 0; t < nthreads && ok; t++)
      if (!buf[t].empty()) ok = std::fwrite(buf[t].data(), 1, buf[t].size(), fo) == buf[t].size();
  }
  if (std::fclose(fo) != 0) ok = false;
  if (!ok) return dger::host_fail(DGER_E_IO, std::string("write failed: ") + path);
  return DGER_OK;
}
