// qubo_text.hip -- .qubo text and binary instance ingest on the GPU (os_qubo_*).
//
// The text grammar is oneSolver's (include/model/qubo.hpp:392-417):
//   file        := comment* header? (comment | coefficient)*
//   comment     := 'c' print* (eol | eoi)
//   header      := "p qubo" uint uint uint uint (eol | eoi)
//   coefficient := lexeme[uint ' '] lexeme[uint ' '] real (eol | eoi)
// with blanks skipped before every token.  Every item ends at its line's
// end-of-line, so the text is cut into lines (ends: '\n', or '\r' not followed
// by '\n') and each line is parsed by its own thread:
//
//   line_count     tiles of 8 KiB: line ends per tile
//   scan_tiles     exclusive prefix over the tiles (one work-group)
//   line_starts    the start offset of every line
//   parse_lines    one thread per line: kind, (i, j, value); decimal strings
//                  are converted exactly in-kernel (a 128-bit product with a
//                  truncated power of five, gen_pow5.py) or, when the product
//                  cannot decide the rounding (ties, > 19 digits, subnormal or
//                  overflow range), handed back to the host's strtod
//   dedupe         first value of a repeated key wins (QUBOBuilder::add_element
//                  inserts, qubo.hpp:304-318): the lowest line index per key,
//                  through a dense n x n owner table or an open-addressing hash
//   compact        the winning terms in file order (tile counts, scan, scatter)
//
// A line the kernel rejects (syntax, a lower-triangle key, a header out of
// place, a "nan(" whose ')' is not on its line, an index above 2^31 - 1)
// makes the whole text go through the sequential reader of
// include/onesolver/model/qubo.hpp, which throws the reference's exception or,
// for the last two cases, returns the model.  The builder's checks
// (qubo.hpp:353-377) run on the host from the kernel's counts.
//
// Binary format (little-endian), written by os_qubo_write_binary:
//   char magic[4] = "OSQB"; uint32 version = 1; int32 n; int32 max_nodes;
//   int64 num_linear; int64 num_quadratic;
//   int32 i[T]; int32 j[T]; float64 v[T];   T = num_linear + num_quadratic
// the terms in the model's order (first appearance in the text), i <= j.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/onesolver.h"
#include "../../include/onesolver/model/qubo.hpp"
#include "os_ctx.h"
#include "os_dense.h"
#include "pow5_table.h"

using os::Ctx;
using os::DevBuf;

namespace {

constexpr int kThreads = 256;
constexpr int kIndexRounds = 8;              // line index: 32 bytes per thread and round
constexpr int kTileBytes = kThreads * 32 * kIndexRounds;
constexpr int kLinesPerThread = 16;          // compaction tiles
constexpr int kTileLines = kThreads * kLinesPerThread;
constexpr int kScanThreads = 1024;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint64_t kEmptyKey = ~0ull;

enum : uint8_t {
  K_COMMENT = 0,
  K_COEF = 1,
  K_HEADER = 2,
  K_BLANK_END = 3,  // the text's last line, blanks only (or empty)
  K_ERROR = 4,
  K_LOWER = 5,      // i > j: add_element throws before the line end is checked
  K_FALLBACK = 6,   // the sequential reader decides
};

struct Stats {
  uint32_t max_index;    // largest j of a coefficient line
  uint32_t n_coef;       // coefficient lines
  uint32_t n_bad;        // K_ERROR, K_LOWER, K_FALLBACK lines
  uint32_t n_header;
  uint32_t header_line;  // lowest header line
  uint32_t first_other;  // lowest line that is neither a comment nor the blank end
  uint32_t n_hostnum;    // values left to strtod
  uint32_t n_lin, n_quad;
  uint32_t hdr[4];
  uint32_t n_long;       // rounds longer than the LDS span (the global-memory kernel's)
  uint32_t n_runs;       // coefficient lines whose previous line is not one
  uint32_t n_unsorted;   // coefficient lines whose key is not above the previous line's
};

struct HostNum {
  uint64_t start;
  uint32_t line, len;
};

// ---------------------------------------------------------------- reduction

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o));
  return v;
}

// exclusive prefix of v over the work-group; *total receives the sum
__device__ uint32_t block_exclusive(uint32_t v, uint32_t *total) {
  __shared__ uint32_t wsum[kScanThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  uint32_t base = 0, all = 0;
  for (int w = 0; w < nw; ++w) {
    if (w < wave) base += wsum[w];
    all += wsum[w];
  }
  __syncthreads();
  *total = all;
  return base + x - v;
}

// ---------------------------------------------------------------- line index

__device__ __forceinline__ uint32_t byte_at(const uint4 &w, int k) {
  const uint32_t d = (k >> 2) == 0 ? w.x : (k >> 2) == 1 ? w.y : (k >> 2) == 2 ? w.z : w.w;
  return (d >> ((k & 3) * 8)) & 0xFFu;
}

__device__ __forceinline__ uint32_t byte_or_zero(const uint8_t *text, uint64_t len, uint64_t p) {
  return p < len ? (uint32_t)text[p] : 0u;
}

// 16 bytes at 16-aligned `p` (text 16-aligned); bytes at or past len read 0
__device__ __forceinline__ uint4 load16(const uint8_t *text, uint64_t len, uint64_t p) {
  if (p + 16 <= len) return *reinterpret_cast<const uint4 *>(text + p);
  uint32_t d[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    uint32_t x = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) x |= byte_or_zero(text, len, p + 4 * k + t) << (8 * t);
    d[k] = x;
  }
  return make_uint4(d[0], d[1], d[2], d[3]);
}

// Line ends among this thread's 32 bytes; `ends` receives their bit mask.
__device__ __forceinline__ uint32_t thread_ends(const uint8_t *text, uint64_t len,
                                                uint64_t base, uint32_t *ends) {
  const uint4 a = load16(text, len, base), b = load16(text, len, base + 16);
  const uint32_t after = byte_or_zero(text, len, base + 32);
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    const uint32_t c = k < 16 ? byte_at(a, k) : byte_at(b, k - 16);
    if (c == '\n') m |= 1u << k;
    if (c == '\r') {
      // a '\r' ends a line unless a '\n' follows
      const uint32_t nx = k < 15 ? byte_at(a, k + 1) : k < 31 ? byte_at(b, k - 15) : after;
      if (nx != '\n') m |= 1u << k;
    }
  }
  // bytes at or past len read as zero, never an end
  *ends = m;
  return __popc(m);
}

// A tile is kIndexRounds rounds of kThreads x 32 contiguous bytes.
__device__ __forceinline__ uint64_t round_base(uint32_t tile, int r) {
  return (uint64_t)tile * kTileBytes + (uint64_t)r * (kThreads * 32) + threadIdx.x * 32u;
}

__global__ void __launch_bounds__(kThreads) line_count_kernel(const uint8_t *text, uint64_t len,
                                                              uint32_t *tile_count) {
  uint32_t c = 0, m;
  for (int r = 0; r < kIndexRounds; ++r) c += thread_ends(text, len, round_base(blockIdx.x, r), &m);
  uint32_t total;
  block_exclusive(c, &total);
  if (threadIdx.x == 0) tile_count[blockIdx.x] = total;
}

// In-place exclusive prefix of cnt[0, n); cnt[n] receives the total.
__global__ void __launch_bounds__(kScanThreads) scan_tiles_kernel(uint32_t *cnt, uint32_t n) {
  const uint32_t per = (n + kScanThreads - 1) / kScanThreads;
  const uint32_t b0 = min(n, threadIdx.x * per), b1 = min(n, b0 + per);
  uint32_t s = 0;
  for (uint32_t k = b0; k < b1; ++k) s += cnt[k];
  uint32_t total;
  uint32_t run = block_exclusive(s, &total);
  for (uint32_t k = b0; k < b1; ++k) {
    const uint32_t c = cnt[k];
    cnt[k] = run;
    run += c;
  }
  if (threadIdx.x == 0) cnt[n] = total;
}

// starts[0] = 0, starts[1 + e] = (position of line end e) + 1, starts[L] = len
__global__ void __launch_bounds__(kThreads) line_starts_kernel(const uint8_t *text, uint64_t len,
                                                               const uint32_t *tile_off,
                                                               uint32_t num_tiles,
                                                               uint64_t *starts) {
  uint32_t run = tile_off[blockIdx.x];
  for (int r = 0; r < kIndexRounds; ++r) {
    const uint64_t base = round_base(blockIdx.x, r);
    uint32_t m;
    const uint32_t c = thread_ends(text, len, base, &m);
    uint32_t total;
    uint32_t idx = run + block_exclusive(c, &total);
    while (m) {
      const int k = __ffs(m) - 1;
      m &= m - 1;
      starts[1 + idx++] = base + k + 1;
    }
    run += total;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    starts[0] = 0;
    starts[(uint64_t)tile_off[num_tiles] + 1] = len;
  }
}

// ---------------------------------------------------------------- line parser

// The bytes of one line, at offsets p < e from its start s; -1 past the
// line.  kLds: the line is in LDS at `lds`; otherwise it is read from the
// global text through an aligned 16-byte window.
template <bool kLds>
struct Cursor {
  const uint8_t *text;  // global text (!kLds)
  const uint8_t *lds;   // the line in LDS (kLds)
  uint64_t len, s, wpos;
  uint32_t p, e;
  uint4 w;
  __device__ __forceinline__ int peek() {
    if (p >= e) return -1;
    if constexpr (kLds) {
      return (int)lds[p];
    } else {
      const uint64_t at = s + p, a = at >> 4;
      if (a != wpos) {
        w = load16(text, len, a << 4);
        wpos = a;
      }
      return (int)byte_at(w, (int)(at & 15));
    }
  }
  __device__ __forceinline__ void skip_blanks() {
    for (int c = peek(); c == ' ' || c == '\t'; c = peek()) ++p;
  }
  // qi::eol | qi::eoi (eoi only on the text's last line)
  __device__ __forceinline__ bool eol(bool last) {
    skip_blanks();
    const int c = peek();
    if (c < 0) return last;
    if (c == '\n') return true;
    if (c == '\r') return true;  // a following '\n' belongs to the same line end
    return false;
  }
  // qi::uint_: digits, at most 2^32 - 1
  __device__ __forceinline__ bool uint_(uint32_t *out) {
    int c = peek();
    if (c < '0' || c > '9') return false;
    uint64_t v = 0;
    while ((c = peek()) >= '0' && c <= '9') {
      v = v * 10 + (uint64_t)(c - '0');
      if (v > 0xFFFFFFFFull) return false;
      ++p;
    }
    *out = (uint32_t)v;
    return true;
  }
  // case-insensitive word; consumed only when it matches whole
  __device__ __forceinline__ bool word_ci(const char *wd) {
    const uint32_t save = p;
    for (; *wd; ++wd, ++p) {
      int c = peek();
      if (c >= 'A' && c <= 'Z') c += 32;
      if (c != *wd) {
        p = save;
        return false;
      }
    }
    return true;
  }
};

// w * 10^q rounded to nearest double, or false when the rounding is not
// decided by the top 128 bits of w * T(q) (the host then uses strtod).
// T(q) = 5^q * 2^(127 - f) rounded down, f = floor(log2 5^q), so
// w * 10^q = (w << lz) * T(q) * 2^(f - 127 - lz + q); the top 64 bits `hi` of
// that 192-bit product give 54 bits (mantissa and round bit) and
// floor(log2 10^q) = f + q fixes the exponent.  Truncating T(q) and the
// lowest product word leaves the product low by less than one unit of `mid`.
__device__ __forceinline__ bool decimal_to_double(uint64_t w, int q, bool neg, uint64_t *bits) {
  const uint64_t sign = neg ? 0x8000000000000000ull : 0ull;
  if (w == 0) {
    *bits = sign;
    return true;
  }
  if (q < kPow5QMin || q > kPow5QMax) return false;
  const int lz = __clzll(w);
  const uint64_t wn = w << lz;
  const uint64_t th = kPow5[q - kPow5QMin][0], tl = kPow5[q - kPow5QMin][1];
  const uint64_t a_lo = wn * th, a_hi = __umul64hi(wn, th);
  const uint64_t b_hi = __umul64hi(wn, tl);
  const uint64_t mid = a_lo + b_hi;
  const uint64_t hi = a_hi + (mid < a_lo ? 1ull : 0ull);
  const int upper = (int)(hi >> 63);
  const int shift = 9 + upper;
  const uint64_t mask = (1ull << shift) - 1;
  const uint64_t below = hi & mask;
  uint64_t m = hi >> shift;  // 2^53 <= m < 2^54
  if (below == mask && mid == ~0ull) return false;     // the tail could carry into m
  if ((m & 1) && below == 0 && mid == 0) return false;  // a tie, or within the tail of one
  int e = (int)kLog2Pow10[q - kPow5QMin] + 63 + upper - lz + 1023;
  m = (m + 1) >> 1;  // round half up (ties excluded above)
  if (m == (1ull << 53)) {
    m >>= 1;
    ++e;
  }
  if (e <= 0 || e >= 0x7FF) return false;
  *bits = sign | ((uint64_t)e << 52) | (m & ((1ull << 52) - 1));
  return true;
}

constexpr uint64_t kQNaN = 0x7FF8000000000000ull;
constexpr uint64_t kInf = 0x7FF0000000000000ull;

// qi::double_ at the cursor.  Returns 0 (no number), 1 (value in *bits),
// 2 (a number for the host's strtod: [*ts, cursor)), 3 (a "nan(" without
// ')' on this line: the sequential reader decides).
template <class Cur>
__device__ __forceinline__ int real_(Cur &c, uint64_t *bits, uint32_t *ts) {
  const uint32_t start = c.p;
  bool neg = false;
  int ch = c.peek();
  if (ch == '+' || ch == '-') {
    neg = ch == '-';
    ++c.p;
    ch = c.peek();
  }
  const uint64_t sign = neg ? 0x8000000000000000ull : 0ull;
  if (ch == 'n' || ch == 'N') {
    if (!c.word_ci("nan")) return 0;
    if (c.peek() == '(') {
      const uint32_t save = c.p;
      int x;
      while ((x = c.peek()) >= 0 && x != ')') ++c.p;
      if (x < 0) {
        c.p = save;
        return 3;
      }
      ++c.p;
    }
    *bits = sign | kQNaN;
    return 1;
  }
  if (ch == 'i' || ch == 'I') {
    if (!c.word_ci("inf")) return 0;
    c.word_ci("inity");
    *bits = sign | kInf;
    return 1;
  }
  uint64_t w = 0;
  int nd = 0;  // significant digits in w
  int q = 0;
  bool digits = false, many = false;
  while ((ch = c.peek()) >= '0' && ch <= '9') {
    digits = true;
    if (w != 0 || ch != '0') {
      if (nd < 19) {
        w = w * 10 + (uint64_t)(ch - '0');
        ++nd;
      } else {
        many = true;
      }
    }
    ++c.p;
  }
  if (c.peek() == '.') {
    const uint32_t dot = c.p;
    ++c.p;
    bool frac = false;
    while ((ch = c.peek()) >= '0' && ch <= '9') {
      frac = true;
      if (w != 0 || ch != '0') {
        if (nd < 19) {
          w = w * 10 + (uint64_t)(ch - '0');
          ++nd;
          --q;
        } else {
          many = true;
        }
      } else {
        --q;  // a leading zero of the fraction
      }
      ++c.p;
    }
    if (!digits && !frac) c.p = dot;
    digits = digits || frac;
  }
  if (!digits) return 0;
  ch = c.peek();
  if (ch == 'e' || ch == 'E') {
    const uint32_t save = c.p;
    ++c.p;
    bool eneg = false;
    ch = c.peek();
    if (ch == '+' || ch == '-') {
      eneg = ch == '-';
      ++c.p;
    }
    int ev = 0;
    bool ed = false;
    while ((ch = c.peek()) >= '0' && ch <= '9') {
      ed = true;
      if (ev < 100000) ev = ev * 10 + (ch - '0');
      ++c.p;
    }
    if (ed)
      q += eneg ? -ev : ev;
    else
      c.p = save;
  }
  *ts = start;
  if (many) return 2;
  return decimal_to_double(w, q, neg, bits) ? 1 : 2;
}

struct LineOut {
  uint32_t i, j;
  uint64_t v;
  uint32_t ts, te;  // the number's text (offsets in the line), for the host's strtod
  uint32_t hdr[4];
  bool hostnum;
};

template <class Cur>
__device__ __forceinline__ uint8_t parse_line(Cur &c, bool last, LineOut &o) {
  c.skip_blanks();
  const int ch = c.peek();
  if (ch < 0) return last ? K_BLANK_END : K_ERROR;
  if (ch == 'c') {
    ++c.p;
    for (;;) {
      c.skip_blanks();
      const int x = c.peek();
      if (x >= 0x20 && x < 0x7F)
        ++c.p;
      else
        break;
    }
    return c.eol(last) ? K_COMMENT : K_ERROR;
  }
  if (ch == 'p') {
    const char *lit = "p qubo";
    for (; *lit; ++lit, ++c.p)
      if (c.peek() != *lit) return K_ERROR;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      c.skip_blanks();
      if (!c.uint_(&o.hdr[k])) return K_ERROR;
    }
    return c.eol(last) ? K_HEADER : K_ERROR;
  }
  if (ch < '0' || ch > '9') return K_ERROR;
  if (!c.uint_(&o.i) || c.peek() != ' ') return K_ERROR;
  ++c.p;
  c.skip_blanks();
  if (!c.uint_(&o.j) || c.peek() != ' ') return K_ERROR;
  ++c.p;
  c.skip_blanks();
  const int r = real_(c, &o.v, &o.ts);
  o.te = c.p;
  if (r == 0) return K_ERROR;
  if (r == 3) return K_FALLBACK;
  if (o.i > 0x7FFFFFFFu || o.j > 0x7FFFFFFFu) return K_FALLBACK;
  if (o.i > o.j) return K_LOWER;
  o.hostnum = r == 2;
  if (o.hostnum) o.v = 0;
  return c.eol(last) ? K_COEF : K_ERROR;
}

constexpr int kParseRounds = 8;  // lines per thread; consecutive lanes take consecutive lines

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o));
  return v;
}

constexpr int kSpanBytes = 16384;  // LDS staging of one round's text

template <bool kLds>
__device__ __forceinline__ void parse_one(const uint8_t *text, uint64_t len, const uint8_t *lds,
                                          uint64_t s, uint64_t e, uint32_t l, uint32_t num_lines,
                                          uint8_t *kind, uint32_t *li, uint32_t *lj, uint64_t *lv,
                                          Stats *st, HostNum *hostnum, uint32_t hostnum_cap,
                                          uint32_t &coef, uint32_t &mx, uint32_t &first) {
  Cursor<kLds> c;
  c.text = text;
  c.lds = lds;
  c.len = len;
  c.s = s;
  c.wpos = ~0ull;
  c.p = 0;
  c.e = (uint32_t)(e - s);
  LineOut o;
  o.hostnum = false;
  const uint8_t k = e - s > 0xFFFFFFFFull ? (uint8_t)K_FALLBACK : parse_line(c, l + 1 == num_lines, o);
  kind[l] = k;
  if (k == K_COEF) {
    li[l] = o.i;
    lj[l] = o.j;
    lv[l] = o.v;
    ++coef;
    mx = max(mx, o.j);
    if (o.hostnum) {
      const uint32_t q = atomicAdd(&st->n_hostnum, 1u);
      if (q < hostnum_cap) hostnum[q] = HostNum{s + o.ts, l, o.te - o.ts};
    }
  } else if (k >= K_ERROR) {
    atomicAdd(&st->n_bad, 1u);
  }
  if (k == K_HEADER) {
    atomicAdd(&st->n_header, 1u);
    atomicMin(&st->header_line, l);
#pragma unroll
    for (int q = 0; q < 4; ++q) st->hdr[q] = o.hdr[q];
  }
  // the lowest line that is neither a comment nor the blank end
  if (k != K_COMMENT && k != K_BLANK_END) first = min(first, l);
}

// Round r of a work-group parses lines l0 + [0, kThreads), one per thread.
// kStaged: the rounds whose text fits kSpanBytes, from LDS (staged with
// 16-byte loads); otherwise the other rounds, from global memory.  The
// staged kernel counts the rounds it leaves; the global one, launched with
// the same grid only when there are any, takes them.  Without the
// global-memory cursor the staged kernel, which handles almost all text,
// keeps its register count low.
template <bool kStaged>
__global__ void __launch_bounds__(kThreads) parse_lines_kernel(
    const uint8_t *text, uint64_t len, const uint64_t *starts, uint32_t num_lines, uint8_t *kind,
    uint32_t *li, uint32_t *lj, uint64_t *lv, Stats *st, HostNum *hostnum,
    uint32_t hostnum_cap) {
  __shared__ uint4 span[kSpanBytes / 16];
  uint32_t coef = 0, mx = 0, first = kNone;
  for (int r = 0; r < kParseRounds; ++r) {
    const uint32_t l0 = (blockIdx.x * kParseRounds + r) * kThreads;
    if (l0 >= num_lines) break;
    const uint32_t l1 = min(l0 + (uint32_t)kThreads, num_lines);
    const uint64_t a0 = starts[l0] & ~15ull, a1 = starts[l1];
    const bool staged = a1 - a0 <= (uint64_t)kSpanBytes;
    if (kStaged && !staged && threadIdx.x == 0) atomicAdd(&st->n_long, 1u);
    if (staged != kStaged) continue;  // uniform over the work-group
    if (kStaged) {
      for (uint32_t k = threadIdx.x; (uint64_t)k * 16 < a1 - a0; k += kThreads)
        span[k] = load16(text, len, a0 + 16ull * k);
      __syncthreads();
    }
    const uint32_t l = l0 + threadIdx.x;
    if (l < num_lines) {
      const uint64_t s = starts[l], e = starts[l + 1];
      if (kStaged)
        parse_one<true>(text, len, reinterpret_cast<const uint8_t *>(span) + (s - a0), s, e, l,
                        num_lines, kind, li, lj, lv, st, hostnum, hostnum_cap, coef, mx, first);
      else
        parse_one<false>(text, len, nullptr, s, e, l, num_lines, kind, li, lj, lv, st, hostnum,
                         hostnum_cap, coef, mx, first);
    }
    if (kStaged) __syncthreads();
  }
  // one atomic per statistic and work-group
  __shared__ uint32_t red[3][kThreads / 64];
  coef = wave_sum(coef);
  mx = wave_max(mx);
  first = wave_min(first);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = coef;
    red[1][wave] = mx;
    red[2][wave] = first;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w) {
      coef += red[0][w];
      mx = max(mx, red[1][w]);
      first = min(first, red[2][w]);
    }
    if (coef) {
      atomicAdd(&st->n_coef, coef);
      atomicMax(&st->max_index, mx);
    }
    if (first != kNone) atomicMin(&st->first_other, first);
  }
}

constexpr uint32_t kTokBytes = 64;

// the text of hostnum[k] (at most kTokBytes long) into out[k * kTokBytes ...]
__global__ void gather_tokens_kernel(const uint8_t *text, const HostNum *hostnum, uint32_t count,
                                     uint8_t *out) {
  const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= count) return;
  const HostNum h = hostnum[k];
  for (uint32_t t = 0; t < h.len && t < kTokBytes; ++t) out[(uint64_t)k * kTokBytes + t] = text[h.start + t];
}

// hostnum[k].line receives the strtod value vals[k]
__global__ void fix_values_kernel(const HostNum *hostnum, const uint64_t *vals, uint32_t count,
                                  uint64_t *lv) {
  const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
  if (k < count) lv[hostnum[k].line] = vals[k];
}

__global__ void init_stats_kernel(Stats *st) {
  if (threadIdx.x == 0) {
    Stats s{};
    s.header_line = kNone;
    s.first_other = kNone;
    *st = s;
  }
}

// ---------------------------------------------------------------- dedupe

// Counts runs of consecutive coefficient lines and the lines whose key (i, j)
// does not rise above the previous line's: one run with rising keys has no
// repeated key, so the dedupe can keep every coefficient line.
__global__ void __launch_bounds__(kThreads) order_check_kernel(const uint8_t *kind,
                                                               const uint32_t *li,
                                                               const uint32_t *lj,
                                                               uint32_t num_lines, Stats *st) {
  const uint32_t l = blockIdx.x * kThreads + threadIdx.x;
  uint32_t runs = 0, unsorted = 0;
  if (l < num_lines && kind[l] == K_COEF) {
    if (l == 0 || kind[l - 1] != K_COEF)
      runs = 1;
    else
      unsorted = li[l - 1] > li[l] || (li[l - 1] == li[l] && lj[l - 1] >= lj[l]);
  }
  runs = wave_sum(runs);
  unsorted = wave_sum(unsorted);
  if ((threadIdx.x & 63) == 0) {
    if (runs) atomicAdd(&st->n_runs, runs);
    if (unsorted) atomicAdd(&st->n_unsorted, unsorted);
  }
}

__device__ __forceinline__ uint64_t hash_key(uint64_t k) { return k * 0x9E3779B97F4A7C15ull; }

__global__ void owner_min_kernel(const uint8_t *kind, const uint32_t *li, const uint32_t *lj,
                                 uint32_t num_lines, uint32_t n, uint32_t *owner) {
  const uint32_t l = blockIdx.x * kThreads + threadIdx.x;
  if (l < num_lines && kind[l] == K_COEF) atomicMin(&owner[(uint64_t)li[l] * n + lj[l]], l);
}

__global__ void hash_insert_kernel(const uint8_t *kind, const uint32_t *li, const uint32_t *lj,
                                   uint32_t num_lines, uint64_t *keys, uint32_t *vals,
                                   uint64_t mask) {
  const uint32_t l = blockIdx.x * kThreads + threadIdx.x;
  if (l >= num_lines || kind[l] != K_COEF) return;
  const uint64_t key = ((uint64_t)li[l] << 32) | lj[l];
  uint64_t h = (hash_key(key) >> 20) & mask;
  // at most half the slots are used, so an empty or matching slot is found
  for (uint64_t probe = 0; probe <= mask; ++probe) {
    const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long *>(&keys[h]),
                                              (unsigned long long)kEmptyKey,
                                              (unsigned long long)key);
    if (prev == kEmptyKey || prev == key) {
      atomicMin(&vals[h], l);
      return;
    }
    h = (h + 1) & mask;
  }
}

struct Dedupe {
  bool all;               // no repeated key: every coefficient line is kept
  const uint32_t *owner;  // dense: [n x n], else null
  uint32_t n;
  const uint64_t *keys;
  const uint32_t *vals;
  uint64_t mask;
  __device__ bool winner(uint32_t l, uint32_t i, uint32_t j) const {
    if (all) return true;
    if (owner) return owner[(uint64_t)i * n + j] == l;
    const uint64_t key = ((uint64_t)i << 32) | j;
    uint64_t h = (hash_key(key) >> 20) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
      if (keys[h] == key) return vals[h] == l;
      h = (h + 1) & mask;
    }
    return false;
  }
};

// ---------------------------------------------------------------- compaction

// A compaction tile is kTileLines lines in kLinesPerThread rounds; round r
// covers lines tile + r * kThreads + [0, kThreads), one per thread.
__device__ __forceinline__ bool is_winner(const uint8_t *kind, const uint32_t *li,
                                          const uint32_t *lj, uint32_t num_lines,
                                          const Dedupe &dd, uint32_t l) {
  return l < num_lines && kind[l] == K_COEF && dd.winner(l, li[l], lj[l]);
}

__global__ void __launch_bounds__(kThreads) compact_count_kernel(
    const uint8_t *kind, const uint32_t *li, const uint32_t *lj, uint32_t num_lines, Dedupe dd,
    uint32_t *tile_count) {
  uint32_t c = 0;
  for (int r = 0; r < kLinesPerThread; ++r)
    c += is_winner(kind, li, lj, num_lines, dd,
                   (blockIdx.x * kLinesPerThread + r) * kThreads + threadIdx.x);
  uint32_t total;
  block_exclusive(c, &total);
  if (threadIdx.x == 0) tile_count[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kThreads) compact_scatter_kernel(
    const uint8_t *kind, const uint32_t *li, const uint32_t *lj, const uint64_t *lv,
    uint32_t num_lines, Dedupe dd, const uint32_t *tile_off, int32_t *ti, int32_t *tj,
    uint64_t *tv, Stats *st) {
  __shared__ uint32_t wcnt[kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t base = tile_off[blockIdx.x], lin = 0;
  for (int r = 0; r < kLinesPerThread; ++r) {
    const uint32_t l = (blockIdx.x * kLinesPerThread + r) * kThreads + threadIdx.x;
    const bool win = is_winner(kind, li, lj, num_lines, dd, l);
    const uint64_t b = __ballot(win);
    if (lane == 0) wcnt[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
      before += w < wave ? wcnt[w] : 0u;
      all += wcnt[w];
    }
    __syncthreads();
    if (win) {
      const uint32_t pos = base + before + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
      const uint32_t i = li[l], j = lj[l];
      ti[pos] = (int32_t)i;
      tj[pos] = (int32_t)j;
      tv[pos] = lv[l];
      lin += i == j;
    }
    base += all;
  }
  lin = wave_sum(lin);
  if (lane == 0 && lin) atomicAdd(&st->n_lin, lin);
}

// ---------------------------------------------------------------- consumers

// flatten_qubo (include/helpers/qubo_helpers.hpp:26-44): the diagonal takes the
// linear value, both off-diagonal cells 0.0 + value (so -0.0 becomes +0.0)
__global__ void scatter_dense_kernel(const int32_t *ti, const int32_t *tj, const uint64_t *tv,
                                     uint64_t count, uint32_t n, uint64_t *q) {
  const uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (k >= count) return;
  const uint64_t i = (uint32_t)ti[k], j = (uint32_t)tj[k];
  uint64_t v = tv[k];
  if (i == j) {
    q[i * n + i] = v;
  } else {
    if (v == 0x8000000000000000ull) v = 0;
    q[i * n + j] = v;
    q[j * n + i] = v;
  }
}

// binary input: every term has 0 <= i <= j < n; counts linear terms and the
// largest index
__global__ void check_terms_kernel(const int32_t *ti, const int32_t *tj, uint32_t count,
                                   uint32_t n, Stats *st) {
  const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
  uint32_t bad = 0, lin = 0, mx = 0;
  if (k < count) {
    const int32_t i = ti[k], j = tj[k];
    bad = i < 0 || j < 0 || i > j || (uint32_t)j >= n;
    lin = !bad && i == j;
    mx = bad ? 0 : (uint32_t)j;
  }
  bad = wave_sum(bad);
  lin = wave_sum(lin);
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) {
    if (bad) atomicAdd(&st->n_bad, bad);
    if (lin) atomicAdd(&st->n_lin, lin);
    atomicMax(&st->max_index, mx);
  }
}

__global__ void terms_as_lines_kernel(const int32_t *ti, const int32_t *tj, uint32_t count,
                                      uint8_t *kind, uint32_t *li, uint32_t *lj) {
  const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
  if (k < count) {
    kind[k] = K_COEF;
    li[k] = (uint32_t)ti[k];
    lj[k] = (uint32_t)tj[k];
  }
}

inline uint32_t blocks_for(uint64_t n, uint64_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace

// ---------------------------------------------------------------- the model

namespace os {

struct Qubo {
  Ctx *ctx = nullptr;
  int32_t n = 0, max_nodes = 0;
  int64_t num_linear = 0, num_quadratic = 0;
  // the term arrays: views (never released themselves) into one allocation
  DevBuf mem, ti, tj, tv;
  // timings of the last text parse (ms): line index, parse, dedupe + compaction
  float ms_index = 0, ms_parse = 0, ms_dedupe = 0;
  ~Qubo() { mem.release(); }
  int64_t terms() const { return num_linear + num_quadratic; }
  // room for T terms: int32 i[T], int32 j[T], 64-bit values[T], 256-byte aligned
  void alloc_terms(size_t T) {
    const size_t a = (4 * T + 4 + 255) & ~(size_t)255, b = (8 * T + 8 + 255) & ~(size_t)255;
    mem.ensure(2 * a + b);
    char *base = static_cast<char *>(mem.ptr);
    ti.ptr = base;
    tj.ptr = base + a;
    tv.ptr = base + 2 * a;
    ti.cap = tj.cap = a;
    tv.cap = b;
  }
};

}  // namespace os

namespace {

// QUBOBuilder semantics (include/model/qubo.hpp:282-379) with the terms kept
// in order of first appearance, for texts the line parser hands back.
struct OrderedBuilder {
  std::vector<int32_t> ti, tj;
  std::vector<double> tv;
  std::unordered_map<uint64_t, int64_t> seen;
  int64_t nlin = 0, nquad = 0;
  int num_quadratic = -1, num_linear = -1, max_nodes = -1;
  bool used = false;
  int max_used = 0;
  void add_element(int i, int j, double coef) {
    if (i > j)
      throw std::invalid_argument(
          "Incorrect file, encountered coefficient from lower triangle of QUBO matrix.");
    const uint64_t key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
    if (seen.emplace(key, (int64_t)ti.size()).second) {
      ti.push_back(i);
      tj.push_back(j);
      tv.push_back(coef);
      ++(i == j ? nlin : nquad);
    }
    max_used = used ? std::max(max_used, j) : j;
    used = true;
  }
  void set_num_quadratic(int v) { num_quadratic = v; }
  void set_num_linear(int v) { num_linear = v; }
  void set_max_nodes(int v) { max_nodes = v; }
};

// QUBOBuilder::build_qubo's checks, qubo.hpp:353-377
void builder_checks(int64_t nlin, int64_t nquad, bool header, int max_nodes, int num_linear,
                    int num_quadratic) {
  if (nlin == 0 && nquad == 0)
    throw os::Error(OS_ERR_INVALID, "An empty input, no coefficients defined.");
  if (!header || num_quadratic == -1 || num_linear == -1 || max_nodes == -1)
    throw os::Error(OS_ERR_INVALID, "No header line or header line misformatted.");
  if (num_linear > max_nodes)
    throw os::Error(OS_ERR_INVALID, "Number of linear terms is greater than num nodes.");
  if ((uint64_t)nquad != (uint64_t)(int64_t)num_quadratic)
    throw os::Error(OS_ERR_INVALID, "Number of quadratic terms is not equal to the declared one.");
  if ((uint64_t)nlin != (uint64_t)(int64_t)num_linear)
    throw os::Error(OS_ERR_INVALID, "Number of linear terms is not equal to the declared one.");
}

void upload_terms(os::Qubo &q, const int32_t *ti, const int32_t *tj, const void *tv) {
  Ctx &c = *q.ctx;
  const size_t T = (size_t)q.terms();
  q.alloc_terms(T);
  OS_HIP_CHECK(hipMemcpyAsync(q.ti.ptr, ti, 4 * T, hipMemcpyHostToDevice, c.stream));
  OS_HIP_CHECK(hipMemcpyAsync(q.tj.ptr, tj, 4 * T, hipMemcpyHostToDevice, c.stream));
  OS_HIP_CHECK(hipMemcpyAsync(q.tv.ptr, tv, 8 * T, hipMemcpyHostToDevice, c.stream));
}

// The whole text through the sequential reader: throws the reference's
// exception, or fills q in file order.
void sequential_parse(os::Qubo &q, const char *text, size_t len) {
  OrderedBuilder b;
  bool ok;
  try {
    qubo::detail::QuboReader<const char *, OrderedBuilder> reader(text, text + len, b);
    ok = reader.run();
  } catch (const std::invalid_argument &e) {
    throw os::Error(OS_ERR_INVALID, e.what());
  }
  if (!ok) throw os::Error(OS_ERR_INVALID, "Parsing failed. Incorrect file format.");
  builder_checks(b.nlin, b.nquad, true, b.max_nodes, b.num_linear, b.num_quadratic);
  // the reference converts indices above 2^31 - 1 to negative ints
  for (size_t k = 0; k < b.ti.size(); ++k)
    if (b.ti[k] < 0 || b.tj[k] < 0)
      throw os::Error(OS_ERR_UNSUPPORTED, "variable indices above 2^31 - 1 are not supported");
  q.n = b.max_used + 1;
  q.max_nodes = b.max_nodes;
  q.num_linear = b.nlin;
  q.num_quadratic = b.nquad;
  upload_terms(q, b.ti.data(), b.tj.data(), b.tv.data());
  q.ctx->sync();
}

// Work arrays of a parse: the context's, kept for the next parse (the
// sequential fallback releases them).
struct Scratch {
  DevBuf &text, &tiles, &starts, &kind, &li, &lj, &lv, &stats, &hostnum, &hvals, &owner, &keys,
      &vals;
  explicit Scratch(Ctx &c)
      : text(c.ingest[0]), tiles(c.ingest[1]), starts(c.ingest[2]), kind(c.ingest[3]),
        li(c.ingest[4]), lj(c.ingest[5]), lv(c.ingest[6]), stats(c.ingest[7]),
        hostnum(c.ingest[8]), hvals(c.ingest[9]), owner(c.ingest[10]), keys(c.ingest[11]),
        vals(c.ingest[12]) {}
  void release() {
    for (auto *b : {&text, &tiles, &starts, &kind, &li, &lj, &lv, &stats, &hostnum, &hvals,
                    &owner, &keys, &vals})
      b->release();
  }
};

// First-wins dedupe and file-order compaction of `num_lines` parsed lines
// into q's term arrays; returns the number of terms, linear ones in *nlin.
uint32_t dedupe_compact(os::Qubo &q, Scratch &s, uint32_t num_lines, uint32_t n, uint32_t n_coef,
                        Stats *st, uint32_t *nlin, bool rising = false) {
  Ctx &c = *q.ctx;
  hipStream_t stream = c.stream;
  Dedupe dd{};
  dd.n = n;
  const uint32_t gl = blocks_for(num_lines, kThreads);
  const uint64_t nn = (uint64_t)n * n;
  if (rising) {
    dd.all = true;
  } else if (n <= 16384 && nn <= 16ull * n_coef + (1ull << 22)) {
    s.owner.ensure(4 * nn);
    OS_HIP_CHECK(hipMemsetAsync(s.owner.ptr, 0xFF, 4 * nn, stream));
    owner_min_kernel<<<gl, kThreads, 0, stream>>>(s.kind.as<uint8_t>(), s.li.as<uint32_t>(),
                                                  s.lj.as<uint32_t>(), num_lines, n,
                                                  s.owner.as<uint32_t>());
    dd.owner = s.owner.as<uint32_t>();
  } else {
    uint64_t cap = 1024;
    while (cap < 2ull * n_coef) cap <<= 1;
    s.keys.ensure(8 * cap);
    s.vals.ensure(4 * cap);
    OS_HIP_CHECK(hipMemsetAsync(s.keys.ptr, 0xFF, 8 * cap, stream));
    OS_HIP_CHECK(hipMemsetAsync(s.vals.ptr, 0xFF, 4 * cap, stream));
    hash_insert_kernel<<<gl, kThreads, 0, stream>>>(s.kind.as<uint8_t>(), s.li.as<uint32_t>(),
                                                    s.lj.as<uint32_t>(), num_lines,
                                                    s.keys.as<uint64_t>(), s.vals.as<uint32_t>(),
                                                    cap - 1);
    dd.keys = s.keys.as<uint64_t>();
    dd.vals = s.vals.as<uint32_t>();
    dd.mask = cap - 1;
  }
  OS_HIP_CHECK(hipGetLastError());
  const uint32_t tiles = blocks_for(num_lines, kTileLines);
  s.tiles.ensure(4ull * (tiles + 1));
  compact_count_kernel<<<tiles, kThreads, 0, stream>>>(s.kind.as<uint8_t>(), s.li.as<uint32_t>(),
                                                       s.lj.as<uint32_t>(), num_lines, dd,
                                                       s.tiles.as<uint32_t>());
  scan_tiles_kernel<<<1, kScanThreads, 0, stream>>>(s.tiles.as<uint32_t>(), tiles);
  OS_HIP_CHECK(hipGetLastError());
  uint32_t T = 0;
  OS_HIP_CHECK(hipMemcpyAsync(&T, s.tiles.as<uint32_t>() + tiles, 4, hipMemcpyDeviceToHost,
                              stream));
  c.sync();
  q.alloc_terms(T);
  compact_scatter_kernel<<<tiles, kThreads, 0, stream>>>(
      s.kind.as<uint8_t>(), s.li.as<uint32_t>(), s.lj.as<uint32_t>(), s.lv.as<uint64_t>(),
      num_lines, dd, s.tiles.as<uint32_t>(), q.ti.as<int32_t>(), q.tj.as<int32_t>(),
      q.tv.as<uint64_t>(), st);
  OS_HIP_CHECK(hipGetLastError());
  Stats h;
  OS_HIP_CHECK(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, stream));
  c.sync();
  *nlin = h.n_lin;
  return T;
}

// Host text to the GPU through the context's two pinned staging buffers:
// while one buffer's copy to the device runs, the other is filled (by
// kCopyThreads host threads).
constexpr size_t kPinnedBytes = 32u << 20;
constexpr int kCopyThreads = 4;

void upload_text(Ctx &c, uint8_t *dst, const char *text, size_t len) {
  if (len == 0) return;
  if (len <= (4u << 20)) {  // small texts: one direct copy
    OS_HIP_CHECK(hipMemcpyAsync(dst, text, len, hipMemcpyHostToDevice, c.stream));
    return;
  }
  for (int k = 0; k < 2; ++k) {
    if (!c.pinned[k]) OS_HIP_CHECK(hipHostMalloc(&c.pinned[k], kPinnedBytes, hipHostMallocDefault));
    if (!c.pinned_ev[k])
      OS_HIP_CHECK(hipEventCreateWithFlags(&c.pinned_ev[k], hipEventDisableTiming));
  }
  int k = 0;
  for (size_t off = 0; off < len; off += kPinnedBytes, k ^= 1) {
    const size_t n = std::min(kPinnedBytes, len - off);
    // the copy that last read this buffer has finished
    if (off >= 2 * kPinnedBytes) OS_HIP_CHECK(hipEventSynchronize(c.pinned_ev[k]));
    char *buf = static_cast<char *>(c.pinned[k]);
    const size_t part = (n + kCopyThreads - 1) / kCopyThreads;
    std::vector<std::thread> pool;
    for (int t = 1; t < kCopyThreads; ++t) {
      const size_t a = std::min(n, t * part), b = std::min(n, a + part);
      if (a < b) pool.emplace_back([=] { std::memcpy(buf + a, text + off + a, b - a); });
    }
    std::memcpy(buf, text + off, std::min(n, part));
    for (auto &th : pool) th.join();
    OS_HIP_CHECK(hipMemcpyAsync(dst + off, buf, n, hipMemcpyHostToDevice, c.stream));
    OS_HIP_CHECK(hipEventRecord(c.pinned_ev[k], c.stream));
  }
}

// Text from the host (`text`) or already on the GPU (`dev`, 16-byte aligned).
void parse_text(os::Qubo &q, const char *text, const uint8_t *dev, size_t len) {
  Ctx &c = *q.ctx;
  hipStream_t stream = c.stream;
  OS_REQUIRE(len < (1ull << 40), "text longer than 2^40 bytes");
  Scratch s(c);
  const uint64_t tiles64 = (len + kTileBytes) / kTileBytes;  // one more byte: a final '\r'
  OS_REQUIRE(tiles64 < (1ull << 31), "text too long");
  const uint32_t tiles = (uint32_t)tiles64;
  if (!dev) {
    s.text.ensure(std::max<size_t>(len, 16));
    upload_text(c, s.text.as<uint8_t>(), text, len);
    dev = s.text.as<uint8_t>();
  }
  // the host copy of a device text, made only when the host needs it
  std::vector<char> host_copy;
  auto host_text = [&]() -> const char * {
    if (text) return text;
    if (host_copy.size() != len) {
      host_copy.resize(len);
      if (len)
        OS_HIP_CHECK(hipMemcpyAsync(host_copy.data(), dev, len, hipMemcpyDeviceToHost, stream));
      c.sync();
    }
    return host_copy.data();
  };
  s.tiles.ensure(4ull * (tiles + 1));
  s.stats.ensure(sizeof(Stats));
  Stats *st = s.stats.as<Stats>();
  init_stats_kernel<<<1, 64, 0, stream>>>(st);

  OS_HIP_CHECK(hipEventRecord(c.ev_start, stream));
  line_count_kernel<<<tiles, kThreads, 0, stream>>>(dev, len, s.tiles.as<uint32_t>());
  scan_tiles_kernel<<<1, kScanThreads, 0, stream>>>(s.tiles.as<uint32_t>(), tiles);
  OS_HIP_CHECK(hipGetLastError());
  uint32_t ends = 0;
  OS_HIP_CHECK(hipMemcpyAsync(&ends, s.tiles.as<uint32_t>() + tiles, 4, hipMemcpyDeviceToHost,
                              stream));
  c.sync();
  OS_REQUIRE(ends < 0xFFFFFFF0u, "more than 2^32 - 16 lines");
  const uint32_t L = ends + 1;
  s.starts.ensure(8ull * (L + 1));
  line_starts_kernel<<<tiles, kThreads, 0, stream>>>(dev, len, s.tiles.as<uint32_t>(), tiles,
                                                    s.starts.as<uint64_t>());
  OS_HIP_CHECK(hipEventRecord(c.ev_stop, stream));
  OS_HIP_CHECK(hipGetLastError());

  s.kind.ensure(L);
  s.li.ensure(4ull * L);
  s.lj.ensure(4ull * L);
  s.lv.ensure(8ull * L);
  const uint32_t hcap = std::min<uint32_t>(L, 1u << 22);
  s.hostnum.ensure(sizeof(HostNum) * hcap);
  for (hipEvent_t &e : c.ingest_ev)
    if (!e) OS_HIP_CHECK(hipEventCreate(&e));
  hipEvent_t e1 = c.ingest_ev[0], e2 = c.ingest_ev[1];
  OS_HIP_CHECK(hipEventRecord(e1, stream));
  const uint32_t pgrid = blocks_for(L, (uint64_t)kThreads * kParseRounds);
  parse_lines_kernel<true><<<pgrid, kThreads, 0, stream>>>(
      dev, len, s.starts.as<uint64_t>(), L, s.kind.as<uint8_t>(), s.li.as<uint32_t>(),
      s.lj.as<uint32_t>(), s.lv.as<uint64_t>(), st, s.hostnum.as<HostNum>(), hcap);
  OS_HIP_CHECK(hipEventRecord(e2, stream));
  order_check_kernel<<<blocks_for(L, kThreads), kThreads, 0, stream>>>(
      s.kind.as<uint8_t>(), s.li.as<uint32_t>(), s.lj.as<uint32_t>(), L, st);
  OS_HIP_CHECK(hipGetLastError());
  Stats h;
  OS_HIP_CHECK(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, stream));
  c.sync();
  // the order check saw the lines of the staged rounds only
  bool rising = h.n_runs == 1 && h.n_unsorted == 0;
  if (h.n_long) {
    rising = false;
    // the rounds too long for LDS, from global memory
    parse_lines_kernel<false><<<pgrid, kThreads, 0, stream>>>(
        dev, len, s.starts.as<uint64_t>(), L, s.kind.as<uint8_t>(), s.li.as<uint32_t>(),
        s.lj.as<uint32_t>(), s.lv.as<uint64_t>(), st, s.hostnum.as<HostNum>(), hcap);
    OS_HIP_CHECK(hipEventRecord(e2, stream));
    OS_HIP_CHECK(hipGetLastError());
    OS_HIP_CHECK(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, stream));
    c.sync();
  }
  (void)hipEventElapsedTime(&q.ms_index, c.ev_start, c.ev_stop);
  (void)hipEventElapsedTime(&q.ms_parse, e1, e2);

  const bool header = h.n_header == 1;
  if (h.n_bad || h.n_header > 1 || (header && h.header_line != h.first_other) ||
      h.n_hostnum > hcap) {
    const char *t = host_text();
    s.release();
    sequential_parse(q, t, len);
    return;
  }
  if (h.n_hostnum) {
    // decimal strings the kernel left open: strtod, as the sequential reader
    std::vector<HostNum> hn(h.n_hostnum);
    OS_HIP_CHECK(hipMemcpyAsync(hn.data(), s.hostnum.ptr, sizeof(HostNum) * hn.size(),
                                hipMemcpyDeviceToHost, stream));
    c.sync();
    std::vector<char> toks;
    bool all_short = true;
    for (const HostNum &x : hn) all_short = all_short && x.len <= kTokBytes;
    if (!text && all_short) {
      // only the tokens come back from a device text
      s.hvals.ensure(kTokBytes * hn.size());
      gather_tokens_kernel<<<blocks_for(hn.size(), kThreads), kThreads, 0, stream>>>(
          dev, s.hostnum.as<HostNum>(), (uint32_t)hn.size(), s.hvals.as<uint8_t>());
      OS_HIP_CHECK(hipGetLastError());
      toks.resize(kTokBytes * hn.size());
      OS_HIP_CHECK(hipMemcpyAsync(toks.data(), s.hvals.ptr, toks.size(), hipMemcpyDeviceToHost,
                                  stream));
      c.sync();
    }
    std::vector<double> vals(hn.size());
    std::string tok;
    for (size_t k = 0; k < hn.size(); ++k) {
      if (!toks.empty())
        tok.assign(toks.data() + kTokBytes * k, hn[k].len);
      else
        tok.assign(host_text() + hn[k].start, hn[k].len);
      vals[k] = std::strtod(tok.c_str(), nullptr);
    }
    s.hvals.ensure(8 * vals.size());
    OS_HIP_CHECK(hipMemcpyAsync(s.hvals.ptr, vals.data(), 8 * vals.size(), hipMemcpyHostToDevice,
                                stream));
    fix_values_kernel<<<blocks_for(hn.size(), kThreads), kThreads, 0, stream>>>(
        s.hostnum.as<HostNum>(), s.hvals.as<uint64_t>(), (uint32_t)hn.size(),
        s.lv.as<uint64_t>());
    OS_HIP_CHECK(hipGetLastError());
  }
  const int max_nodes = (int)h.hdr[1], num_linear = (int)h.hdr[2], num_quadratic = (int)h.hdr[3];
  if (h.n_coef == 0)
    builder_checks(0, 0, header, max_nodes, num_linear, num_quadratic);
  const uint32_t n = h.max_index + 1;
  hipEvent_t e3 = c.ingest_ev[2], e4 = c.ingest_ev[3];
  OS_HIP_CHECK(hipEventRecord(e3, stream));
  uint32_t nlin = 0;
  const uint32_t T = dedupe_compact(q, s, L, n, h.n_coef, st, &nlin, rising);
  OS_HIP_CHECK(hipEventRecord(e4, stream));
  c.sync();
  (void)hipEventElapsedTime(&q.ms_dedupe, e3, e4);
  builder_checks(nlin, (int64_t)T - nlin, header, max_nodes, num_linear, num_quadratic);
  q.n = (int32_t)n;
  q.max_nodes = max_nodes;
  q.num_linear = nlin;
  q.num_quadratic = (int64_t)T - nlin;
}

constexpr char kMagic[4] = {'O', 'S', 'Q', 'B'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeaderBytes = 32;

void read_binary(os::Qubo &q, const uint8_t *buf, size_t len) {
  Ctx &c = *q.ctx;
  hipStream_t stream = c.stream;
  OS_REQUIRE(len >= kHeaderBytes && std::memcmp(buf, kMagic, 4) == 0,
             "not a binary QUBO (magic OSQB)");
  uint32_t version;
  int32_t n, max_nodes;
  int64_t nl, nq;
  std::memcpy(&version, buf + 4, 4);
  std::memcpy(&n, buf + 8, 4);
  std::memcpy(&max_nodes, buf + 12, 4);
  std::memcpy(&nl, buf + 16, 8);
  std::memcpy(&nq, buf + 24, 8);
  OS_REQUIRE(version == kVersion, "binary QUBO: unsupported version " + std::to_string(version));
  OS_REQUIRE(nl >= 0 && nq >= 0 && nl + nq >= 1 && nl + nq < (1ll << 32),
             "binary QUBO: bad term counts");
  OS_REQUIRE(n >= 1, "binary QUBO: bad n");
  const uint64_t T = (uint64_t)(nl + nq);
  OS_REQUIRE(len == kHeaderBytes + 16 * T, "binary QUBO: size does not match the term counts");
  q.n = n;
  q.max_nodes = max_nodes;
  q.num_linear = nl;
  q.num_quadratic = nq;
  upload_terms(q, reinterpret_cast<const int32_t *>(buf + kHeaderBytes),
               reinterpret_cast<const int32_t *>(buf + kHeaderBytes + 4 * T),
               buf + kHeaderBytes + 8 * T);
  Scratch s(c);
  s.stats.ensure(sizeof(Stats));
  Stats *st = s.stats.as<Stats>();
  init_stats_kernel<<<1, 64, 0, stream>>>(st);
  check_terms_kernel<<<blocks_for(T, kThreads), kThreads, 0, stream>>>(
      q.ti.as<int32_t>(), q.tj.as<int32_t>(), (uint32_t)T, (uint32_t)n, st);
  OS_HIP_CHECK(hipGetLastError());
  Stats h;
  OS_HIP_CHECK(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, stream));
  c.sync();
  OS_REQUIRE(h.n_bad == 0, "binary QUBO: a term outside 0 <= i <= j < n");
  OS_REQUIRE(h.n_lin == (uint64_t)nl, "binary QUBO: linear term count does not match");
  OS_REQUIRE(h.max_index + 1 == (uint32_t)n, "binary QUBO: n is not the largest index + 1");
  // no repeated key: dedupe must keep every term
  s.kind.ensure(T);
  s.li.ensure(4 * T);
  s.lj.ensure(4 * T);
  terms_as_lines_kernel<<<blocks_for(T, kThreads), kThreads, 0, stream>>>(
      q.ti.as<int32_t>(), q.tj.as<int32_t>(), (uint32_t)T, s.kind.as<uint8_t>(),
      s.li.as<uint32_t>(), s.lj.as<uint32_t>());
  OS_HIP_CHECK(hipGetLastError());
  s.lv.ensure(8 * T);
  OS_HIP_CHECK(hipMemcpyAsync(s.lv.ptr, q.tv.ptr, 8 * T, hipMemcpyDeviceToDevice, stream));
  os::Qubo tmp;
  tmp.ctx = &c;
  init_stats_kernel<<<1, 64, 0, stream>>>(st);
  uint32_t nlin = 0;
  const uint32_t kept = dedupe_compact(tmp, s, (uint32_t)T, (uint32_t)n, (uint32_t)T, st, &nlin);
  OS_REQUIRE(kept == T, "binary QUBO: a repeated (i, j) key");
}

}  // namespace

// ---------------------------------------------------------------- C ABI

extern "C" {

int os_qubo_parse(os_ctx *ctx_, const char *text, size_t len, os_qubo **out) {
  return os::guarded([&] {
    OS_REQUIRE(ctx_ && out && (text || len == 0), "null argument");
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    std::unique_ptr<os::Qubo> q(new os::Qubo());
    q->ctx = &c;
    parse_text(*q, text ? text : "", nullptr, len);
    *out = reinterpret_cast<os_qubo *>(q.release());
  });
}

int os_qubo_parse_device(os_ctx *ctx_, const void *text, size_t len, os_qubo **out) {
  return os::guarded([&] {
    OS_REQUIRE(ctx_ && out && (text || len == 0), "null argument");
    OS_REQUIRE(((uintptr_t)text & 15) == 0, "device text must be 16-byte aligned");
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    std::unique_ptr<os::Qubo> q(new os::Qubo());
    q->ctx = &c;
    DevBuf empty;
    if (!text) {
      empty.ensure(16);
      text = empty.ptr;
    }
    try {
      parse_text(*q, nullptr, static_cast<const uint8_t *>(text), len);
    } catch (...) {
      empty.release();
      throw;
    }
    empty.release();
    *out = reinterpret_cast<os_qubo *>(q.release());
  });
}

int os_qubo_read_binary(os_ctx *ctx_, const void *data, size_t len, os_qubo **out) {
  return os::guarded([&] {
    OS_REQUIRE(ctx_ && data && out, "null argument");
    Ctx &c = *reinterpret_cast<Ctx *>(ctx_);
    c.activate();
    std::unique_ptr<os::Qubo> q(new os::Qubo());
    q->ctx = &c;
    read_binary(*q, static_cast<const uint8_t *>(data), len);
    *out = reinterpret_cast<os_qubo *>(q.release());
  });
}

int os_qubo_destroy(os_qubo *q) {
  return os::guarded([&] { delete reinterpret_cast<os::Qubo *>(q); });
}

int os_qubo_info(os_qubo *q_, int32_t *n, int32_t *max_nodes, int64_t *num_linear,
                 int64_t *num_quadratic) {
  return os::guarded([&] {
    OS_REQUIRE(q_, "null argument");
    const os::Qubo &q = *reinterpret_cast<os::Qubo *>(q_);
    if (n) *n = q.n;
    if (max_nodes) *max_nodes = q.max_nodes;
    if (num_linear) *num_linear = q.num_linear;
    if (num_quadratic) *num_quadratic = q.num_quadratic;
  });
}

int os_qubo_parse_times(os_qubo *q_, float *ms) {
  return os::guarded([&] {
    OS_REQUIRE(q_ && ms, "null argument");
    const os::Qubo &q = *reinterpret_cast<os::Qubo *>(q_);
    ms[0] = q.ms_index;
    ms[1] = q.ms_parse;
    ms[2] = q.ms_dedupe;
  });
}

int os_qubo_terms(os_qubo *q_, int32_t *i, int32_t *j, double *v) {
  return os::guarded([&] {
    OS_REQUIRE(q_, "null argument");
    os::Qubo &q = *reinterpret_cast<os::Qubo *>(q_);
    q.ctx->activate();
    const size_t T = (size_t)q.terms();
    hipStream_t st = q.ctx->stream;
    if (i) OS_HIP_CHECK(hipMemcpyAsync(i, q.ti.ptr, 4 * T, hipMemcpyDeviceToHost, st));
    if (j) OS_HIP_CHECK(hipMemcpyAsync(j, q.tj.ptr, 4 * T, hipMemcpyDeviceToHost, st));
    if (v) OS_HIP_CHECK(hipMemcpyAsync(v, q.tv.ptr, 8 * T, hipMemcpyDeviceToHost, st));
    q.ctx->sync();
  });
}

int os_qubo_flatten(os_qubo *q_, double *q_sym) {
  return os::guarded([&] {
    OS_REQUIRE(q_ && q_sym, "null argument");
    os::Qubo &q = *reinterpret_cast<os::Qubo *>(q_);
    Ctx &c = *q.ctx;
    c.activate();
    const uint64_t nn = (uint64_t)q.n * q.n;
    OS_REQUIRE(nn <= (1ull << 33), "n x n matrix larger than 64 GiB");
    DevBuf buf;
    try {
      buf.ensure(8 * nn);
      OS_HIP_CHECK(hipMemsetAsync(buf.ptr, 0, 8 * nn, c.stream));
      scatter_dense_kernel<<<blocks_for(q.terms(), kThreads), kThreads, 0, c.stream>>>(
          q.ti.as<int32_t>(), q.tj.as<int32_t>(), q.tv.as<uint64_t>(), (uint64_t)q.terms(),
          (uint32_t)q.n, buf.as<uint64_t>());
      OS_HIP_CHECK(hipGetLastError());
      OS_HIP_CHECK(hipMemcpyAsync(q_sym, buf.ptr, 8 * nn, hipMemcpyDeviceToHost, c.stream));
      c.sync();
    } catch (...) {
      buf.release();
      throw;
    }
    buf.release();
  });
}

int os_qubo_dense(os_qubo *q_, os_dense **out) {
  return os::guarded([&] {
    OS_REQUIRE(q_ && out, "null argument");
    os::Qubo &q = *reinterpret_cast<os::Qubo *>(q_);
    struct Fill {
      os::Qubo *q;
      static void run(void *arg, double *q64) {
        os::Qubo &m = *static_cast<Fill *>(arg)->q;
        scatter_dense_kernel<<<blocks_for(m.terms(), kThreads), kThreads, 0, m.ctx->stream>>>(
            m.ti.as<int32_t>(), m.tj.as<int32_t>(), m.tv.as<uint64_t>(), (uint64_t)m.terms(),
            (uint32_t)m.n, reinterpret_cast<uint64_t *>(q64));
        OS_HIP_CHECK(hipGetLastError());
      }
    } f{&q};
    *out = reinterpret_cast<os_dense *>(os::dense_create_device(*q.ctx, q.n, &Fill::run, &f));
  });
}

int64_t os_qubo_binary_size(os_qubo *q_) {
  if (!q_) return -1;
  return (int64_t)kHeaderBytes + 16 * reinterpret_cast<os::Qubo *>(q_)->terms();
}

int os_qubo_write_binary(os_qubo *q_, void *buf, size_t len) {
  return os::guarded([&] {
    OS_REQUIRE(q_ && buf, "null argument");
    os::Qubo &q = *reinterpret_cast<os::Qubo *>(q_);
    const uint64_t T = (uint64_t)q.terms();
    OS_REQUIRE(len >= kHeaderBytes + 16 * T, "buffer smaller than os_qubo_binary_size");
    uint8_t *b = static_cast<uint8_t *>(buf);
    std::memcpy(b, kMagic, 4);
    std::memcpy(b + 4, &kVersion, 4);
    std::memcpy(b + 8, &q.n, 4);
    std::memcpy(b + 12, &q.max_nodes, 4);
    std::memcpy(b + 16, &q.num_linear, 8);
    std::memcpy(b + 24, &q.num_quadratic, 8);
    q.ctx->activate();
    hipStream_t st = q.ctx->stream;
    OS_HIP_CHECK(hipMemcpyAsync(b + kHeaderBytes, q.ti.ptr, 4 * T, hipMemcpyDeviceToHost, st));
    OS_HIP_CHECK(hipMemcpyAsync(b + kHeaderBytes + 4 * T, q.tj.ptr, 4 * T, hipMemcpyDeviceToHost,
                                st));
    OS_HIP_CHECK(hipMemcpyAsync(b + kHeaderBytes + 8 * T, q.tv.ptr, 8 * T, hipMemcpyDeviceToHost,
                                st));
    q.ctx->sync();
  });
}

}  // extern "C"
