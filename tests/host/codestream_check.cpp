// Test infrastructure: a JPEG XL codestream reader used to check the
// library's emitted streams.  It restates the reading rules of the
// reference's decoder (jpxl-decode, jpxl-entropy, jpxl-bitstream,
// jpxl-core/src/modular_weighted.rs): headers and colour encodings
// (jpxl-decode/src/headers/*.rs), the frame header and TOC (frame/header.rs,
// frame/toc.rs), the entropy streams with prefix codes or ANS, cluster maps
// and LZ77 (jpxl-entropy/src/decoder.rs, dist.rs, prefix.rs, hybrid.rs,
// ans.rs, lz77.rs), modular sub-bitstreams with local or global MA trees and
// every predictor including the weighted one (modular/mod.rs, tree.rs,
// predictor.rs), and the VarDCT sections (vardct/block_ctx.rs, order.rs,
// hf_coeff.rs): custom HF block contexts, custom coefficient orders and every
// DctSelect type's placement and HF coefficients.  It returns the quantized
// plan it read (for DCT8x8/16x16/32x32 varblocks the HF coefficients in the
// footprint layout), the LF planes, the block map, HfMul, the CfL tiles and
// the sharpness.  It never renders.  Features neither the library nor the
// reference's VarDCT fixtures use (patches, splines, noise, custom
// dequantization matrices, transforms of modular images, extra channels,
// progressive passes, ICC profiles) are refused with a message naming the
// clause.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace {

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};
[[noreturn]] void fail(const std::string& m) { throw Fail(m); }

// ---- bits (B.2) ------------------------------------------------------------
struct Reader {
  const uint8_t* d;
  size_t n;
  uint64_t pos = 0;  // bits
  uint32_t bits(uint32_t k) {
    if (k == 0) return 0;
    if (pos + k > (uint64_t)n * 8) fail("read past the end of the stream");
    uint64_t v = 0;
    const size_t byte = pos >> 3;
    const uint32_t sh = pos & 7;
    for (uint32_t i = 0; i < 8 && byte + i < n; i++) v |= (uint64_t)d[byte + i] << (8 * i);
    pos += k;
    return (uint32_t)((v >> sh) & ((k == 32) ? 0xFFFFFFFFull : ((1ull << k) - 1)));
  }
  bool b() { return bits(1) != 0; }
  void pad() {
    const uint32_t r = pos & 7;
    if (r) {
      if (bits(8 - r) != 0) fail("B.2.7: ZeroPadToByte bits are not zero");
    }
  }
};

struct Dist {
  bool val;
  uint32_t bits, off;
};
constexpr Dist V(uint32_t v) { return {true, 0, v}; }
constexpr Dist B(uint32_t b, uint32_t o = 0) { return {false, b, o}; }
uint32_t u32(Reader& r, const Dist (&d)[4]) {
  const Dist& x = d[r.bits(2)];
  return x.val ? x.off : x.off + r.bits(x.bits);
}
uint64_t u64(Reader& r) {
  const uint32_t s = r.bits(2);
  if (s == 0) return 0;
  if (s == 1) return 1 + r.bits(4);
  if (s == 2) return 17 + r.bits(8);
  uint64_t v = r.bits(12);
  uint32_t shift = 12;
  while (r.b()) {
    if (shift == 60) {
      v |= (uint64_t)r.bits(4) << 60;
      break;
    }
    v |= (uint64_t)r.bits(8) << shift;
    shift += 8;
  }
  return v;
}
float f16(Reader& r) {
  const uint32_t h = r.bits(16);
  const uint32_t sign = h >> 15, e = (h >> 10) & 31, m = h & 1023;
  if (e == 31) fail("B.2.4: F16 is not finite");
  float v = e == 0 ? std::ldexp((float)m, -24) : std::ldexp((float)(m + 1024), (int)e - 25);
  return sign ? -v : v;
}
uint32_t bit_width(uint32_t n) { return n ? 32 - __builtin_clz(n) : 0; }
uint32_t ceil_log2(uint64_t n) {
  uint32_t b = 0;
  while ((1ull << b) < n) b++;
  return b;
}
int32_t unpack_signed(uint32_t u) { return (u & 1) ? -(int32_t)(u >> 1) - 1 : (int32_t)(u >> 1); }

// ---- entropy (C.2, C.3) -----------------------------------------------------
struct Hybrid {
  uint32_t se = 0, msb = 0, lsb = 0;
  void read(Reader& r, uint32_t log_alpha) {
    se = r.bits(bit_width(log_alpha));
    if (se > log_alpha) fail("C.2.3: split_exponent exceeds log_alphabet_size");
    if (se == log_alpha) { msb = lsb = 0; return; }
    msb = r.bits(bit_width(se));
    if (msb > se) fail("C.2.3: msb_in_token exceeds split_exponent");
    lsb = r.bits(bit_width(se - msb));
    if (lsb > se - msb) fail("C.2.3: lsb_in_token too large");
  }
  uint32_t value(Reader& r, uint32_t tok) const {
    const uint32_t split = 1u << se;
    if (tok < split) return tok;
    const uint32_t in = msb + lsb;
    const uint32_t nb = se - in + ((tok - split) >> in);
    if (nb >= 32) fail("C.3.3: extra-bit count too large");
    const uint32_t low = tok & ((1u << lsb) - 1);
    uint32_t mid = (tok >> lsb) & ((1u << msb) - 1);
    mid |= 1u << msb;
    const uint64_t v = ((((uint64_t)mid << nb) | r.bits(nb)) << lsb) | low;
    if (v >> 32) fail("C.3.3: value exceeds 32 bits");
    return (uint32_t)v;
  }
};

struct Prefix {
  int single = -1;
  uint32_t counts[16] = {0};
  std::vector<uint32_t> sorted;
  void from_lengths(const std::vector<uint8_t>& len) {
    int present = 0, last = 0;
    memset(counts, 0, sizeof counts);
    for (size_t s = 0; s < len.size(); s++)
      if (len[s]) {
        if (len[s] > 15) fail("RFC 7932 3.2: code length over 15");
        counts[len[s]]++;
        present++;
        last = (int)s;
      }
    if (present == 0) fail("RFC 7932 3.2: no symbol has a code");
    if (present == 1) { single = last; return; }
    uint64_t used = 0;
    for (int l = 1; l <= 15; l++) used += (uint64_t)counts[l] << (15 - l);
    if (used != (1u << 15)) fail("RFC 7932 3.2: prefix code incomplete or over-subscribed");
    sorted.clear();
    for (int l = 1; l <= 15; l++)
      for (size_t s = 0; s < len.size(); s++)
        if (len[s] == l) sorted.push_back((uint32_t)s);
  }
  uint32_t decode(Reader& r) const {
    if (single >= 0) return (uint32_t)single;
    uint32_t code = 0, first = 0, index = 0;
    for (int l = 1; l <= 15; l++) {
      code |= r.bits(1);
      if (code < first + counts[l]) return sorted[index + code - first];
      index += counts[l];
      first = (first + counts[l]) << 1;
      code <<= 1;
    }
    fail("RFC 7932 3.2: no code matched");
  }
};

const int kClcOrder[18] = {1, 2, 3, 4, 0, 5, 17, 6, 16, 7, 8, 9, 10, 11, 12, 13, 14, 15};

Prefix read_prefix(Reader& r, uint32_t alphabet) {
  Prefix p;
  if (alphabet == 1) { p.single = 0; return p; }
  const uint32_t hskip = r.bits(2);
  if (hskip == 1) {
    const uint32_t abits = bit_width(alphabet - 1);
    const uint32_t nsym = r.bits(2) + 1;
    uint32_t sym[4];
    for (uint32_t i = 0; i < nsym; i++) {
      sym[i] = r.bits(abits);
      if (sym[i] >= alphabet) fail("RFC 7932 3.4: symbol outside the alphabet");
      for (uint32_t j = 0; j < i; j++)
        if (sym[j] == sym[i]) fail("RFC 7932 3.4: duplicate symbol");
    }
    if (nsym == 1) { p.single = (int)sym[0]; return p; }
    static const uint8_t L2[] = {1, 1}, L3[] = {1, 2, 2}, L4a[] = {2, 2, 2, 2}, L4b[] = {1, 2, 3, 3};
    const uint8_t* lens = nsym == 2 ? L2 : nsym == 3 ? L3 : (r.b() ? L4b : L4a);
    std::vector<uint8_t> len(alphabet, 0);
    for (uint32_t i = 0; i < nsym; i++) len[sym[i]] = lens[i];
    p.from_lengths(len);
    return p;
  }
  Prefix fixed;
  fixed.from_lengths({2, 4, 3, 2, 2, 4});
  std::vector<uint8_t> clc(18, 0);
  int space = 32, nonzero = 0;
  for (int k = (int)hskip; k < 18; k++) {
    const uint32_t l = fixed.decode(r);
    clc[kClcOrder[k]] = (uint8_t)l;
    if (l) {
      nonzero++;
      space -= 32 >> l;
      if (space <= 0) break;
    }
  }
  if (nonzero != 1 && space != 0) fail("RFC 7932 3.5: code-length code incomplete");
  Prefix cl;
  if (nonzero == 1) {
    for (int s = 0; s < 18; s++)
      if (clc[s]) cl.single = s;
  } else {
    cl.from_lengths(clc);
  }
  std::vector<uint8_t> len(alphabet, 0);
  int64_t sp = 32768;
  uint32_t idx = 0, repeat = 0;
  uint8_t prev_nz = 8, repeat_len = 0;
  while (idx < alphabet && sp > 0) {
    const uint32_t s = cl.decode(r);
    if (s < 16) {
      repeat = 0;
      len[idx++] = (uint8_t)s;
      if (s) {
        prev_nz = (uint8_t)s;
        sp -= 32768 >> s;
      }
      continue;
    }
    const uint32_t eb = s == 16 ? 2 : 3;
    const uint8_t nl = s == 16 ? prev_nz : 0;
    if (repeat_len != nl) {
      repeat = 0;
      repeat_len = nl;
    }
    const uint32_t before = repeat;
    if (repeat > 0) repeat = (repeat - 2) << eb;
    repeat += 3 + r.bits(eb);
    const uint32_t cnt = repeat - before;
    if (idx + cnt > alphabet) fail("RFC 7932 3.5: repeat past the alphabet");
    for (uint32_t i = 0; i < cnt; i++) len[idx++] = nl;
    if (nl) sp -= (int64_t)cnt * (32768 >> nl);
  }
  if (sp != 0) fail("RFC 7932 3.5: prefix code incomplete");
  p.from_lengths(len);
  return p;
}

struct SymbolDecoder;
std::vector<uint8_t> read_cluster_map(Reader& r, uint32_t num_dist, int depth);

// ---- ANS (C.2.5, C.2.6, C.3.2; jpxl-entropy/src/ans.rs) ----------------------
const uint32_t kLogCountCode[14][2] = {{5, 0x11}, {4, 0xD}, {4, 0xF}, {4, 0xC}, {4, 0x9}, {4, 0xE}, {3, 0x1},
                                       {3, 0x2}, {3, 0x5}, {3, 0x3}, {3, 0x0}, {6, 0x21}, {7, 0x40}, {7, 0x41}};
uint32_t read_logcount(Reader& r) {
  uint32_t code = 0;
  for (uint32_t len = 1; len <= 7; len++) {
    code = (code << 1) | r.bits(1);
    for (uint32_t n = 0; n < 14; n++)
      if (kLogCountCode[n][0] == len && kLogCountCode[n][1] == code) return n;
  }
  fail("C.2.5: no logcount code matched");
}
uint32_t read_u8(Reader& r) {
  if (!r.b()) return 0;
  const uint32_t n = r.bits(3);
  return r.bits(n) + (1u << n);
}

struct AnsDist {
  uint32_t log_bucket = 0, bucket = 0;
  std::vector<uint32_t> prob, sym, off, cut;
  void read(Reader& r, uint32_t log_alpha) {
    if (log_alpha > 12) fail("C.2.5: log_alphabet_size above 12");
    const uint32_t table = 1u << log_alpha;
    std::vector<uint32_t> d(table, 0);
    uint32_t alpha = 0;
    auto bounded = [&](uint32_t v) {
      if (v >= table) fail("C.2.5: symbol outside the table");
      return v;
    };
    if (r.b()) {
      if (r.b()) {
        const uint32_t v1 = bounded(read_u8(r)), v2 = bounded(read_u8(r));
        if (v1 == v2) fail("C.2.5: two-symbol form with equal symbols");
        const uint32_t p = r.bits(12);
        d[v1] = p;
        d[v2] = 4096 - p;
        alpha = 1 + std::max(v1, v2);
      } else {
        const uint32_t x = bounded(read_u8(r));
        d[x] = 4096;
        alpha = 1 + x;
      }
    } else if (r.b()) {
      alpha = read_u8(r) + 1;
      if (alpha > table) fail("C.2.5: flat alphabet too large");
      for (uint32_t i = 0; i < alpha; i++) d[i] = 4096 / alpha + (i < 4096 % alpha ? 1 : 0);
    } else {
      uint32_t len = 0;
      while (len < 3 && r.b()) len++;
      const int shift = (int)(r.bits(len) + (1u << len) - 1);
      if (shift > 13) fail("C.2.5: shift above 13");
      alpha = read_u8(r) + 3;
      if (alpha > table) fail("C.2.5: alphabet too large");
      std::vector<uint32_t> lc(alpha, 0), same(alpha, 0);
      int omit_log = -1, omit_pos = -1;
      for (uint32_t i = 0; i < alpha;) {
        const uint32_t c = read_logcount(r);
        lc[i] = c;
        if (c == 13) {
          const uint32_t rle = read_u8(r);
          same[i] = rle + 5;
          i += rle + 4;
          continue;
        }
        if ((int)c > omit_log) {
          omit_log = (int)c;
          omit_pos = (int)i;
        }
        i++;
      }
      if (omit_pos < 0) fail("C.2.5: no omitted position");
      uint32_t total = 0, numsame = 0, prev = 0;
      for (uint32_t i = 0; i < alpha; i++) {
        if (same[i]) {
          numsame = same[i] - 1;
          prev = i > 0 ? d[i - 1] : 0;
        }
        if (numsame > 0) {
          d[i] = prev;
          numsame--;
        } else {
          const uint32_t c = lc[i];
          if ((int)i == omit_pos || c == 0) continue;
          if (c == 1) {
            d[i] = 1;
          } else {
            const int bc = std::min(std::max(shift - ((12 - (int)c + 1) >> 1), 0), (int)c - 1);
            const uint32_t extra = r.bits((uint32_t)bc);
            d[i] = (1u << (c - 1)) + (extra << (c - 1 - bc));
          }
        }
        total += d[i];
      }
      if (total > 4096) fail("C.2.5: probabilities over 4096");
      d[omit_pos] = 4096 - total;
    }
    build(d, alpha, log_alpha);
  }
  void build(const std::vector<uint32_t>& d, uint32_t alpha, uint32_t log_alpha) {
    const uint32_t table = 1u << log_alpha;
    uint32_t total = 0;
    for (uint32_t p : d) total += p;
    if (total != 4096) fail("C.2.5: probabilities do not sum to 4096");
    prob = d;
    log_bucket = 12 - log_alpha;
    bucket = 1u << log_bucket;
    sym.assign(table, 0);
    std::vector<int64_t> offs(table, 0), cuts(table, 0);
    int nz = 0, last = 0;
    for (uint32_t i = 0; i < table; i++)
      if (d[i]) nz++, last = (int)i;
    if (nz == 1) {
      for (uint32_t i = 0; i < table; i++) {
        sym[i] = (uint32_t)last;
        offs[i] = (int64_t)bucket * i;
      }
    } else {
      std::vector<uint32_t> over, under;
      for (uint32_t i = 0; i < alpha; i++) {
        cuts[i] = d[i];
        sym[i] = i;
        if (d[i] > bucket) over.push_back(i);
        else if (d[i] < bucket) under.push_back(i);
      }
      for (uint32_t i = alpha; i < table; i++) {
        cuts[i] = 0;
        under.push_back(i);
      }
      while (!over.empty()) {
        const uint32_t o = over.back();
        over.pop_back();
        if (under.empty()) fail("C.2.6: alias table cannot be balanced");
        const uint32_t u = under.back();
        under.pop_back();
        const int64_t by = (int64_t)bucket - cuts[u];
        cuts[o] -= by;
        sym[u] = o;
        offs[u] = cuts[o];
        if (cuts[o] < bucket) under.push_back(o);
        else if (cuts[o] > bucket) over.push_back(o);
      }
      if (!under.empty()) fail("C.2.6: underfull buckets remain");
      for (uint32_t i = 0; i < table; i++) {
        if (cuts[i] == bucket) {
          sym[i] = i;
          offs[i] = 0;
          cuts[i] = 0;
        } else {
          offs[i] -= cuts[i];
        }
      }
    }
    off.resize(table);
    cut.resize(table);
    for (uint32_t i = 0; i < table; i++) {
      if (offs[i] < 0 || cuts[i] < 0) fail("C.2.6: alias entry out of range");
      off[i] = (uint32_t)offs[i];
      cut[i] = (uint32_t)cuts[i];
    }
  }
  uint32_t decode(Reader& r, uint32_t* state) const {
    const uint32_t x = *state & 0xFFF;
    const uint32_t i = x >> log_bucket, pos = x & (bucket - 1);
    uint32_t s, o;
    if (pos >= cut[i]) {
      s = sym[i];
      o = off[i] + pos;
    } else {
      s = i;
      o = pos;
    }
    const uint64_t next = (uint64_t)prob[s] * (*state >> 12) + o;
    if (next >> 32) fail("C.3.2: ANS state overflow");
    uint32_t n = (uint32_t)next;
    if (n < (1u << 16)) n = (n << 16) | r.bits(16);
    *state = n;
    return s;
  }
};

// C.3.3 LZ77 special distances (jpxl-entropy/src/lz77.rs: the 120-entry table).
const int8_t kSpecialDist[120][2] = {
    {0, 1},  {1, 0},  {1, 1},  {-1, 1}, {0, 2},  {2, 0},  {1, 2},  {-1, 2}, {2, 1},  {-2, 1}, {2, 2},  {-2, 2},
    {0, 3},  {3, 0},  {1, 3},  {-1, 3}, {3, 1},  {-3, 1}, {2, 3},  {-2, 3}, {3, 2},  {-3, 2}, {0, 4},  {4, 0},
    {1, 4},  {-1, 4}, {4, 1},  {-4, 1}, {3, 3},  {-3, 3}, {2, 4},  {-2, 4}, {4, 2},  {-4, 2}, {0, 5},  {3, 4},
    {-3, 4}, {4, 3},  {-4, 3}, {5, 0},  {1, 5},  {-1, 5}, {5, 1},  {-5, 1}, {2, 5},  {-2, 5}, {5, 2},  {-5, 2},
    {4, 4},  {-4, 4}, {3, 5},  {-3, 5}, {5, 3},  {-5, 3}, {0, 6},  {6, 0},  {1, 6},  {-1, 6}, {6, 1},  {-6, 1},
    {2, 6},  {-2, 6}, {6, 2},  {-6, 2}, {4, 5},  {-4, 5}, {5, 4},  {-5, 4}, {3, 6},  {-3, 6}, {6, 3},  {-6, 3},
    {0, 7},  {7, 0},  {1, 7},  {-1, 7}, {5, 5},  {-5, 5}, {7, 1},  {-7, 1}, {4, 6},  {-4, 6}, {6, 4},  {-6, 4},
    {2, 7},  {-2, 7}, {7, 2},  {-7, 2}, {3, 7},  {-3, 7}, {7, 3},  {-7, 3}, {5, 6},  {-5, 6}, {6, 5},  {-6, 5},
    {8, 0},  {4, 7},  {-4, 7}, {7, 4},  {-7, 4}, {8, 1},  {8, 2},  {6, 6},  {-6, 6}, {8, 3},  {5, 7},  {-5, 7},
    {7, 5},  {-7, 5}, {8, 4},  {6, 7},  {-6, 7}, {7, 6},  {-7, 6}, {8, 5},  {7, 7},  {-7, 7}, {8, 6},  {8, 7}};

// An entropy stream (C.2.1, decoder.rs): LZ77 parameters, cluster map,
// hybrid configurations, and prefix codes or ANS distributions.
struct SymbolDecoder {
  std::vector<uint8_t> map;
  uint32_t num_clusters = 1;
  std::vector<Hybrid> cfg;
  std::vector<Prefix> codes;
  std::vector<AnsDist> dists;
  bool ans = false;
  uint32_t state = 0;
  bool lz = false;
  uint32_t lz_min_symbol = 0, lz_min_length = 0, lz_dist_ctx = 0, dist_mult = 0;
  Hybrid lz_len;
  std::vector<uint32_t> window;
  uint64_t num_decoded = 0, copy_pos = 0;
  uint32_t to_copy = 0;

  void open(Reader& r, uint32_t num_dist, int depth = 0, bool forbid_lz77 = false, bool seed = true) {
    if (depth > 4) fail("C.2.1: nested too deep");
    lz = r.b();
    if (lz) {
      static const Dist kMinSym[4] = {V(224), V(512), V(4096), B(15, 8)};
      static const Dist kMinLen[4] = {V(3), V(4), B(2, 5), B(8, 9)};
      lz_min_symbol = u32(r, kMinSym);
      lz_min_length = u32(r, kMinLen);
      if (forbid_lz77) fail("C.2.2: a two-context nested stream enables LZ77");
      lz_dist_ctx = num_dist;
      num_dist += 1;
      lz_len.read(r, 8);
      window.assign(1u << 20, 0);
    }
    map = read_cluster_map(r, num_dist, depth);
    num_clusters = 0;
    for (uint8_t c : map) num_clusters = std::max<uint32_t>(num_clusters, c + 1u);
    const bool prefix = r.b();
    ans = !prefix;
    const uint32_t log_alpha = prefix ? 15 : 5 + r.bits(2);
    cfg.resize(num_clusters);
    for (auto& c : cfg) c.read(r, log_alpha);
    codes.clear();
    dists.clear();
    if (prefix) {
      std::vector<uint32_t> count(num_clusters);
      for (auto& c : count) {
        if (r.b()) {
          const uint32_t nb = r.bits(4);
          c = 1 + (1u << nb) + r.bits(nb);
          if (c > 32768) fail("C.2.1: alphabet too large");
        } else {
          c = 1;
        }
      }
      for (uint32_t k = 0; k < num_clusters; k++) codes.push_back(read_prefix(r, count[k]));
    } else {
      dists.resize(num_clusters);
      for (auto& d : dists) d.read(r, log_alpha);
      if (seed) state = r.bits(32);
    }
  }
  // C.3.2: a section's ANS state starts here (the HF pass groups).
  void restart(Reader& r) {
    if (ans) state = r.bits(32);
    if (lz) {
      std::fill(window.begin(), window.end(), 0u);
      num_decoded = copy_pos = 0;
      to_copy = 0;
    }
  }
  void finish() const {
    if (ans && state != 0x130000) fail("C.3.2: ANS stream did not end in its final state");
  }
  uint32_t token(Reader& r, uint32_t cluster) {
    return ans ? dists[cluster].decode(r, &state) : codes[cluster].decode(r);
  }
  void push(uint32_t v) {
    window[num_decoded & ((1u << 20) - 1)] = v;
    num_decoded++;
  }
  uint32_t copied() {
    const uint32_t v = window[copy_pos & ((1u << 20) - 1)];
    copy_pos++;
    to_copy--;
    push(v);
    return v;
  }
  uint32_t read(Reader& r, uint32_t ctx) {
    if (lz && to_copy > 0) return copied();
    if (ctx >= map.size()) fail("C.2.2: context outside the stream");
    const uint32_t c = map[ctx];
    const uint32_t t = token(r, c);
    if (lz && t >= lz_min_symbol) {
      const uint32_t length = lz_len.value(r, t - lz_min_symbol) + lz_min_length;
      const uint32_t dc = map[lz_dist_ctx];
      const uint32_t raw = cfg[dc].value(r, token(r, dc));
      uint64_t dist;
      if (dist_mult == 0) {
        dist = (uint64_t)raw + 1;
      } else if (raw < 120) {
        const int64_t dd = kSpecialDist[raw][0] + (int64_t)dist_mult * kSpecialDist[raw][1];
        dist = (uint64_t)std::max<int64_t>(dd, 1);
      } else {
        dist = raw - 119;
      }
      dist = std::min<uint64_t>(std::min<uint64_t>(dist, num_decoded), 1u << 20);
      to_copy = length;
      copy_pos = num_decoded - dist;
      if (to_copy == 0) fail("C.3.3: an LZ77 copy of length zero");
      return copied();
    }
    const uint32_t v = cfg[c].value(r, t);
    if (lz) push(v);
    return v;
  }
};

std::vector<uint8_t> read_cluster_map(Reader& r, uint32_t num_dist, int depth) {
  std::vector<uint8_t> m(num_dist, 0);
  if (num_dist == 1) return m;
  if (r.b()) {
    const uint32_t nb = r.bits(2);
    for (auto& v : m) v = (uint8_t)r.bits(nb);
  } else {
    const bool mtf = r.b();
    SymbolDecoder nested;
    nested.open(r, 1, depth + 1, num_dist == 2, true);
    for (auto& v : m) {
      const uint32_t x = nested.read(r, 0);
      if (x > 255) fail("C.2.2: cluster index over 255");
      v = (uint8_t)x;
    }
    nested.finish();
    if (mtf) {
      uint8_t t[256];
      for (int i = 0; i < 256; i++) t[i] = (uint8_t)i;
      for (auto& v : m) {
        const uint8_t idx = v, val = t[idx];
        v = val;
        for (int j = idx; j > 0; j--) t[j] = t[j - 1];
        t[0] = val;
      }
    }
  }
  uint32_t nc = 0;
  for (uint8_t c : m) nc = std::max<uint32_t>(nc, c + 1u);
  std::vector<bool> seen(nc, false);
  for (uint8_t c : m) seen[c] = true;
  for (bool s : seen)
    if (!s) fail("C.2.2: cluster map is not dense");
  return m;
}

// ---- modular (Annex H) -------------------------------------------------------
struct Channel {
  uint32_t w = 0, h = 0;
  std::vector<int32_t> v;
  int32_t at(uint32_t x, uint32_t y) const { return v[(size_t)y * w + x]; }
};

struct Node {
  bool leaf;
  uint32_t prop;
  int32_t value;
  uint32_t left, right;
  uint32_t ctx, pred;
  int32_t offset;
  int64_t mul;
};

std::vector<Node> read_tree(Reader& r, SymbolDecoder& sd, uint32_t* num_leaves) {
  std::vector<Node> t;
  uint32_t left = 1, ctx = 0;
  while (left > 0) {
    left--;
    if (t.size() >= (1u << 20)) fail("H.4.2: tree too large");
    Node n{};
    const uint32_t p = sd.read(r, 1);
    if (p > 0) {
      n.leaf = false;
      n.prop = p - 1;
      n.value = unpack_signed(sd.read(r, 0));
      n.left = (uint32_t)t.size() + left + 1;
      n.right = n.left + 1;
      left += 2;
    } else {
      n.leaf = true;
      n.ctx = ctx++;
      n.pred = sd.read(r, 2);
      if (n.pred > 13) fail("H.4.2: predictor out of range");
      n.offset = unpack_signed(sd.read(r, 3));
      const uint32_t ml = sd.read(r, 4);
      const uint32_t mb = sd.read(r, 5);
      if (ml > 30 || mb > (1u << (31 - ml)) - 2) fail("H.4.2: multiplier out of range");
      n.mul = (int64_t)(mb + 1) << ml;
    }
    t.push_back(n);
  }
  *num_leaves = ctx;
  return t;
}

struct Nb {
  int64_t w, n, nw, ne, nn, nee, ww;
};
Nb gather(const Channel& c, uint32_t x, uint32_t y) {
  Nb b;
  auto at = [&](uint32_t X, uint32_t Y) { return (int64_t)c.at(X, Y); };
  b.w = x > 0 ? at(x - 1, y) : (y > 0 ? at(x, y - 1) : 0);
  b.n = y > 0 ? at(x, y - 1) : b.w;
  b.nw = (x > 0 && y > 0) ? at(x - 1, y - 1) : b.w;
  b.ne = (x + 1 < c.w && y > 0) ? at(x + 1, y - 1) : b.n;
  b.nn = y > 1 ? at(x, y - 2) : b.n;
  b.nee = (x + 2 < c.w && y > 0) ? at(x + 2, y - 1) : b.ne;
  b.ww = x > 1 ? at(x - 2, y) : b.w;
  return b;
}
int64_t gradient(int64_t w, int64_t n, int64_t nw) {
  const int64_t g = w + n - nw;
  return std::min(std::max(g, std::min(w, n)), std::max(w, n));
}
// NarrowToI32 (18181-1 4.2): the low 32 bits.
int32_t narrow(int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; }

// H.5 the self-correcting (weighted) predictor, as the reference evaluates it
// (jpxl-core/src/modular_weighted.rs): the true error of the clamped
// prediction, the last column's error sum counting W twice, the max-error
// property by largest magnitude.
struct WpParams {
  int64_t p1 = 16, p2 = 10, p3[5] = {7, 7, 7, 0, 0}, w[4] = {13, 12, 12, 12};
};
struct WpEntry {
  int32_t true_err = 0;
  uint64_t err[4] = {0, 0, 0, 0};
};
struct WpPrediction {
  int64_t prediction;
  int32_t max_error;
  int64_t sub[4];
};
int64_t error2weight(uint32_t err_sum, int64_t maxweight) {
  const uint32_t shift = std::max<int>(0, 63 - __builtin_clzll((uint64_t)err_sum + 1) - 5);
  const int64_t den = (int64_t)(err_sum >> shift) + 1;
  return 4 + ((maxweight * ((1ll << 24) / den)) >> shift);
}
struct WpState {
  uint32_t width;
  std::vector<WpEntry> prev, curr;
  bool has_prev = false;
  explicit WpState(uint32_t w) : width(w), prev(w), curr(w) {}
  int64_t te_w(uint32_t x) const { return x > 0 ? curr[x - 1].true_err : 0; }
  int64_t te_n(uint32_t x) const { return has_prev ? prev[x].true_err : 0; }
  int64_t te_nw(uint32_t x) const { return x > 0 && has_prev ? prev[x - 1].true_err : te_n(x); }
  int64_t te_ne(uint32_t x) const { return x + 1 < width && has_prev ? prev[x + 1].true_err : te_n(x); }
  uint64_t e_w(uint32_t x, int i) const { return x > 0 ? curr[x - 1].err[i] : 0; }
  uint64_t e_ww(uint32_t x, int i) const { return x > 1 ? curr[x - 2].err[i] : 0; }
  uint64_t e_n(uint32_t x, int i) const { return has_prev ? prev[x].err[i] : 0; }
  uint64_t e_nw(uint32_t x, int i) const { return x > 0 && has_prev ? prev[x - 1].err[i] : e_n(x, i); }
  uint64_t e_ne(uint32_t x, int i) const { return x + 1 < width && has_prev ? prev[x + 1].err[i] : e_n(x, i); }
  WpPrediction predict(const WpParams& h, int64_t W, int64_t N, int64_t NW, int64_t NE, int64_t NN, uint32_t x) const {
    const int64_t n3 = N << 3, nw3 = NW << 3, ne3 = NE << 3, w3 = W << 3, nn3 = NN << 3;
    const int64_t tw = te_w(x), tn = te_n(x), tnw = te_nw(x), tne = te_ne(x);
    WpPrediction p;
    p.sub[0] = w3 + ne3 - n3;
    p.sub[1] = n3 - (((tw + tn + tne) * h.p1) >> 5);
    p.sub[2] = w3 - (((tw + tn + tnw) * h.p2) >> 5);
    p.sub[3] = n3 - ((tnw * h.p3[0] + tn * h.p3[1] + tne * h.p3[2] + (nn3 - n3) * h.p3[3] + (nw3 - w3) * h.p3[4]) >> 5);
    int64_t wt[4];
    for (int i = 0; i < 4; i++) {
      uint64_t sum = e_n(x, i) + e_w(x, i) + e_nw(x, i) + e_ww(x, i) + e_ne(x, i);
      if (x + 1 == width) sum += e_w(x, i);
      wt[i] = error2weight((uint32_t)sum, h.w[i]);
    }
    int64_t sw = wt[0] + wt[1] + wt[2] + wt[3];
    const int log_weight = 63 - __builtin_clzll((uint64_t)std::max<int64_t>(sw, 1)) + 1;
    const int shift = std::max(0, log_weight - 5);
    for (auto& v : wt) v >>= shift;
    sw = wt[0] + wt[1] + wt[2] + wt[3];
    const int64_t div = std::max<int64_t>(sw, 1);
    int64_t acc = (sw >> 1) - 1;
    for (int i = 0; i < 4; i++) acc += p.sub[i] * wt[i];
    const __int128 wide = ((__int128)acc * (__int128)((1ll << 24) / div)) >> 24;
    int64_t pred = (int64_t)wide;
    if (((tn ^ tw) | (tn ^ tnw)) <= 0)
      pred = std::min(std::max(pred, std::min(w3, std::min(n3, ne3))), std::max(w3, std::max(n3, ne3)));
    p.prediction = pred;
    int64_t me = tw;
    for (int64_t c : {tn, tnw, tne})
      if ((c < 0 ? -c : c) > (me < 0 ? -me : me)) me = c;
    p.max_error = narrow(me);
    return p;
  }
  void update(uint32_t x, const WpPrediction& p, int32_t value) {
    const int64_t shifted = (int64_t)value << 3;
    WpEntry e;
    e.true_err = narrow(p.prediction - shifted);
    for (int i = 0; i < 4; i++) {
      const int64_t m = p.sub[i] - shifted;
      e.err[i] = (uint64_t)(((m < 0 ? -m : m) + 3) >> 3);
    }
    curr[x] = e;
  }
  void advance_row() {
    std::swap(prev, curr);
    for (auto& e : curr) e = WpEntry{};
    has_prev = true;
  }
};

// The global MA tree of LfGlobal (G.1.3, modular/mod.rs:183-198) and its
// distributions, re-seeded by each sub-bitstream that uses it.
struct GlobalTree {
  bool present = false;
  std::vector<Node> tree;
  uint32_t leaves = 0;
  SymbolDecoder dist;
};

void read_modular(Reader& r, std::vector<Channel>& ch, uint32_t stream_id, const GlobalTree* gt = nullptr) {
  if (ch.empty()) return;
  const bool use_global = r.b();
  WpParams wp;
  if (!r.b()) {
    wp.p1 = r.bits(5);
    wp.p2 = r.bits(5);
    for (auto& v : wp.p3) v = r.bits(5);
    for (auto& v : wp.w) v = r.bits(4);
  }
  static const Dist kNbT[4] = {V(0), V(1), B(4, 2), B(8, 18)};
  if (u32(r, kNbT) != 0) fail("H.6: modular transforms are not read by this checker");
  std::vector<Node> local;
  const std::vector<Node>* tree;
  SymbolDecoder ds;
  if (use_global) {
    if (!gt || !gt->present) fail("H.2: use_global_tree without a global tree");
    tree = &gt->tree;
    ds = gt->dist;
    ds.restart(r);
  } else {
    SymbolDecoder ts;
    ts.open(r, 6);
    uint32_t leaves = 0;
    local = read_tree(r, ts, &leaves);
    ts.finish();
    ds.open(r, leaves);
    tree = &local;
  }
  uint32_t maxw = 0;
  for (auto& c : ch)
    if (c.w && c.h) maxw = std::max(maxw, c.w);
  ds.dist_mult = maxw;
  for (size_t i = 0; i < ch.size(); i++) {
    Channel& c = ch[i];
    if (!c.w || !c.h) continue;
    // previous channels of identical shape (H.4.1)
    std::vector<size_t> prevs;
    for (size_t j = i; j-- > 0;)
      if (ch[j].w == c.w && ch[j].h == c.h) prevs.push_back(j);
    std::vector<int32_t> props(16 + 4 * prevs.size());
    WpState ws(c.w);
    for (uint32_t y = 0; y < c.h; y++) {
      for (uint32_t x = 0; x < c.w; x++) {
        const Nb nb = gather(c, x, y);
        const WpPrediction wpp = ws.predict(wp, nb.w, nb.n, nb.nw, nb.ne, nb.nn, x);
        int64_t wge = nb.w;
        if (x > 0) {
          const Nb l = gather(c, x - 1, y);
          wge = nb.w - (l.w + l.n - l.nw);
        }
        const int64_t st[16] = {(int64_t)i, stream_id, y, x, nb.n < 0 ? -nb.n : nb.n, nb.w < 0 ? -nb.w : nb.w, nb.n,
                                nb.w, wge, nb.w + nb.n - nb.nw, nb.w - nb.nw, nb.nw - nb.n, nb.n - nb.ne,
                                nb.n - nb.nn, nb.w - nb.ww, wpp.max_error};
        for (int k = 0; k < 16; k++) props[k] = narrow(st[k]);
        size_t k = 16;
        for (size_t j : prevs) {
          const Channel& o = ch[j];
          const int64_t rc = o.at(x, y);
          const int64_t rw = x > 0 ? o.at(x - 1, y) : 0;
          const int64_t rn = y > 0 ? o.at(x, y - 1) : rw;
          const int64_t rnw = (x > 0 && y > 0) ? o.at(x - 1, y - 1) : rw;
          const int64_t rg = gradient(rw, rn, rnw);
          const int64_t d = rc - rg;
          for (int64_t v : {rc < 0 ? -rc : rc, rc, d < 0 ? -d : d, d}) props[k++] = narrow(v);
        }
        uint32_t at = 0;
        while (!(*tree)[at].leaf) {
          const Node& n = (*tree)[at];
          const int32_t p = n.prop < props.size() ? props[n.prop] : 0;
          at = p > n.value ? n.left : n.right;
          if (at >= tree->size()) fail("H.4.1: tree walk out of range");
        }
        const Node& lf = (*tree)[at];
        int64_t pred;
        switch (lf.pred) {
          case 0: pred = 0; break;
          case 1: pred = nb.w; break;
          case 2: pred = nb.n; break;
          case 3: pred = (nb.w + nb.n) / 2; break;
          case 4: {
            const int64_t a = nb.n - nb.nw, b2 = nb.w - nb.nw;
            pred = (a < 0 ? -a : a) < (b2 < 0 ? -b2 : b2) ? nb.w : nb.n;
            break;
          }
          case 5: pred = gradient(nb.w, nb.n, nb.nw); break;
          case 6: pred = (wpp.prediction + 3) >> 3; break;
          case 7: pred = nb.ne; break;
          case 8: pred = nb.nw; break;
          case 9: pred = nb.ww; break;
          case 10: pred = (nb.w + nb.nw) / 2; break;
          case 11: pred = (nb.n + nb.nw) / 2; break;
          case 12: pred = (nb.n + nb.ne) / 2; break;
          default: pred = (6 * nb.n - 2 * nb.nn + 7 * nb.w + nb.ww + nb.nee + 3 * nb.ne + 8) / 16; break;
        }
        const int64_t diff = (int64_t)unpack_signed(ds.read(r, lf.ctx)) * lf.mul + lf.offset;
        const int32_t v = narrow(diff + pred);
        c.v[(size_t)y * c.w + x] = v;
        ws.update(x, wpp, v);
      }
      ws.advance_row();
    }
  }
  ds.finish();
}

// ---- VarDCT tables -------------------------------------------------------------
const uint8_t kDefaultBlockCtx[39] = {0, 1, 2, 2, 3, 3, 4, 5, 6, 6, 6, 6, 6, 7, 8, 9, 9, 10, 11, 12, 13, 14,
                                      14, 14, 14, 14, 7, 8, 9, 9, 10, 11, 12, 13, 14, 14, 14, 14, 14};
const uint32_t kFreq[64] = {0,  0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14,
                            15, 15, 16, 16, 17, 17, 18, 18, 19, 19, 20, 20, 21, 21, 22, 22,
                            23, 23, 23, 23, 24, 24, 24, 24, 25, 25, 25, 25, 26, 26, 26, 26,
                            27, 27, 27, 27, 28, 28, 28, 28, 29, 29, 29, 29, 30, 30, 30, 30};
const uint32_t kNzc[64] = {0,   0,   31,  62,  62,  93,  93,  93,  93,  123, 123, 123, 123,
                           152, 152, 152, 152, 152, 152, 152, 152, 180, 180, 180, 180, 180,
                           180, 180, 180, 180, 180, 180, 180, 206, 206, 206, 206, 206, 206,
                           206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206,
                           206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206};

// I.3.2 natural order of a square DCT of side s (varblock.rs:665-745):
// the LLF cells first, then the anti-diagonals, alternating direction.
// I.3.2 natural order of a bwidth x bheight coefficient array
// (varblock.rs:665-745): the LLF cells first, then the scaled anti-diagonals,
// alternating direction.
std::vector<uint32_t> natural_order_wh(uint32_t bw, uint32_t bh) {
  const int64_t cx = bw / 8, cy = bh / 8, scale = std::max(cx, cy);
  std::vector<std::pair<int64_t, uint32_t>> llf;
  std::vector<std::tuple<int64_t, int64_t, uint32_t>> hf;
  for (uint32_t y = 0; y < bh; y++)
    for (uint32_t x = 0; x < bw; x++) {
      const uint32_t lin = y * bw + x;
      if ((int64_t)x < cx && (int64_t)y < cy) {
        llf.push_back({(int64_t)y * cx + x, lin});
        continue;
      }
      const int64_t sx = (int64_t)x * scale / cx, sy = (int64_t)y * scale / cy;
      const int64_t k1 = sx + sy;
      int64_t k2 = sx - sy;
      if (k1 % 2 == 1) k2 = -k2;
      hf.push_back({k1, k2, lin});
    }
  std::stable_sort(llf.begin(), llf.end(), [](auto& a, auto& b) { return a.first < b.first; });
  std::stable_sort(hf.begin(), hf.end(), [](auto& a, auto& b) {
    return std::get<0>(a) != std::get<0>(b) ? std::get<0>(a) < std::get<0>(b) : std::get<1>(a) < std::get<1>(b);
  });
  std::vector<uint32_t> o;
  for (auto& e : llf) o.push_back(e.second);
  for (auto& e : hf) o.push_back(std::get<2>(e));
  return o;
}
std::vector<uint32_t> natural_order(uint32_t s) { return natural_order_wh(s, s); }

// Table I.1 / I.7 per DctSelect: block rows, block columns, Order ID
// (jpxl-core/src/varblock.rs:188-251).
const uint8_t kTransformInfo[27][3] = {
    {1, 1, 0},  {1, 1, 1},  {1, 1, 1},   {1, 1, 1},   {2, 2, 2},   {4, 4, 3},   {2, 1, 4},
    {1, 2, 4},  {4, 1, 5},  {1, 4, 5},   {4, 2, 6},   {2, 4, 6},   {1, 1, 1},   {1, 1, 1},
    {1, 1, 1},  {1, 1, 1},  {1, 1, 1},   {1, 1, 1},   {8, 8, 7},   {8, 4, 8},   {4, 8, 8},
    {16, 16, 9}, {16, 8, 10}, {8, 16, 10}, {32, 32, 11}, {32, 16, 12}, {16, 32, 12}};

// A.4 ColourEncoding (headers/colour.rs:203-260): returns the colour space
// (0 RGB, 1 grey, 2 XYB, 3 unknown); refuses ICC profiles.
uint32_t read_enum(Reader& r);
void read_custom_xy(Reader& r) {
  static const Dist kXy[4] = {B(19), B(19, 524288), B(20, 1048576), B(21, 2097152)};
  u32(r, kXy);
  u32(r, kXy);
}
uint32_t read_colour_encoding(Reader& r) {
  if (r.b()) return 0;  // all_default: sRGB
  const bool want_icc = r.b();
  const uint32_t space = read_enum(r);
  if (want_icc) fail("A.4: ICC profiles are not read by this checker");
  const bool not_xyb = space != 2;
  if (not_xyb) {
    const uint32_t wp = read_enum(r);
    if (wp == 2) read_custom_xy(r);
  }
  if (not_xyb && space != 1) {
    const uint32_t pr = read_enum(r);
    if (pr == 2) {
      read_custom_xy(r);
      read_custom_xy(r);
      read_custom_xy(r);
    }
  }
  if (r.b()) r.bits(24);  // have_gamma, gamma
  else read_enum(r);      // transfer function
  read_enum(r);           // rendering intent
  return space;
}

// A.3 ToneMapping (headers/colour.rs:299-320): returns intensity_target.
float read_tone_mapping(Reader& r) {
  if (r.b()) return 255.0f;
  const float target = f16(r);
  f16(r);
  r.b();
  f16(r);
  return target;
}

// B.2.6 Enum(): U32(0, 1, 2 + u(4), 18 + u(6)).
uint32_t read_enum(Reader& r) {
  static const Dist kEnum[4] = {V(0), V(1), B(4, 2), B(6, 18)};
  const uint32_t v = u32(r, kEnum);
  if (v > 63) fail("B.2.6: enum value above 63");
  return v;
}

// A.3 Extensions: a U64 of set bits, then a U64 length per bit, then the bits.
void skip_extensions(Reader& r) {
  const uint64_t ext = u64(r);
  uint64_t total = 0;
  for (int i = 0; i < 64; i++)
    if (ext >> i & 1) total += u64(r);
  while (total > 0) {
    const uint32_t k = (uint32_t)std::min<uint64_t>(total, 32);
    r.bits(k);
    total -= k;
  }
}

}  // namespace

extern "C" {

typedef struct {
  uint32_t width, height, bits, xyb;
  uint32_t encoding, frame_type, x_qm_scale, b_qm_scale, gaborish, epf_iters, num_passes, is_last;
  uint64_t flags;
  uint32_t global_scale, quant_lf;
  int32_t x_factor_lf, b_factor_lf;
  uint32_t lf_corr_default, num_sections, num_groups, num_lf_groups;
  uint64_t header_bytes, toc_sum;
  uint32_t stage;  // 1 headers, 2 frame header, 3 TOC, 4 LfGlobal, 5 LF groups, 6 HfGlobal, 7 pass groups
  uint32_t colour_space;  // 0 RGB, 1 grey, 2 XYB
  uint32_t colour_factor;
  float base_corr_x, base_corr_b;
  uint32_t global_tree, used_orders, extra_precision;
  uint32_t dct_select_histogram[27];
  float m_lf[3];            // G.1.2 LfChannelDequantization (m_x_lf, m_y_lf, m_b_lf)
  float intensity_target;   // A.3 tone mapping
} cs_info;

// Every varblock of a frame with its full quantized coefficients: for
// varblock i, DctSelect, top-left block, HfMul and the arena offset of its X,
// Y, B coefficient arrays (each in the landscape layout of I.3.2: 8 min(N, M)
// rows by 8 max(N, M) columns, row-major, LLF cells 0), in the order of the
// LF groups and, within one, the raster order of the top-left blocks.
typedef struct {
  uint32_t dct_select, bx, by, hf_mul;
  uint64_t start;
} cs_varblock;

typedef struct {
  int32_t* coeffs;       // arena, capacity 3 x 64 x blocks
  cs_varblock* varblocks;
  uint32_t capacity;     // entries of varblocks
  uint32_t count;        // filled
  uint8_t* extra_precision;  // per LF group
} cs_full;

// Reads `data`; fills info as far as it gets.  The plan outputs (any may be
// NULL): quant 3 x (8 bh) x (8 bw) footprint layout, HF cells only (LLF
// cells 0); lf 3 x bh x bw; bmap bh x bw (8/16/32 at a varblock's top-left,
// 0 covered, 100 + DctSelect for other transforms); hf_mul bh x bw at the
// top-left; cfl_x/cfl_b ceil(bh/8) x ceil(bw/8); sharp bh x bw.
static int decode_impl(const uint8_t* data, size_t len, cs_info* info, int32_t* quant, int32_t* lf, uint8_t* bmap,
                       uint32_t* hf_mul, int8_t* cfl_x, int8_t* cfl_b, uint8_t* sharp, cs_full* full, char* err,
                       int errlen) {
  memset(info, 0, sizeof(*info));
  info->m_lf[0] = 1.0f / 32.0f;
  info->m_lf[1] = 1.0f / 4.0f;
  info->m_lf[2] = 1.0f / 2.0f;
  info->intensity_target = 255.0f;
  if (full) full->count = 0;
  try {
    Reader r{data, len};
    // Headers: signature, SizeHeader, ImageMetadata (headers/size.rs, metadata.rs)
    if (r.bits(16) != 0x0AFF) fail("A.1: not a naked codestream (FF 0A)");
    static const Dist kDim[4] = {B(9, 1), B(13, 1), B(18, 1), B(30, 1)};
    uint32_t h, w;
    const bool small = r.b();
    if (small) h = (r.bits(5) + 1) * 8;
    else h = u32(r, kDim);
    const uint32_t ratio = r.bits(3);
    static const uint32_t kRn[8] = {1, 1, 12, 4, 3, 16, 5, 2}, kRd[8] = {1, 1, 10, 3, 2, 9, 4, 1};
    if (ratio) w = (uint32_t)((uint64_t)h * kRn[ratio] / kRd[ratio]);
    else w = small ? (r.bits(5) + 1) * 8 : u32(r, kDim);
    info->width = w;
    info->height = h;
    info->bits = 8;
    info->xyb = 1;
    const bool all_default = r.b();
    if (!all_default) {
      const bool extra = r.b();
      if (extra) {
        r.bits(3);  // orientation - 1
        if (r.b()) fail("A.3: intrinsic size is not read by this checker");
        if (r.b()) fail("A.3: previews are not read by this checker");
        if (r.b()) fail("A.3: animation is not read by this checker");
      }
      if (r.b()) fail("A.3: float samples are not read by this checker");
      static const Dist kBps[4] = {V(8), V(10), V(12), B(6, 1)};
      info->bits = u32(r, kBps);
      r.b();  // modular_16_bit_buffer_sufficient
      static const Dist kNe[4] = {V(0), V(1), B(4, 2), B(12, 1)};
      if (u32(r, kNe) != 0) fail("A.3: extra channels are not read by this checker");
      info->xyb = r.b();
      info->colour_space = read_colour_encoding(r);
      if (extra) info->intensity_target = read_tone_mapping(r);
      skip_extensions(r);
    }
    if (!r.b()) fail("A.3: custom transform data (default_m false) is not read by this checker");
    r.pad();
    info->stage = 1;
    // Frame header (frame/header.rs)
    if (r.b()) fail("F.2: an all-default frame header is not what the checks expect");
    info->frame_type = r.bits(2);
    info->encoding = r.bits(1);
    info->flags = u64(r);
    if (info->frame_type != 0 || info->encoding != 0) fail("F.2: only regular VarDCT frames are read");
    if (info->flags & ~0x80ull) fail("F.2: frame flags other than skip-adaptive-LF-smoothing are not read");
    if (!info->xyb) fail("F.2: non-XYB VarDCT frames are not read");
    static const Dist kUp[4] = {V(1), V(2), V(4), V(8)};
    if (u32(r, kUp) != 1) fail("F.2: upsampling is not read");
    info->x_qm_scale = r.bits(3);
    info->b_qm_scale = r.bits(3);
    static const Dist kNp[4] = {V(1), V(2), V(3), B(3, 4)};
    info->num_passes = u32(r, kNp);
    if (info->num_passes != 1) fail("F.2: progressive passes are not read");
    if (r.b()) fail("F.2: cropped frames are not read");
    static const Dist kBm[4] = {V(0), V(1), V(2), B(2, 3)};
    if (u32(r, kBm) != 0) fail("F.2: blending is not read");
    info->is_last = r.b();
    if (!info->is_last) r.bits(2);  // save_as_reference
    static const Dist kNl[4] = {V(0), B(4), B(5, 16), B(10, 48)};
    const uint32_t name_len = u32(r, kNl);
    for (uint32_t i = 0; i < name_len; i++) r.bits(8);
    // RestorationFilter
    if (r.b()) {
      info->gaborish = 1;
      info->epf_iters = 1;
    } else {
      info->gaborish = r.b();
      if (info->gaborish && r.b()) fail("J.1: custom Gaborish weights are not read");
      info->epf_iters = r.bits(2);
      if (info->epf_iters) {
        if (r.b()) fail("J.1: custom EPF sharpness is not read");
        if (r.b()) fail("J.1: custom EPF weights are not read");
        if (r.b()) fail("J.1: custom EPF sigma is not read");
      }
      if (u64(r) != 0) fail("J.1: restoration extensions are not read");
    }
    if (u64(r) != 0) fail("F.2: frame extensions are not read");
    info->stage = 2;
    // TOC (frame/toc.rs)
    const uint32_t bw = (w + 7) / 8, bh = (h + 7) / 8;
    const uint32_t gx = (w + 255) / 256, gy = (h + 255) / 256, lgx = (w + 2047) / 2048, lgy = (h + 2047) / 2048;
    const uint32_t ng = gx * gy, nlf = lgx * lgy;
    info->num_groups = ng;
    info->num_lf_groups = nlf;
    const bool single = ng == 1;
    const uint32_t nsec = single ? 1 : 2 + nlf + ng;
    info->num_sections = nsec;
    if (r.b()) fail("F.3.2: permuted TOCs are not read");
    r.pad();
    static const Dist kToc[4] = {B(10), B(14, 1024), B(22, 17408), B(30, 4211712)};
    std::vector<uint64_t> sec(nsec);
    for (auto& s : sec) {
      s = u32(r, kToc);
      info->toc_sum += s;
    }
    r.pad();
    info->header_bytes = r.pos / 8;
    if (info->header_bytes + info->toc_sum != len)
      fail("F.3: the TOC's section lengths do not add up to the stream's length");
    info->stage = 3;
    std::vector<uint64_t> start(nsec);
    uint64_t at = info->header_bytes;
    for (uint32_t i = 0; i < nsec; i++) {
      start[i] = at;
      at += sec[i];
    }
    auto section = [&](uint32_t i) {
      Reader s{data + start[i], (size_t)sec[i]};
      return s;
    };
    Reader one = section(0);
    // LfGlobal (G.1)
    struct BlockModel {
      std::vector<int32_t> lf_thr[3];
      std::vector<uint32_t> qf_thr;
      std::vector<uint8_t> map;
      uint32_t nb_ctx = 15;
    } bm_model;
    GlobalTree gtree;
    {
      Reader& s = one;
      if (!s.b()) {  // G.1.2 LfChannelDequantization
        info->m_lf[0] = f16(s);
        info->m_lf[1] = f16(s);
        info->m_lf[2] = f16(s);
      }
      static const Dist kGs[4] = {B(11, 1), B(11, 2049), B(12, 4097), B(16, 8193)};
      static const Dist kQl[4] = {V(16), B(5, 1), B(8, 1), B(16, 1)};
      info->global_scale = u32(s, kGs);
      info->quant_lf = u32(s, kQl);
      // I.2.2 HF block context (block_ctx.rs:131-200)
      if (s.b()) {
        bm_model.map.assign(kDefaultBlockCtx, kDefaultBlockCtx + 39);
      } else {
        static const Dist kThr[4] = {B(4), B(8, 16), B(16, 272), B(32, 65808)};
        static const Dist kQf[4] = {B(2), B(3, 4), B(5, 12), B(8, 44)};
        uint32_t factors = 1;
        for (int c = 0; c < 3; c++) {
          const uint32_t n = s.bits(4);
          for (uint32_t k = 0; k < n; k++) bm_model.lf_thr[c].push_back(unpack_signed(u32(s, kThr)));
          factors *= n + 1;
        }
        const uint32_t nq = s.bits(4);
        for (uint32_t k = 0; k < nq; k++) bm_model.qf_thr.push_back(u32(s, kQf) + 1);
        factors *= nq + 1;
        const uint32_t bsize = 39 * factors;
        if (bsize > 39 * 64) fail("I.2.2: block context map too large");
        bm_model.map = read_cluster_map(s, bsize, 0);
        uint32_t nc = 0;
        for (uint8_t c : bm_model.map) nc = std::max<uint32_t>(nc, c + 1u);
        if (nc > 16) fail("I.2.2: more than 16 block contexts");
        bm_model.nb_ctx = nc;
      }
      // I.2.3 LF channel correlation
      info->lf_corr_default = s.b();
      if (!info->lf_corr_default) {
        static const Dist kCf[4] = {V(84), V(256), B(8, 2), B(16, 258)};
        info->colour_factor = u32(s, kCf);
        info->base_corr_x = f16(s);
        info->base_corr_b = f16(s);
        info->x_factor_lf = (int32_t)s.bits(8) - 128;
        info->b_factor_lf = (int32_t)s.bits(8) - 128;
      } else {
        info->colour_factor = 84;
        info->base_corr_x = 0.0f;
        info->base_corr_b = 1.0f;
      }
      // G.1.3 GlobalModular: the global tree, then no channels
      if (s.b()) {
        SymbolDecoder ts;
        ts.open(s, 6);
        gtree.tree = read_tree(s, ts, &gtree.leaves);
        ts.finish();
        gtree.dist.open(s, gtree.leaves, 0, false, false);
        gtree.present = true;
        info->global_tree = 1;
      }
      if (!single) s.pad();
    }
    info->stage = 4;
    const uint32_t tx = (bw + 7) / 8;
    std::vector<uint8_t> bm((size_t)bw * bh, 255);
    std::vector<uint8_t> sel((size_t)bw * bh, 255);
    std::vector<uint32_t> mul((size_t)bw * bh, 0);
    std::vector<int32_t> lfq((size_t)3 * bw * bh, 0);
    std::vector<uint64_t> vstart((size_t)bw * bh, 0);
    uint64_t arena_at = 0;
    // LF groups (G.2)
    for (uint32_t id = 0; id < nlf; id++) {
      Reader sep = single ? Reader{nullptr, 0} : section(1 + id);
      Reader& s = single ? one : sep;
      const uint32_t lx = id % lgx, ly = id / lgx, b0x = lx * 256, b0y = ly * 256;
      const uint32_t gw = std::min(256u, bw - b0x), gh = std::min(256u, bh - b0y);
      info->extra_precision = s.bits(2);
      if (full && full->extra_precision) full->extra_precision[id] = (uint8_t)info->extra_precision;
      std::vector<Channel> c(3);
      for (auto& x : c) {
        x.w = gw;
        x.h = gh;
        x.v.assign((size_t)gw * gh, 0);
      }
      read_modular(s, c, 1 + id, &gtree);
      const int plane_of[3] = {1, 0, 2};
      for (int i = 0; i < 3; i++)
        for (uint32_t y = 0; y < gh; y++)
          for (uint32_t x = 0; x < gw; x++)
            lfq[(size_t)plane_of[i] * bw * bh + (size_t)(b0y + y) * bw + b0x + x] = c[i].at(x, y);
      const uint32_t nb = s.bits(ceil_log2(std::max<uint64_t>((uint64_t)gw * gh, 1))) + 1;
      const uint32_t tw = (gw + 7) / 8, th = (gh + 7) / 8;
      std::vector<Channel> m(4);
      m[0].w = m[1].w = tw;
      m[0].h = m[1].h = th;
      m[2].w = nb;
      m[2].h = 2;
      m[3].w = gw;
      m[3].h = gh;
      for (auto& x : m) x.v.assign((size_t)x.w * x.h, 0);
      read_modular(s, m, 1 + 2 * nlf + id, &gtree);
      for (uint32_t y = 0; y < th; y++)
        for (uint32_t x = 0; x < tw; x++) {
          const int32_t vx = m[0].at(x, y), vb = m[1].at(x, y);
          if (vx < -128 || vx > 127 || vb < -128 || vb > 127) fail("G.2.4: CfL factor outside int8");
          const size_t t = (size_t)(ly * 32 + y) * tx + lx * 32 + x;
          if (cfl_x) cfl_x[t] = (int8_t)vx;
          if (cfl_b) cfl_b[t] = (int8_t)vb;
        }
      // each varblock goes to the first uncovered block in raster order
      uint32_t cur = 0;
      for (uint32_t i = 0; i < nb; i++) {
        const int32_t ds = m[2].at(i, 0), q = m[2].at(i, 1);
        if (ds < 0 || ds > 26) fail("G.2.4: DctSelect out of range");
        if (q < 0) fail("G.2.4: negative HfMul");
        const uint32_t rows = kTransformInfo[ds][0], cols = kTransformInfo[ds][1];
        while (cur < (uint32_t)gw * gh && bm[(size_t)(b0y + cur / gw) * bw + b0x + cur % gw] != 255) cur++;
        if (cur >= (uint32_t)gw * gh) fail("G.2.4: more varblocks than blocks");
        const uint32_t X = cur % gw, Y = cur / gw;
        if (X + cols > gw || Y + rows > gh) fail("G.2.4: varblock crosses the LF group");
        if ((b0x + X) / 32 != (b0x + X + cols - 1) / 32 || (b0y + Y) / 32 != (b0y + Y + rows - 1) / 32)
          fail("G.2.4: varblock crosses a group");
        for (uint32_t j = 0; j < rows; j++)
          for (uint32_t k = 0; k < cols; k++) {
            uint8_t& cell = bm[(size_t)(b0y + Y + j) * bw + b0x + X + k];
            if (cell != 255) fail("G.2.4: varblocks overlap");
            cell = 0;
          }
        const size_t at = (size_t)(b0y + Y) * bw + b0x + X;
        bm[at] = (ds == 0 || ds == 4 || ds == 5) ? (uint8_t)(8 * rows) : (uint8_t)(100 + ds);
        sel[at] = (uint8_t)ds;
        mul[at] = (uint32_t)q + 1;
        info->dct_select_histogram[ds]++;
        if (full) {
          if (full->count >= full->capacity) fail("cs_full: varblock table too small");
          cs_varblock& v = full->varblocks[full->count++];
          v.dct_select = (uint32_t)ds;
          v.bx = b0x + X;
          v.by = b0y + Y;
          v.hf_mul = (uint32_t)q + 1;
          v.start = arena_at;
          vstart[at] = arena_at;
          arena_at += (uint64_t)3 * 64 * rows * cols;
        }
      }
      for (uint32_t y = 0; y < gh; y++)
        for (uint32_t x = 0; x < gw; x++) {
          if (bm[(size_t)(b0y + y) * bw + b0x + x] == 255) fail("G.2.4: blocks left uncovered");
          const int32_t sv = m[3].at(x, y);
          if (sv < 0 || sv > 7) fail("G.2.4: sharpness out of range");
          if (sharp) sharp[(size_t)(b0y + y) * bw + b0x + x] = (uint8_t)sv;
        }
      if (!single) s.pad();
    }
    if (bmap) memcpy(bmap, bm.data(), bm.size());
    if (hf_mul) memcpy(hf_mul, mul.data(), mul.size() * 4);
    if (lf) memcpy(lf, lfq.data(), lfq.size() * 4);
    info->stage = 5;
    // HfGlobal (G.3): matrices, presets, coefficient orders, histograms
    const uint32_t nctx = 495 * bm_model.nb_ctx;
    SymbolDecoder hf;
    uint32_t presets;
    std::vector<std::vector<uint32_t>> natural(13), orders(13 * 3);
    for (int b = 0; b < 13; b++) {
      int t = 0;
      while (kTransformInfo[t][2] != b) t++;
      const uint32_t r8 = kTransformInfo[t][0], c8 = kTransformInfo[t][1];
      natural[b] = natural_order_wh(8 * std::max(r8, c8), 8 * std::min(r8, c8));
      for (int c = 0; c < 3; c++) orders[b * 3 + c] = natural[b];
    }
    {
      Reader sep = single ? Reader{nullptr, 0} : section(1 + nlf);
      Reader& s = single ? one : sep;
      if (!s.b()) fail("I.2.4: custom dequantization matrices are not read by this checker");
      presets = s.bits(ceil_log2(ng)) + 1;
      static const Dist kUo[4] = {V(0x5F), V(0x13), V(0), B(13)};
      const uint32_t used = u32(s, kUo);
      info->used_orders = used;
      if (used) {
        auto get_context = [](uint32_t x) { return std::min<uint32_t>(7, x ? 32 - __builtin_clz(x) : 0); };
        SymbolDecoder od;
        od.open(s, 8);
        for (int b = 0; b < 13; b++) {
          if (!(used >> b & 1)) continue;
          const uint32_t size = (uint32_t)natural[b].size(), skip = size / 64;
          for (int c = 0; c < 3; c++) {
            const uint32_t end = od.read(s, get_context(size));
            if (end > size - skip) fail("F.3.2: permutation end out of range");
            std::vector<uint32_t> lehmer(size, 0);
            uint32_t prev = 0;
            for (uint32_t i = skip; i < skip + end; i++) {
              const uint32_t v = od.read(s, get_context(i > skip ? prev : 0));
              if (v >= size - i) fail("F.3.2: Lehmer code out of range");
              lehmer[i] = v;
              prev = v;
            }
            std::vector<uint32_t> temp(size);
            for (uint32_t i = 0; i < size; i++) temp[i] = i;
            std::vector<uint32_t>& ord = orders[b * 3 + c];
            for (uint32_t i = 0; i < size; i++) {
              const uint32_t p = temp[lehmer[i]];
              temp.erase(temp.begin() + lehmer[i]);
              ord[i] = natural[b][p];
            }
          }
        }
        od.finish();
      }
      hf.open(s, nctx * presets, 0, false, false);
      if (!single) s.pad();
    }
    info->stage = 6;
    // Pass groups (G.4, I.4)
    const size_t cs = (size_t)8 * bw, cpl = cs * 8 * bh;
    if (quant) memset(quant, 0, 3 * cpl * 4);
    if (full && full->coeffs) memset(full->coeffs, 0, arena_at * 4);
    for (uint32_t gid = 0; gid < ng; gid++) {
      Reader sep = single ? Reader{nullptr, 0} : section(2 + nlf + gid);
      Reader& s = single ? one : sep;
      const uint32_t hfp = s.bits(ceil_log2(presets));
      if (hfp >= presets) fail("I.4: hfp out of range");
      const uint32_t off = nctx * hfp;
      hf.restart(s);
      const uint32_t g0x = (gid % gx) * 32, g0y = (gid / gx) * 32;
      const uint32_t gw = std::min(32u, bw - g0x), gh = std::min(32u, bh - g0y);
      std::vector<uint32_t> grid[3];
      for (auto& g : grid) g.assign((size_t)gw * gh, 0);
      for (uint32_t y = 0; y < gh; y++)
        for (uint32_t x = 0; x < gw; x++) {
          const size_t at = (size_t)(g0y + y) * bw + g0x + x;
          const uint8_t ds = sel[at];
          if (ds == 255) continue;
          const uint32_t rows = kTransformInfo[ds][0], cols = kTransformInfo[ds][1];
          const uint32_t order_id = kTransformInfo[ds][2];
          const uint32_t nblocks = rows * cols, size = 64 * nblocks;
          const uint32_t bwd = 8 * std::max(rows, cols);  // coefficient array width
          for (int c : {1, 0, 2}) {
            // I.4 block context (hf_coeff.rs:480-525)
            uint32_t idx = (uint32_t)(c < 2 ? c ^ 1 : 2) * 13 + order_id;
            idx *= (uint32_t)bm_model.qf_thr.size() + 1;
            for (uint32_t t : bm_model.qf_thr)
              if (mul[at] > t) idx++;
            for (int i = 0; i < 3; i++) idx *= (uint32_t)bm_model.lf_thr[i].size() + 1;
            uint32_t lf_idx = 0;
            const int32_t qdc[3] = {lfq[at], lfq[(size_t)bw * bh + at], lfq[(size_t)2 * bw * bh + at]};
            for (int32_t t : bm_model.lf_thr[0])
              if (qdc[0] > t) lf_idx++;
            lf_idx *= (uint32_t)bm_model.lf_thr[2].size() + 1;
            for (int32_t t : bm_model.lf_thr[2])
              if (qdc[2] > t) lf_idx++;
            lf_idx *= (uint32_t)bm_model.lf_thr[1].size() + 1;
            for (int32_t t : bm_model.lf_thr[1])
              if (qdc[1] > t) lf_idx++;
            if (idx + lf_idx >= bm_model.map.size()) fail("I.4: block context index out of range");
            const uint32_t bctx = bm_model.map[idx + lf_idx];
            std::vector<uint32_t>& g = grid[c];
            uint32_t pred;
            if (x == 0 && y == 0) pred = 32;
            else if (x == 0) pred = g[(y - 1) * gw + x];
            else if (y == 0) pred = g[y * gw + x - 1];
            else pred = (g[(y - 1) * gw + x] + g[y * gw + x - 1] + 1) >> 1;
            pred = std::min(pred, 64u);
            const uint32_t nb_ctx = bm_model.nb_ctx;
            const uint32_t nz = hf.read(s, bctx + nb_ctx * (pred < 8 ? pred : 4 + pred / 2) + off);
            if (nz > size - nblocks) fail("I.4: more non-zeros than HF positions");
            const uint32_t per = (nz + nblocks - 1) / nblocks;
            for (uint32_t j = 0; j < rows; j++)
              for (uint32_t i = 0; i < cols; i++) g[(y + j) * gw + x + i] = per;
            if (!nz) continue;
            const std::vector<uint32_t>& order = orders[order_id * 3 + c];
            uint32_t rem = nz, prev = nz <= size / 16 ? 1 : 0;
            for (uint32_t k = nblocks; k < size; k++) {
              const uint32_t ctx = (kNzc[(rem + nblocks - 1) / nblocks] + kFreq[k / nblocks]) * 2 + prev +
                                   bctx * 458 + 37 * nb_ctx + off;
              const uint32_t u = hf.read(s, ctx);
              const int32_t v = unpack_signed(u);
              if (quant && (ds == 0 || ds == 4 || ds == 5)) {
                const uint32_t cell = order[k];
                quant[(size_t)c * cpl + (size_t)(8 * (g0y + y) + cell / bwd) * cs + 8 * (g0x + x) + cell % bwd] = v;
              }
              if (full && full->coeffs) full->coeffs[vstart[at] + (uint64_t)c * size + order[k]] = v;
              prev = u != 0;
              if (u) {
                if (--rem == 0) break;
              }
            }
            if (rem) fail("I.4: stream ended the varblock with non-zeros undelivered");
          }
        }
      hf.finish();
      if (!single) {
        // G.4.2 carries no modular group data here; the rest is padding
        s.pad();
        if (s.pos != (uint64_t)s.n * 8) fail("G.4: bytes left over in a pass group");
      }
    }
    if (single) {
      one.pad();
      if (one.pos != (uint64_t)one.n * 8) fail("F.3: bytes left over in the single section");
    }
    info->stage = 7;
    return 0;
  } catch (const Fail& e) {
    if (err && errlen > 0) snprintf(err, errlen, "%s", e.what());
    return -1;
  }
}

int cs_decode(const uint8_t* data, size_t len, cs_info* info, int32_t* quant, int32_t* lf, uint8_t* bmap,
              uint32_t* hf_mul, int8_t* cfl_x, int8_t* cfl_b, uint8_t* sharp, char* err, int errlen) {
  return decode_impl(data, len, info, quant, lf, bmap, hf_mul, cfl_x, cfl_b, sharp, nullptr, err, errlen);
}

// cs_decode plus every varblock's full coefficient arrays (all DctSelect
// types), for reconstructing a frame's samples.
int cs_decode_full(const uint8_t* data, size_t len, cs_info* info, int32_t* lf, uint8_t* bmap, int8_t* cfl_x,
                   int8_t* cfl_b, uint8_t* sharp, cs_full* full, char* err, int errlen) {
  return decode_impl(data, len, info, nullptr, lf, bmap, nullptr, cfl_x, cfl_b, sharp, full, err, errlen);
}

}  // extern "C"

// Entry points for the known-answer checks of the reading rules.
extern "C" {

// Decodes `count` symbols with the canonical prefix code of `lengths`.
int cs_prefix_decode(const uint8_t* lengths, int n, const uint8_t* data, size_t len, uint32_t* syms, int count,
                     uint64_t* bits_read) {
  try {
    Prefix p;
    p.from_lengths(std::vector<uint8_t>(lengths, lengths + n));
    Reader r{data, len};
    for (int i = 0; i < count; i++) syms[i] = p.decode(r);
    *bits_read = r.pos;
    return 0;
  } catch (const Fail&) {
    return -1;
  }
}

// C.2.2 cluster map of num_dist contexts.
int cs_cluster_map(const uint8_t* data, size_t len, uint32_t num_dist, uint8_t* out, uint64_t* bits_read) {
  try {
    Reader r{data, len};
    std::vector<uint8_t> m = read_cluster_map(r, num_dist, 0);
    memcpy(out, m.data(), m.size());
    *bits_read = r.pos;
    return 0;
  } catch (const Fail&) {
    return -1;
  }
}

// C.3.3 hybrid integer of `token` under (se, msb, lsb), extra bits from data.
int cs_hybrid(uint32_t se, uint32_t msb, uint32_t lsb, uint32_t token, const uint8_t* data, size_t len,
              uint32_t* value, uint64_t* bits_read) {
  try {
    Hybrid h;
    h.se = se;
    h.msb = msb;
    h.lsb = lsb;
    Reader r{data, len};
    *value = h.value(r, token);
    *bits_read = r.pos;
    return 0;
  } catch (const Fail&) {
    return -1;
  }
}

// I.3.2 natural order of a square DCT of side s.
void cs_natural_order(uint32_t side, uint32_t* out) {
  const std::vector<uint32_t> o = natural_order(side);
  memcpy(out, o.data(), o.size() * 4);
}

}  // extern "C"
