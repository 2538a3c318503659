// The CPU entropy and bitstream stage: the quantized plan (the QuantizedFrameIr
// hand-off of jpxlg_plan_download_ir) in, a JPEG XL codestream out.
//
// Replaces attach_entropy (jpxl-encode-policy/src/lib.rs:1012) followed by
// emit_codestream_with_executor (jpxl-encode/src/vardct/write.rs:350-455):
// the image headers and frame header as the reference writes them
// (jpxl-encode/src/vardct/headers.rs:104-233), the TOC (F.3), LfGlobal (G.1,
// write.rs:643-682), LfGroups (G.2, write.rs:688-785: the LF planes in Y, X, B
// order, the HF metadata channels), HfGlobal (G.3, write.rs:792-823: default
// matrices, one HF preset, natural coefficient orders) and the pass groups
// (G.4 / I.4: the HF coefficients with the decoder's context model,
// jpxl-decode/src/vardct/hf_coeff.rs:480-860).
//
// The entropy coding is this library's own (the reference's clustering,
// tree learning and ANS tables are not restated): every stream uses prefix
// codes (C.2.4, RFC 7932 3.4-3.5) with length-limited Huffman lengths,
// contexts are clustered greedily by coding cost, control images (LF planes,
// CfL tiles, BlockInfo, sharpness) are modular streams (Annex H) under a
// channel-chain MA tree with the cheapest of four fixed predictors per
// channel.  Pass groups are tokenised and written in parallel; the sections
// are independent given the HF code tables, as in write.rs:476-511.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <unistd.h>
#include <vector>

#include "../../include/jpxlg.h"

namespace jpxlg {
void natural_order_table(int side, uint16_t* order);  // tables.cpp (I.3.2)
}

namespace {

// ---------------------------------------------------------------------------
// Bits (B.2): LSB first within bytes.
// ---------------------------------------------------------------------------
class BitWriter {
 public:
  void put(uint32_t n, uint64_t v) {  // n <= 32, v < 2^n
    acc_ |= v << nacc_;
    nacc_ += n;
    if (nacc_ >= 32) {
      if (len_ + 4 > cap_) grow(4);
      const uint32_t w = (uint32_t)acc_;
      memcpy(buf_.get() + len_, &w, 4);
      len_ += 4;
      acc_ >>= 32;
      nacc_ -= 32;
    }
  }
  void put_bool(bool b) { put(1, b ? 1u : 0u); }
  uint64_t bits() const { return (uint64_t)len_ * 8 + nacc_; }
  void pad_to_byte() {
    nacc_ = (nacc_ + 7) & ~7u;
    if (len_ + 8 > cap_) grow(8);
    while (nacc_ >= 8) {
      buf_[len_++] = (uint8_t)acc_;
      acc_ >>= 8;
      nacc_ -= 8;
    }
  }
  // Appends another writer's bits at this writer's bit position.
  void append(const BitWriter& o) {
    size_t i = 0;
    for (; i + 4 <= o.len_; i += 4) {
      uint32_t v;
      memcpy(&v, o.buf_.get() + i, 4);
      put(32, v);
    }
    for (; i < o.len_; i++) put(8, o.buf_[i]);
    if (o.nacc_) put(o.nacc_, o.acc_ & ((1ull << o.nacc_) - 1));
  }
  // The bytes written (whole after pad_to_byte).
  const uint8_t* data() const { return buf_.get(); }
  size_t size() const { return len_; }
  void reserve(size_t n) {
    if (n > cap_) grow(n - len_);
  }

 private:
  void grow(size_t need) {
    size_t cap = std::max<size_t>(cap_ * 2, 256);
    while (cap < len_ + need) cap *= 2;
    std::unique_ptr<uint8_t[]> b(new uint8_t[cap]);
    if (len_) memcpy(b.get(), buf_.get(), len_);
    buf_ = std::move(b);
    cap_ = cap;
  }
  std::unique_ptr<uint8_t[]> buf_;
  size_t len_ = 0, cap_ = 0;
  uint64_t acc_ = 0;
  uint32_t nacc_ = 0;
};

// ---------------------------------------------------------------------------
// A small persistent worker pool (the executor of the sections).
// ---------------------------------------------------------------------------
class Pool {
 public:
  explicit Pool(int n) {
    for (int i = 1; i < n; i++) threads_.emplace_back([this, i] { loop(i); });
    size_ = n;
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }
  int size() const { return size_; }
  // fn(index, worker) for index in [0, n); returns when all are done.
  void run(int n, const std::function<void(int, int)>& fn) {
    if (n <= 0) return;
    if (size_ == 1 || n == 1) {
      for (int i = 0; i < n; i++) fn(i, 0);
      return;
    }
    std::unique_lock<std::mutex> lk(m_);
    fn_ = &fn;
    n_ = n;
    next_.store(0);
    pending_ = size_ - 1;
    gen_++;
    lk.unlock();
    cv_.notify_all();
    work(0);
    lk.lock();
    done_cv_.wait(lk, [this] { return pending_ == 0; });
    fn_ = nullptr;
  }

 private:
  void work(int id) {
    for (;;) {
      const int i = next_.fetch_add(1);
      if (i >= n_) break;
      (*fn_)(i, id);
    }
  }
  void loop(int id) {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(m_);
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      lk.unlock();
      work(id);
      lk.lock();
      if (--pending_ == 0) done_cv_.notify_all();
    }
  }
  std::vector<std::thread> threads_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int, int)>* fn_ = nullptr;
  std::atomic<int> next_{0};
  int n_ = 0, pending_ = 0, size_ = 1;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// U32(d0, d1, d2, d3) (B.2.2): the first distribution that holds the value.
struct Dist {
  bool is_val;
  uint32_t bits, off;
};
constexpr Dist Val(uint32_t v) { return {true, 0, v}; }
constexpr Dist Bits(uint32_t b, uint32_t off = 0) { return {false, b, off}; }

bool put_u32(BitWriter& w, const Dist (&d)[4], uint32_t value) {
  for (uint32_t s = 0; s < 4; s++) {
    if (d[s].is_val) {
      if (d[s].off != value) continue;
      w.put(2, s);
      return true;
    }
    if (value < d[s].off) continue;
    const uint64_t p = (uint64_t)value - d[s].off;
    if (d[s].bits < 32 && (p >> d[s].bits) != 0) continue;
    w.put(2, s);
    w.put(d[s].bits, p);
    return true;
  }
  return false;
}

// U64 (B.2.3), the shortest form (bitstream/src/writer.rs:249-277).
void put_u64(BitWriter& w, uint64_t v) {
  if (v == 0) { w.put(2, 0); return; }
  if (v <= 16) { w.put(2, 1); w.put(4, v - 1); return; }
  if (v <= 272) { w.put(2, 2); w.put(8, v - 17); return; }
  w.put(2, 3);
  w.put(12, v & 0xFFF);
  uint64_t rest = v >> 12;
  uint32_t shift = 12;
  while (rest) {
    w.put(1, 1);
    if (shift == 60) { w.put(4, rest & 0xF); return; }
    w.put(8, rest & 0xFF);
    rest >>= 8;
    shift += 8;
  }
  w.put(1, 0);
}

uint32_t ceil_log2(uint64_t n) {
  uint32_t b = 0;
  while ((1ull << b) < n) b++;
  return b;
}
uint32_t bit_width(uint32_t n) { return n ? 32 - __builtin_clz(n) : 0; }
inline uint32_t pack_signed(int32_t v) { return v >= 0 ? 2u * (uint32_t)v : 2u * (uint32_t)(-(int64_t)v) - 1u; }

// ---------------------------------------------------------------------------
// Hybrid unsigned integers (C.2.3 / C.3.3).
// ---------------------------------------------------------------------------
struct Hybrid {
  uint32_t se, msb, lsb;
};

inline void hybrid_token(const Hybrid& c, uint32_t v, uint32_t* tok, uint32_t* nb, uint32_t* bits) {
  if (v < (1u << c.se)) { *tok = v; *nb = 0; *bits = 0; return; }
  const uint32_t n = 31 - __builtin_clz(v);
  const uint32_t m = (v >> (n - c.msb)) & ((1u << c.msb) - 1);
  const uint32_t l = v & ((1u << c.lsb) - 1);
  *nb = n - c.msb - c.lsb;
  *tok = (1u << c.se) + ((n - c.se) << (c.msb + c.lsb)) + (m << c.lsb) + l;
  *bits = (v >> c.lsb) & ((*nb >= 32) ? 0xFFFFFFFFu : ((1u << *nb) - 1));
}

void put_hybrid_config(BitWriter& w, const Hybrid& c, uint32_t log_alpha) {
  w.put(bit_width(log_alpha), c.se);
  if (c.se == log_alpha) return;
  w.put(bit_width(c.se), c.msb);
  w.put(bit_width(c.se - c.msb), c.lsb);
}

// ---------------------------------------------------------------------------
// Prefix codes (C.2.4 -> RFC 7932 3.2, 3.4, 3.5).
// ---------------------------------------------------------------------------
// Huffman lengths of `n` symbols limited to `limit` bits: counts below a
// floor are raised to it, the floor doubling until the tree fits.
void huffman_lengths(const uint32_t* counts, int n, int limit, uint8_t* len) {
  memset(len, 0, n);
  std::vector<int> used;
  for (int i = 0; i < n; i++)
    if (counts[i]) used.push_back(i);
  if (used.empty()) return;
  if (used.size() == 1) { len[used[0]] = 1; return; }
  struct Node { uint64_t w; int left, right; };
  std::vector<Node> nodes;
  std::vector<std::pair<uint64_t, int>> heap;
  for (uint64_t floor_count = 1;; floor_count *= 2) {
    nodes.clear();
    heap.clear();
    for (int s : used) {
      nodes.push_back({std::max<uint64_t>(counts[s], floor_count), -1, s});
      heap.push_back({nodes.back().w, (int)nodes.size() - 1});
    }
    auto cmp = [](const std::pair<uint64_t, int>& a, const std::pair<uint64_t, int>& b) {
      return a.first != b.first ? a.first > b.first : a.second > b.second;
    };
    std::make_heap(heap.begin(), heap.end(), cmp);
    while (heap.size() > 1) {
      std::pop_heap(heap.begin(), heap.end(), cmp);
      auto a = heap.back(); heap.pop_back();
      std::pop_heap(heap.begin(), heap.end(), cmp);
      auto b = heap.back(); heap.pop_back();
      nodes.push_back({a.first + b.first, a.second, b.second});
      heap.push_back({nodes.back().w, (int)nodes.size() - 1});
      std::push_heap(heap.begin(), heap.end(), cmp);
    }
    // depths by walking from the root
    std::vector<std::pair<int, int>> stack{{(int)nodes.size() - 1, 0}};
    int maxd = 0;
    while (!stack.empty()) {
      auto [id, d] = stack.back();
      stack.pop_back();
      const Node& nd = nodes[id];
      if (nd.left < 0) {
        len[nd.right] = (uint8_t)d;
        maxd = std::max(maxd, d);
      } else {
        stack.push_back({nd.left, d + 1});
        stack.push_back({nd.right, d + 1});
      }
    }
    if (maxd <= limit) return;
  }
}

// Canonical codes (shorter first, then by symbol), stored bit-reversed so
// that the LSB-first writer emits the code's most significant bit first.
void canonical_codes(const uint8_t* len, int n, uint16_t* rev) {
  uint32_t count[16] = {0}, next[16] = {0};
  for (int i = 0; i < n; i++) count[len[i]]++;
  count[0] = 0;
  uint32_t code = 0;
  for (int l = 1; l < 16; l++) {
    code = (code + count[l - 1]) << 1;
    next[l] = code;
  }
  for (int i = 0; i < n; i++) {
    rev[i] = 0;
    if (!len[i]) continue;
    const uint32_t c = next[len[i]]++;
    uint32_t r = 0;
    for (int b = 0; b < len[i]; b++) r |= ((c >> b) & 1u) << (len[i] - 1 - b);
    rev[i] = (uint16_t)r;
  }
}

struct PrefixCode {
  std::vector<uint8_t> len;
  std::vector<uint16_t> rev;
};

// The code-length code's order and the fixed code its lengths are written
// with (RFC 7932 3.5; jpxl-entropy/src/prefix.rs:10-15).
constexpr int kClcOrder[18] = {1, 2, 3, 4, 0, 5, 17, 6, 16, 7, 8, 9, 10, 11, 12, 13, 14, 15};
constexpr uint8_t kClcFixedLen[6] = {2, 4, 3, 2, 2, 4};

// A run of n >= 3 equal code lengths as repeat codes (16 or 17): each code
// of a consecutive chain extends the run to (run - 2) << extra + 3 + bits.
void repeat_codes(uint32_t n, uint32_t extra_bits, int sym, std::vector<std::pair<int, int>>* out) {
  const uint32_t e = (n - 3) & ((1u << extra_bits) - 1), q = (n - 3) >> extra_bits;
  if (q > 0) repeat_codes(q + 2, extra_bits, sym, out);
  out->push_back({sym, (int)e});
}

// Writes a prefix code over alphabet `n` (n >= 2) with lengths `len`
// (a complete code, or exactly one used symbol).
void write_prefix_code(BitWriter& w, const uint8_t* len, int n) {
  int used[4], nused = 0;
  for (int i = 0; i < n; i++)
    if (len[i]) {
      if (nused < 4) used[nused] = i;
      nused++;
    }
  if (nused <= 4) {
    // simple prefix code: symbols listed shortest first
    std::sort(used, used + nused, [&](int a, int b) { return len[a] != len[b] ? len[a] < len[b] : a < b; });
    w.put(2, 1);
    w.put(2, nused - 1);
    const uint32_t abits = bit_width((uint32_t)n - 1);
    for (int i = 0; i < nused; i++) w.put(abits, used[i]);
    if (nused == 4) w.put_bool(len[used[0]] == 1);  // tree-select: 1,2,3,3 rather than 2,2,2,2
    return;
  }
  // complex prefix code: the lengths as code-length symbols with repeats
  int last = n - 1;
  while (last > 0 && !len[last]) last--;
  std::vector<std::pair<int, int>> seq;  // (code-length symbol, extra bits)
  for (int i = 0; i <= last;) {
    int j = i;
    while (j <= last && len[j] == len[i]) j++;
    uint32_t run = j - i;
    if (len[i] == 0) {
      if (run >= 3) repeat_codes(run, 3, 17, &seq);
      else for (uint32_t k = 0; k < run; k++) seq.push_back({0, 0});
    } else {
      seq.push_back({len[i], 0});
      run -= 1;
      if (run >= 3) repeat_codes(run, 2, 16, &seq);
      else for (uint32_t k = 0; k < run; k++) seq.push_back({len[i], 0});
    }
    i = j;
  }
  uint32_t clc_count[18] = {0};
  for (auto& s : seq) clc_count[s.first]++;
  uint8_t clc_len[18];
  huffman_lengths(clc_count, 18, 5, clc_len);
  int clc_nonzero = 0;
  for (int i = 0; i < 18; i++) clc_nonzero += clc_len[i] != 0;
  uint16_t clc_rev[18];
  canonical_codes(clc_len, 18, clc_rev);
  uint16_t fixed_rev[6];
  canonical_codes(kClcFixedLen, 6, fixed_rev);
  w.put(2, 0);  // HSKIP 0
  int space = 32;
  for (int k = 0; k < 18; k++) {
    const int sym = kClcOrder[k];
    const int l = clc_len[sym];
    w.put(kClcFixedLen[l], fixed_rev[l]);
    if (l) {
      space -= 32 >> l;
      if (space <= 0) break;
    }
  }
  const bool single = clc_nonzero == 1;
  for (auto& s : seq) {
    if (!single) w.put(clc_len[s.first], clc_rev[s.first]);
    if (s.first == 16) w.put(2, s.second);
    else if (s.first == 17) w.put(3, s.second);
  }
}

// ---------------------------------------------------------------------------
// Entropy streams (C.2): contexts, clusters, one prefix code per cluster.
// ---------------------------------------------------------------------------
// A token tape entry: context, token, extra-bit count and extra bits.
inline uint64_t tape_entry(uint32_t ctx, uint32_t tok, uint32_t nb, uint32_t bits) {
  return (uint64_t)bits | ((uint64_t)nb << 32) | ((uint64_t)tok << 37) | ((uint64_t)ctx << 45);
}
inline uint32_t tape_ctx(uint64_t e) { return (uint32_t)(e >> 45); }
inline uint32_t tape_tok(uint64_t e) { return (uint32_t)(e >> 37) & 0xFF; }
inline uint32_t tape_nb(uint64_t e) { return (uint32_t)(e >> 32) & 0x1F; }
inline uint32_t tape_bits(uint64_t e) { return (uint32_t)e; }

constexpr int kAlphabet = 128;  // token bound: hybrid (4,2,0) of a 31-bit value is < 128

double entropy_bits(const uint32_t* h, int n) {
  double tot = 0, s = 0;
  for (int i = 0; i < n; i++) tot += h[i];
  if (tot <= 0) return 0;
  for (int i = 0; i < n; i++)
    if (h[i]) s -= h[i] * std::log2(h[i] / tot);
  return s;
}

// Per-symbol code lengths (bits) of a distribution with counts q; unseen
// symbols are priced as half a count.
void cost_table(const uint32_t* q, double* L) {
  double qt = 0;
  for (int i = 0; i < kAlphabet; i++) qt += q[i];
  const double lt = std::log2(qt + 0.5 * kAlphabet);
  for (int i = 0; i < kAlphabet; i++) L[i] = lt - std::log2(q[i] ? (double)q[i] : 0.5);
}

struct Sparse {
  std::vector<std::pair<uint16_t, uint32_t>> e;  // (token, count)
  double cost(const double* L) const {
    double s = 0;
    for (auto& p : e) s += p.second * L[p.first];
    return s;
  }
};

// Greedy clustering of per-context token histograms: seeds by largest
// coding-cost gain, then two rounds of nearest-cluster assignment (spread
// over the pool when one is given).  Clusters are numbered in order of first
// use; unused contexts join the cluster of the context before them.
void cluster_contexts(const std::vector<uint32_t>& hist, int num_ctx, int max_clusters, double min_gain,
                      std::vector<uint8_t>* map, int* num_clusters, Pool* pool = nullptr) {
  map->assign(num_ctx, 0);
  std::vector<Sparse> sp(num_ctx);
  std::vector<double> tot(num_ctx, 0), self(num_ctx, 0);
  const int pchunks = pool && num_ctx > 1024 ? 32 : 1;
  auto scan = [&](int ci, int) {
    for (int c = num_ctx * ci / pchunks; c < num_ctx * (ci + 1) / pchunks; c++) {
      const uint32_t* h = &hist[(size_t)c * kAlphabet];
      double t = 0, e = 0;
      for (int s = 0; s < kAlphabet; s++)
        if (h[s]) {
          t += h[s];
          sp[c].e.push_back({(uint16_t)s, h[s]});
        }
      if (t > 0) {
        for (auto& q : sp[c].e) e -= q.second * std::log2(q.second / t);
        tot[c] = t;
        self[c] = e;
      }
    }
  };
  if (pchunks > 1) pool->run(pchunks, scan);
  else scan(0, 0);
  std::vector<int> used;
  for (int c = 0; c < num_ctx; c++)
    if (tot[c] > 0) used.push_back(c);
  if (used.empty()) { *num_clusters = 1; return; }
  std::vector<int> assign(num_ctx, -1);
  int k = 0;
  if ((int)used.size() <= max_clusters) {
    for (int c : used) assign[c] = k++;
  } else {
    std::vector<std::vector<uint32_t>> cent;
    std::vector<std::vector<double>> L;
    int first = used[0];
    for (int c : used)
      if (tot[c] > tot[first]) first = c;
    auto add_seed = [&](int c) {
      cent.emplace_back(hist.begin() + (size_t)c * kAlphabet, hist.begin() + (size_t)(c + 1) * kAlphabet);
      L.emplace_back(kAlphabet);
      cost_table(cent.back().data(), L.back().data());
    };
    add_seed(first);
    // Seeds are drawn from the heaviest contexts only (their coding-cost
    // gaps are the ones kept up to date); every used context is assigned.
    std::vector<int> cand(used);
    if (cand.size() > 384) {
      std::nth_element(cand.begin(), cand.begin() + 384, cand.end(),
                       [&](int a, int b) { return tot[a] != tot[b] ? tot[a] > tot[b] : a < b; });
      cand.resize(384);
      std::sort(cand.begin(), cand.end());
    }
    std::vector<double> gap(num_ctx, 0);
    for (int c : cand) gap[c] = sp[c].cost(L[0].data()) - self[c];
    while ((int)cent.size() < max_clusters) {
      int best = -1;
      for (int c : cand)
        if (best < 0 || gap[c] > gap[best]) best = c;
      if (gap[best] < min_gain) break;
      add_seed(best);
      const double* Lj = L.back().data();
      for (int c : cand) gap[c] = std::min(gap[c], sp[c].cost(Lj) - self[c]);
    }
    const int nu = (int)used.size(), chunks = pool && nu > 256 ? 32 : 1;
    for (int round = 0; round < 2; round++) {
      auto body = [&](int ci, int) {
        for (int u = nu * ci / chunks; u < nu * (ci + 1) / chunks; u++) {
          const int c = used[u];
          double bestb = 1e300;
          for (size_t j = 0; j < cent.size(); j++) {
            const double b = sp[c].cost(L[j].data());
            if (b < bestb) { bestb = b; assign[c] = (int)j; }
          }
        }
      };
      if (chunks > 1) pool->run(chunks, body);
      else body(0, 0);
      for (auto& v : cent) std::fill(v.begin(), v.end(), 0u);
      for (int c : used)
        for (auto& q : sp[c].e) cent[assign[c]][q.first] += q.second;
      for (size_t j = 0; j < cent.size(); j++) cost_table(cent[j].data(), L[j].data());
    }
    k = (int)cent.size();
  }
  // dense renumbering in order of first use
  std::vector<int> renum(std::max(k, 1), -1);
  int next = 0, prev = 0;
  for (int c = 0; c < num_ctx; c++) {
    if (assign[c] < 0) { (*map)[c] = (uint8_t)prev; continue; }
    if (renum[assign[c]] < 0) renum[assign[c]] = next++;
    prev = renum[assign[c]];
    (*map)[c] = (uint8_t)prev;
  }
  *num_clusters = std::max(next, 1);
}

// The stream's header and codes, built from its tokens' histograms.
struct EntropyCode {
  int num_ctx = 0, num_clusters = 1;
  Hybrid cfg{4, 2, 0};
  std::vector<uint8_t> map;
  std::vector<PrefixCode> codes;
  std::vector<uint32_t> alphabet;

  // hist: num_ctx x kAlphabet token counts; or, with ids, one row per
  // context of ids (ascending; the contexts with tokens), the other contexts
  // of the nctx having none.
  void build(const std::vector<uint32_t>& hist, int nctx, Hybrid c, int max_clusters, Pool* pool = nullptr,
             const std::vector<uint32_t>* ids = nullptr) {
    num_ctx = nctx;
    cfg = c;
    const int rows = ids ? (int)ids->size() : nctx;
    std::vector<uint8_t> row_map;
    if (nctx == 1) {
      row_map.assign(rows, 0);
      num_clusters = 1;
    } else {
      cluster_contexts(hist, rows, max_clusters, 64.0, &row_map, &num_clusters, pool);
    }
    if (ids) {
      // contexts without tokens join the cluster of the context before them
      map.assign(nctx, 0);
      uint8_t prev = 0;
      size_t u = 0;
      for (int ctx = 0; ctx < nctx; ctx++) {
        if (u < ids->size() && (*ids)[u] == (uint32_t)ctx) prev = row_map[u++];
        map[ctx] = prev;
      }
    } else {
      map = row_map;
    }
    std::vector<uint32_t> ch((size_t)num_clusters * kAlphabet, 0);
    for (int r = 0; r < rows; r++) {
      const uint32_t* h = &hist[(size_t)r * kAlphabet];
      uint32_t* d = &ch[(size_t)row_map[r] * kAlphabet];
      int last = kAlphabet;
      if (rows > 1024)  // most of a large stream's contexts are sparse: stop at the last count
        while (last > 0 && !h[last - 1]) last--;
      for (int s = 0; s < last; s++) d[s] += h[s];
    }
    codes.assign(num_clusters, {});
    alphabet.assign(num_clusters, 1);
    single_symbol_.assign(num_clusters, -1);
    packed_.assign((size_t)num_clusters * kAlphabet, 0u);
    auto code_of = [&](int k, int) {
      const uint32_t* h = &ch[(size_t)k * kAlphabet];
      int last = 0;
      for (int s = 0; s < kAlphabet; s++)
        if (h[s]) last = s;
      alphabet[k] = last + 1;
      PrefixCode& pc = codes[k];
      pc.len.assign(alphabet[k], 0);
      pc.rev.assign(alphabet[k], 0);
      huffman_lengths(h, alphabet[k], 15, pc.len.data());
      int nused = 0;
      for (uint32_t s = 0; s < alphabet[k]; s++) nused += pc.len[s] != 0;
      if (nused <= 1) {
        std::fill(pc.len.begin(), pc.len.end(), 0);  // one symbol: zero bits each
        single_symbol_[k] = last;
      } else {
        canonical_codes(pc.len.data(), alphabet[k], pc.rev.data());
      }
      for (uint32_t s = 0; s < alphabet[k]; s++)
        packed_[(size_t)k * kAlphabet + s] = pc.len[s] | (uint32_t)pc.rev[s] << 5;
    };
    if (pool && num_clusters > 8) pool->run(num_clusters, code_of);
    else
      for (int k = 0; k < num_clusters; k++) code_of(k, 0);
  }

  void write_header(BitWriter& w) const;

  uint32_t code_len(uint32_t ctx, uint32_t tok) const {
    const PrefixCode& pc = codes[map[ctx]];
    return tok < pc.len.size() ? pc.len[tok] : 0;
  }

  inline void put(BitWriter& w, uint64_t e) const {
    const PrefixCode& pc = codes[map[tape_ctx(e)]];
    const uint32_t t = tape_tok(e);
    if (!pc.len.empty() && pc.len[t]) w.put(pc.len[t], pc.rev[t]);
    const uint32_t nb = tape_nb(e);
    if (nb) w.put(nb, tape_bits(e));
  }

  // A tape's codes and extra bits, in order (the hot loop of the pass
  // groups): one table load per token, one write when both fit 32 bits.
  void put_tape(BitWriter& w, const uint64_t* e, size_t n, uint64_t ctx_bits = 0) const {
    const uint8_t* m = map.data();
    const uint32_t* pk = packed_.data();
    for (size_t i = 0; i < n; i++) {
      const uint64_t x = e[i] | ctx_bits;
      const uint32_t c = pk[(size_t)m[tape_ctx(x)] * kAlphabet + tape_tok(x)];
      const uint32_t len = c & 31, nb = tape_nb(x);
      if (len + nb <= 32) {
        w.put(len + nb, (uint64_t)(c >> 5) | ((uint64_t)tape_bits(x) << len));
      } else {
        w.put(len, c >> 5);
        w.put(nb, tape_bits(x));
      }
    }
  }

  std::vector<int> single_symbol_;
  std::vector<uint32_t> packed_;  // per cluster and token: code length | bit-reversed code << 5
};

void write_cluster_map(BitWriter& w, const std::vector<uint8_t>& map, int num_clusters);

void EntropyCode::write_header(BitWriter& w) const {
  w.put_bool(false);  // LZ77 off
  if (num_ctx > 1) write_cluster_map(w, map, num_clusters);
  w.put_bool(true);  // prefix codes: log_alphabet_size 15
  for (int k = 0; k < num_clusters; k++) put_hybrid_config(w, cfg, 15);
  for (int k = 0; k < num_clusters; k++) {
    const uint32_t n = alphabet[k];
    if (n == 1) { w.put_bool(false); continue; }
    w.put_bool(true);
    const uint32_t nb = 31 - __builtin_clz(n - 1);
    w.put(4, nb);
    w.put(nb, n - 1 - (1u << nb));
  }
  for (int k = 0; k < num_clusters; k++) {
    const uint32_t n = alphabet[k];
    if (n == 1) continue;
    if (single_symbol_[k] >= 0) {
      // one used symbol: a simple code of one symbol
      w.put(2, 1);
      w.put(2, 0);
      w.put(bit_width(n - 1), (uint32_t)single_symbol_[k]);
      continue;
    }
    write_prefix_code(w, codes[k].len.data(), (int)n);
  }
}

// A whole stream from (context, value) pairs: header then symbols.
void write_stream(BitWriter& w, const std::vector<std::pair<uint32_t, uint32_t>>& syms, int num_ctx, Hybrid cfg,
                  int max_clusters) {
  std::vector<uint32_t> hist((size_t)num_ctx * kAlphabet, 0);
  std::vector<uint64_t> tape;
  tape.reserve(syms.size());
  for (auto& s : syms) {
    uint32_t tok, nb, bits;
    hybrid_token(cfg, s.second, &tok, &nb, &bits);
    hist[(size_t)s.first * kAlphabet + tok]++;
    tape.push_back(tape_entry(s.first, tok, nb, bits));
  }
  EntropyCode ec;
  ec.build(hist, num_ctx, cfg, max_clusters);
  ec.write_header(w);
  for (uint64_t e : tape) ec.put(w, e);
}

// C.2.2: a simple map when at most eight clusters, else move-to-front values
// through a nested one-context stream.
void write_cluster_map(BitWriter& w, const std::vector<uint8_t>& map, int num_clusters) {
  const uint32_t nbits = ceil_log2((uint64_t)num_clusters);
  if (nbits <= 3 && nbits * map.size() <= 64) {
    w.put_bool(true);
    w.put(2, nbits);
    for (uint8_t c : map) w.put(nbits, c);
    return;
  }
  w.put_bool(false);
  w.put_bool(true);  // use_mtf
  uint8_t table[256];
  for (int i = 0; i < 256; i++) table[i] = (uint8_t)i;
  std::vector<std::pair<uint32_t, uint32_t>> syms;
  syms.reserve(map.size());
  for (uint8_t v : map) {
    int idx = 0;
    while (table[idx] != v) idx++;
    syms.push_back({0u, (uint32_t)idx});
    for (int j = idx; j > 0; j--) table[j] = table[j - 1];
    table[0] = v;
  }
  write_stream(w, syms, 1, Hybrid{4, 0, 0}, 1);
}

// ---------------------------------------------------------------------------
// Modular control images (Annex H): the LF planes and the HF metadata.
// ---------------------------------------------------------------------------
struct Channel {
  uint32_t w, h;
  std::vector<int32_t> v;
};

// Table H.3 rows 0, 1, 2 and 5 with H.3's edge rules (predictor.rs:172-213).
inline int64_t predict(int pred, const int32_t* row, const int32_t* up, uint32_t x, uint32_t y) {
  const int64_t W = x > 0 ? row[x - 1] : (y > 0 ? up[x] : 0);
  const int64_t N = y > 0 ? up[x] : W;
  switch (pred) {
    case 0: return 0;
    case 1: return W;
    case 2: return N;
    default: {
      const int64_t NW = (x > 0 && y > 0) ? up[x - 1] : W;
      const int64_t g = W + N - NW;
      return std::min(std::max(g, std::min(W, N)), std::max(W, N));
    }
  }
}

// Calls f on the packed residual of every sample of rows 0, step, 2 step, ...
// under predictor P (a template argument, so the switch folds away).
template <int P, class F>
void for_residuals(const Channel& c, uint32_t step, F&& f) {
  for (uint32_t y = 0; y < c.h; y += step) {
    const int32_t* row = &c.v[(size_t)y * c.w];
    const int32_t* up = y ? row - c.w : row;
    for (uint32_t x = 0; x < c.w; x++) f(pack_signed((int32_t)(row[x] - predict(P, row, up, x, y))));
  }
}

template <class F>
void for_residuals(int pred, const Channel& c, uint32_t step, F&& f) {
  switch (pred) {
    case 0: for_residuals<0>(c, step, f); break;
    case 1: for_residuals<1>(c, step, f); break;
    case 2: for_residuals<2>(c, step, f); break;
    default: for_residuals<5>(c, step, f); break;
  }
}

// A channel ready for a modular stream: its leaf's predictor and offset and
// its packed residuals.  A constant channel predicts zero with the constant
// as offset (its samples cost no bits); any other carries the cheapest of
// four predictors, priced on a subsample of rows.
struct PreparedChannel {
  uint32_t w = 0, h = 0;
  int pred = 5;
  int32_t offset = 0;
  bool constant = false;              // every residual is token 0 with no extra bits
  std::vector<uint64_t> tape;         // its residuals' tokens (context added when written); empty if constant
  std::vector<uint32_t> hist;         // their counts (kAlphabet)
};

constexpr Hybrid kModularHybrid{4, 1, 0};

void prepare_channel(const Channel& ch, PreparedChannel* pc) {
  pc->w = ch.w;
  pc->h = ch.h;
  pc->pred = 5;
  pc->offset = 0;
  pc->tape.clear();
  pc->hist.assign(kAlphabet, 0);
  if (!ch.w || !ch.h) return;
  const size_t n = (size_t)ch.w * ch.h;
  if (std::all_of(ch.v.begin(), ch.v.end(), [&](int32_t x) { return x == ch.v[0]; })) {
    pc->pred = 0;
    pc->offset = ch.v[0];
    pc->constant = true;
    pc->hist[0] = (uint32_t)n;
    return;
  }
  // price each predictor on about 4096 samples (whole rows)
  const uint32_t rows = std::max<uint32_t>(1, (4096 + ch.w - 1) / ch.w);
  const uint32_t step = std::max<uint32_t>(1, ch.h / rows);
  double best = 1e300;
  for (int p : {5, 0, 1, 2}) {
    uint32_t h[kAlphabet] = {0};
    double extra = 0;
    for_residuals(p, ch, step, [&](uint32_t r) {
      uint32_t tok, nb, bits;
      hybrid_token(kModularHybrid, r, &tok, &nb, &bits);
      h[tok]++;
      extra += nb;
    });
    const double cost = entropy_bits(h, kAlphabet) + extra;
    if (cost < best) { best = cost; pc->pred = p; }
  }
  pc->tape.resize(n);
  uint64_t* t = pc->tape.data();
  uint32_t* hist = pc->hist.data();
  for_residuals(pc->pred, ch, 1, [&](uint32_t r) {
    uint32_t tok, nb, bits;
    hybrid_token(kModularHybrid, r, &tok, &nb, &bits);
    hist[tok]++;
    *t++ = tape_entry(0, tok, nb, bits);
  });
}

// Writes a modular sub-bitstream of prepared channels (H.2): local tree,
// default WP header, no transforms; the tree splits on the channel index
// (property 0) so that every channel has a leaf.  Zero channels write
// nothing (H.1).
void write_modular_prepared(BitWriter& w, const std::vector<PreparedChannel>& pcs) {
  const int C = (int)pcs.size();
  if (C == 0) return;
  w.put_bool(false);  // use_global_tree
  w.put_bool(true);   // default WP header
  w.put(2, 0);        // nb_transforms = 0 (U32 Val(0))
  // The tree, breadth first: channel > C-2 ? leaf(C-1) : (channel > C-3 ? ...).
  std::vector<std::pair<uint32_t, uint32_t>> tree;
  auto leaf = [&](int ch) {
    tree.push_back({1u, 0u});
    tree.push_back({2u, (uint32_t)pcs[ch].pred});
    tree.push_back({3u, pack_signed(pcs[ch].offset)});
    tree.push_back({4u, 0u});
    tree.push_back({5u, 0u});
  };
  if (C == 1) {
    leaf(0);
  } else {
    for (int t = C - 2; t >= 0; t--) {
      tree.push_back({1u, 1u});  // property 0 (channel index) + 1
      tree.push_back({0u, pack_signed(t)});
      leaf(t + 1);
    }
    leaf(0);
  }
  write_stream(w, tree, 6, Hybrid{4, 0, 0}, 6);
  // the data: leaf context of channel i is C-1-i (leaves are numbered in decode order)
  std::vector<uint32_t> hist((size_t)C * kAlphabet);
  for (int i = 0; i < C; i++)
    std::copy(pcs[i].hist.begin(), pcs[i].hist.end(), hist.begin() + (size_t)(C - 1 - i) * kAlphabet);
  EntropyCode ec;
  ec.build(hist, C, kModularHybrid, C);
  ec.write_header(w);
  for (int i = 0; i < C; i++) {
    const uint32_t c = (uint32_t)(C - 1 - i);
    const uint64_t ctx = (uint64_t)c << 45;
    if (pcs[i].constant) {
      // token 0 without extra bits, w x h times: nothing to write when its
      // code is the zero-length one of a single-symbol cluster
      if (ec.code_len(c, 0))
        for (size_t k = 0; k < (size_t)pcs[i].w * pcs[i].h; k++) ec.put(w, tape_entry(c, 0, 0, 0));
      continue;
    }
    ec.put_tape(w, pcs[i].tape.data(), pcs[i].tape.size(), ctx);
  }
}

void write_modular(BitWriter& w, const std::vector<Channel>& chans) {
  std::vector<PreparedChannel> pcs(chans.size());
  for (size_t i = 0; i < chans.size(); i++) prepare_channel(chans[i], &pcs[i]);
  write_modular_prepared(w, pcs);
}

// ---------------------------------------------------------------------------
// The HF context model (I.4; hf_coeff.rs:480-575, block_ctx.rs:16-20).
// ---------------------------------------------------------------------------
constexpr uint8_t kDefaultBlockCtx[39] = {0, 1, 2, 2, 3, 3, 4, 5, 6, 6, 6, 6, 6,  //
                                          7, 8, 9, 9, 10, 11, 12, 13, 14, 14, 14, 14, 14,
                                          7, 8, 9, 9, 10, 11, 12, 13, 14, 14, 14, 14, 14};
constexpr uint32_t kCoeffFreqCtx[64] = {
    0,  0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 15, 16, 16, 17, 17,
    18, 18, 19, 19, 20, 20, 21, 21, 22, 22, 23, 23, 23, 23, 24, 24, 24, 24, 25, 25, 25, 25,
    26, 26, 26, 26, 27, 27, 27, 27, 28, 28, 28, 28, 29, 29, 29, 29, 30, 30, 30, 30};
constexpr uint32_t kCoeffNumNonzeroCtx[64] = {
    0,   0,   31,  62,  62,  93,  93,  93,  93,  123, 123, 123, 123, 152, 152, 152, 152, 152, 152, 152, 152, 180,
    180, 180, 180, 180, 180, 180, 180, 180, 180, 180, 180, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206,
    206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206, 206};
constexpr int kNbBlockCtx = 15;
constexpr int kNonZeroBuckets = 37, kZeroDensityCtx = 458;
constexpr int kHfContexts = (kNonZeroBuckets + kZeroDensityCtx) * kNbBlockCtx;  // 7425

struct NaturalOrders {
  std::vector<uint16_t> o8, o16, o32;    // position -> cell
  std::vector<uint16_t> i8, i16, i32;    // cell -> position
  NaturalOrders() : o8(64), o16(256), o32(1024), i8(64), i16(256), i32(1024) {
    jpxlg::natural_order_table(8, o8.data());
    jpxlg::natural_order_table(16, o16.data());
    jpxlg::natural_order_table(32, o32.data());
    for (int k = 0; k < 64; k++) i8[o8[k]] = (uint16_t)k;
    for (int k = 0; k < 256; k++) i16[o16[k]] = (uint16_t)k;
    for (int k = 0; k < 1024; k++) i32[o32[k]] = (uint16_t)k;
  }
  const uint16_t* of(uint32_t side) const { return side == 8 ? o8.data() : side == 16 ? o16.data() : o32.data(); }
  const uint16_t* inv(uint32_t side) const { return side == 8 ? i8.data() : side == 16 ? i16.data() : i32.data(); }
};
const NaturalOrders& natural_orders() {
  static const NaturalOrders o;
  return o;
}

// One pool per process, rebuilt when the worker count changes.  A process
// forked from one that had a pool inherits the object but not its threads:
// the child drops it (it cannot join threads it does not have) and starts
// its own.  Calls are serialised by g_pool_mu.
std::mutex g_pool_mu;
Pool* pool_for(int threads) {
  static Pool* pool = nullptr;
  static pid_t owner = 0;
  if (pool && owner != getpid()) pool = nullptr;  // inherited across fork: left unjoined
  if (!pool || pool->size() != threads) {
    delete pool;
    pool = new Pool(threads);
    owner = getpid();
  }
  return pool;
}

// ---------------------------------------------------------------------------
// The frame.
// ---------------------------------------------------------------------------
struct Geometry {
  uint32_t w, h, bw, bh, gx, gy, lgx, lgy, tx, ty;
  uint32_t num_groups() const { return gx * gy; }
  uint32_t num_lf_groups() const { return lgx * lgy; }
};

Geometry geometry_of(uint32_t w, uint32_t h) {
  Geometry g;
  g.w = w;
  g.h = h;
  g.bw = (w + 7) / 8;
  g.bh = (h + 7) / 8;
  g.gx = (w + 255) / 256;
  g.gy = (h + 255) / 256;
  g.lgx = (w + 2047) / 2048;
  g.lgy = (h + 2047) / 2048;
  g.tx = (g.bw + 7) / 8;
  g.ty = (g.bh + 7) / 8;
  return g;
}

const Dist kSizeDim[4] = {Bits(9, 1), Bits(13, 1), Bits(18, 1), Bits(30, 1)};
const Dist kIntBps[4] = {Val(8), Val(10), Val(12), Bits(6, 1)};
const Dist kNumExtra[4] = {Val(0), Val(1), Bits(4, 2), Bits(12, 1)};
const Dist kUpsampling[4] = {Val(1), Val(2), Val(4), Val(8)};
const Dist kNumPasses[4] = {Val(1), Val(2), Val(3), Bits(3, 4)};
const Dist kBlendMode[4] = {Val(0), Val(1), Val(2), Bits(2, 3)};
const Dist kNameLen[4] = {Val(0), Bits(4), Bits(5, 16), Bits(10, 48)};
const Dist kGlobalScale[4] = {Bits(11, 1), Bits(11, 2049), Bits(12, 4097), Bits(16, 8193)};
const Dist kQuantLf[4] = {Val(16), Bits(5, 1), Bits(8, 1), Bits(16, 1)};
const Dist kColourFactor[4] = {Val(84), Val(256), Bits(8, 2), Bits(16, 258)};
const Dist kUsedOrders[4] = {Val(0x5F), Val(0x13), Val(0), Bits(13)};
const Dist kTocEntry[4] = {Bits(10), Bits(14, 1024), Bits(22, 17408), Bits(30, 4211712)};

// Headers (headers.rs:104-152): signature, SizeHeader with explicit width,
// ImageMetadata all default at 8 bits, else the bit depth row.
void write_image_headers(BitWriter& w, const jpxlg_frame_desc& f) {
  w.put(16, 0x0AFF);
  w.put_bool(false);  // not div8
  put_u32(w, kSizeDim, f.height);
  w.put(3, 0);  // ratio: width explicit
  put_u32(w, kSizeDim, f.width);
  if (f.bits_per_sample == 8) {
    w.put_bool(true);  // all_default
    w.put_bool(true);  // default_m
  } else {
    w.put_bool(false);  // all_default
    w.put_bool(false);  // extra_fields
    w.put_bool(false);  // float_sample
    put_u32(w, kIntBps, f.bits_per_sample);
    w.put_bool(true);  // modular_16_bit_buffer_sufficient
    put_u32(w, kNumExtra, 0);
    w.put_bool(true);  // xyb_encoded
    w.put_bool(true);  // colour encoding all_default (sRGB)
    put_u64(w, 0);     // extensions
    w.put_bool(true);  // default_m
  }
  w.pad_to_byte();
}

// Frame header (headers.rs:159-233): a regular, last VarDCT frame without
// adaptive LF smoothing, one pass, the restoration filter as rendered.
void write_frame_header(BitWriter& w, const jpxlg_frame_desc& f) {
  w.put_bool(false);  // all_default
  w.put(2, 0);        // kRegularFrame
  w.put(1, 0);        // kVarDCT
  put_u64(w, 0x80);   // flags: skip adaptive LF smoothing
  put_u32(w, kUpsampling, 1);
  w.put(3, f.x_qm_scale);
  w.put(3, f.b_qm_scale);
  put_u32(w, kNumPasses, 1);
  w.put_bool(false);  // have_crop
  put_u32(w, kBlendMode, 0);
  w.put_bool(true);  // is_last
  put_u32(w, kNameLen, 0);
  // RestorationFilter
  w.put_bool(false);  // all_default
  w.put_bool(f.gaborish != 0);
  if (f.gaborish) w.put_bool(false);  // gab_custom
  w.put(2, f.epf_iters);
  if (f.epf_iters) {
    w.put_bool(false);  // epf_sharp_custom
    w.put_bool(false);  // epf_weight_custom
    w.put_bool(false);  // epf_sigma_custom
  }
  put_u64(w, 0);  // restoration extensions
  put_u64(w, 0);  // frame extensions
}

struct Frame {
  const jpxlg_frame_desc* f;
  Geometry g;
  const int32_t* arena;
  const int32_t* lf;
  const jpxlg_varblock* vb;
  uint32_t nvb;
  const int8_t* cfl_x;
  const int8_t* cfl_b;
  const jpxlg_run* runs = nullptr;  // with runs, arena holds the run arena of runs_len entries
  size_t runs_len = 0;
  std::vector<std::vector<uint32_t>> group_vbs;  // per pass group, IR order
  std::vector<std::vector<uint32_t>> lf_group_vbs;
  std::vector<size_t> lf_offset;  // start of each LF group's planes in lf
};

// G.1 (write.rs:643-682).
void write_lf_global(BitWriter& w, const Frame& fr) {
  const jpxlg_frame_desc& f = *fr.f;
  w.put_bool(true);  // LfChannelDequantization all default
  put_u32(w, kGlobalScale, f.global_scale);
  put_u32(w, kQuantLf, f.quant_lf);
  w.put_bool(true);  // HF block context: default
  if (f.x_factor_lf == 0 && f.b_factor_lf == 0) {
    w.put_bool(true);  // LfChannelCorrelation all default
  } else {
    w.put_bool(false);
    put_u32(w, kColourFactor, 84);
    w.put(16, 0x0000);  // base_correlation_x = 0 (binary16)
    w.put(16, 0x3C00);  // base_correlation_b = 1 (binary16)
    w.put(8, (uint32_t)(f.x_factor_lf + 128));
    w.put(8, (uint32_t)(f.b_factor_lf + 128));
  }
  w.put_bool(false);  // GlobalModular: no global tree; no channels follow
}

// G.2 (write.rs:688-785).  G.2.2 (extra precision, the LF planes as Y, X, B)
// is prepared per channel and written with the section; G.2.4 (nb_blocks,
// then XFromY, BFromY, BlockInfo and Sharpness) is written whole into its own
// writer and appended bitwise.  G.2.3 (no extra channels) writes nothing.
struct LfGroupBox {
  uint32_t b0x, b0y, gw, gh;
};
LfGroupBox lf_group_box(const Geometry& g, uint32_t id) {
  const uint32_t lx = id % g.lgx, ly = id / g.lgx;
  return {lx * 256, ly * 256, std::min(256u, g.bw - lx * 256), std::min(256u, g.bh - ly * 256)};
}

void prepare_lf_plane(const Frame& fr, uint32_t id, int i, PreparedChannel* pc) {
  const LfGroupBox b = lf_group_box(fr.g, id);
  const size_t area = (size_t)b.gw * b.gh;
  const int plane_of[3] = {1, 0, 2};
  Channel ch;
  ch.w = b.gw;
  ch.h = b.gh;
  const int32_t* src = fr.lf + fr.lf_offset[id] + (size_t)plane_of[i] * area;
  ch.v.assign(src, src + area);
  prepare_channel(ch, pc);
}

void write_hf_metadata(BitWriter& w, const Frame& fr, uint32_t id) {
  const Geometry& g = fr.g;
  const LfGroupBox b = lf_group_box(g, id);
  const uint32_t gw = b.gw, gh = b.gh;
  const size_t area = (size_t)gw * gh;
  const std::vector<uint32_t>& list = fr.lf_group_vbs[id];
  const uint32_t nb = (uint32_t)list.size();
  w.put(ceil_log2(std::max<uint64_t>(area, 1)), nb - 1);
  const uint32_t t0x = b.b0x / 8, t0y = b.b0y / 8;
  const uint32_t tw = (gw * 8 + 63) / 64, th = (gh * 8 + 63) / 64;
  std::vector<Channel> ch(4);
  ch[0].w = ch[1].w = tw;
  ch[0].h = ch[1].h = th;
  ch[0].v.resize((size_t)tw * th);
  ch[1].v.resize((size_t)tw * th);
  for (uint32_t y = 0; y < th; y++)
    for (uint32_t x = 0; x < tw; x++) {
      const size_t t = (size_t)(t0y + y) * g.tx + (t0x + x);
      ch[0].v[(size_t)y * tw + x] = fr.cfl_x ? fr.cfl_x[t] : 0;
      ch[1].v[(size_t)y * tw + x] = fr.cfl_b ? fr.cfl_b[t] : 0;
    }
  ch[2].w = nb;
  ch[2].h = 2;
  ch[2].v.resize((size_t)nb * 2);
  for (uint32_t i = 0; i < nb; i++) {
    const jpxlg_varblock& v = fr.vb[list[i]];
    ch[2].v[i] = v.side == 8 ? 0 : v.side == 16 ? 4 : 5;  // DctSelect of DCT8x8, DCT16x16, DCT32x32
    ch[2].v[nb + i] = (int32_t)v.hf_mul - 1;
  }
  ch[3].w = gw;
  ch[3].h = gh;
  ch[3].v.assign(area, (int32_t)fr.f->sharpness);
  write_modular(w, ch);
}

// Per-worker HF token counts in a dense table sized for the common tokens:
// the non-zero-count contexts (the first kNonZeroBuckets * kNbBlockCtx) keep
// kNzTok tokens each, the coefficient contexts kCoefTok; rarer, larger
// tokens are listed apart.  Small tables keep each worker's counts in its
// own cache and the merge short.
constexpr int kNzContexts = kNonZeroBuckets * kNbBlockCtx;
constexpr int kNzTok = 48, kCoefTok = 24;
constexpr size_t kDenseCounts = (size_t)kNzContexts * kNzTok + (size_t)(kHfContexts - kNzContexts) * kCoefTok;
inline size_t dense_slot(uint32_t ctx) {
  return ctx < (uint32_t)kNzContexts ? (size_t)ctx * kNzTok
                                     : (size_t)kNzContexts * kNzTok + (size_t)(ctx - kNzContexts) * kCoefTok;
}
inline uint32_t dense_width(uint32_t ctx) { return ctx < (uint32_t)kNzContexts ? kNzTok : kCoefTok; }
constexpr int kTouchedWords = (kHfContexts + 63) / 64;
struct HfCounts {
  std::vector<uint32_t> dense;    // kDenseCounts; all zero between calls
  std::vector<uint64_t> touched;  // kTouchedWords: the contexts counted; all zero between calls
  std::vector<uint32_t> escape;   // ctx << 8 | token
};

// The pool workers' count tables, kept across calls (and zero between them:
// the merge clears every slot it reads).  Guarded by g_pool_mu.
std::vector<HfCounts>& worker_counts(int nw) {
  static std::vector<HfCounts> counts;
  static pid_t owner = 0;
  if (owner != getpid()) counts.clear();
  owner = getpid();
  if ((int)counts.size() < nw) counts.resize(nw);
  for (auto& c : counts) {
    if (c.dense.size() != kDenseCounts) c.dense.assign(kDenseCounts, 0);
    if (c.touched.size() != (size_t)kTouchedWords) c.touched.assign(kTouchedWords, 0);
    c.escape.clear();
  }
  return counts;
}

// A pass group's token tape (uninitialised storage sized by a bound).
struct Tape {
  std::unique_ptr<uint64_t[]> e;
  size_t n = 0;
};

// I.4 tokens of one pass group, appended to its tape; counted into hc.
// Returns false when a run breaks its definition (the tokens are then void).
bool tokenize_group(const Frame& fr, uint32_t gid, Tape* tape, HfCounts* hc) {
  bool ok = true;
  const Geometry& g = fr.g;
  const uint32_t g0x = (gid % g.gx) * 32, g0y = (gid / g.gx) * 32;
  size_t bound = 0;
  for (uint32_t idx : fr.group_vbs[gid]) {
    const uint32_t s = fr.vb[idx].side;
    bound += 3 * (s * s - (s / 8) * (s / 8) + 1);
  }
  tape->e.reset(new uint64_t[std::max<size_t>(bound, 1)]);
  uint64_t* out = tape->e.get();
  uint32_t nzgrid[3][32 * 32];
  const NaturalOrders& nat = natural_orders();
  uint32_t* dense = hc->dense.data();
  uint64_t* touched = hc->touched.data();
  auto emit = [&](uint32_t ctx, uint32_t v) {
    touched[ctx >> 6] |= 1ull << (ctx & 63);
    uint32_t tok, nb, bits;
    if (v < 16) {
      tok = v;
      nb = bits = 0;
    } else {
      hybrid_token(Hybrid{4, 2, 0}, v, &tok, &nb, &bits);
    }
    if (tok < dense_width(ctx)) dense[dense_slot(ctx) + tok]++;
    else hc->escape.push_back(ctx << 8 | tok);
    *out++ = tape_entry(ctx, tok, nb, bits);
  };
  for (uint32_t idx : fr.group_vbs[gid]) {
    const jpxlg_varblock& v = fr.vb[idx];
    const uint32_t side = v.side, nblk1 = side / 8, nblocks = nblk1 * nblk1, size = side * side;
    const uint32_t lb = nblocks == 1 ? 0 : nblocks == 4 ? 2 : 4;  // log2 of covered blocks
    const uint32_t order_id = side == 8 ? 0 : side == 16 ? 2 : 3;
    const uint16_t* order = nat.of(side);
    const uint16_t* inv = nat.inv(side);
    const uint32_t x = v.bx - g0x, y = v.by - g0y;
    for (int c : {1, 0, 2}) {
      const uint32_t block_ctx = kDefaultBlockCtx[(c < 2 ? c ^ 1 : 2) * 13 + order_id];
      uint32_t* grid = nzgrid[c];
      uint32_t pred;
      if (x == 0 && y == 0) pred = 32;
      else if (x == 0) pred = grid[(y - 1) * 32 + x];
      else if (y == 0) pred = grid[y * 32 + x - 1];
      else pred = (grid[(y - 1) * 32 + x] + grid[y * 32 + x - 1] + 1) >> 1;
      // non-zeros among the HF positions and the last one's position
      uint32_t nz = 0, last = 0;
      const int32_t* a;     // coefficient of natural-order position k: a[order[k]], or run[k - nblocks]
      const int32_t* run = nullptr;
      if (fr.runs) {
        const jpxlg_run& r = fr.runs[3 * (size_t)idx + c];
        nz = r.nonzeros;
        last = r.count ? nblocks + r.count - 1 : 0;
        run = fr.arena + r.start;
        a = nullptr;
        if (r.count && run[r.count - 1] == 0) ok = false;
      } else {
        // over the array in raster order (LLF cells are the positions below nblocks)
        a = fr.arena + v.start + (uint64_t)c * size;
        for (uint32_t cell = 0; cell < size; cell++) {
          const uint32_t k = inv[cell];
          const uint32_t m = (uint32_t)(a[cell] != 0) & (uint32_t)(k >= nblocks);
          nz += m;
          last = std::max(last, m * k);
        }
      }
      const uint32_t p = std::min(pred, 64u);
      emit(block_ctx + kNbBlockCtx * (p < 8 ? p : 4 + p / 2), nz);
      const uint32_t per = (nz + nblocks - 1) >> lb;
      for (uint32_t j = 0; j < nblk1; j++)
        for (uint32_t i = 0; i < nblk1; i++) grid[(y + j) * 32 + x + i] = per;
      if (!nz) continue;
      const uint32_t base = block_ctx * kZeroDensityCtx + kNonZeroBuckets * kNbBlockCtx;
      uint32_t prev = nz <= size / 16 ? 1 : 0, rem = nz;
      if (run) {
        for (uint32_t k = nblocks; k <= last; k++) {
          const int32_t q = run[k - nblocks];
          if (q != 0 && rem == 0) {  // more non-zeros than the run declares
            ok = false;
            break;
          }
          const uint32_t ctx =
              base + (kCoeffNumNonzeroCtx[(rem + nblocks - 1) >> lb] + kCoeffFreqCtx[k >> lb]) * 2 + prev;
          emit(ctx, pack_signed(q));
          prev = q != 0;
          rem -= prev;
        }
        if (rem) ok = false;
        continue;
      }
      for (uint32_t k = nblocks; k <= last; k++) {
        const int32_t q = a[order[k]];
        const uint32_t ctx = base + (kCoeffNumNonzeroCtx[(rem + nblocks - 1) >> lb] + kCoeffFreqCtx[k >> lb]) * 2 + prev;
        emit(ctx, pack_signed(q));
        prev = q != 0;
        rem -= prev;
      }
    }
  }
  tape->n = (size_t)(out - tape->e.get());
  return ok;
}

int32_t validate(const jpxlg_frame_desc* f, const jpxlg_varblock* vb, uint32_t n, Frame* fr) {
  if (!f || !vb || !fr->arena || !fr->lf) return JPXLG_E_INVALID;
  if (f->width == 0 || f->height == 0 || f->width > (1u << 30) || f->height > (1u << 30)) return JPXLG_E_RANGE;
  if (f->bits_per_sample < 1 || f->bits_per_sample > 16) return JPXLG_E_RANGE;
  if (f->global_scale < 1 || f->global_scale > 65536 + 8192 || f->quant_lf < 1 || f->quant_lf > 65536)
    return JPXLG_E_RANGE;
  if (f->x_qm_scale > 7 || f->b_qm_scale > 7 || f->epf_iters > 3 || f->sharpness > 7 || f->gaborish > 1)
    return JPXLG_E_RANGE;
  if (f->x_factor_lf < -128 || f->x_factor_lf > 127 || f->b_factor_lf < -128 || f->b_factor_lf > 127)
    return JPXLG_E_RANGE;
  const Geometry& g = fr->g;
  const uint64_t arena_len = 192ull * g.bw * g.bh;
  std::vector<uint8_t> covered((size_t)g.bw * g.bh, 0);
  fr->group_vbs.assign(g.num_groups(), {});
  fr->lf_group_vbs.assign(g.num_lf_groups(), {});
  for (uint32_t i = 0; i < n; i++) {
    const jpxlg_varblock& v = vb[i];
    if (v.side != 8 && v.side != 16 && v.side != 32) return JPXLG_E_INVALID;
    const uint32_t s = v.side / 8;
    if (v.bx % s || v.by % s) return JPXLG_E_INVALID;
    if (v.hf_mul < 1 || v.hf_mul > (1u << 20)) return JPXLG_E_RANGE;
    if (fr->runs) {
      const uint32_t hf = v.side * v.side - s * s;
      for (int c = 0; c < 3; c++) {
        const jpxlg_run& r = fr->runs[3 * (size_t)i + c];
        if (r.count > hf || r.nonzeros > r.count || (r.count == 0) != (r.nonzeros == 0)) return JPXLG_E_INVALID;
        if (r.start > fr->runs_len || r.count > fr->runs_len - r.start) return JPXLG_E_INVALID;
      }
    } else if (v.start + 3ull * v.side * v.side > arena_len) {
      return JPXLG_E_INVALID;
    }
    const uint32_t lgid = (v.by / 256) * g.lgx + v.bx / 256;
    if (v.lf_group != lgid) return JPXLG_E_INVALID;
    for (uint32_t j = 0; j < s; j++)
      for (uint32_t k = 0; k < s; k++) {
        const uint32_t X = v.bx + k, Y = v.by + j;
        if (X >= g.bw || Y >= g.bh) return JPXLG_E_INVALID;  // varblocks stay inside the block grid
        if (covered[(size_t)Y * g.bw + X]++) return JPXLG_E_INVALID;
      }
    fr->group_vbs[(v.by / 32) * g.gx + v.bx / 32].push_back(i);
    fr->lf_group_vbs[lgid].push_back(i);
  }
  for (uint8_t c : covered)
    if (!c) return JPXLG_E_INVALID;
  // Each group's varblocks in block-raster order of their top-left blocks
  // (the IR's order already is; sort only when it is not).
  auto raster = [&](std::vector<uint32_t>& l) {
    auto before = [&](uint32_t a, uint32_t b) { return vb[a].by != vb[b].by ? vb[a].by < vb[b].by : vb[a].bx < vb[b].bx; };
    if (!std::is_sorted(l.begin(), l.end(), before)) std::stable_sort(l.begin(), l.end(), before);
  };
  for (auto& l : fr->group_vbs) raster(l);
  for (auto& l : fr->lf_group_vbs) raster(l);
  fr->lf_offset.assign(g.num_lf_groups(), 0);
  size_t off = 0;
  for (uint32_t id = 0; id < g.num_lf_groups(); id++) {
    const uint32_t lx = id % g.lgx, ly = id / g.lgx;
    fr->lf_offset[id] = off;
    off += 3ull * std::min(256u, g.bw - lx * 256) * std::min(256u, g.bh - ly * 256);
  }
  return JPXLG_OK;
}

}  // namespace

namespace {

int32_t emit_frame(Frame& fr, int32_t threads, uint8_t* out, size_t capacity, size_t* out_len,
                   jpxlg_emit_sizes* sizes) {
  const jpxlg_frame_desc* f = fr.f;
  if (int32_t rc = validate(f, fr.vb, fr.nvb, &fr)) return rc;
  const Geometry& g = fr.g;
  const uint32_t ng = g.num_groups(), nlf = g.num_lf_groups();
  std::lock_guard<std::mutex> lk(g_pool_mu);
  if (threads <= 0) threads = (int32_t)std::max(1u, std::thread::hardware_concurrency());
  Pool* pool = pool_for(std::max(1, std::min(threads, 64)));
  const int nw = pool->size();

  // Phase 1: LF-group parts (independent of the HF tables) and the pass
  // groups' tokens with per-worker counts, all at once.
  std::vector<Tape> tapes(ng);
  std::vector<HfCounts>& counts = worker_counts(nw);
  std::vector<BitWriter> hf_meta(nlf);
  std::vector<std::vector<PreparedChannel>> lf_prep(nlf, std::vector<PreparedChannel>(3));
  std::atomic<bool> runs_ok{true};
  pool->run((int)(4 * nlf + ng), [&](int t, int worker) {
    if (t < (int)(4 * nlf)) {
      if (t % 4 == 3) write_hf_metadata(hf_meta[t / 4], fr, t / 4);
      else prepare_lf_plane(fr, t / 4, t % 4, &lf_prep[t / 4][t % 4]);
      return;
    }
    if (!tokenize_group(fr, t - 4 * nlf, &tapes[t - 4 * nlf], &counts[worker])) runs_ok = false;
  });
  // Phase 2: the frame's HF histograms (over the contexts with tokens),
  // clusters and codes; the workers' tables are cleared as they are read.
  uint64_t any[kTouchedWords] = {0};
  for (int k = 0; k < nw; k++)
    for (int i = 0; i < kTouchedWords; i++) any[i] |= counts[k].touched[i];
  std::vector<uint32_t> used;
  std::vector<int32_t> row_of(kHfContexts, -1);
  for (int i = 0; i < kTouchedWords; i++)
    for (uint64_t b = any[i]; b; b &= b - 1) {
      const uint32_t ctx = (uint32_t)(i * 64 + __builtin_ctzll(b));
      row_of[ctx] = (int32_t)used.size();
      used.push_back(ctx);
    }
  std::vector<uint32_t> hist(used.size() * kAlphabet, 0u);
  const int chunks = 32, nu = (int)used.size();
  pool->run(chunks, [&](int ci, int) {
    for (int u = nu * ci / chunks; u < nu * (ci + 1) / chunks; u++) {
      const uint32_t ctx = used[u];
      uint32_t* d = &hist[(size_t)u * kAlphabet];
      const size_t slot = dense_slot(ctx);
      const uint32_t wdt = dense_width(ctx);
      for (int k = 0; k < nw; k++) {
        if (!(counts[k].touched[ctx >> 6] >> (ctx & 63) & 1)) continue;
        uint32_t* c = &counts[k].dense[slot];
        for (uint32_t t = 0; t < wdt; t++) d[t] += c[t];
        std::fill(c, c + wdt, 0u);
      }
    }
  });
  for (auto& c : counts) {
    for (uint32_t e : c.escape) hist[(size_t)row_of[e >> 8] * kAlphabet + (e & 0xFF)]++;
    c.escape.clear();
    std::fill(c.touched.begin(), c.touched.end(), 0ull);
  }
  if (!runs_ok) return JPXLG_E_INVALID;  // the workers' tables are clear again
  EntropyCode hf;
  hf.build(hist, kHfContexts, Hybrid{4, 2, 0}, 64, pool, &used);

  // Phase 3: sections in TOC order: LfGlobal, LF groups, HfGlobal, pass groups.
  const bool single = ng == 1;
  const uint32_t nsec = single ? 1 : 2 + nlf + ng;
  std::vector<BitWriter> sec(single ? 1 : nsec);
  auto hf_global = [&](BitWriter& w) {
    w.put_bool(true);  // default dequantization matrices
    w.put(ceil_log2(ng), 0);  // num_hf_presets - 1
    put_u32(w, kUsedOrders, 0);  // natural coefficient orders
    hf.write_header(w);
  };
  auto pass_group = [&](BitWriter& w, uint32_t gi) {
    // hfp takes ceil(log2(1)) = 0 bits; no modular group data follows
    const Tape& tp = tapes[gi];
    hf.put_tape(w, tp.e.get(), tp.n);
  };
  if (single) {
    BitWriter& w = sec[0];
    write_lf_global(w, fr);
    w.put(2, 0);  // LF extra precision
    write_modular_prepared(w, lf_prep[0]);
    w.append(hf_meta[0]);
    hf_global(w);
    pass_group(w, 0);
    w.pad_to_byte();
  } else {
    pool->run((int)nsec, [&](int s, int) {
      BitWriter& w = sec[s];
      if (s == 0) {
        write_lf_global(w, fr);
      } else if (s <= (int)nlf) {
        w.put(2, 0);  // LF extra precision
        write_modular_prepared(w, lf_prep[s - 1]);
        w.append(hf_meta[s - 1]);
      } else if (s == (int)nlf + 1) {
        hf_global(w);
      } else {
        w.reserve(tapes[s - nlf - 2].n * 2);
        pass_group(w, s - nlf - 2);
      }
      w.pad_to_byte();
    });
  }

  BitWriter head;
  write_image_headers(head, *f);
  const uint64_t image_headers = head.bits() / 8;
  write_frame_header(head, *f);
  const uint64_t frame_header_bits = head.bits() - image_headers * 8;
  const uint64_t toc0 = head.bits();
  head.put_bool(false);  // TOC not permuted
  head.pad_to_byte();
  for (auto& s : sec)
    if (!put_u32(head, kTocEntry, (uint32_t)s.size())) return JPXLG_E_RANGE;
  head.pad_to_byte();
  const uint64_t toc_bits = head.bits() - toc0;
  size_t total = head.size();
  for (auto& s : sec) total += s.size();
  if (sizes) {
    memset(sizes, 0, sizeof(*sizes));
    sizes->total = total;
    sizes->image_headers = image_headers;
    sizes->frame_header_bits = frame_header_bits;
    sizes->toc_bits = toc_bits;
    if (single) {
      sizes->pass_groups = sec[0].size();
    } else {
      sizes->lf_global = sec[0].size();
      for (uint32_t i = 0; i < nlf; i++) sizes->lf_groups += sec[1 + i].size();
      sizes->hf_global = sec[1 + nlf].size();
      for (uint32_t i = 0; i < ng; i++) sizes->pass_groups += sec[2 + nlf + i].size();
    }
  }
  if (out_len) *out_len = total;
  if (!out) return JPXLG_OK;
  if (capacity < total) return JPXLG_E_TOO_SMALL;
  size_t at = 0;
  memcpy(out, head.data(), head.size());
  at += head.size();
  for (auto& s : sec) {
    memcpy(out + at, s.data(), s.size());
    at += s.size();
  }
  return JPXLG_OK;
}

}  // namespace

extern "C" int32_t jpxlg_emit_codestream(const jpxlg_frame_desc* f, const int32_t* coeff_arena,
                                         const int32_t* lf_groups, const jpxlg_varblock* varblocks,
                                         uint32_t n_varblocks, const int8_t* cfl_x, const int8_t* cfl_b,
                                         int32_t threads, uint8_t* out, size_t capacity, size_t* out_len,
                                         jpxlg_emit_sizes* sizes) {
  if (!f) return JPXLG_E_INVALID;
  Frame fr;
  fr.f = f;
  fr.g = geometry_of(f->width, f->height);
  fr.arena = coeff_arena;
  fr.lf = lf_groups;
  fr.vb = varblocks;
  fr.nvb = n_varblocks;
  fr.cfl_x = cfl_x;
  fr.cfl_b = cfl_b;
  return emit_frame(fr, threads, out, capacity, out_len, sizes);
}

extern "C" int32_t jpxlg_emit_codestream_runs(const jpxlg_frame_desc* f, const int32_t* run_arena, size_t arena_len,
                                              const jpxlg_run* runs, const int32_t* lf_groups,
                                              const jpxlg_varblock* varblocks, uint32_t n_varblocks,
                                              const int8_t* cfl_x, const int8_t* cfl_b, int32_t threads,
                                              uint8_t* out, size_t capacity, size_t* out_len,
                                              jpxlg_emit_sizes* sizes) {
  if (!f || !runs || (!run_arena && arena_len)) return JPXLG_E_INVALID;
  static const int32_t kNone = 0;
  Frame fr;
  fr.f = f;
  fr.g = geometry_of(f->width, f->height);
  fr.arena = run_arena ? run_arena : &kNone;
  fr.runs = runs;
  fr.runs_len = arena_len;
  fr.lf = lf_groups;
  fr.vb = varblocks;
  fr.nvb = n_varblocks;
  fr.cfl_x = cfl_x;
  fr.cfl_b = cfl_b;
  return emit_frame(fr, threads, out, capacity, out_len, sizes);
}
