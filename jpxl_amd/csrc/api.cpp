// Host runtime behind include/jpxlg.h: contexts, resident images, the probe
// chain (quantize -> reconstruct -> score) and the quality controller.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/jpxlg.h"
#include "controller.h"
#include "internal.h"
#include "metric_host.h"
#include "tables.h"

using namespace jpxlg;

// The scratch of jpxlg_score_pair_noalloc: a file-scope device array, one
// per GPU, held until the process ends.  Host code reaches it only through
// the device address hipGetSymbolAddress gives at context creation
// (jpxlg_ctx::static_scratch); its name is a host-side stand-in.  Every
// call on one GPU shares it.
static __device__ __align__(256) char scratch[JPXLG_NOALLOC_SCRATCH_BYTES];

struct DepthSet {
  float* thr = nullptr;
  float* lut = nullptr;
  int levels = 0;
};

struct ProfRec {
  const char* name;
  hipEvent_t a, b;
};

// Released image arenas kept for reuse per context: up to half of the device
// memory (set at context creation; at least 16 GiB), so a batch or a served
// stream that cycles many images does not allocate and free them again.
constexpr size_t kSpareCapMin = (size_t)16 << 30;

struct jpxlg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // A second stream for work that can overlap the main chain (forked and
  // joined with events; both at the default priority).
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  bool prof = false;
  std::string prof_only;  // when non-empty, only launches of this name are timed
  // Self-check of the certified block-map search: every region through the
  // exact pass (JPXLG_COVER_ALL_EXACT=1 at context creation).
  int cover_all_exact = 0;
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, uint64_t>> totals;
  ConstTables* d_tab = nullptr;
  DctMats* d_mats = nullptr;
  float* d_lut8 = nullptr;
  // Order-statistic scratch of the first-guess features (one image at a
  // time on the context's stream).
  SelectScratch* d_select = nullptr;
  std::map<int, float*> d_lut16;
  std::map<int, DepthSet> depth;
  std::string last_error;
  // Image arenas released by jpxlg_image_destroy, kept for the next image
  // of a similar size (size -> device pointer); freed with the context.
  std::multimap<size_t, void*> spare;
  size_t spare_bytes = 0;
  size_t spare_cap = kSpareCapMin;
  // Ring of upload slots for the probe chain's descriptors (device buffer,
  // host staging, completion event).  A slot is claimed from the chain's
  // enqueue until its results are read back (a server lane's round can stay
  // claimed across many other chains); the ring hands out only unclaimed
  // slots, waits for a slot's last use before reusing it, and grows when
  // every slot is claimed.  Slots are never moved (chains hold pointers).
  struct Slot {
    char* dev = nullptr;
    char* host = nullptr;  // pinned staging of the same size
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
    bool claimed = false;
  };
  std::vector<std::unique_ptr<Slot>> slots;
  size_t next_slot = 0;
  // Pinned staging: a ring for small uploads (probe parameters of the
  // structure passes) and a block for structure readbacks.
  char* up_host = nullptr;
  size_t up_cap = 0, up_off = 0;
  char* rb_host = nullptr;
  size_t rb_cap = 0;
  // Pinned staging of the QuantizedFrameIr hand-off read back for the
  // codestream stage (arena, LF planes, varblock table), grown on demand.
  char* ir_host = nullptr;
  size_t ir_cap = 0;
  char* runs_host = nullptr;  // pinned staging of the runs' count and start tables
  size_t runs_cap = 0;
  // This device's address of the no-allocation scratch array (from
  // hipGetSymbolAddress at creation) and pinned space for its sums.
  char* static_scratch = nullptr;
  double* static_sums = nullptr;
  // Launch shapes of the streaming SSIMULACRA2 pass per size class (class =
  // log2 of a chain's total pixels - 16, clamped): units per workgroup and
  // the XCD swizzle, as picked by jpxlg_autotune (waves per workgroup fixed
  // at 4; swizzle default off).  No choice changes a result.
  struct Shape {
    int upw = 4, swizzle = 0;
  };
  Shape tune[12];
};

namespace {


// Arena sizes are rounded up to classes of 2^k x {1, 1.25, 1.5, 1.75} so a
// released arena fits the next image of a similar size (at most 25% slack).
size_t arena_class(size_t bytes) {
  size_t p = 1;
  while (p * 2 <= bytes) p *= 2;
  for (int q = 4; q <= 8; q++) {
    const size_t c = p / 4 * q;
    if (c >= bytes) return c;
  }
  return p * 2;
}

// A released arena of at least `bytes` (and at most twice that), or null.
void* take_spare(jpxlg_ctx* ctx, size_t bytes, size_t* got) {
  auto it = ctx->spare.lower_bound(bytes);
  if (it == ctx->spare.end() || it->first > 2 * bytes) return nullptr;
  void* p = it->second;
  *got = it->first;
  ctx->spare_bytes -= it->first;
  ctx->spare.erase(it);
  return p;
}

void release_arena(jpxlg_ctx* ctx, void* base, size_t bytes) {
  if (!base) return;
  while (ctx->spare_bytes + bytes > ctx->spare_cap && !ctx->spare.empty()) {
    auto it = ctx->spare.begin();
    (void)hipFree(it->second);
    ctx->spare_bytes -= it->first;
    ctx->spare.erase(it);
  }
  if (bytes > ctx->spare_cap) {
    (void)hipFree(base);
    return;
  }
  ctx->spare.emplace(bytes, base);
  ctx->spare_bytes += bytes;
}

// hipMalloc that, when the device is out of memory, frees the context's
// spare arenas (largest first) and tries again before giving up.  A failed
// attempt's error is cleared so later launch checks do not report it.
bool device_alloc(jpxlg_ctx* ctx, void** p, size_t bytes) {
  *p = nullptr;
  for (;;) {
    if (hipMalloc(p, bytes) == hipSuccess) return true;
    (void)hipGetLastError();
    *p = nullptr;
    if (ctx->spare.empty()) return false;
    auto it = std::prev(ctx->spare.end());
    (void)hipFree(it->second);
    ctx->spare_bytes -= it->first;
    ctx->spare.erase(it);
  }
}

}  // namespace

namespace {

constexpr int kMaxScales = 6;

hipEvent_t take_event(jpxlg_ctx* ctx) {
  if (!ctx->pool.empty()) {
    hipEvent_t e = ctx->pool.back();
    ctx->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Brackets one launch with events on the context stream when profiling.
struct Prof {
  jpxlg_ctx* ctx;
  const char* name;
  hipStream_t st;
  hipEvent_t a = nullptr;
  Prof(jpxlg_ctx* c, const char* n, hipStream_t s = nullptr) : ctx(c), name(n), st(s ? s : c->stream) {
    if (ctx->prof && (ctx->prof_only.empty() || ctx->prof_only == name)) {
      a = take_event(ctx);
      (void)hipEventRecord(a, st);
    }
  }
  ~Prof() {
    if (a) {
      hipEvent_t b = take_event(ctx);
      (void)hipEventRecord(b, st);
      ctx->recs.push_back({name, a, b});
    }
  }
};

int fail(jpxlg_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->last_error = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(ctx, JPXLG_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

const char* const kDctNames[3] = {"dct_fwd_8", "dct_fwd_16", "dct_fwd_32"};

// One device allocation carved into the image's buffers.
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
  void* take(size_t bytes) {
    size_t off = (used + 255) & ~(size_t)255;
    if (off + bytes > size) return nullptr;
    used = off + bytes;
    return base + off;
  }
};

void scale_dims(int w, int h, int* n, int sw[kMaxScales], int sh[kMaxScales]) {
  int s = 0;
  for (; s < kMaxScales; s++) {
    if (s > 0) {
      if (w < 8 || h < 8) break;
      w = (w + 1) / 2;
      h = (h + 1) / 2;
    }
    sw[s] = w;
    sh[s] = h;
  }
  *n = s;
}

size_t image_bytes(int w, int h) {
  const size_t px = (size_t)w * h;
  const size_t bw = (w + 7) / 8, bh = (h + 7) / 8;
  const size_t cpx = 64 * bw * bh;
  int n, sw[kMaxScales], sh[kMaxScales];
  scale_dims(w, h, &n, sw, sh);
  size_t pyr = 0, pyrT = 0;
  for (int s = 0; s < n; s++) {
    pyr += (size_t)sw[s] * sh[s];
    pyrT += transposed_floats(sw[s], sh[s]);
  }
  size_t b = 0;
  b += 6 * px;                       // rgb (up to 16-bit)
  b += 3 * px * 4;                   // source XYB
  b += 3 * pyr * 4 + 9 * pyrT * 4;  // reference pyramid: XYB', and transposed: XYB', mu1, s11
  b += 3 * cpx * 4 * 2;              // coefficients f32 + quantized i32
  b += 3 * bw * bh * 4 + bw * bh * 4 + bw * bh;  // LF planes, sigma, block map
  b += 3 * bw * bh * 4;              // work lists
  b += 2 * ((bw + 255) / 256) * ((bh + 255) / 256) * 256 * 16;  // list-building row counts and bases
  b += 6 * bw * bh * 8;              // cover candidate costs (f64) and their bounds
  b += (((bw + 3) / 4) * ((bh + 3) / 4) + 1) * 4;  // cover's exact-pass queue
  b += cpx * 18 + bw * bh * (16 + 3 * 16) + 2 * ((bw + 7) / 8 + 1) * ((bh + 7) / 8 + 1) * 4;  // CfL samples
  b += 6 * px * 4;                   // render planes + ping-pong
  b += 3 * pyr * 4 + 12 * pyrT * 4; // candidate XYB' pyramid, and transposed: XYB', three horizontal blurs
  size_t units = 0;
  for (int s = 0; s < n; s++) units += (size_t)pool_units(sw[s], sh[s]);
  b += units * 18 * 8 + kMaxScales * 18 * 8 + 4096;
  b += 3 * kMaxScales * sizeof(LineSet) + 2 * (3 * kMaxScales + 1) * 4;  // the reference's line sets
  b += 2 * ((bw + 7) / 8) * ((bh + 7) / 8) + 4096;
  b += bw * bh * 4 + 256;             // AQ field + scratch
  b += bw * bh * 4;                   // a caller-supplied plan's per-varblock HfMul
  b += 64 + 2 * ((bw + 7) / 8) * ((bh + 7) / 8) + bw * bh + 256;  // structure readback block
  b += (size_t)preanalysis_cells(w, h) * sizeof(CellStats);  // first-guess features: cell sums,
  b += (bw * bh / 64 + 2) * sizeof(SummaryPartial) + kSummaryBins * 4;  // DCT8 summary partials
  b += 512 * 256;  // alignment of every carved buffer
  return b;
}

}  // namespace

struct jpxlg_img {
  jpxlg_ctx* ctx = nullptr;
  Geom g{};
  int bits = 8;
  int gray = 0;
  int nscales = 0;
  int sw[kMaxScales], sh[kMaxScales];
  Arena arena;
  void* d_rgb = nullptr;
  float* xyb[3];
  float* ref_x[kMaxScales][3];
  float* ref_xT[kMaxScales][3];   // the metric reference planes transposed (column x at x * h)
  float* ref_mu[kMaxScales][3];   // mu1 and s11 of every scale, transposed
  float* ref_s11[kMaxScales][3];
  float* coeff[3];
  int32_t* quant[3];
  int32_t* lf[3];
  float* sigma = nullptr;
  uint8_t* d_bmap = nullptr;
  uint32_t* d_list[3];
  int list_count[3] = {0, 0, 0};
  int8_t* d_cfl_x = nullptr;
  int8_t* d_cfl_b = nullptr;
  CflParams* d_cflp = nullptr;
  ProbeParams* d_pp = nullptr;
  uint32_t* d_err = nullptr;
  int* d_gray = nullptr;
  float* rx[3];
  float* aux[3];
  float* cand_x[kMaxScales][3];   // [0] also the reference's linear RGB while the image is built
  float* cand_xT[kMaxScales][3];  // the candidate planes transposed
  float* cand_h[kMaxScales][3][3];  // the horizontal blurs of the moments, transposed
  LineSet* d_ref_sets = nullptr;  // the reference's blur line sets, then their unit prefixes (device)
  double* d_partials = nullptr;
  CflLfScratch* d_lfc = nullptr;
  double* d_sums = nullptr;
  double* d_cost[3] = {nullptr, nullptr, nullptr};
  double* d_rad[3] = {nullptr, nullptr, nullptr};  // bounds of the certified cover costs
  uint32_t* d_cover_queue = nullptr;               // regions the exact cover pass redoes
  float4* d_lf_s = nullptr;
  float4* d_hf_s = nullptr;
  uint16_t* d_hf_cell = nullptr;
  StructInfo* d_si = nullptr;        // counts of the structure built on the device
  uint32_t* d_lf_off = nullptr;      // per block: the varblock's first CfL LF sample
  uint32_t* d_hf_local = nullptr;    // per block: the varblock's first HF sample inside its tile
  uint32_t* d_tile_count = nullptr;  // HF samples per CfL tile
  char* d_readback = nullptr;        // the structure's gathered readback block
  bool gray_known = false;           // img->gray read back from the device
  uint32_t* d_tile_base = nullptr;
  uint4* d_rows = nullptr;           // list-building scratch: per-row counts and bases
  // Adaptive quantization: the per-atom field (bw x bh), its scratch, and
  // whether it is active (a mode other than off whose field is not neutral).
  float* aq_adj = nullptr;
  AqScratch* d_aq = nullptr;
  // A caller-supplied plan (jpxlg_set_plan): each varblock's HfMul at its
  // top-left block, on the device and the host, used while explicit_muls.
  uint32_t* d_vb_mul = nullptr;
  std::vector<uint32_t> vb_mul;
  bool explicit_muls = false;
  int aq_mode = kAqOff;
  bool aq_active = false;
  bool preconditioned = false;
  // Frame features of the first-guess predictors (quality_features.rs,
  // preanalysis.rs, the DCT8 summary of lib.rs:2769), computed on first use
  // from the source planes and samples.
  bool features_known = false;
  jpxlg_source_features sf{};
  jpxlg_transform_summary ts{};
  CellStats* d_cells = nullptr;
  SummaryPartial* d_sum_parts = nullptr;
  uint32_t* d_hist = nullptr;
  // Bumped whenever the varblock lists, the DCT form or the source planes
  // change: a captured chain is valid only for the generation it saw.
  uint64_t generation = 0;
  // host mirrors of the structure
  std::vector<uint8_t> bmap;
  std::vector<int8_t> cfl_x, cfl_b;
  int32_t lf_factors[2] = {0, 0};
  bool has_structure = false;
  bool has_quant = false;
  bool has_render = false;
  jpxlg_quantizer last_q{};
  int last_sharpness = 7;
  jpxlg_restoration last_r{1, 0, 7};  // the restoration of the last render (the frame header signals it)
  int dct_form = kDctExact;  // DctForm of the structure's forward DCT and the render's IDCT
  // The arena goes back to the context (or is freed) when the image dies.
  ~jpxlg_img() {
    if (arena.base) (void)hipStreamSynchronize(ctx->stream);
    release_arena(ctx, arena.base, arena.size);
  }
};

// ----------------------------------------------------------------------------
// Context
// ----------------------------------------------------------------------------

extern "C" int32_t jpxlg_ctx_create(int32_t device, jpxlg_ctx** out) {
  if (!out) return JPXLG_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return JPXLG_E_DEVICE;
  std::unique_ptr<jpxlg_ctx> ctx(new jpxlg_ctx());
  {
    const char* e = getenv("JPXLG_COVER_ALL_EXACT");
    ctx->cover_all_exact = e && e[0] == '1';
  }
  ctx->device = device;
  HIP_TRY(ctx.get(), hipSetDevice(device));
  {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) ctx->spare_cap = std::max(kSpareCapMin, total_b / 2);
  }
  HIP_TRY(ctx.get(), hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HIP_TRY(ctx.get(), hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
  HIP_TRY(ctx.get(), hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
  HIP_TRY(ctx.get(), hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
  ConstTables host;
  build_const_tables(&host);
  HIP_TRY(ctx.get(), hipMalloc(&ctx->d_tab, sizeof(ConstTables)));
  HIP_TRY(ctx.get(), hipMemcpy(ctx->d_tab, &host, sizeof(ConstTables), hipMemcpyHostToDevice));
  {
    DctMats mats;
    build_dct_mats(&mats);
    HIP_TRY(ctx.get(), hipMalloc(&ctx->d_mats, sizeof(DctMats)));
    HIP_TRY(ctx.get(), hipMemcpy(ctx->d_mats, &mats, sizeof(DctMats), hipMemcpyHostToDevice));
  }
  float lut8[256];
  srgb8_lut(lut8);
  HIP_TRY(ctx.get(), hipMalloc(&ctx->d_lut8, sizeof(lut8)));
  HIP_TRY(ctx.get(), hipMemcpy(ctx->d_lut8, lut8, sizeof(lut8), hipMemcpyHostToDevice));
  HIP_TRY(ctx.get(), hipMalloc(&ctx->d_select, sizeof(SelectScratch)));
  for (int i = 0; i < 8; i++) ctx->slots.emplace_back(new jpxlg_ctx::Slot());
  {
    // The device address of the no-allocation scratch array on this device
    // (the array's name is a host-side stand-in, not a device address).
    void* p = nullptr;
    HIP_TRY(ctx.get(), hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ctx->static_scratch = (char*)p;
    HIP_TRY(ctx.get(), hipHostMalloc((void**)&ctx->static_sums, sizeof(double) * 18 * kMaxScales,
                                     hipHostMallocDefault));
  }
  ctx->up_cap = (size_t)256 << 10;
  HIP_TRY(ctx.get(), hipHostMalloc((void**)&ctx->up_host, ctx->up_cap, hipHostMallocDefault));
  *out = ctx.release();
  return JPXLG_OK;
}

extern "C" void jpxlg_ctx_destroy(jpxlg_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(ctx->d_tab);
  (void)hipFree(ctx->d_mats);
  (void)hipFree(ctx->d_lut8);
  (void)hipFree(ctx->d_select);
  for (auto& kv : ctx->d_lut16) (void)hipFree(kv.second);
  for (auto& kv : ctx->depth) {
    (void)hipFree(kv.second.thr);
    (void)hipFree(kv.second.lut);
  }
  for (auto& kv : ctx->spare) (void)hipFree(kv.second);
  for (auto& sl : ctx->slots) {
    if (sl->dev) (void)hipFree(sl->dev);
    if (sl->host) (void)hipHostFree(sl->host);
    if (sl->done) (void)hipEventDestroy(sl->done);
  }
  if (ctx->up_host) (void)hipHostFree(ctx->up_host);
  if (ctx->rb_host) (void)hipHostFree(ctx->rb_host);
  if (ctx->ir_host) (void)hipHostFree(ctx->ir_host);
  if (ctx->runs_host) (void)hipHostFree(ctx->runs_host);
  if (ctx->static_sums) (void)hipHostFree(ctx->static_sums);
  for (auto& r : ctx->recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : ctx->pool) (void)hipEventDestroy(e);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  if (ctx->fork) (void)hipEventDestroy(ctx->fork);
  if (ctx->join) (void)hipEventDestroy(ctx->join);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char* jpxlg_last_error(const jpxlg_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : "no context";
}

extern "C" uint64_t jpxlg_image_footprint(uint32_t width, uint32_t height) {
  return (uint64_t)arena_class(image_bytes((int)width, (int)height));
}

static int get_lut16(jpxlg_ctx* ctx, int bits, float** out) {
  auto it = ctx->d_lut16.find(bits);
  if (it != ctx->d_lut16.end()) { *out = it->second; return JPXLG_OK; }
  std::vector<float> lut(65536);
  srgb16_lut(bits, lut.data());
  float* d = nullptr;
  HIP_TRY(ctx, hipMalloc(&d, 65536 * sizeof(float)));
  HIP_TRY(ctx, hipMemcpy(d, lut.data(), 65536 * sizeof(float), hipMemcpyHostToDevice));
  ctx->d_lut16[bits] = d;
  *out = d;
  return JPXLG_OK;
}

static int get_depth(jpxlg_ctx* ctx, int bits, DepthSet* out) {
  auto it = ctx->depth.find(bits);
  if (it != ctx->depth.end()) { *out = it->second; return JPXLG_OK; }
  std::vector<float> thr, lut;
  DepthSet d;
  d.levels = depth_tables(bits, &thr, &lut);
  HIP_TRY(ctx, hipMalloc(&d.thr, thr.size() * sizeof(float) + 4));
  HIP_TRY(ctx, hipMalloc(&d.lut, lut.size() * sizeof(float)));
  HIP_TRY(ctx, hipMemcpy(d.thr, thr.data(), thr.size() * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(d.lut, lut.data(), lut.size() * sizeof(float), hipMemcpyHostToDevice));
  ctx->depth[bits] = d;
  *out = d;
  return JPXLG_OK;
}

// ----------------------------------------------------------------------------
// Images
// ----------------------------------------------------------------------------

static int upload_small(jpxlg_ctx* ctx, hipStream_t st, void* dev, const void* src, size_t n);

static int image_create(jpxlg_ctx* ctx, uint32_t width, uint32_t height, uint32_t bits,
                        const void* rgb, bool device_src, jpxlg_img** out) {
  if (!ctx || !out || !rgb) return fail(ctx, JPXLG_E_INVALID, "null argument");
  *out = nullptr;
  if (width == 0 || height == 0) return fail(ctx, JPXLG_E_INVALID, "a zero frame dimension");
  if (width < 8 || height < 8)
    return fail(ctx, JPXLG_E_TOO_SMALL, "image is smaller than the metric's 8x8 floor");
  if (width > 65535 * 8 || height > 65535 * 8) return fail(ctx, JPXLG_E_UNSUPPORTED, "frame too large");
  if (bits < 8 || bits > 16) return fail(ctx, JPXLG_E_UNSUPPORTED, "bit depth outside 8..16");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<jpxlg_img> img(new jpxlg_img());
  img->ctx = ctx;
  img->bits = (int)bits;
  Geom& g = img->g;
  g.w = (int)width;
  g.h = (int)height;
  g.bw = (g.w + 7) / 8;
  g.bh = (g.h + 7) / 8;
  g.cstride = 8 * g.bw;
  g.crows = 8 * g.bh;
  g.tw = (g.bw + 7) / 8;
  g.th = (g.bh + 7) / 8;
  scale_dims(g.w, g.h, &img->nscales, img->sw, img->sh);
  size_t bytes = arena_class(image_bytes(g.w, g.h));
  void* base = take_spare(ctx, bytes, &bytes);
  if (!base && !device_alloc(ctx, &base, bytes)) return fail(ctx, JPXLG_E_OOM, "device allocation failed");
  img->arena.base = (char*)base;
  img->arena.size = bytes;
  Arena& A = img->arena;
  const size_t px = (size_t)g.w * g.h, bpx = (size_t)g.bw * g.bh, cpx = (size_t)g.cstride * g.crows;
  const int tiles_cfl = g.tw * g.th;
  const int sample_bytes = bits == 8 ? 1 : 2;
  img->d_rgb = A.take(px * 3 * sample_bytes);
  for (int c = 0; c < 3; c++) {
    img->xyb[c] = (float*)A.take(px * 4);
    img->coeff[c] = (float*)A.take(cpx * 4);
    img->quant[c] = (int32_t*)A.take(cpx * 4);
    img->lf[c] = (int32_t*)A.take(bpx * 4);
    img->d_list[c] = (uint32_t*)A.take(bpx * 4);
    img->rx[c] = (float*)A.take(px * 4);
    img->aux[c] = (float*)A.take(px * 4);
  }
  for (int s = 0; s < img->nscales; s++) {
    size_t sp = (size_t)img->sw[s] * img->sh[s];
    for (int c = 0; c < 3; c++) {
      const size_t tp = transposed_floats(img->sw[s], img->sh[s]);
      img->ref_x[s][c] = (float*)A.take(sp * 4);
      img->ref_xT[s][c] = (float*)A.take(tp * 4);
      img->ref_mu[s][c] = (float*)A.take(tp * 4);
      img->ref_s11[s][c] = (float*)A.take(tp * 4);
      img->cand_x[s][c] = (float*)A.take(sp * 4);
      img->cand_xT[s][c] = (float*)A.take(tp * 4);
      for (int m = 0; m < 3; m++) img->cand_h[s][c][m] = (float*)A.take(tp * 4);
    }
  }
  img->d_ref_sets = (LineSet*)A.take(3 * kMaxScales * sizeof(LineSet) + 2 * (3 * kMaxScales + 1) * 4);
  img->sigma = (float*)A.take(bpx * 4);
  for (int i = 0; i < 3; i++) img->d_cost[i] = (double*)A.take(bpx * 8);
  for (int i = 0; i < 3; i++) img->d_rad[i] = (double*)A.take(bpx * 8);
  img->d_cover_queue = (uint32_t*)A.take((size_t)(((g.bw + 3) / 4) * ((g.bh + 3) / 4) + 1) * 4);
  img->d_lf_s = (float4*)A.take(bpx * 16);
  img->d_hf_s = (float4*)A.take(cpx * 16);
  img->d_hf_cell = (uint16_t*)A.take(cpx * 2);
  img->d_si = (StructInfo*)A.take(sizeof(StructInfo));
  img->d_lf_off = (uint32_t*)A.take(bpx * 4);
  img->d_hf_local = (uint32_t*)A.take(bpx * 4);
  img->d_tile_count = (uint32_t*)A.take((size_t)tiles_cfl * 4);
  img->d_readback = (char*)A.take(structure_readback_bytes(tiles_cfl, (int)bpx));
  img->d_tile_base = (uint32_t*)A.take((size_t)(tiles_cfl + 1) * 4);
  img->d_rows = (uint4*)A.take((size_t)structure_rows(g) * 2 * sizeof(uint4));
  img->d_bmap = (uint8_t*)A.take(bpx);
  img->d_cfl_x = (int8_t*)A.take(tiles_cfl);
  img->d_cfl_b = (int8_t*)A.take(tiles_cfl);
  img->d_cflp = (CflParams*)A.take(sizeof(CflParams));
  img->d_lfc = (CflLfScratch*)A.take(sizeof(CflLfScratch));
  img->d_pp = (ProbeParams*)A.take(sizeof(ProbeParams));
  img->d_err = (uint32_t*)A.take(4);
  img->d_gray = (int*)A.take(4);
  img->aq_adj = (float*)A.take(bpx * 4);
  img->d_aq = (AqScratch*)A.take(sizeof(AqScratch));
  img->d_vb_mul = (uint32_t*)A.take(bpx * 4);
  const int ncells = preanalysis_cells(g.w, g.h);
  img->d_cells = ncells ? (CellStats*)A.take((size_t)ncells * sizeof(CellStats)) : nullptr;
  img->d_sum_parts = (SummaryPartial*)A.take((size_t)summary_groups(g) * sizeof(SummaryPartial));
  img->d_hist = (uint32_t*)A.take(kSummaryBins * 4);
  size_t punits = 0;
  for (int s = 0; s < img->nscales; s++) punits += (size_t)pool_units(img->sw[s], img->sh[s]);
  img->d_partials = (double*)A.take(punits * 18 * 8);
  img->d_sums = (double*)A.take(kMaxScales * 18 * 8);
  bool carved = img->d_cost[2] && img->d_rad[2] && img->d_cover_queue && img->d_lf_s && img->d_hf_s && img->d_hf_cell &&
                img->d_si && img->d_lf_off && img->d_hf_local && img->d_tile_count && img->d_readback && img->d_tile_base && img->d_rows && img->d_rgb && img->sigma && img->d_bmap && img->d_cfl_x && img->d_cfl_b &&
                img->d_cflp && img->d_lfc && img->d_pp && img->d_err && img->d_gray && img->d_partials && img->d_sums &&
                img->aq_adj && img->d_aq && img->d_vb_mul && img->d_ref_sets && img->d_sum_parts && img->d_hist &&
                (img->d_cells || ncells == 0);
  for (int c = 0; c < 3; c++) {
    carved = carved && img->xyb[c] && img->coeff[c] && img->quant[c] && img->lf[c] &&
             img->d_list[c] && img->rx[c] && img->aux[c];
    for (int s = 0; s < img->nscales; s++)
      carved = carved && img->ref_x[s][c] && img->ref_xT[s][c] && img->ref_mu[s][c] && img->ref_s11[s][c] &&
               img->cand_x[s][c] && img->cand_xT[s][c] && img->cand_h[s][c][0] && img->cand_h[s][c][1] &&
               img->cand_h[s][c][2];
  }
  if (!carved) return fail(ctx, JPXLG_E_OOM, "arena sizing");
  hipStream_t st = ctx->stream;
  HIP_TRY(ctx, hipMemcpyAsync(img->d_rgb, rgb, px * 3 * sample_bytes,
                              device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, hipMemsetD32Async((hipDeviceptr_t)img->d_gray, 1, 1, st));
  const float* lut = ctx->d_lut8;
  if (bits != 8) {
    float* l16 = nullptr;
    int rc = get_lut16(ctx, (int)bits, &l16);
    if (rc) return rc;
    lut = l16;
  }
  // Source XYB + linear RGB (scale 0 of the metric reference).
  float** lin0 = img->cand_x[0];
  { Prof p(ctx, "xyb_fwd"); launch_xyb_fwd(st, g, (int)bits, img->d_rgb, lut, img->xyb, lin0, img->d_gray); }
  // The reference pyramid (reference.rs:176-283): positive XYB of every
  // scale in one launch, then mu1 and s11.
  {
    Prof p(ctx, "ref_colour");
    launch_colour_pyr_ref(st, img->nscales, img->sw, img->sh, lin0, img->ref_x);
  }
  {
    // mu1 and s11 of every scale and channel: the horizontal pass leaves the
    // moments' row blurs (in the candidate's blur planes, free until the
    // first probe) and the transposed reference; the vertical pass finishes.
    LineSet sets[3 * kMaxScales];
    int off_b[3 * kMaxScales + 1], off_c[3 * kMaxScales + 1];
    int ns = 0, nb = 0, ncol = 0;
    for (int s = 0; s < img->nscales; s++)
      for (int c = 0; c < 3; c++, ns++) {
        LineSet& L = sets[ns];
        L = LineSet{};
        L.w = img->sw[s];
        L.h = img->sh[s];
        L.pitch = line_pitch(L.h);
        L.x1 = img->ref_x[s][c];
        L.hT[0] = img->cand_h[s][c][0];
        L.hT[1] = img->cand_h[s][c][1];
        L.xT = img->ref_xT[s][c];
        L.out[0] = img->ref_mu[s][c];
        L.out[1] = img->ref_s11[s][c];
        off_b[ns] = nb;
        off_c[ns] = ncol;
        nb += blur_bands(L.w, L.h);
        ncol += blur_columns(L.w, L.h);
      }
    off_b[ns] = nb;
    off_c[ns] = ncol;
    // one upload: the sets, then the band prefixes, then the column prefixes
    static_assert(sizeof(LineSet) % 8 == 0, "line sets stay 8-byte aligned");
    char blob[sizeof(sets) + sizeof(off_b) + sizeof(off_c)];
    const size_t sb = sizeof(LineSet) * ns, ob = 4 * (size_t)(ns + 1);
    memcpy(blob, sets, sb);
    memcpy(blob + sb, off_b, ob);
    memcpy(blob + sb + ob, off_c, ob);
    int rc = upload_small(ctx, st, img->d_ref_sets, blob, sb + 2 * ob);
    if (rc) return rc;
    const int* d_ob = (const int*)((char*)img->d_ref_sets + sb);
    const int* d_oc = (const int*)((char*)img->d_ref_sets + sb + ob);
    Prof p(ctx, "ref_blur");
    launch_blur_h(st, 0, ctx->d_tab, img->d_ref_sets, d_ob, ns, nb);
    launch_blur_v(st, 0, ctx->d_tab, img->d_ref_sets, d_oc, ns, ncol, 0);
  }
  HIP_TRY(ctx, hipGetLastError());
  // The grayscale flag stays on the device (the CfL kernels read it); the
  // host reads it back only when asked (jpxlg_image_info).
  *out = img.release();
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_image_upload_srgb(jpxlg_ctx* ctx, uint32_t width, uint32_t height,
                                           uint32_t bits, const void* rgb, jpxlg_img** out) {
  return image_create(ctx, width, height, bits, rgb, false, out);
}

extern "C" int32_t jpxlg_image_from_device(jpxlg_ctx* ctx, uint32_t width, uint32_t height,
                                           uint32_t bits, const void* device_rgb, jpxlg_img** out) {
  return image_create(ctx, width, height, bits, device_rgb, true, out);
}

extern "C" void jpxlg_image_destroy(jpxlg_img* img) {
  if (!img) return;
  (void)hipSetDevice(img->ctx->device);
  (void)hipStreamSynchronize(img->ctx->stream);
  delete img;
}

extern "C" int32_t jpxlg_image_info(const jpxlg_img* img, uint32_t* bx, uint32_t* by,
                                    int32_t* gray, uint32_t* scales) {
  if (!img) return JPXLG_E_INVALID;
  if (bx) *bx = (uint32_t)img->g.bw;
  if (by) *by = (uint32_t)img->g.bh;
  if (gray) {
    if (!img->gray_known) {
      jpxlg_img* m = const_cast<jpxlg_img*>(img);
      HIP_TRY(m->ctx, hipSetDevice(m->ctx->device));
      HIP_TRY(m->ctx, hipMemcpyAsync(&m->gray, m->d_gray, 4, hipMemcpyDeviceToHost, m->ctx->stream));
      HIP_TRY(m->ctx, hipStreamSynchronize(m->ctx->stream));
      m->gray_known = true;
    }
    *gray = img->gray;
  }
  if (scales) *scales = (uint32_t)img->nscales;
  return JPXLG_OK;
}

// ----------------------------------------------------------------------------
// Structure
// ----------------------------------------------------------------------------

// Small host-to-device upload on st through the context's pinned ring (an
// asynchronous copy, unlike one from pageable memory).  The ring wraps after
// the stream has drained.
static int upload_small(jpxlg_ctx* ctx, hipStream_t st, void* dev, const void* src, size_t n) {
  const size_t need = (n + 255) & ~(size_t)255;
  if (need > ctx->up_cap) return fail(ctx, JPXLG_E_INVALID, "upload larger than the staging ring");
  if (ctx->up_off + need > ctx->up_cap) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->side));
    ctx->up_off = 0;
  }
  char* h = ctx->up_host + ctx->up_off;
  memcpy(h, src, n);
  ctx->up_off += need;
  HIP_TRY(ctx, hipMemcpyAsync(dev, h, n, hipMemcpyHostToDevice, st));
  return JPXLG_OK;
}

static int check_quantizer(jpxlg_ctx* ctx, const jpxlg_quantizer* q);
static void probe_params_for(const jpxlg_img* img, const jpxlg_quantizer& q, int sharpness, ProbeParams* p);

static int validate_map(const Geom& g, const uint8_t* bmap, std::string* why) {
  std::vector<uint8_t> covered((size_t)g.bw * g.bh, 0);
  for (int by = 0; by < g.bh; by++)
    for (int bx = 0; bx < g.bw; bx++) {
      int side = bmap[by * g.bw + bx];
      if (side == 0) continue;
      if (side != 8 && side != 16 && side != 32) { *why = "block map entry not 0/8/16/32"; return 0; }
      int n = side / 8;
      if (bx % n || by % n) { *why = "varblock not aligned to its size"; return 0; }
      if (bx + n > g.bw || by + n > g.bh) { *why = "varblock leaves the block grid"; return 0; }
      if ((bx / 8) != (bx + n - 1) / 8 || (by / 8) != (by + n - 1) / 8) { *why = "varblock crosses a 64x64 tile"; return 0; }
      for (int dy = 0; dy < n; dy++)
        for (int dx = 0; dx < n; dx++) {
          uint8_t& c = covered[(by + dy) * g.bw + bx + dx];
          if (c) { *why = "overlapping varblocks"; return 0; }
          c = 1;
          if ((dx || dy) && bmap[(by + dy) * g.bw + bx + dx] != 0) { *why = "covered block carries a size"; return 0; }
        }
    }
  for (uint8_t c : covered)
    if (!c) { *why = "block map leaves blocks uncovered"; return 0; }
  return 1;
}

// The block map in d_bmap (and, without a CfL estimate, the factors already
// on the device) becomes the image's structure: the work lists and CfL
// sample slots are built on the device, the forward DCT of every varblock
// follows, and with q the CfL estimate; one synchronisation at the end
// brings back the counts and the host mirrors (block map, factors).
static int finish_structure(jpxlg_img* img, const jpxlg_quantizer* q) {
  jpxlg_ctx* ctx = img->ctx;
  const Geom& g = img->g;
  hipStream_t st = ctx->stream;
  img->explicit_muls = false;  // a new structure drops a supplied plan
  const size_t bpx = (size_t)g.bw * g.bh, tiles = (size_t)g.tw * g.th;
  {
    Prof p(ctx, "structure_lists");
    launch_structure_lists(st, g, img->d_bmap, img->d_list, img->d_lf_off, img->d_hf_local, img->d_tile_count,
                           img->d_tile_base, img->d_si, img->d_rows);
  }
  for (int i = 0; i < 3; i++) {
    Prof p(ctx, kDctNames[i]);
    launch_dct_fwd(st, kSides[i], img->dct_form, g, img->d_list[i], 0, img->d_si, ctx->d_tab, ctx->d_mats,
                   img->xyb, img->coeff);
  }
  if (q) {
    ProbeParams pp;
    probe_params_for(img, *q, img->last_sharpness, &pp);
    int rc = upload_small(ctx, st, img->d_pp, &pp, sizeof(pp));
    if (rc) return rc;
    {
      Prof p(ctx, "cfl_collect");
      launch_cfl_collect(st, g, ctx->d_tab, img->d_pp, img->d_list, img->d_si, img->d_lf_off, img->d_hf_local,
                         img->d_tile_base, img->coeff, img->d_lf_s, img->d_hf_s, img->d_hf_cell, img->d_err);
    }
    // The frame-LF regression (one long sequential sum) runs on the side
    // stream beside the per-tile regressions.
    HIP_TRY(ctx, hipEventRecord(ctx->fork, st));
    HIP_TRY(ctx, hipStreamWaitEvent(ctx->side, ctx->fork, 0));
    {
      Prof p(ctx, "cfl_lf", ctx->side);
      launch_cfl_lf(ctx->side, img->d_pp, img->d_si, img->d_lf_s, img->d_lfc, img->d_cflp, img->d_gray);
    }
    HIP_TRY(ctx, hipEventRecord(ctx->join, ctx->side));
    {
      Prof p(ctx, "cfl_tiles");
      launch_cfl_tiles(st, g, ctx->d_tab, img->d_pp, img->d_tile_base, img->d_hf_s, img->d_hf_cell, img->d_cfl_x,
                       img->d_cfl_b, img->d_gray);
    }
    HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->join, 0));
  }
  launch_gather_structure(st, img->d_si, img->d_cflp, img->d_err, img->d_cfl_x, img->d_cfl_b, img->d_bmap,
                          (int)tiles, (int)bpx, img->d_readback);
  HIP_TRY(ctx, hipGetLastError());
  const size_t rb = structure_readback_bytes((int)tiles, (int)bpx);
  if (ctx->rb_cap < rb) {
    if (ctx->rb_host) (void)hipHostFree(ctx->rb_host);
    ctx->rb_host = nullptr;
    ctx->rb_cap = 0;
    if (hipHostMalloc((void**)&ctx->rb_host, rb, hipHostMallocDefault) != hipSuccess)
      return fail(ctx, JPXLG_E_OOM, "readback staging allocation failed");
    ctx->rb_cap = rb;
  }
  HIP_TRY(ctx, hipMemcpyAsync(ctx->rb_host, img->d_readback, rb, hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  const char* R = ctx->rb_host;
  StructInfo si;
  CflParams cp;
  uint32_t err = 0;
  memcpy(&si, R, sizeof(si));
  memcpy(&cp, R + 32, sizeof(cp));
  memcpy(&err, R + 48, 4);
  img->cfl_x.assign((const int8_t*)R + 64, (const int8_t*)R + 64 + tiles);
  img->cfl_b.assign((const int8_t*)R + 64 + tiles, (const int8_t*)R + 64 + 2 * tiles);
  img->bmap.assign((const uint8_t*)R + 64 + 2 * tiles, (const uint8_t*)R + 64 + 2 * tiles + bpx);
  for (int i = 0; i < 3; i++) img->list_count[i] = si.count[i];
  img->lf_factors[0] = cp.x_factor_lf;
  img->lf_factors[1] = cp.b_factor_lf;
  img->has_structure = !err;
  img->generation++;
  img->has_quant = false;
  img->has_render = false;
  if (err) return fail(ctx, JPXLG_E_RANGE, "a coefficient outside the quantizer's working range");
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_set_structure(jpxlg_img* img, const uint8_t* bmap, const int8_t* cfl_x,
                                       const int8_t* cfl_b, const int32_t* lf_factors) {
  if (!img || !bmap) return JPXLG_E_INVALID;
  const Geom& g = img->g;
  std::string why;
  if (!validate_map(g, bmap, &why)) return fail(img->ctx, JPXLG_E_INVALID, why);
  img->bmap.assign(bmap, bmap + (size_t)g.bw * g.bh);
  const size_t nt = (size_t)g.tw * g.th;
  img->cfl_x.assign(nt, 0);
  img->cfl_b.assign(nt, 0);
  if (cfl_x) img->cfl_x.assign(cfl_x, cfl_x + nt);
  if (cfl_b) img->cfl_b.assign(cfl_b, cfl_b + nt);
  img->lf_factors[0] = lf_factors ? lf_factors[0] : 0;
  img->lf_factors[1] = lf_factors ? lf_factors[1] : 0;
  if (img->lf_factors[0] < -128 || img->lf_factors[0] > 127 || img->lf_factors[1] < -128 ||
      img->lf_factors[1] > 127)
    return fail(img->ctx, JPXLG_E_RANGE, "LF CfL factor outside u8 after bias");
  jpxlg_ctx* ctx = img->ctx;
  hipStream_t st = ctx->stream;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMemcpyAsync(img->d_bmap, img->bmap.data(), img->bmap.size(), hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(img->d_cfl_x, img->cfl_x.data(), img->cfl_x.size(), hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(img->d_cfl_b, img->cfl_b.data(), img->cfl_b.size(), hipMemcpyHostToDevice, st));
  CflParams cp{img->lf_factors[0], img->lf_factors[1]};
  HIP_TRY(ctx, hipMemcpyAsync(img->d_cflp, &cp, sizeof(cp), hipMemcpyHostToDevice, st));
  const uint32_t zero = 0;
  HIP_TRY(ctx, hipMemcpyAsync(img->d_err, &zero, 4, hipMemcpyHostToDevice, st));
  return finish_structure(img, nullptr);
}

// PerceptualEvaluator::evaluate(&ValidatedPixelPlan) (quality.rs:87-157):
// a plan built elsewhere becomes the image's structure and levels, so the
// next render renders exactly it.
extern "C" int32_t jpxlg_set_plan(jpxlg_img* img, const jpxlg_quantizer* wire, const uint8_t* bmap,
                                  const int8_t* cfl_x, const int8_t* cfl_b, const int32_t* lf_factors,
                                  const uint32_t* hf_mul, const int32_t* quant, const int32_t* lf) {
  if (!img || !wire || !bmap || !quant || !lf) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  int rc = check_quantizer(ctx, wire);
  if (rc) return rc;
  const Geom& g = img->g;
  const size_t bpx = (size_t)g.bw * g.bh, cpx = (size_t)g.cstride * g.crows;
  for (size_t i = 0; i < 3 * cpx; i++)
    if (quant[i] > kMaxQuant || quant[i] < -kMaxQuant)
      return fail(ctx, JPXLG_E_RANGE, "a quantized coefficient outside the quantizer's working range");
  std::vector<uint32_t> muls(bpx, 0u);
  for (size_t i = 0; i < bpx; i++)
    if (bmap[i]) {
      muls[i] = hf_mul ? hf_mul[i] : wire->hf_mul;
      if (muls[i] < 1 || muls[i] > 0x7fffffffu) return fail(ctx, JPXLG_E_RANGE, "a varblock HfMul outside 1..2^31-1");
    }
  rc = jpxlg_set_structure(img, bmap, cfl_x, cfl_b, lf_factors);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  HIP_TRY(ctx, hipMemcpyAsync(img->d_vb_mul, muls.data(), bpx * 4, hipMemcpyHostToDevice, st));
  for (int c = 0; c < 3; c++) {
    HIP_TRY(ctx, hipMemcpyAsync(img->quant[c], quant + c * cpx, cpx * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipMemcpyAsync(img->lf[c], lf + c * bpx, bpx * 4, hipMemcpyHostToDevice, st));
  }
  HIP_TRY(ctx, hipStreamSynchronize(st));
  img->vb_mul = std::move(muls);
  img->explicit_muls = true;
  img->last_q = *wire;
  img->has_quant = true;
  img->has_render = false;
  img->generation++;
  return JPXLG_OK;
}

// prepare_gaborish_frame (jpxl-encode-policy/src/lib.rs:2256-2268): the
// source XYB planes replaced by their 12-step Jacobi Gaborish inverse, so a
// plan rendered with Gaborish restores them.  Any structure is dropped.
extern "C" int32_t jpxlg_precondition_gaborish(jpxlg_img* img) {
  if (!img) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t bytes = (size_t)img->g.w * img->g.h * 4;
  float* cur[3] = {img->rx[0], img->rx[1], img->rx[2]};
  float* nxt[3] = {img->aux[0], img->aux[1], img->aux[2]};
  for (int c = 0; c < 3; c++) HIP_TRY(ctx, hipMemcpyAsync(cur[c], img->xyb[c], bytes, hipMemcpyDeviceToDevice, st));
  {
    Prof p(ctx, "gaborish_sharpen");
    for (int it = 0; it < 12; it++) {
      launch_gaborish_jacobi(st, img->g.w, img->g.h, img->xyb, cur, nxt);
      std::swap(cur, nxt);
    }
  }
  HIP_TRY(ctx, hipGetLastError());
  // the result (in cur) becomes the source planes; the old planes become scratch
  for (int c = 0; c < 3; c++) {
    float* old = img->xyb[c];
    img->xyb[c] = cur[c];
    if (img->rx[c] == cur[c]) img->rx[c] = old;
    else img->aux[c] = old;
  }
  img->has_structure = false;
  img->has_quant = false;
  img->has_render = false;
  img->preconditioned = true;
  img->generation++;
  return JPXLG_OK;
}

// AqSetup::build's factorization (jpxl-encode-policy/src/lib.rs:1859-1908):
// global_scale / factor (snapped down to a multiple), quant_lf and the
// baseline HfMul times the factor (2, or 16 on the fine lattice); the
// request's values and refine-only when a factored value is not on the wire.
// Returns refine_only.
static int aq_wire(const jpxlg_quantizer& q, bool fine, jpxlg_quantizer* w) {
  const uint32_t factor = fine ? 16u : 2u;
  const uint32_t snapped = q.global_scale - q.global_scale % factor;
  const uint64_t g2 = std::max<uint64_t>(snapped / factor, 1);
  const uint64_t q2 = std::min<uint64_t>((uint64_t)q.quant_lf * factor, 0xffffffffu);
  const uint64_t m2 = std::min<uint64_t>((uint64_t)q.hf_mul * factor, 0xffffffffu);
  *w = q;
  const bool ok = snapped >= factor && g2 <= 8193u + 65535u && q2 >= 1 && q2 <= 65536u && m2 >= 1 &&
                  m2 <= 0x7fffffffu;
  if (!ok) return 1;
  w->global_scale = (uint32_t)g2;
  w->quant_lf = (uint32_t)q2;
  w->hf_mul = (uint32_t)m2;
  return 0;
}

// The probe parameters of quantizer q on this image: the wire quantizer and
// the per-varblock HfMul field when the image's AQ field is active.
static void probe_params_for(const jpxlg_img* img, const jpxlg_quantizer& q, int sharpness, ProbeParams* p) {
  if (img->explicit_muls) {
    // a caller-supplied plan: its wire quantizer and per-varblock HfMul
    build_probe_params(q, sharpness, p);
    p->vb_mul = img->d_vb_mul;
    return;
  }
  const bool fine = img->aq_mode >= kAqFineMasking;
  jpxlg_quantizer w = q;
  const int refine_only = img->aq_active ? aq_wire(q, fine, &w) : 0;
  build_probe_params(w, sharpness, p);
  if (img->aq_active) {
    p->aq_adj = img->aq_adj;
    p->aq_refine_only = refine_only;
    p->aq_fine = fine ? 1 : 0;
  }
}

extern "C" int32_t jpxlg_set_aq(jpxlg_img* img, int32_t mode, int32_t* active) {
  if (!img) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  if (mode < kAqOff || mode > kAqEdgeRefine) return fail(ctx, JPXLG_E_UNSUPPORTED, "unknown AQ mode");
  if (img->preconditioned)
    return fail(ctx, JPXLG_E_INVALID, "the AQ field reads the source planes: set it before the Gaborish precondition");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  img->aq_mode = mode;
  img->aq_active = false;
  if (mode != kAqOff) {
    hipStream_t st = ctx->stream;
    HIP_TRY(ctx, hipMemsetAsync(img->d_aq, 0, sizeof(AqScratch), st));
    {
      Prof p(ctx, "aq_field");
      launch_aq_field(st, img->g, mode, img->xyb, (float*)img->d_cost[0], (float*)img->d_cost[1], img->d_aq,
                      img->aq_adj);
    }
    HIP_TRY(ctx, hipGetLastError());
    AqScratch sc;
    HIP_TRY(ctx, hipMemcpyAsync(&sc, img->d_aq, sizeof(sc), hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    img->aq_active = sc.active != 0;
  }
  if (active) *active = img->aq_active ? 1 : 0;
  // the cover and the CfL estimate depend on the field
  img->generation++;
  img->has_structure = false;
  img->has_quant = false;
  img->has_render = false;
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_plan_aq(jpxlg_img* img, uint32_t* hf_mul, jpxlg_quantizer* wire, int32_t* refine_only) {
  if (!img) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  if (!img->has_quant) return fail(ctx, JPXLG_E_INVALID, "nothing quantized yet");
  if (img->explicit_muls) {
    if (wire) *wire = img->last_q;
    if (refine_only) *refine_only = 0;
    if (hf_mul) memcpy(hf_mul, img->vb_mul.data(), img->vb_mul.size() * 4);
    return JPXLG_OK;
  }
  const bool fine = img->aq_mode >= kAqFineMasking;
  jpxlg_quantizer w = img->last_q;
  const int ro = img->aq_active ? aq_wire(img->last_q, fine, &w) : 0;
  if (wire) *wire = w;
  if (refine_only) *refine_only = ro;
  if (!hf_mul) return JPXLG_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t bpx = (size_t)img->g.bw * img->g.bh;
  ProbeParams pp;
  probe_params_for(img, img->last_q, img->last_sharpness, &pp);
  uint32_t* d_muls = (uint32_t*)img->d_cost[2];
  HIP_TRY(ctx, hipMemcpyAsync(img->d_pp, &pp, sizeof(pp), hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, hipMemsetAsync(d_muls, 0, bpx * 4, st));
  launch_aq_block_muls(st, img->g, img->d_pp, img->d_list, img->list_count, d_muls);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipMemcpyAsync(hf_mul, d_muls, bpx * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_set_dct_form(jpxlg_img* img, int32_t form) {
  if (!img) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  if (form != JPXLG_DCT_EXACT && form != JPXLG_DCT_FP32 && form != JPXLG_DCT_BF16X3)
    return fail(ctx, JPXLG_E_UNSUPPORTED, "unknown DCT form");
  if (form == img->dct_form) return JPXLG_OK;
  img->generation++;
  img->dct_form = form;
  img->has_quant = false;
  img->has_render = false;
  if (!img->has_structure) return JPXLG_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  for (int i = 0; i < 3; i++) {
    if (!img->list_count[i]) continue;
    Prof p(ctx, kDctNames[i]);
    launch_dct_fwd(st, kSides[i], img->dct_form, img->g, img->d_list[i], img->list_count[i], nullptr, ctx->d_tab,
                   ctx->d_mats, img->xyb, img->coeff);
  }
  HIP_TRY(ctx, hipGetLastError());
  return JPXLG_OK;
}

// The chroma-from-luma estimate of build_pixel_plan (lib.rs:3437-3763) on
// the structure just uploaded: the host lays out the sample slots in the
// reference's order, the device gathers, regresses and prices.
extern "C" int32_t jpxlg_plan_structure(jpxlg_img* img, const jpxlg_quantizer* q, int32_t cover_mode,
                                        int32_t enable_cfl) {
  if (!img || !q) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  const Geom& g = img->g;
  if (cover_mode != JPXLG_COVER_FIXED_DCT8 && cover_mode != JPXLG_COVER_HIERARCHICAL &&
      cover_mode != JPXLG_COVER_HIERARCHICAL_LEGACY)
    return fail(ctx, JPXLG_E_UNSUPPORTED, "unknown cover mode");
  int rc = check_quantizer(ctx, q);
  if (rc) return rc;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t bpx = (size_t)g.bw * g.bh, tiles = (size_t)g.tw * g.th;
  HIP_TRY(ctx, hipMemsetAsync(img->d_err, 0, 4, st));
  if (cover_mode != JPXLG_COVER_FIXED_DCT8) {
    ProbeParams pp;
    probe_params_for(img, *q, img->last_sharpness, &pp);
    set_cover_objective(&pp, cover_mode == JPXLG_COVER_HIERARCHICAL_LEGACY);
    rc = upload_small(ctx, st, img->d_pp, &pp, sizeof(pp));
    if (rc) return rc;
    Prof p(ctx, "cover");
    HIP_TRY(ctx, hipMemsetAsync(img->d_cover_queue, 0, 4, st));
    launch_cover(st, g, ctx->d_tab, img->d_pp, img->xyb, img->d_cost[0], img->d_cost[1], img->d_cost[2],
                 img->d_rad[0], img->d_rad[1], img->d_rad[2], img->d_cover_queue, img->d_bmap, img->d_err,
                 ctx->cover_all_exact);
  } else {
    HIP_TRY(ctx, hipMemsetAsync(img->d_bmap, 8, bpx, st));  // every block its own DCT8
  }
  if (!enable_cfl) {
    HIP_TRY(ctx, hipMemsetAsync(img->d_cfl_x, 0, tiles, st));
    HIP_TRY(ctx, hipMemsetAsync(img->d_cfl_b, 0, tiles, st));
    HIP_TRY(ctx, hipMemsetAsync(img->d_cflp, 0, sizeof(CflParams), st));
  }
  return finish_structure(img, enable_cfl ? q : nullptr);
}

// ----------------------------------------------------------------------------
// Probe chain
// ----------------------------------------------------------------------------

static int check_quantizer(jpxlg_ctx* ctx, const jpxlg_quantizer* q) {
  if (q->global_scale < 1 || q->global_scale > 73728) return fail(ctx, JPXLG_E_RANGE, "global_scale outside 1..73728");
  if (q->hf_mul < 1) return fail(ctx, JPXLG_E_RANGE, "hf_mul below 1");
  if (q->quant_lf < 1) return fail(ctx, JPXLG_E_RANGE, "quant_lf below 1");
  if (q->x_qm_scale > 7 || q->b_qm_scale > 7) return fail(ctx, JPXLG_E_RANGE, "qm_scale outside u(3)");
  return JPXLG_OK;
}

// ----------------------------------------------------------------------------
// The probe chain over a batch of images (one image is a batch of one).
// ----------------------------------------------------------------------------

namespace {

// Per-slot layout: descs, probe params, error words, workgroup offsets, and
// the 18-per-scale metric sums of every image.
struct ChainLayout {
  size_t desc = 0, pp = 0, sets = 0, err = 0, off = 0, sums = 0, total = 0;
  int n = 0;
  // offset arrays (each n+1 ints, except the blur passes': 18n+1, one entry
  // per line set)
  int q_off[3], r_off[3], g_off, e_off, c_off[kDescScales], b_off, v_off;
  int noffs = 0;
};

ChainLayout chain_layout(int n) {
  ChainLayout L;
  L.n = n;
  auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
  size_t o = 0;
  L.desc = o; o = al(o + sizeof(ProbeDesc) * n);
  L.pp = o; o = al(o + sizeof(ProbeParams) * n);
  L.sets = o; o = al(o + sizeof(LineSet) * 3 * kDescScales * n);
  L.err = o; o = al(o + 4 * n);
  int k = 0;
  for (int i = 0; i < 3; i++) { L.q_off[i] = k; k += n + 1; }
  for (int i = 0; i < 3; i++) { L.r_off[i] = k; k += n + 1; }
  L.g_off = k; k += n + 1;
  L.e_off = k; k += n + 1;
  for (int i = 0; i < kDescScales; i++) { L.c_off[i] = k; k += n + 1; }
  L.b_off = k; k += 3 * kDescScales * n + 1;
  L.v_off = k; k += 3 * kDescScales * n + 1;
  L.noffs = k;
  L.off = o; o = al(o + 4 * (size_t)k);
  L.sums = o; o = al(o + sizeof(double) * 18 * kDescScales * n);
  L.total = o;
  return L;
}

}  // namespace

// Takes the next unclaimed upload slot with room for `bytes` (waiting for
// its last use) and claims it; read_chain releases it.  With every slot
// claimed the ring grows by one.
static int take_slot(jpxlg_ctx* ctx, size_t bytes, jpxlg_ctx::Slot** out) {
  const size_t n = ctx->slots.size();
  jpxlg_ctx::Slot* pick = nullptr;
  for (size_t i = 0; i < n && !pick; i++) {
    jpxlg_ctx::Slot* c = ctx->slots[(ctx->next_slot + i) % n].get();
    if (!c->claimed) {
      pick = c;
      ctx->next_slot = (ctx->next_slot + i + 1) % n;
    }
  }
  if (!pick) {
    ctx->slots.emplace_back(new jpxlg_ctx::Slot());
    pick = ctx->slots.back().get();
    ctx->next_slot = 0;
  }
  jpxlg_ctx::Slot& sl = *pick;
  if (sl.pending) HIP_TRY(ctx, hipEventSynchronize(sl.done));
  sl.pending = false;
  if (!sl.done) HIP_TRY(ctx, hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  if (sl.cap < bytes) {
    if (sl.dev) (void)hipFree(sl.dev);
    if (sl.host) (void)hipHostFree(sl.host);
    sl.dev = sl.host = nullptr;
    sl.cap = 0;
    const size_t cap = std::max(bytes, (size_t)64 << 10);
    if (!device_alloc(ctx, (void**)&sl.dev, cap)) return fail(ctx, JPXLG_E_OOM, "chain slot allocation failed");
    if (hipHostMalloc(&sl.host, cap, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return fail(ctx, JPXLG_E_OOM, "chain slot staging allocation failed");
    }
    sl.cap = cap;
  }
  sl.claimed = true;
  *out = &sl;
  return JPXLG_OK;
}

struct ChainRun {
  jpxlg_ctx::Slot* slot = nullptr;  // null for a graph's own block (already complete)
  const char* host = nullptr;
  ChainLayout L;
};

// Workgroup totals per stage of a filled chain.
struct ChainTotals {
  int q[3] = {0, 0, 0}, r[3] = {0, 0, 0}, g = 0, e = 0, c[kDescScales] = {0, 0, 0, 0, 0, 0};
  int nsets = 0, bands = 0, cols = 0;  // blur line sets, horizontal bands, vertical columns
  int upw = 4, swizzle = 0;            // vertical-pass launch shape
};

static int tune_class(double pixels) {
  int c = (int)floor(log2(std::max(pixels, 1.0))) - 16;
  return std::min(std::max(c, 0), 11);
}

static void epf_steps_of(const jpxlg_restoration* r, int* nsteps, int steps[3]) {
  *nsteps = 0;
  steps[0] = steps[1] = steps[2] = 0;
  if (!r) return;
  if (r->epf_iters == 1) { steps[0] = 1; *nsteps = 1; }
  else if (r->epf_iters == 2) { steps[0] = 1; steps[1] = 2; *nsteps = 2; }
  else if (r->epf_iters >= 3) { steps[0] = 0; steps[1] = 1; steps[2] = 2; *nsteps = 3; }
}

// Writes the descriptors, probe parameters and workgroup prefixes of a chain
// over n images into H, the host staging of the device block dev (layout L;
// H zeroed by the caller).
static int fill_chain(jpxlg_ctx* ctx, jpxlg_img* const* imgs, int n, const jpxlg_quantizer* qs,
                      const jpxlg_restoration* r, char* H, char* dev, const ChainLayout& L, ChainTotals* T) {
  ProbeDesc* hd = (ProbeDesc*)(H + L.desc);
  ProbeParams* hp = (ProbeParams*)(H + L.pp);
  LineSet* hs = (LineSet*)(H + L.sets);
  int* ho = (int*)(H + L.off);
  ChainTotals t;
  double pixels = 0.0;
  for (int k = 0; k < n; k++) pixels += (double)imgs[k]->g.w * imgs[k]->g.h;
  const jpxlg_ctx::Shape shape = ctx->tune[tune_class(pixels)];
  t.upw = shape.upw;
  t.swizzle = shape.swizzle;
  for (int k = 0; k < n; k++) {
    jpxlg_img* img = imgs[k];
    const Geom& g = img->g;
    const jpxlg_quantizer& q = qs ? qs[k] : img->last_q;
    const int sharp = r ? r->sharpness : img->last_sharpness;
    probe_params_for(img, q, sharp, &hp[k]);
    ProbeDesc& P = hd[k];
    P.g = g;
    P.nscales = img->nscales;
    for (int s = 0; s < kDescScales; s++) {
      P.sw[s] = s < img->nscales ? img->sw[s] : 0;
      P.sh[s] = s < img->nscales ? img->sh[s] : 0;
    }
    for (int i = 0; i < 3; i++) {
      P.list[i] = img->d_list[i];
      P.count[i] = img->list_count[i];
      P.coeff[i] = img->coeff[i];
      P.quant[i] = img->quant[i];
      P.lf[i] = img->lf[i];
      P.rx[i] = img->rx[i];
      P.aux[i] = img->aux[i];
    }
    P.sigma = img->sigma;
    P.cfl_x = img->d_cfl_x;
    P.cfl_b = img->d_cfl_b;
    P.cflp = img->d_cflp;
    P.pp = (const ProbeParams*)(dev + L.pp) + k;
    P.err = (uint32_t*)(dev + L.err) + k;
    int pt = 0;
    for (int s = 0; s < img->nscales; s++)
      for (int c = 0; c < 3; c++) P.cand_x[s][c] = img->cand_x[s][c];
    for (int s = 0; s < img->nscales; s++) {
      P.ptile_off[s] = pt;
      P.punits[s] = pool_units(img->sw[s], img->sh[s]);
      pt += P.punits[s];
    }
    P.partials = img->d_partials;
    // the blur passes' line sets: (scale, channel) of this image
    for (int s = 0; s < img->nscales; s++)
      for (int c = 0; c < 3; c++) {
        LineSet& S = hs[t.nsets];
        S.w = img->sw[s];
        S.h = img->sh[s];
        S.pitch = line_pitch(S.h);
        S.x1 = img->ref_x[s][c];
        S.x2 = img->cand_x[s][c];
        for (int m = 0; m < 3; m++) S.hT[m] = img->cand_h[s][c][m];
        S.xT = img->cand_xT[s][c];
        S.x1T = img->ref_xT[s][c];
        S.x2T = img->cand_xT[s][c];
        S.mu1T = img->ref_mu[s][c];
        S.s11T = img->ref_s11[s][c];
        S.partials = img->d_partials + (size_t)P.ptile_off[s] * 18 + c * 6;
        ho[L.b_off + t.nsets] = t.bands;
        ho[L.v_off + t.nsets] = t.cols;
        t.bands += blur_bands(S.w, S.h);
        t.cols += blur_columns(S.w, S.h);
        t.nsets++;
      }
    P.sums = (double*)(dev + L.sums) + (size_t)k * 18 * kDescScales;
    DepthSet dd;
    const int rc = get_depth(ctx, img->bits, &dd);
    if (rc) return rc;
    P.thr = dd.thr;
    P.lut = dd.lut;
    P.levels = dd.levels;
    // workgroup prefixes
    for (int i = 0; i < 3; i++) {
      ho[L.q_off[i] + k] = t.q[i];
      t.q[i] += img->list_count[i] ? quant_groups(kSides[i], img->list_count[i]) : 0;
      ho[L.r_off[i] + k] = t.r[i];
      t.r[i] += img->list_count[i] ? render_groups(kSides[i], img->list_count[i]) : 0;
    }
    ho[L.g_off + k] = t.g;
    t.g += gaborish_blocks(g.w, g.h);
    ho[L.e_off + k] = t.e;
    t.e += epf_tiles(g.w, g.h);
    ho[L.c_off[0] + k] = t.c[0];  // the colour pyramid's 64x64 tiles
    t.c[0] += pyramid_tiles(g.w, g.h);
  }
  for (int i = 0; i < 3; i++) {
    ho[L.q_off[i] + n] = t.q[i];
    ho[L.r_off[i] + n] = t.r[i];
  }
  ho[L.g_off + n] = t.g;
  ho[L.e_off + n] = t.e;
  ho[L.c_off[0] + n] = t.c[0];
  ho[L.b_off + t.nsets] = t.bands;
  ho[L.v_off + t.nsets] = t.cols;
  *T = t;
  return JPXLG_OK;
}

// Issues a filled chain on st: one upload of the descriptor block, the
// stage launches (quantize when `quantize`, reconstruction + restoration +
// SSIMULACRA2 when r), and the download of the error words (and sums).
// *src = 1 when the restored planes end in the descriptors' aux planes.
// Whether the image's device levels are those of quantizer q (its last
// quantization, on the current structure).
static bool holds_levels_of(const jpxlg_img* img, const jpxlg_quantizer& q) {
  return img->has_quant && memcmp(&img->last_q, &q, sizeof(q)) == 0;
}

static int issue_chain(jpxlg_ctx* ctx, hipStream_t st, int n, bool quantize, const jpxlg_restoration* r, int form,
                       char* H, char* dev, const ChainLayout& L, const ChainTotals& T, int* src_out) {
  int nsteps, steps[3];
  epf_steps_of(r, &nsteps, steps);
  const int gab = r && r->gaborish ? 1 : 0;
  HIP_TRY(ctx, hipMemcpyAsync(dev, H, L.off + 4 * (size_t)L.noffs, hipMemcpyHostToDevice, st));
  const ProbeDesc* D = (const ProbeDesc*)(dev + L.desc);
  const int* O = (const int*)(dev + L.off);
  // quantize and the exact-form reconstruction in one launch when both run
  const bool fused = quantize && r && form == kDctExact;
  if (quantize) {
    Prof p(ctx, fused ? "quantize_render" : "quantize", st);
    const int* offs[3] = {O + L.q_off[0], O + L.q_off[1], O + L.q_off[2]};
    launch_quantize_all(st, D, offs, n, T.q, ctx->d_tab, fused);
  }
  int src = 0;
  if (r) {
    if (!fused) {
      Prof p(ctx, "render", st);
      const int* offs[3] = {O + L.r_off[0], O + L.r_off[1], O + L.r_off[2]};
      launch_render_all(st, form, D, offs, n, T.r, ctx->d_tab, ctx->d_mats);
    }
    if (gab) {
      { Prof p(ctx, "gaborish", st); launch_gaborish_b(st, D, O + L.g_off, n, T.g, src); }
      src ^= 1;
    }
    static const char* kEpfNames[3] = {"epf_step0", "epf_step1", "epf_step2"};
    for (int i = 0; i < nsteps; i++) {
      { Prof p(ctx, kEpfNames[steps[i]], st); launch_epf_b(st, steps[i], D, O + L.e_off, n, T.e, src); }
      src ^= 1;
    }
    {
      Prof p(ctx, "colour_pyramid", st);
      launch_colour_pyr_b(st, D, O + L.c_off[0], n, T.c[0], src);
    }
    const LineSet* S = (const LineSet*)(dev + L.sets);
    {
      Prof p(ctx, "blur_h", st);
      launch_blur_h(st, 1, ctx->d_tab, S, O + L.b_off, T.nsets, T.bands);
    }
    {
      Prof p(ctx, "blur_pool", st);
      launch_blur_v(st, 1, ctx->d_tab, S, O + L.v_off, T.nsets, T.cols, T.swizzle);
    }
    { Prof p(ctx, "pool_reduce", st); launch_reduce_b(st, D, n); }
  }
  HIP_TRY(ctx, hipGetLastError());
  // error words and sums come back to the host staging
  HIP_TRY(ctx, hipMemcpyAsync(H + L.err, dev + L.err, L.off - L.err, hipMemcpyDeviceToHost, st));
  if (r) HIP_TRY(ctx, hipMemcpyAsync(H + L.sums, dev + L.sums, L.total - L.sums, hipMemcpyDeviceToHost, st));
  *src_out = src;
  return JPXLG_OK;
}

// Enqueues quantize (with each image's quantizer in qs, when qs is non-null)
// and/or reconstruction + restoration + SSIMULACRA2 (when r is non-null) for
// n images of one context, with one descriptor upload; the sums and error
// words come back with read_chain.
static int enqueue_chain(jpxlg_ctx* ctx, hipStream_t st, jpxlg_img* const* imgs, int n, const jpxlg_quantizer* qs,
                         const jpxlg_restoration* r, ChainRun* run) {
  const ChainLayout L = chain_layout(n);
  if (qs)
    for (int k = 0; k < n; k++)
      if (imgs[k]->explicit_muls) {  // quantizing replaces a supplied plan
        imgs[k]->explicit_muls = false;
        imgs[k]->generation++;
      }
  jpxlg_ctx::Slot* sl = nullptr;
  int rc = take_slot(ctx, L.total, &sl);
  if (rc) return rc;
  char* H = sl->host;
  memset(H, 0, L.total);
  ChainTotals T;
  rc = fill_chain(ctx, imgs, n, qs, r, H, sl->dev, L, &T);
  int src = 0;
  if (!rc) rc = issue_chain(ctx, st, n, qs != nullptr, r, imgs[0]->dct_form, H, sl->dev, L, T, &src);
  if (rc) {
    // Nothing reads this chain back: wait for what was issued, free the slot.
    (void)hipStreamSynchronize(st);
    sl->claimed = false;
    return rc;
  }
  // Keep the restored planes in rx (downloads read them there).
  if (src)
    for (int k = 0; k < n; k++)
      for (int c = 0; c < 3; c++) std::swap(imgs[k]->rx[c], imgs[k]->aux[c]);
  HIP_TRY(ctx, hipEventRecord(sl->done, st));
  sl->pending = true;
  for (int k = 0; k < n; k++) {
    if (qs) {
      imgs[k]->last_q = qs[k];
      imgs[k]->has_quant = true;
    }
    if (r) {
      imgs[k]->last_sharpness = r->sharpness;
      imgs[k]->last_r = *r;
    }
  }
  run->slot = sl;
  run->host = H;
  run->L = L;
  return JPXLG_OK;
}

// Waits for a chain and reads back each image's error word and, when it
// scored, its metric (scores[k], optional raw/terms of image 0).
static int read_chain(jpxlg_ctx* ctx, jpxlg_img* const* imgs, int n, const ChainRun& run, bool scored,
                      double* scores, double* raw0, double* terms0, uint32_t* nterms0, int* rcs = nullptr) {
  if (rcs)
    for (int k = 0; k < n; k++) rcs[k] = JPXLG_E_DEVICE;
  if (run.slot) {
    HIP_TRY(ctx, hipEventSynchronize(run.slot->done));
    run.slot->pending = false;
  }
  // The slot is free for the next chain once this function has read it.
  struct Unclaim {
    jpxlg_ctx::Slot* s;
    ~Unclaim() { if (s) s->claimed = false; }
  } unclaim{run.slot};
  const char* H = run.host;
  const uint32_t* err = (const uint32_t*)(H + run.L.err);
  int bad = 0;
  for (int k = 0; k < n; k++) {
    bad |= err[k] != 0;
    if (rcs) rcs[k] = err[k] ? JPXLG_E_RANGE : JPXLG_OK;
  }
  if (scored) {
    for (int k = 0; k < n; k++) {
      jpxlg_img* img = imgs[k];
      const double* sums = (const double*)(H + run.L.sums) + (size_t)k * 18 * kDescScales;
      MetricResult m = pool_terms(sums, img->nscales, img->sw, img->sh);
      if (scores) scores[k] = m.score;
      if (k == 0) {
        if (raw0) *raw0 = m.raw_error;
        if (terms0) memcpy(terms0, m.terms, sizeof(double) * 18 * img->nscales);
        if (nterms0) *nterms0 = (uint32_t)(18 * img->nscales);
      }
      img->has_render = true;
    }
  }
  if (bad) return fail(ctx, JPXLG_E_RANGE, "a coefficient outside the quantizer's working range");
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_quantize(jpxlg_img* img, const jpxlg_quantizer* q) {
  if (!img || !q) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  if (!img->has_structure) return fail(ctx, JPXLG_E_INVALID, "no plan structure yet");
  int rc = check_quantizer(ctx, q);
  if (rc) return rc;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  ChainRun run;
  rc = enqueue_chain(ctx, ctx->stream, &img, 1, q, nullptr, &run);
  if (rc) return rc;
  img->has_render = false;
  return read_chain(ctx, &img, 1, run, false, nullptr, nullptr, nullptr, nullptr);
}

extern "C" int32_t jpxlg_render_score(jpxlg_img* img, const jpxlg_restoration* r, double* score,
                                      double* raw_error, double* terms, uint32_t* nterms) {
  if (!img || !r) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  if (!img->has_quant) return fail(ctx, JPXLG_E_INVALID, "nothing quantized yet");
  if (r->epf_iters < 0 || r->epf_iters > 3) return fail(ctx, JPXLG_E_UNSUPPORTED, "epf_iters above 3");
  if (r->sharpness < 0 || r->sharpness > 7) return fail(ctx, JPXLG_E_RANGE, "sharpness outside 0..7");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  ChainRun run;
  int rc = enqueue_chain(ctx, ctx->stream, &img, 1, nullptr, r, &run);
  if (rc) return rc;
  return read_chain(ctx, &img, 1, run, true, score, raw_error, terms, nterms);
}

// One probe: quantize + reconstruct + score, one host synchronisation.
int jpxlg_probe(jpxlg_img* img, const jpxlg_quantizer* q, const jpxlg_restoration* r, double* score) {
  jpxlg_ctx* ctx = img->ctx;
  int rc = check_quantizer(ctx, q);
  if (rc) return rc;
  ChainRun run;
  rc = enqueue_chain(ctx, ctx->stream, &img, 1, q, r, &run);
  if (rc) return rc;
  return read_chain(ctx, &img, 1, run, true, score, nullptr, nullptr, nullptr);
}

// ----------------------------------------------------------------------------
// Downloads
// ----------------------------------------------------------------------------

extern "C" int32_t jpxlg_plan_download(jpxlg_img* img, int32_t* quant, int32_t* lf, uint8_t* bmap,
                                       int8_t* cfl_x, int8_t* cfl_b, int32_t* lf_factors) {
  if (!img) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  if (!img->has_structure) return fail(ctx, JPXLG_E_INVALID, "no plan structure yet");
  if ((quant || lf) && !img->has_quant) return fail(ctx, JPXLG_E_INVALID, "nothing quantized yet");
  const Geom& g = img->g;
  const size_t cpx = (size_t)g.cstride * g.crows, bpx = (size_t)g.bw * g.bh;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  for (int c = 0; c < 3; c++) {
    if (quant) HIP_TRY(ctx, hipMemcpyAsync(quant + c * cpx, img->quant[c], cpx * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (lf) HIP_TRY(ctx, hipMemcpyAsync(lf + c * bpx, img->lf[c], bpx * 4, hipMemcpyDeviceToHost, ctx->stream));
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (bmap) memcpy(bmap, img->bmap.data(), img->bmap.size());
  if (cfl_x) memcpy(cfl_x, img->cfl_x.data(), img->cfl_x.size());
  if (cfl_b) memcpy(cfl_b, img->cfl_b.data(), img->cfl_b.size());
  if (lf_factors) { lf_factors[0] = img->lf_factors[0]; lf_factors[1] = img->lf_factors[1]; }
  return JPXLG_OK;
}

// What the frame signals besides the QuantizedFrameIr: the wire quantizer
// (jpxlg_plan_aq), the structure's LF factors, the source depth and the
// restoration the plan was rendered with.
extern "C" int32_t jpxlg_plan_frame_desc(jpxlg_img* img, const jpxlg_restoration* r, jpxlg_frame_desc* f) {
  if (!img || !f) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  if (!r) r = &img->last_r;
  if (r->epf_iters < 0 || r->epf_iters > 3 || r->gaborish < 0 || r->gaborish > 1 || r->sharpness < 0 ||
      r->sharpness > 7)
    return fail(ctx, JPXLG_E_RANGE, "restoration outside its ranges");
  jpxlg_quantizer wire;
  const int32_t rc = jpxlg_plan_aq(img, nullptr, &wire, nullptr);
  if (rc) return rc;
  f->width = (uint32_t)img->g.w;
  f->height = (uint32_t)img->g.h;
  f->bits_per_sample = (uint32_t)img->bits;
  f->global_scale = wire.global_scale;
  f->quant_lf = wire.quant_lf;
  f->x_qm_scale = wire.x_qm_scale;
  f->b_qm_scale = wire.b_qm_scale;
  f->x_factor_lf = img->lf_factors[0];
  f->b_factor_lf = img->lf_factors[1];
  f->gaborish = (uint32_t)r->gaborish;
  f->epf_iters = (uint32_t)r->epf_iters;
  f->sharpness = (uint32_t)r->sharpness;
  return JPXLG_OK;
}

// attach_entropy (jpxl-encode-policy/src/lib.rs:1012) and the emit
// (jpxl-encode/src/vardct/write.rs:350) over the image's last quantized plan:
// the IR regathered on the device is read back into the context's pinned
// staging, and the CPU codestream stage writes the frame with the quantizer
// the frame signals (jpxlg_plan_aq's wire), the structure's CfL and LF
// factors, and the restoration the plan was rendered with.
extern "C" int32_t jpxlg_encode_emit(jpxlg_img* img, const jpxlg_restoration* r, int32_t threads, uint8_t* out,
                                     size_t capacity, size_t* out_len, jpxlg_emit_sizes* sizes) {
  if (!img) return JPXLG_E_INVALID;
  if (!r) r = &img->last_r;
  jpxlg_ctx* ctx = img->ctx;
  if (!img->has_structure || !img->has_quant) return fail(ctx, JPXLG_E_INVALID, "nothing quantized yet");
  const Geom& g = img->g;
  const size_t bpx = (size_t)g.bw * g.bh;
  const size_t a_bytes = 189 * bpx * 4, l_bytes = 3 * bpx * 4, v_bytes = bpx * sizeof(jpxlg_varblock);
  const size_t r_bytes = 3 * bpx * sizeof(jpxlg_run);
  const size_t need = a_bytes + l_bytes + v_bytes + r_bytes;
  if (need > ctx->ir_cap) {
    if (ctx->ir_host) (void)hipHostFree(ctx->ir_host);
    ctx->ir_host = nullptr;
    ctx->ir_cap = 0;
    if (hipHostMalloc((void**)&ctx->ir_host, need, hipHostMallocDefault) != hipSuccess)
      return fail(ctx, JPXLG_E_OOM, "pinned staging for the IR hand-off");
    ctx->ir_cap = need;
  }
  int32_t* arena = (int32_t*)ctx->ir_host;
  int32_t* lfg = (int32_t*)(ctx->ir_host + a_bytes);
  jpxlg_varblock* vbs = (jpxlg_varblock*)(ctx->ir_host + a_bytes + l_bytes);
  jpxlg_run* runs = (jpxlg_run*)(ctx->ir_host + a_bytes + l_bytes + v_bytes);
  uint32_t n = 0;
  size_t arena_len = 0;
  int32_t rc = jpxlg_plan_download_runs(img, arena, 189 * bpx, &arena_len, runs, lfg, vbs, &n);
  if (rc) return rc;
  jpxlg_frame_desc f;
  rc = jpxlg_plan_frame_desc(img, r, &f);
  if (rc) return rc;
  rc = jpxlg_emit_codestream_runs(&f, arena, arena_len, runs, lfg, vbs, n, img->cfl_x.data(), img->cfl_b.data(),
                                  threads, out, capacity, out_len, sizes);
  if (rc == JPXLG_E_TOO_SMALL) return fail(ctx, rc, "output buffer smaller than the codestream");
  if (rc) return fail(ctx, rc, "codestream stage refused the plan");
  return JPXLG_OK;
}

static int32_t ir_table(jpxlg_img* img, jpxlg_varblock* varblocks, uint32_t* n_varblocks,
                        std::vector<uint32_t>* offsets);

// The QuantizedFrameIr hand-off (jpxl-encode/src/vardct/plan.rs:519-636):
// the quantized plan regathered on the device from the footprint layout
// into the IR's order (LF groups in raster order; inside a group, varblocks
// in block-raster order of their top-left blocks; per varblock its X, Y and
// B coefficient arrays, each in the array's own raster order), and each
// group's LF planes.
extern "C" int32_t jpxlg_plan_download_ir(jpxlg_img* img, int32_t* coeff_arena, int32_t* lf_groups,
                                          jpxlg_varblock* varblocks, uint32_t* n_varblocks) {
  if (!img) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  if (!img->has_structure) return fail(ctx, JPXLG_E_INVALID, "no plan structure yet");
  if (!img->has_quant) return fail(ctx, JPXLG_E_INVALID, "nothing quantized yet");
  const Geom& g = img->g;
  const size_t bpx = (size_t)g.bw * g.bh;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (coeff_arena || lf_groups) {
    // the CfL sample buffers are free outside a structure build
    int32_t* d_arena = (int32_t*)img->d_hf_s;
    int32_t* d_lfg = (int32_t*)img->d_lf_s;
    launch_gather_ir(st, g, img->d_list, img->list_count, img->d_lf_off, img->quant, img->lf, d_arena, d_lfg);
    HIP_TRY(ctx, hipGetLastError());
    if (coeff_arena)
      HIP_TRY(ctx, hipMemcpyAsync(coeff_arena, d_arena, 192 * bpx * 4, hipMemcpyDeviceToHost, st));
    if (lf_groups) HIP_TRY(ctx, hipMemcpyAsync(lf_groups, d_lfg, 3 * bpx * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
  }
  return ir_table(img, varblocks, n_varblocks, nullptr);
}

// The IR's varblock table (LF groups in raster order, inside a group the
// block-raster order of the top-left blocks) and, when offsets is given,
// each varblock's offset in 8x8 blocks.
static int32_t ir_table(jpxlg_img* img, jpxlg_varblock* varblocks, uint32_t* n_varblocks,
                        std::vector<uint32_t>* offsets) {
  const Geom& g = img->g;
  const size_t bpx = (size_t)g.bw * g.bh;
  if (varblocks || n_varblocks || offsets) {
    std::vector<uint32_t> muls(bpx, 0u);
    if (varblocks) {
      const int rc = jpxlg_plan_aq(img, muls.data(), nullptr, nullptr);
      if (rc) return rc;
    }
    if (offsets) offsets->clear();
    const int gw = (g.bw + 255) / 256, gh = (g.bh + 255) / 256;
    uint64_t off = 0;
    uint32_t n = 0;
    for (int gy = 0; gy < gh; gy++)
      for (int gx = 0; gx < gw; gx++)
        for (int by = 256 * gy; by < std::min(g.bh, 256 * gy + 256); by++)
          for (int bx = 256 * gx; bx < std::min(g.bw, 256 * gx + 256); bx++) {
            const size_t i = (size_t)by * g.bw + bx;
            const int side = img->bmap[i];
            if (!side) continue;
            if (varblocks) {
              jpxlg_varblock& v = varblocks[n];
              v.lf_group = (uint32_t)(gy * gw + gx);
              v.bx = (uint32_t)bx;
              v.by = (uint32_t)by;
              v.side = (uint32_t)side;
              v.hf_mul = muls[i];
              v.reserved = 0;
              v.start = 192 * off;
            }
            if (offsets) offsets->push_back((uint32_t)off);
            off += (uint64_t)(side / 8) * (side / 8);
            n++;
          }
    if (n_varblocks) *n_varblocks = n;
  }
  return JPXLG_OK;
}

// Token-ready runs (no reference counterpart; the QuantizedFrameIr content
// trimmed to what attach_entropy codes): computed on the device from the
// footprint planes, so only the runs cross PCIe.
extern "C" int32_t jpxlg_plan_download_runs(jpxlg_img* img, int32_t* run_arena, size_t capacity, size_t* arena_len,
                                            jpxlg_run* runs, int32_t* lf_groups, jpxlg_varblock* varblocks,
                                            uint32_t* n_varblocks) {
  if (!img) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  if (!img->has_structure) return fail(ctx, JPXLG_E_INVALID, "no plan structure yet");
  if (!img->has_quant) return fail(ctx, JPXLG_E_INVALID, "nothing quantized yet");
  if ((run_arena == nullptr) != (runs == nullptr)) return fail(ctx, JPXLG_E_INVALID, "run_arena and runs go together");
  const Geom& g = img->g;
  const size_t bpx = (size_t)g.bw * g.bh;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (lf_groups) {
    int32_t* d_lfg = (int32_t*)img->d_lf_s;
    launch_gather_ir(st, g, img->d_list, img->list_count, img->d_lf_off, img->quant, img->lf, nullptr, d_lfg);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipMemcpyAsync(lf_groups, d_lfg, 3 * bpx * 4, hipMemcpyDeviceToHost, st));
  }
  if (runs) {
    // the CfL sample buffer (256 words per block) holds the count table
    // (3 per block), the start table (3 per block + the total + the scan's
    // workgroup totals, one per 1024 entries) and the runs (at most 189 per
    // block)
    uint32_t* d_len = (uint32_t*)img->d_hf_s;
    uint32_t* d_start = d_len + 3 * bpx;
    int32_t* d_arena = (int32_t*)(d_len + 8 * bpx);
    launch_runs(st, g, img->d_list, img->list_count, img->d_lf_off, img->quant, ctx->d_tab, d_len, d_start, d_arena);
    HIP_TRY(ctx, hipGetLastError());
    const size_t t_bytes = (6 * bpx + 1) * 4;
    if (t_bytes > ctx->runs_cap) {
      if (ctx->runs_host) (void)hipHostFree(ctx->runs_host);
      ctx->runs_host = nullptr;
      ctx->runs_cap = 0;
      if (hipHostMalloc((void**)&ctx->runs_host, t_bytes, hipHostMallocDefault) != hipSuccess)
        return fail(ctx, JPXLG_E_OOM, "pinned staging for the runs' tables");
      ctx->runs_cap = t_bytes;
    }
    uint32_t* h_len = (uint32_t*)ctx->runs_host;
    uint32_t* h_start = h_len + 3 * bpx;
    HIP_TRY(ctx, hipMemcpyAsync(h_len, d_len, t_bytes, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    const size_t total = h_start[3 * bpx];
    if (arena_len) *arena_len = total;
    if (total > capacity) return fail(ctx, JPXLG_E_TOO_SMALL, "run arena larger than its buffer");
    if (total) HIP_TRY(ctx, hipMemcpyAsync(run_arena, d_arena, total * 4, hipMemcpyDeviceToHost, st));
    std::vector<uint32_t> offsets;
    const int32_t rc = ir_table(img, varblocks, n_varblocks, &offsets);
    if (rc) return rc;
    for (size_t i = 0; i < offsets.size(); i++)
      for (int c = 0; c < 3; c++) {
        const size_t at = 3 * (size_t)offsets[i] + c;
        runs[3 * i + c] = {h_start[at], h_len[at] & 0xffffu, h_len[at] >> 16};
      }
    HIP_TRY(ctx, hipStreamSynchronize(st));
    return JPXLG_OK;
  }
  HIP_TRY(ctx, hipStreamSynchronize(st));
  return ir_table(img, varblocks, n_varblocks, nullptr);
}

extern "C" int32_t jpxlg_image_download(jpxlg_img* img, int32_t which, void* host, size_t bytes) {
  if (!img || !host) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  const Geom& g = img->g;
  const size_t px = (size_t)g.w * g.h, cpx = (size_t)g.cstride * g.crows, bpx = (size_t)g.bw * g.bh;
  float* const* planes = nullptr;
  size_t per = 0;
  int count = 3;
  float* single = nullptr;
  switch (which) {
    case JPXLG_BUF_XYB: planes = img->xyb; per = px; break;
    case JPXLG_BUF_COEFF:
      if (!img->has_structure) return fail(ctx, JPXLG_E_INVALID, "no plan structure yet");
      planes = img->coeff; per = cpx; break;
    case JPXLG_BUF_RENDER_XYB:
      if (!img->has_render) return fail(ctx, JPXLG_E_INVALID, "nothing rendered yet");
      planes = img->rx; per = px; break;
    case JPXLG_BUF_REF_XYB0: planes = img->ref_x[0]; per = px; break;
    case JPXLG_BUF_SIGMA:
      if (!img->has_render) return fail(ctx, JPXLG_E_INVALID, "nothing rendered yet");
      single = img->sigma; per = bpx; count = 1; break;
    case JPXLG_BUF_AQ_FIELD:
      if (!img->aq_active) return fail(ctx, JPXLG_E_INVALID, "no active AQ field");
      single = img->aq_adj; per = bpx; count = 1; break;
    case JPXLG_BUF_LINEAR: {
      if (!img->has_render) return fail(ctx, JPXLG_E_INVALID, "nothing rendered yet");
      // Linear RGB at depth is recomputed from the restored planes.
      if (bytes < 3 * px * 4) return fail(ctx, JPXLG_E_INVALID, "buffer too small");
      DepthSet d;
      int rc = get_depth(ctx, img->bits, &d);
      if (rc) return rc;
      HIP_TRY(ctx, hipSetDevice(ctx->device));
      // aux holds scratch after a render; it receives the linear planes here.
      launch_linear_at_depth(ctx->stream, g.w, g.h, img->rx, d.thr, d.lut, d.levels, img->aux);
      for (int c = 0; c < 3; c++)
        HIP_TRY(ctx, hipMemcpyAsync((char*)host + c * px * 4, img->aux[c], px * 4, hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      return JPXLG_OK;
    }
    default: return fail(ctx, JPXLG_E_INVALID, "unknown buffer");
  }
  if (bytes < count * per * 4) return fail(ctx, JPXLG_E_INVALID, "buffer too small");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  for (int c = 0; c < count; c++) {
    const void* src = single ? (const void*)single : (const void*)planes[c];
    HIP_TRY(ctx, hipMemcpyAsync((char*)host + c * per * 4, src, per * 4, hipMemcpyDeviceToHost, ctx->stream));
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return JPXLG_OK;
}

// ----------------------------------------------------------------------------
// Quality controller
// ----------------------------------------------------------------------------

// The frame features of the first-guess predictor: every atom's Y variance
// on the device, the median and the flat fraction on the host.
// The first-guess features (quality.rs:2449-2474): per-atom variances of
// the XYB planes (quality_features.rs:173-213), the 32x32 preanalysis cells
// of the source samples (preanalysis.rs:277-484) and the DCT8 summary
// (lib.rs:2769-2955) on the device; their order statistics and reductions on
// the host.  Computed once per image (the source planes do not change).
// Where an image's feature results land in its share of the pinned
// staging: the order-statistic tail, the grayscale flag, the cell sums, the
// summary partials and the histogram.
struct FeatureStaging {
  size_t sel_off, sel_bytes, o_gray, o_cells, o_parts, o_hist, bytes;
  explicit FeatureStaging(const Geom& g) {
    sel_off = offsetof(SelectScratch, flat);
    sel_bytes = sizeof(SelectScratch) - sel_off;
    o_gray = 192;
    o_cells = 256;
    o_parts = o_cells + (size_t)preanalysis_cells(g.w, g.h) * sizeof(CellStats);
    o_hist = o_parts + (size_t)summary_groups(g) * sizeof(SummaryPartial);
    bytes = (o_hist + kSummaryBins * 4 + 255) & ~(size_t)255;
  }
};

static int readback_staging(jpxlg_ctx* ctx, size_t need) {
  if (ctx->rb_cap >= need) return JPXLG_OK;
  if (ctx->rb_host) (void)hipHostFree(ctx->rb_host);
  ctx->rb_host = nullptr;
  ctx->rb_cap = 0;
  if (hipHostMalloc((void**)&ctx->rb_host, need, hipHostMallocDefault) != hipSuccess)
    return fail(ctx, JPXLG_E_OOM, "readback staging allocation failed");
  ctx->rb_cap = need;
  return JPXLG_OK;
}

// Launches an image's feature kernels on the context's stream and the copies
// of their results into h (FeatureStaging(g).bytes of pinned memory).
static int features_enqueue(jpxlg_img* img, char* h) {
  jpxlg_ctx* ctx = img->ctx;
  const Geom& g = img->g;
  const int n = g.bw * g.bh;
  const int ncells = preanalysis_cells(g.w, g.h), nparts = summary_groups(g);
  const FeatureStaging L(g);
  hipStream_t st = ctx->stream;
  // cover scratch, free between structure builds
  float* d_luma = (float*)img->d_cost[0];
  float* d_chroma = (float*)img->d_cost[1];
  {
    Prof p(ctx, "features");
    launch_block_features(st, g, ctx->d_tab, img->xyb, d_luma, d_chroma, img->d_sum_parts, img->d_hist);
    // quality_features.rs:141-152: ranks floor((n - 1) q) for the luma
    // q10/q50/q90 and the chroma q50
    SelectRanks r{};
    const float qs[kSelectQueries] = {0.1f, 0.5f, 0.9f, 0.5f};
    for (int q = 0; q < kSelectQueries; q++) {
      r.rank[q] = (uint32_t)std::min<double>(floor((double)(n - 1) * (double)qs[q]), (double)(n - 1));
      r.set[q] = q < 3 ? 0 : 1;
    }
    launch_select(st, d_luma, d_chroma, n, kFlatVariance, r, ctx->d_select);
    launch_preanalysis_cells(st, g.w, g.h, img->bits, img->d_rgb, img->d_cells);
  }
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipMemcpyAsync(h, (const char*)ctx->d_select + L.sel_off, L.sel_bytes, hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipMemcpyAsync(h + L.o_gray, img->d_gray, 4, hipMemcpyDeviceToHost, st));
  if (ncells)
    HIP_TRY(ctx, hipMemcpyAsync(h + L.o_cells, img->d_cells, (size_t)ncells * sizeof(CellStats),
                                hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipMemcpyAsync(h + L.o_parts, img->d_sum_parts, (size_t)nparts * sizeof(SummaryPartial),
                              hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipMemcpyAsync(h + L.o_hist, img->d_hist, kSummaryBins * 4, hipMemcpyDeviceToHost, st));
  return JPXLG_OK;
}

// The host half, once the copies of features_enqueue have landed.
static void features_finish(jpxlg_img* img, const char* h) {
  const Geom& g = img->g;
  const FeatureStaging L(g);
  if (!img->gray_known) {
    memcpy(&img->gray, h + L.o_gray, 4);
    img->gray_known = true;
  }
  SelectScratch sel;  // only its tail (flat count onward) was read back
  memcpy((char*)&sel + L.sel_off, h, L.sel_bytes);
  jpxlg_source_features& f = img->sf;
  f = jpxlg_source_features{};
  f.width = (uint32_t)g.w;
  f.height = (uint32_t)g.h;
  f.grayscale = img->gray ? 1 : 0;
  source_features_set(sel.value[0], sel.value[1], sel.value[2], sel.value[3], sel.flat, (uint32_t)(g.bw * g.bh),
                      &f);
  preanalysis_from_cells((const CellStats*)(h + L.o_cells), preanalysis_cells(g.w, g.h), &f);
  transform_summary_from((const SummaryPartial*)(h + L.o_parts), summary_groups(g),
                         (const uint32_t*)(h + L.o_hist), &img->ts);
  img->features_known = true;
}

// The first-guess features (quality.rs:2449-2474) of several images of one
// context: per-atom variances of the XYB planes (quality_features.rs:173-213),
// the 32x32 preanalysis cells of the source samples (preanalysis.rs:277-484)
// and the DCT8 summary (lib.rs:2769-2955) on the device, one synchronisation
// for all; their reductions on the host.  Computed once per image (the
// source planes do not change).
static int images_features(jpxlg_img* const* imgs, int n) {
  constexpr size_t kChunkBytes = (size_t)16 << 20;  // pinned staging per synchronisation
  if (n == 0) return JPXLG_OK;
  jpxlg_ctx* ctx = imgs[0]->ctx;
  for (int k0 = 0; k0 < n;) {
    // a run of images whose staging fits the chunk (at least one)
    std::vector<size_t> off(1, 0);
    int k1 = k0;
    while (k1 < n) {
      const size_t b = imgs[k1]->features_known ? 0 : FeatureStaging(imgs[k1]->g).bytes;
      if (k1 > k0 && off.back() + b > kChunkBytes) break;
      off.push_back(off.back() + b);
      k1++;
    }
    if (off.back() > 0) {
      int rc = readback_staging(ctx, off.back());
      if (rc) return rc;
      char* h = (char*)ctx->rb_host;
      for (int k = k0; k < k1; k++)
        if (!imgs[k]->features_known && (rc = features_enqueue(imgs[k], h + off[k - k0]))) return rc;
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      for (int k = k0; k < k1; k++)
        if (!imgs[k]->features_known) features_finish(imgs[k], h + off[k - k0]);
    }
    k0 = k1;
  }
  return JPXLG_OK;
}

static int image_features(jpxlg_img* img) { return images_features(&img, 1); }

// The first probe and the navigator's prior exponent (quality.rs:2449-2530)
// from an image's features: the crossing predictor's rung where it is
// trusted, else the legacy table's (quality.rs:2452-2460).
static void start_from_features(const jpxlg_img* img, double target, int32_t effort, uint32_t* rung,
                                double* prior_beta) {
  const uint32_t legacy =
      rung_for_scale(predicted_effective_scale(img->sf.luma_variance_q50, img->sf.flat_fraction, target));
  const QualityStart q = quality_start(img->sf, img->ts, target, quality_policy(effort).reserve, legacy);
  *rung = q.rung;
  *prior_beta = q.prior_beta;
}

static int predicted_start(jpxlg_img* img, double target, int32_t effort, uint32_t* rung, double* prior_beta) {
  const int rc = image_features(img);
  if (rc) return rc;
  start_from_features(img, target, effort, rung, prior_beta);
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_image_quality_features(jpxlg_img* img, jpxlg_source_features* source,
                                                jpxlg_transform_summary* transform) {
  if (!img) return JPXLG_E_INVALID;
  HIP_TRY(img->ctx, hipSetDevice(img->ctx->device));
  const int rc = image_features(img);
  if (rc) return rc;
  if (source) *source = img->sf;
  if (transform) *transform = img->ts;
  return JPXLG_OK;
}

// A structure kept on the host (the anchor of a search), to be set again
// after a finalist was rebuilt with its own.
struct SavedStructure {
  std::vector<uint8_t> bmap;
  std::vector<int8_t> cx, cb;
  int32_t lf[2] = {0, 0};
  void save(const jpxlg_img* img) {
    bmap = img->bmap;
    cx = img->cfl_x;
    cb = img->cfl_b;
    lf[0] = img->lf_factors[0];
    lf[1] = img->lf_factors[1];
  }
  int restore(jpxlg_img* img) const {
    return jpxlg_set_structure(img, bmap.data(), cx.data(), cb.data(), lf);
  }
};

// Which structure the image holds during a search, and the builds that put
// it there: a fresh build at a rung (the first is the anchor) or the
// anchor set again.
struct SearchStructure {
  SavedStructure anchor;
  bool anchored = false;
  bool holds_anchor = false;
  uint32_t fresh_rung = 0;  // the rung of the build the image holds, when not the anchor
  int need(jpxlg_img* img, const QualityPolicy& pol, uint32_t rung, bool fresh) {
    if (fresh) {
      if (anchored && !holds_anchor && fresh_rung == rung) return JPXLG_OK;
      const jpxlg_quantizer q = pol.quantizer_at(rung);
      const int rc = jpxlg_plan_structure(img, &q, pol.cover_mode, pol.enable_cfl);
      if (rc) return rc;
      if (!anchored) {
        anchor.save(img);
        anchored = true;
        holds_anchor = true;
      } else {
        holds_anchor = false;
        fresh_rung = rung;
      }
      return JPXLG_OK;
    }
    if (holds_anchor) return JPXLG_OK;
    const int rc = anchor.restore(img);
    if (rc) return rc;
    holds_anchor = true;
    return JPXLG_OK;
  }
  // Whether the probe (rung, fresh) can run on the structure held now.
  bool ready(uint32_t rung, bool fresh) const {
    if (!anchored) return false;
    return fresh ? (!holds_anchor && fresh_rung == rung) : holds_anchor;
  }
};

static void fill_quality_result(const ControllerOutcome& oc, const QualityPolicy& pol, uint32_t predicted,
                                jpxlg_quality_result* o) {
  memset(o, 0, sizeof(*o));
  const jpxlg_quantizer qc = pol.quantizer_at(oc.rung);
  o->score = oc.score;
  o->rung = oc.rung;
  o->global_scale = qc.global_scale;
  o->hf_mul = qc.hf_mul;
  o->quant_lf = qc.quant_lf;
  o->met = oc.met;
  o->probes = (int32_t)oc.probes.size();
  for (size_t i = 0; i < oc.probes.size() && i < 16; i++) {
    o->probe_scores[i] = oc.probes[i].score;
    o->probe_rungs[i] = oc.probes[i].rung;
    o->probe_ms[i] = oc.probes[i].ms;
    o->probe_fresh[i] = oc.probes[i].fresh ? 1 : 0;
  }
  o->status = oc.status;
  o->predicted_rung = predicted;
  o->structural_builds = oc.structural_builds;
  o->structure_rebuilt = oc.chosen_fresh ? 1 : 0;
  o->exact_prices = oc.exact_prices;
  o->bytes = oc.bytes;
}

extern "C" int32_t jpxlg_encode_quality(jpxlg_img* img, const jpxlg_quality_request* req,
                                        jpxlg_quality_result* out) {
  if (!img || !req || !out) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  memset(out, 0, sizeof(*out));
  const QualityPolicy pol = quality_policy(req->effort);
  uint32_t predicted = 0;
  double prior_beta = NAN;
  int rc = predicted_start(img, req->target, req->effort, &predicted, &prior_beta);
  if (rc) return rc;
  SearchStructure ss;
  ProbeFn probe = [&](uint32_t rung, bool fresh, double* score, double* ms) -> int {
    auto t0 = std::chrono::steady_clock::now();
    // Navigator::probe (quality.rs:1153-1190): a fresh probe builds the cover
    // and CfL at its own quantizer, the others reuse the anchor's.
    int prc = ss.need(img, pol, rung, fresh);
    if (prc) return prc;
    const jpxlg_quantizer q = pol.quantizer_at(rung);
    jpxlg_restoration r{pol.epf_iters, 0, pol.sharpness};
    prc = jpxlg_probe(img, &q, &r, score);
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return prc;
  };
  if (req->mode < 0 || req->mode > 2) return fail(ctx, JPXLG_E_INVALID, "unknown quality search mode");
  // The exact price of a finalist (quality.rs:1886-1995): its plan quantized
  // on its structure and written by the codestream stage (sizes only).
  PriceFn price = [&](uint32_t rung, bool fresh, uint64_t* bytes) -> int {
    int prc = ss.need(img, pol, rung, fresh);
    if (prc) return prc;
    const jpxlg_quantizer q = pol.quantizer_at(rung);
    if (!holds_levels_of(img, q)) {
      prc = jpxlg_quantize(img, &q);
      if (prc) return prc;
    }
    jpxlg_restoration r{pol.epf_iters, 0, pol.sharpness};
    size_t n = 0;
    prc = jpxlg_encode_emit(img, &r, 0, nullptr, 0, &n, nullptr);
    *bytes = n;
    return prc;
  };
  ControllerOutcome oc;
  rc = req->mode == 1 ? bisect_controller(req->target, req->max_probes, probe, &oc)
                      : navigate_controller(req->target, req->effort, req->max_probes, probe, &oc, predicted,
                                            prior_beta, req->mode == 2 ? &price : nullptr);
  if (rc) return rc;
  // Leave the chosen plan quantized on the device for the hand-off, on the
  // structure it was scored with (the last probe's levels usually are).
  rc = ss.need(img, pol, oc.rung, oc.chosen_fresh);
  if (rc) return rc;
  const jpxlg_quantizer qc = pol.quantizer_at(oc.rung);
  if (!holds_levels_of(img, qc)) {
    rc = jpxlg_quantize(img, &qc);
    if (rc) return rc;
  }
  fill_quality_result(oc, pol, req->mode == 1 ? (oc.probes.empty() ? 0 : oc.probes[0].rung) : predicted, out);
  return JPXLG_OK;
}

// The byte-rate search (rate.rs:2016-2140, search_ladder :387-921): each
// candidate planned fresh at its rung and priced by the codestream stage.
extern "C" int32_t jpxlg_encode_bytes(jpxlg_img* img, const jpxlg_rate_request* req, jpxlg_rate_result* out) {
  if (!img || !req || !out) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = img->ctx;
  memset(out, 0, sizeof(*out));
  if (req->target_bytes == 0) return fail(ctx, JPXLG_E_RANGE, "a zero byte target");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  // EncodeRequest::for_target (request.rs:883-910): quant_lf 4, AQ off, EPF 1
  // at uniform sharpness 7, trailing truncation at lambda x4, the calibrated
  // cover with quantizer-donor weights and CfL; chroma per the effort.
  QualityPolicy pol = quality_policy(req->effort);
  pol.cover_mode = JPXLG_COVER_HIERARCHICAL;
  pol.enable_cfl = 1;
  int rc = jpxlg_set_aq(img, JPXLG_AQ_OFF, nullptr);
  if (rc) return rc;
  const jpxlg_restoration r{pol.epf_iters, 0, pol.sharpness};
  // RateSearchBudget::default: 40 prices, 4 fill probes; the exhaustive
  // search spends max_prices - full_refinement_reserve on the ladder
  // (rate.rs:146-160, :2050-2056).
  const uint32_t max_prices = req->max_prices ? req->max_prices : 40;
  const uint32_t reserve = max_prices <= 6 ? 1 : std::min(std::max(std::min(16u, max_prices / 2), 8u), max_prices - 1);
  const uint32_t fast_cap = std::max(1u, max_prices - reserve);
  auto plan_at = [&](uint32_t rung) -> int {
    const jpxlg_quantizer q = pol.quantizer_at(rung);
    int prc = jpxlg_plan_structure(img, &q, pol.cover_mode, pol.enable_cfl);
    if (prc) return prc;
    return jpxlg_quantize(img, &q);
  };
  uint32_t last_planned = UINT32_MAX;
  RatePriceFn price = [&](uint32_t rung, uint64_t* bytes) -> int {
    int prc = plan_at(rung);
    if (prc) return prc;
    last_planned = rung;
    size_t n = 0;
    prc = jpxlg_encode_emit(img, &r, 0, nullptr, 0, &n, nullptr);
    *bytes = n;
    return prc;
  };
  RateOutcome oc;
  rc = rate_search_ladder(Rung::for_global_scale(32768), req->target_bytes, 8, 0.01, fast_cap, 4, price, &oc);
  out->prices = (int32_t)oc.trace.size();
  for (size_t i = 0; i < oc.trace.size() && i < 64; i++) {
    out->trace_rungs[i] = oc.trace[i].rung;
    out->trace_bytes[i] = oc.trace[i].bytes;
  }
  if (rc == kRateUnreachable) {
    out->floor_bytes = oc.floor_bytes;
    return fail(ctx, JPXLG_E_RANGE, "the byte target is below the coarsest rung's size");
  }
  if (rc == kRateBudgetExhausted) return fail(ctx, JPXLG_E_RANGE, "no rung fits the target within the price budget");
  if (rc) return rc;
  if (last_planned != oc.rung) {
    rc = plan_at(oc.rung);
    if (rc) return rc;
  }
  img->last_r = r;
  const jpxlg_quantizer qc = pol.quantizer_at(oc.rung);
  out->rung = oc.rung;
  out->global_scale = qc.global_scale;
  out->hf_mul = qc.hf_mul;
  out->quant_lf = qc.quant_lf;
  out->bytes = oc.bytes;
  out->saturated = oc.saturated ? 1 : 0;
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_rung_quantizer(uint32_t rung, int32_t effort, jpxlg_quantizer* out) {
  if (!out || rung > Rung::top()) return JPXLG_E_INVALID;
  *out = quality_policy(effort).quantizer_at(rung);
  return JPXLG_OK;
}

// ----------------------------------------------------------------------------
// Batches: many images of one context advance their probes together, one
// launch per stage per iteration.
// ----------------------------------------------------------------------------

static int batch_check(uint32_t n, jpxlg_img* const* imgs) {
  if (!imgs || n == 0) return JPXLG_E_INVALID;
  for (uint32_t k = 0; k < n; k++)
    if (!imgs[k] || imgs[k]->ctx != imgs[0]->ctx) return JPXLG_E_INVALID;
  for (uint32_t k = 0; k < n; k++)
    if (imgs[k]->dct_form != imgs[0]->dct_form)
      return fail(imgs[0]->ctx, JPXLG_E_UNSUPPORTED, "a batch mixes DCT forms");
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_batch_probe(uint32_t n, jpxlg_img* const* imgs, const jpxlg_quantizer* qs,
                                     const jpxlg_restoration* r, double* scores) {
  int rc = batch_check(n, imgs);
  if (rc) return rc;
  jpxlg_ctx* ctx = imgs[0]->ctx;
  if (!qs || !r || !scores) return fail(ctx, JPXLG_E_INVALID, "null argument");
  for (uint32_t k = 0; k < n; k++) {
    if (!imgs[k]->has_structure) return fail(ctx, JPXLG_E_INVALID, "an image has no plan structure yet");
    rc = check_quantizer(ctx, &qs[k]);
    if (rc) return rc;
  }
  if (r->epf_iters < 0 || r->epf_iters > 3) return fail(ctx, JPXLG_E_UNSUPPORTED, "epf_iters above 3");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  ChainRun run;
  rc = enqueue_chain(ctx, ctx->stream, imgs, (int)n, qs, r, &run);
  if (rc) return rc;
  return read_chain(ctx, imgs, (int)n, run, true, scores, nullptr, nullptr, nullptr);
}

// ----------------------------------------------------------------------------
// Captured probe chains: the quantize -> reconstruct -> score chain of a
// fixed set of images captured once as a HIP graph and replayed per probe
// with new quantizers.  The graph owns its descriptor block (device + pinned
// staging); a replay rewrites the probe parameters in the staging, whose
// upload is the graph's first node, and the error words and metric sums come
// back by its last nodes.
// ----------------------------------------------------------------------------

struct jpxlg_graph {
  jpxlg_ctx* ctx = nullptr;
  std::vector<jpxlg_img*> imgs;
  std::vector<uint64_t> gens;
  jpxlg_restoration r{};
  ChainLayout L;
  char* dev = nullptr;
  char* host = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // where each image's restored planes end after a replay
  std::vector<std::array<float*, 3>> final_rx, final_aux;
  ~jpxlg_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (dev) (void)hipFree(dev);
    if (host) (void)hipHostFree(host);
  }
};

extern "C" int32_t jpxlg_graph_create(uint32_t n, jpxlg_img* const* imgs, const jpxlg_restoration* r,
                                      jpxlg_graph** out) {
  if (!out) return JPXLG_E_INVALID;
  *out = nullptr;
  int rc = batch_check(n, imgs);
  if (rc) return rc;
  jpxlg_ctx* ctx = imgs[0]->ctx;
  if (!r) return fail(ctx, JPXLG_E_INVALID, "null restoration");
  if (r->epf_iters < 0 || r->epf_iters > 3) return fail(ctx, JPXLG_E_UNSUPPORTED, "epf_iters above 3");
  for (uint32_t k = 0; k < n; k++) {
    if (!imgs[k]->has_structure) return fail(ctx, JPXLG_E_INVALID, "an image has no plan structure yet");
    if (imgs[k]->dct_form != imgs[0]->dct_form) return fail(ctx, JPXLG_E_UNSUPPORTED, "mixed DCT forms in one graph");
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  auto G = std::make_unique<jpxlg_graph>();
  G->ctx = ctx;
  G->imgs.assign(imgs, imgs + n);
  for (uint32_t k = 0; k < n; k++) G->gens.push_back(imgs[k]->generation);
  G->r = *r;
  G->L = chain_layout((int)n);
  if (!device_alloc(ctx, (void**)&G->dev, G->L.total)) return fail(ctx, JPXLG_E_OOM, "graph block allocation failed");
  if (hipHostMalloc(&G->host, G->L.total, hipHostMallocDefault) != hipSuccess)
    return fail(ctx, JPXLG_E_OOM, "graph staging allocation failed");
  memset(G->host, 0, G->L.total);
  // Descriptors now; the probe parameters are rewritten before each replay.
  std::vector<jpxlg_quantizer> q0(n);
  for (uint32_t k = 0; k < n; k++) q0[k] = jpxlg_quantizer{32768, 1, 16, 2, 2, 0, 1.0f, 1.0f};
  ChainTotals T;
  rc = fill_chain(ctx, imgs, (int)n, q0.data(), r, G->host, G->dev, G->L, &T);
  if (rc) return rc;
  // Every chain on the stream before the capture has finished (the stream
  // must be idle for a clean capture), and timing events stay out of it.
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const bool prof = ctx->prof;
  ctx->prof = false;
  int src = 0;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = issue_chain(ctx, ctx->stream, (int)n, true, r, imgs[0]->dct_form, G->host, G->dev, G->L, T, &src);
    const hipError_t e2 = hipStreamEndCapture(ctx->stream, &G->graph);
    if (!rc && e2 != hipSuccess) e = e2;
  }
  ctx->prof = prof;
  if (rc) return rc;
  if (e != hipSuccess) return fail(ctx, JPXLG_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
  HIP_TRY(ctx, hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
  const ProbeDesc* hd = (const ProbeDesc*)(G->host + G->L.desc);
  for (uint32_t k = 0; k < n; k++) {
    std::array<float*, 3> a, b;
    for (int c = 0; c < 3; c++) {
      a[c] = src ? hd[k].aux[c] : hd[k].rx[c];
      b[c] = src ? hd[k].rx[c] : hd[k].aux[c];
    }
    G->final_rx.push_back(a);
    G->final_aux.push_back(b);
  }
  *out = G.release();
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_graph_probe(jpxlg_graph* G, const jpxlg_quantizer* qs, double* scores) {
  if (!G) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = G->ctx;
  if (!qs || !scores) return fail(ctx, JPXLG_E_INVALID, "null argument");
  const int n = (int)G->imgs.size();
  for (int k = 0; k < n; k++) {
    if (G->imgs[k]->generation != G->gens[k] || !G->imgs[k]->has_structure)
      return fail(ctx, JPXLG_E_INVALID, "an image's structure changed since the graph was captured");
    const int rc = check_quantizer(ctx, &qs[k]);
    if (rc) return rc;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  // The previous replay has completed (this call is synchronous), so the
  // staging can be rewritten.
  ProbeParams* hp = (ProbeParams*)(G->host + G->L.pp);
  for (int k = 0; k < n; k++) probe_params_for(G->imgs[k], qs[k], G->r.sharpness, &hp[k]);
  memset(G->host + G->L.err, 0, 4 * (size_t)n);
  {
    Prof p(ctx, "graph_chain");
    HIP_TRY(ctx, hipGraphLaunch(G->exec, ctx->stream));
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < n; k++) {
    jpxlg_img* img = G->imgs[k];
    for (int c = 0; c < 3; c++) {
      img->rx[c] = G->final_rx[k][c];
      img->aux[c] = G->final_aux[k][c];
    }
    img->last_q = qs[k];
    img->has_quant = true;
    img->last_sharpness = G->r.sharpness;
    img->last_r = G->r;
  }
  ChainRun run;
  run.host = G->host;
  run.L = G->L;
  return read_chain(ctx, G->imgs.data(), n, run, true, scores, nullptr, nullptr, nullptr);
}

extern "C" void jpxlg_graph_destroy(jpxlg_graph* G) {
  if (!G) return;
  (void)hipSetDevice(G->ctx->device);
  (void)hipStreamSynchronize(G->ctx->stream);
  delete G;
}

// ----------------------------------------------------------------------------
// Continuous batching: a server of quality requests.  Each request runs its
// own navigator thread; step() gathers the pending probe of every active
// request (building the structure of new ones), runs them as one batched
// chain, and completes the requests whose navigators finished.  Requests
// can be submitted between steps and join the next round.
// ----------------------------------------------------------------------------

// New requests whose structure a server step builds before starting an idle
// lane that has no request ready.
constexpr int kServerBootstrap = 8;

struct ServerReq {
  uint64_t id = 0;  // its ticket
  jpxlg_img* img = nullptr;
  // Requests submitted as host samples: the server creates (and destroys)
  // their image when it admits (and releases) them.
  bool owned = false;
  bool admitted = true;
  const void* rgb = nullptr;
  uint32_t w = 0, h = 0, bits = 0;
  size_t footprint = 0;
  jpxlg_quality_request q{};
  QualityPolicy pol{};
  std::thread th;
  SearchStructure ss;     // the structure its image holds (anchor or a rebuilt finalist's)
  uint32_t predicted = 0; // the first probe's rung
  double prior_beta = NAN; // the crossing predictor's loss exponent, when trusted
  bool pending = false;
  uint32_t rung = 0;
  bool fresh = false;     // the pending probe builds its own structure
  double score = 0.0;
  int rc = 0;
  bool in_lane = false;   // its probe is in a lane's running round
  bool finished = false;  // navigator returned
  bool done = false;      // result available
  int nav_rc = 0;
  std::condition_variable cv;  // its navigator waits here for its probe's result
  ControllerOutcome oc;
  jpxlg_quality_result res{};
};

// A lane runs the rounds of one class of requests (small or large images)
// on its own stream, so a large image's round does not hold back the
// rounds of small ones; the GPU overlaps the two.
struct ServerLane {
  hipStream_t stream = nullptr;
  bool busy = false;
  ChainRun run;
  std::vector<ServerReq*> reqs;
};

struct jpxlg_server {
  jpxlg_ctx* ctx = nullptr;
  std::mutex m;
  std::condition_variable cv_req;
  int active = 0, waiting = 0;
  uint64_t next_id = 1;
  std::map<uint64_t, std::unique_ptr<ServerReq>> reqs;
  ServerLane lane[2];
  // Admission of requests submitted as host samples: in arrival order while
  // the projected device footprint of the admitted, unreleased ones fits the
  // budget (one request is always admitted when none holds memory).
  std::deque<uint64_t> queue;
  size_t budget = 0, used = 0;
  // Requests whose image exists and whose navigator has not started yet
  // (a step starts them together, after one batched feature pass).
  std::vector<uint64_t> unstarted;
  // Requests completed since the caller last asked (jpxlg_server_completed).
  std::deque<uint64_t> completed;
  ~jpxlg_server() {
    for (auto& l : lane)
      if (l.stream) {
        (void)hipStreamSynchronize(l.stream);
        (void)hipStreamDestroy(l.stream);
      }
  }
};

// Images above this many pixels run in the large-image lane.
constexpr uint64_t kLargeLanePixels = 4u << 20;

static int server_init(jpxlg_server* srv, jpxlg_ctx* ctx) {
  srv->ctx = ctx;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  for (auto& l : srv->lane) HIP_TRY(ctx, hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking));
  size_t free_b = 0, total_b = 0;
  HIP_TRY(ctx, hipMemGetInfo(&free_b, &total_b));
  srv->budget = free_b / 10 * 9;
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_server_create(jpxlg_ctx* ctx, jpxlg_server** out) {
  if (!ctx || !out) return JPXLG_E_INVALID;
  auto srv = std::make_unique<jpxlg_server>();
  int rc = server_init(srv.get(), ctx);
  if (rc) return rc;
  *out = srv.release();
  return JPXLG_OK;
}

// Starts the navigator of a request whose image is on the device (its
// first guess needs the image's features); called with the lock held.
static void server_start(jpxlg_server* srv, ServerReq* R) {
  srv->active++;
  R->th = std::thread([srv, R]() {
    ProbeFn probe = [srv, R](uint32_t rung, bool fresh, double* score, double* ms) -> int {
      auto t0 = std::chrono::steady_clock::now();
      std::unique_lock<std::mutex> lk(srv->m);
      R->rung = rung;
      R->fresh = fresh;
      R->pending = true;
      srv->waiting++;
      srv->cv_req.notify_all();
      R->cv.wait(lk, [R] { return !R->pending; });
      *score = R->score;
      *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      return R->rc;
    };
    ControllerOutcome oc;
    const int rc = R->q.mode == 1 ? bisect_controller(R->q.target, R->q.max_probes, probe, &oc)
                                  : navigate_controller(R->q.target, R->q.effort, R->q.max_probes, probe, &oc,
                                                        R->predicted, R->prior_beta);
    std::lock_guard<std::mutex> lk(srv->m);
    R->oc = oc;
    R->nav_rc = rc;
    R->finished = true;
    srv->active--;
    srv->cv_req.notify_all();
  });
}

extern "C" int32_t jpxlg_server_submit(jpxlg_server* srv, jpxlg_img* img, const jpxlg_quality_request* q,
                                       uint64_t* ticket) {
  if (!srv || !img || !q || !ticket) return JPXLG_E_INVALID;
  if (img->ctx != srv->ctx) return fail(srv->ctx, JPXLG_E_INVALID, "the image belongs to another context");
  if (q->mode < 0 || q->mode > 1)
    return fail(srv->ctx, JPXLG_E_UNSUPPORTED, "served and batched searches are navigator or bisection only");
  auto r = std::make_unique<ServerReq>();
  ServerReq* R = r.get();
  R->img = img;
  R->q = *q;
  R->pol = quality_policy(q->effort);
  std::lock_guard<std::mutex> lk(srv->m);
  const uint64_t id = srv->next_id++;
  R->id = id;
  srv->reqs.emplace(id, std::move(r));
  srv->unstarted.push_back(id);
  *ticket = id;
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_server_submit_srgb(jpxlg_server* srv, uint32_t width, uint32_t height, uint32_t bits,
                                            const void* rgb, const jpxlg_quality_request* q, uint64_t* ticket) {
  if (!srv || !rgb || !q || !ticket) return JPXLG_E_INVALID;
  if (q->mode < 0 || q->mode > 1)
    return fail(srv->ctx, JPXLG_E_UNSUPPORTED, "served and batched searches are navigator or bisection only");
  if (width < 8 || height < 8) return fail(srv->ctx, JPXLG_E_TOO_SMALL, "image is smaller than the metric's 8x8 floor");
  if (width > 65535 * 8 || height > 65535 * 8) return fail(srv->ctx, JPXLG_E_UNSUPPORTED, "frame too large");
  if (bits < 8 || bits > 16) return fail(srv->ctx, JPXLG_E_UNSUPPORTED, "bit depth outside 8..16");
  auto r = std::make_unique<ServerReq>();
  ServerReq* R = r.get();
  R->owned = true;
  R->admitted = false;
  R->rgb = rgb;
  R->w = width;
  R->h = height;
  R->bits = bits;
  R->footprint = (size_t)jpxlg_image_footprint(width, height);
  R->q = *q;
  R->pol = quality_policy(q->effort);
  std::lock_guard<std::mutex> lk(srv->m);
  const uint64_t id = srv->next_id++;
  R->id = id;
  srv->reqs.emplace(id, std::move(r));
  srv->queue.push_back(id);
  *ticket = id;
  return JPXLG_OK;
}

// Admits queued requests in arrival order while their projected footprint
// fits (each gets its image: the upload, XYB and metric reference), then
// starts every admitted request's navigator, the first guesses of all of
// them from one batched feature pass.  A request whose image cannot be
// created ends with that status.
static void server_admit(jpxlg_server* srv) {
  std::vector<ServerReq*> batch;
  for (;;) {
    ServerReq* R = nullptr;
    {
      std::lock_guard<std::mutex> lk(srv->m);
      if (srv->queue.empty()) break;
      const uint64_t id = srv->queue.front();
      ServerReq* F = srv->reqs.at(id).get();
      if (srv->used > 0 && srv->used + F->footprint > srv->budget) break;
      srv->queue.pop_front();
      srv->used += F->footprint;
      srv->unstarted.push_back(id);
      R = F;
    }
    jpxlg_img* img = nullptr;
    const int rc = image_create(srv->ctx, R->w, R->h, R->bits, R->rgb, false, &img);
    std::lock_guard<std::mutex> lk(srv->m);
    if (rc == JPXLG_E_OOM && srv->used > R->footprint) {
      // The device holds less than the budget assumed: while other admitted
      // requests hold memory, the request goes back to the head of the
      // queue and the budget drops to what is admitted now (releases make
      // room); alone, it fails.
      srv->used -= R->footprint;
      srv->budget = std::min(srv->budget, srv->used);
      const uint64_t id = srv->unstarted.back();
      srv->unstarted.pop_back();
      srv->queue.push_front(id);
      break;
    }
    R->img = img;
    R->admitted = true;
    if (rc) R->nav_rc = rc;
  }
  {
    std::lock_guard<std::mutex> lk(srv->m);
    for (uint64_t id : srv->unstarted) batch.push_back(srv->reqs.at(id).get());
    srv->unstarted.clear();
  }
  if (batch.empty()) return;
  std::vector<jpxlg_img*> imgs;
  for (ServerReq* R : batch)
    if (!R->nav_rc && R->q.mode != 1) imgs.push_back(R->img);
  const int frc = images_features(imgs.data(), (int)imgs.size());
  std::lock_guard<std::mutex> lk(srv->m);
  for (ServerReq* R : batch) {
    if (!R->nav_rc && R->q.mode != 1) {
      if (frc) R->nav_rc = frc;
      else start_from_features(R->img, R->q.target, R->q.effort, &R->predicted, &R->prior_beta);
    }
    if (R->nav_rc) {
      memset(&R->res, 0, sizeof(R->res));
      R->finished = R->done = true;
      srv->completed.push_back(R->id);
      continue;
    }
    server_start(srv, R);
  }
}

// Hands a lane's finished round back to its requests' navigators.
// Each request gets its own image's status (a coefficient out of range in
// one image fails that request only); the return value is a device failure
// of the round, if any.
static int server_collect(jpxlg_server* srv, ServerLane& l) {
  std::vector<jpxlg_img*> imgs;
  for (ServerReq* R : l.reqs) imgs.push_back(R->img);
  std::vector<double> sc(imgs.size(), 0.0);
  std::vector<int> rcs(imgs.size(), JPXLG_E_DEVICE);
  int rc = read_chain(srv->ctx, imgs.data(), (int)imgs.size(), l.run, true, sc.data(), nullptr, nullptr, nullptr,
                      rcs.data());
  if (rc == JPXLG_E_RANGE) rc = JPXLG_OK;
  {
    std::lock_guard<std::mutex> lk(srv->m);
    for (size_t t = 0; t < l.reqs.size(); t++) {
      ServerReq* R = l.reqs[t];
      R->score = sc[t];
      R->rc = rcs[t];
      R->pending = false;
      R->in_lane = false;
      srv->waiting--;
    }
  }
  // wake exactly this round's navigators
  for (ServerReq* R : l.reqs) R->cv.notify_one();
  l.reqs.clear();
  l.busy = false;
  return rc;
}

// One scheduling pass: completes finished searches, starts a round in every
// idle lane with pending probes (building new requests' structures first),
// and waits until at least one lane's round has finished, handing its scores
// back.  Returns the number of requests completed (>= 0) or a status.
extern "C" int32_t jpxlg_server_step(jpxlg_server* srv) {
  if (!srv) return JPXLG_E_INVALID;
  jpxlg_ctx* ctx = srv->ctx;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  server_admit(srv);
  std::vector<ServerReq*> pending[2], finishing;
  {
    std::unique_lock<std::mutex> lk(srv->m);
    srv->cv_req.wait(lk, [srv] { return srv->waiting == srv->active; });
    for (auto& kv : srv->reqs) {
      ServerReq* R = kv.second.get();
      if (!R->admitted) continue;
      if (R->pending && !R->in_lane) {
        const uint64_t px = (uint64_t)R->img->g.w * R->img->g.h;
        pending[px > kLargeLanePixels ? 1 : 0].push_back(R);
      } else if (R->finished && !R->done) {
        finishing.push_back(R);
      }
    }
  }
  int status = JPXLG_OK;
  // Completed searches: the chosen plans quantized in one batched launch.
  std::vector<jpxlg_img*> fimgs;
  std::vector<jpxlg_quantizer> fqs;
  for (ServerReq* R : finishing) {
    memset(&R->res, 0, sizeof(R->res));
    if (R->nav_rc) continue;
    // the chosen plan's own structure (the anchor, or its rebuilt one)
    const int src = R->ss.need(R->img, R->pol, R->oc.rung, R->oc.chosen_fresh);
    if (src) {
      R->nav_rc = src;
      continue;
    }
    const jpxlg_quantizer qc = R->pol.quantizer_at(R->oc.rung);
    if (!holds_levels_of(R->img, qc)) {
      fimgs.push_back(R->img);
      fqs.push_back(qc);
    }
    const uint32_t pred = R->q.mode == 1 ? (R->oc.probes.empty() ? 0 : R->oc.probes[0].rung) : R->predicted;
    fill_quality_result(R->oc, R->pol, pred, &R->res);
  }
  if (!fimgs.empty()) {
    // A request whose plan could not be left quantized fails with that
    // image's status (poll reports it); the others complete.
    std::vector<int> qrcs(fimgs.size(), JPXLG_E_DEVICE);
    ChainRun run;
    int qrc = enqueue_chain(ctx, ctx->stream, fimgs.data(), (int)fimgs.size(), fqs.data(), nullptr, &run);
    if (qrc) std::fill(qrcs.begin(), qrcs.end(), qrc);
    else qrc = read_chain(ctx, fimgs.data(), (int)fimgs.size(), run, false, nullptr, nullptr, nullptr, nullptr,
                          qrcs.data());
    for (ServerReq* R : finishing)
      for (size_t t = 0; t < fimgs.size(); t++)
        if (fimgs[t] == R->img && qrcs[t]) {
          R->nav_rc = qrcs[t];
          memset(&R->res, 0, sizeof(R->res));
        }
    if (qrc && qrc != JPXLG_E_RANGE) status = qrc;
  }
  {
    std::lock_guard<std::mutex> lk(srv->m);
    for (ServerReq* R : finishing) {
      R->done = true;
      srv->completed.push_back(R->id);
    }
  }
  // The structure a request's pending probe needs (a new request's anchor,
  // a finalist rebuilt at its rung, or the anchor set again) is built on the
  // context's stream; it returns false when that failed (the request then
  // ends with the error).
  auto build = [&](ServerReq* R) -> bool {
    const int prc = R->ss.need(R->img, R->pol, R->rung, R->fresh);
    if (prc) {
      std::lock_guard<std::mutex> lk(srv->m);
      R->rc = prc;
      R->pending = false;
      srv->waiting--;
      R->cv.notify_one();
      return false;
    }
    return true;
  };
  auto ready = [](const ServerReq* R) { return R->ss.ready(R->rung, R->fresh); };
  // Start a round in each idle lane with the requests whose structure is
  // built; a lane with none builds a few first.  The other new requests'
  // structures are built below, while the rounds run.
  for (int li = 0; li < 2; li++) {
    ServerLane& l = srv->lane[li];
    if (l.busy || pending[li].empty()) continue;
    std::vector<ServerReq*> ready_reqs;
    for (ServerReq* R : pending[li])
      if (ready(R)) ready_reqs.push_back(R);
    if (ready_reqs.empty()) {
      int built = 0;
      for (ServerReq* R : pending[li]) {
        if (built >= kServerBootstrap) break;
        if (build(R)) ready_reqs.push_back(R);
        built++;
      }
    }
    if (ready_reqs.empty()) continue;
    // one chain per lane round: group by (form, effort) and take the largest
    // group now; the rest wait for the lane's next round
    std::map<std::pair<int, int>, std::vector<ServerReq*>> groups;
    for (ServerReq* R : ready_reqs) groups[{R->img->dct_form, R->q.effort}].push_back(R);
    auto best = groups.begin();
    for (auto it = groups.begin(); it != groups.end(); ++it)
      if (it->second.size() > best->second.size()) best = it;
    std::vector<jpxlg_img*> bimgs;
    std::vector<jpxlg_quantizer> bqs;
    for (ServerReq* R : best->second) {
      bimgs.push_back(R->img);
      bqs.push_back(R->pol.quantizer_at(R->rung));
    }
    const QualityPolicy& pol = best->second[0]->pol;
    jpxlg_restoration r{pol.epf_iters, 0, pol.sharpness};
    const int brc = enqueue_chain(ctx, l.stream, bimgs.data(), (int)bimgs.size(), bqs.data(), &r, &l.run);
    if (brc) {
      std::lock_guard<std::mutex> lk(srv->m);
      for (ServerReq* R : best->second) {
        R->rc = brc;
        R->pending = false;
        srv->waiting--;
        R->cv.notify_one();
      }
      status = brc;
      continue;
    }
    {
      std::lock_guard<std::mutex> lk(srv->m);
      for (ServerReq* R : best->second) R->in_lane = true;
    }
    l.reqs = best->second;
    l.busy = true;
  }
  // Structures of the remaining new requests, while the rounds run on the
  // lanes' streams (the GPU overlaps them); stop once a round has finished.
  for (int li = 0; li < 2; li++)
    for (ServerReq* R : pending[li]) {
      bool lane_done = false;
      for (auto& l : srv->lane)
        if (l.busy && hipEventQuery(l.run.slot->done) != hipErrorNotReady) lane_done = true;
      if (lane_done && (srv->lane[0].busy || srv->lane[1].busy)) break;
      if (!R->pending || R->in_lane || ready(R)) continue;
      build(R);  // a failure ends that request (its navigator is woken)
    }
  // Wait for the first lane round to finish (short host polls), then hand
  // back every finished round.
  if (srv->lane[0].busy || srv->lane[1].busy) {
    for (;;) {
      bool any = false;
      for (auto& l : srv->lane)
        if (l.busy && hipEventQuery(l.run.slot->done) != hipErrorNotReady) any = true;
      if (any) break;
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    for (auto& l : srv->lane)
      if (l.busy && hipEventQuery(l.run.slot->done) != hipErrorNotReady) {
        const int rc = server_collect(srv, l);
        if (rc && !status) status = rc;
      }
  }
  if (status) return status;
  return (int32_t)finishing.size();
}

extern "C" int32_t jpxlg_server_poll(jpxlg_server* srv, uint64_t ticket, jpxlg_quality_result* out) {
  if (!srv) return JPXLG_E_INVALID;
  std::lock_guard<std::mutex> lk(srv->m);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) return fail(srv->ctx, JPXLG_E_INVALID, "unknown ticket");
  ServerReq* R = it->second.get();
  if (!R->done) return 0;
  if (out) *out = R->res;
  return R->nav_rc ? R->nav_rc : 1;
}

extern "C" int32_t jpxlg_server_release(jpxlg_server* srv, uint64_t ticket) {
  if (!srv) return JPXLG_E_INVALID;
  std::unique_ptr<ServerReq> r;
  {
    std::lock_guard<std::mutex> lk(srv->m);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) return fail(srv->ctx, JPXLG_E_INVALID, "unknown ticket");
    if (!it->second->done) return fail(srv->ctx, JPXLG_E_INVALID, "the request is still running");
    r = std::move(it->second);
    srv->reqs.erase(it);
  }
  if (r->th.joinable()) r->th.join();
  if (r->owned) {
    if (r->img) jpxlg_image_destroy(r->img);
    std::lock_guard<std::mutex> lk(srv->m);
    srv->used -= r->footprint;
  }
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_server_image(jpxlg_server* srv, uint64_t ticket, jpxlg_img** out) {
  if (!srv || !out) return JPXLG_E_INVALID;
  std::lock_guard<std::mutex> lk(srv->m);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) return fail(srv->ctx, JPXLG_E_INVALID, "unknown ticket");
  if (!it->second->img) return fail(srv->ctx, JPXLG_E_INVALID, "the request is not admitted yet");
  *out = it->second->img;
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_server_set_budget(jpxlg_server* srv, uint64_t bytes) {
  if (!srv) return JPXLG_E_INVALID;
  std::lock_guard<std::mutex> lk(srv->m);
  srv->budget = (size_t)bytes;
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_server_admission(jpxlg_server* srv, uint64_t* used, uint64_t* budget, uint32_t* queued) {
  if (!srv) return JPXLG_E_INVALID;
  std::lock_guard<std::mutex> lk(srv->m);
  if (used) *used = srv->used;
  if (budget) *budget = srv->budget;
  if (queued) *queued = (uint32_t)srv->queue.size();
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_server_completed(jpxlg_server* srv, uint64_t* tickets, uint32_t cap, uint32_t* n) {
  if (!srv || !n || (cap && !tickets)) return JPXLG_E_INVALID;
  std::lock_guard<std::mutex> lk(srv->m);
  uint32_t k = 0;
  while (k < cap && !srv->completed.empty()) {
    tickets[k++] = srv->completed.front();
    srv->completed.pop_front();
  }
  *n = k;
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_server_pending(jpxlg_server* srv) {
  if (!srv) return JPXLG_E_INVALID;
  std::lock_guard<std::mutex> lk(srv->m);
  int n = 0;
  for (auto& kv : srv->reqs) n += kv.second->done ? 0 : 1;
  return n;
}

// Ends every navigator that is waiting for a probe with `rc` (used when the
// device stops making progress, so no thread is left waiting).
static void server_abort_waiting(jpxlg_server* srv, int rc) {
  std::lock_guard<std::mutex> lk(srv->m);
  for (auto& kv : srv->reqs) {
    ServerReq* R = kv.second.get();
    if (R->pending && !R->in_lane) {
      R->rc = rc;
      R->pending = false;
      srv->waiting--;
      R->cv.notify_one();
    }
  }
}

// Steps until nothing is pending.  Every step either completes requests or
// advances their navigators (a failed round hands each request its error,
// which ends its search); after kMaxFailedSteps failing steps in a row with
// no request completed, the waiting navigators are ended with the error.
static int server_drain(jpxlg_server* srv) {
  constexpr int kMaxFailedSteps = 64;
  int status = JPXLG_OK, failed = 0;
  while (jpxlg_server_pending(srv) > 0) {
    const int st = jpxlg_server_step(srv);
    if (st >= 0) {
      failed = 0;
      continue;
    }
    if (!status) status = st;
    if (++failed >= kMaxFailedSteps) {
      server_abort_waiting(srv, st);
      failed = 0;
    }
  }
  return status;
}

extern "C" void jpxlg_server_destroy(jpxlg_server* srv) {
  if (!srv) return;
  // Requests never admitted end unrun; every admitted one is driven to
  // completion before the threads go.
  {
    std::lock_guard<std::mutex> lk(srv->m);
    for (uint64_t id : srv->queue) {
      ServerReq* R = srv->reqs.at(id).get();
      memset(&R->res, 0, sizeof(R->res));
      R->nav_rc = JPXLG_E_INVALID;
      R->admitted = R->finished = R->done = true;
    }
    srv->queue.clear();
  }
  (void)server_drain(srv);
  for (auto& kv : srv->reqs) {
    if (kv.second->th.joinable()) kv.second->th.join();
    if (kv.second->owned && kv.second->img) jpxlg_image_destroy(kv.second->img);
  }
  delete srv;
}

extern "C" int32_t jpxlg_batch_encode_quality(uint32_t n, jpxlg_img* const* imgs,
                                              const jpxlg_quality_request* reqs, jpxlg_quality_result* out) {
  int rc = batch_check(n, imgs);
  if (rc) return rc;
  jpxlg_ctx* ctx = imgs[0]->ctx;
  if (!reqs || !out) return fail(ctx, JPXLG_E_INVALID, "null argument");
  jpxlg_server srv;
  rc = server_init(&srv, ctx);
  if (rc) return rc;
  std::vector<uint64_t> tickets(n);
  for (uint32_t k = 0; k < n; k++) {
    rc = jpxlg_server_submit(&srv, imgs[k], &reqs[k], &tickets[k]);
    if (rc) break;
  }
  int status = rc;
  const int drc = server_drain(&srv);
  if (!status) status = drc;
  for (uint32_t k = 0; k < n && !rc; k++) {
    const int p = jpxlg_server_poll(&srv, tickets[k], &out[k]);
    if (p < 0 && !status) status = p;
  }
  for (auto& kv : srv.reqs)
    if (kv.second->th.joinable()) kv.second->th.join();
  return status;
}

// ----------------------------------------------------------------------------
// Stand-alone kernels
// ----------------------------------------------------------------------------

template <class T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
};

extern "C" int32_t jpxlg_dct_blocks(jpxlg_ctx* ctx, int32_t side, int32_t inverse, uint32_t count,
                                    const float* in, float* out) {
  if (!ctx || !in || !out) return JPXLG_E_INVALID;
  if (side != 8 && side != 16 && side != 32) return fail(ctx, JPXLG_E_UNSUPPORTED, "side not 8/16/32");
  if (count == 0) return JPXLG_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  size_t bytes = (size_t)count * side * side * 4;
  DevBuf<float> di, dout;
  HIP_TRY(ctx, hipMalloc(&di.p, bytes));
  HIP_TRY(ctx, hipMalloc(&dout.p, bytes));
  HIP_TRY(ctx, hipMemcpyAsync(di.p, in, bytes, hipMemcpyHostToDevice, ctx->stream));
  launch_dct_blocks(ctx->stream, side, inverse, (int)count, ctx->d_tab, di.p, dout.p);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipMemcpyAsync(out, dout.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_epf_step(jpxlg_ctx* ctx, int32_t step, uint32_t width, uint32_t height,
                                  const float* in, const float* sigma, float* out) {
  if (!ctx || !in || !sigma || !out) return JPXLG_E_INVALID;
  if (step < 0 || step > 2) return fail(ctx, JPXLG_E_INVALID, "step outside 0..2");
  if (width == 0 || height == 0) return fail(ctx, JPXLG_E_INVALID, "a zero frame dimension");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t px = (size_t)width * height, bpx = (size_t)((width + 7) / 8) * ((height + 7) / 8);
  DevBuf<float> di, ds, dout;
  HIP_TRY(ctx, hipMalloc(&di.p, 3 * px * 4));
  HIP_TRY(ctx, hipMalloc(&ds.p, bpx * 4));
  HIP_TRY(ctx, hipMalloc(&dout.p, 3 * px * 4));
  HIP_TRY(ctx, hipMemcpyAsync(di.p, in, 3 * px * 4, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ds.p, sigma, bpx * 4, hipMemcpyHostToDevice, ctx->stream));
  float* ip[3] = {di.p, di.p + px, di.p + 2 * px};
  float* op[3] = {dout.p, dout.p + px, dout.p + 2 * px};
  launch_epf(ctx->stream, step, (int)width, (int)height, ip, ds.p, op);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipMemcpyAsync(out, dout.p, 3 * px * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_gaborish(jpxlg_ctx* ctx, uint32_t width, uint32_t height, const float* in,
                                  float* out) {
  if (!ctx || !in || !out) return JPXLG_E_INVALID;
  if (width == 0 || height == 0) return fail(ctx, JPXLG_E_INVALID, "a zero frame dimension");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t px = (size_t)width * height;
  DevBuf<float> di, dout;
  HIP_TRY(ctx, hipMalloc(&di.p, 3 * px * 4));
  HIP_TRY(ctx, hipMalloc(&dout.p, 3 * px * 4));
  HIP_TRY(ctx, hipMemcpyAsync(di.p, in, 3 * px * 4, hipMemcpyHostToDevice, ctx->stream));
  float* ip[3] = {di.p, di.p + px, di.p + 2 * px};
  float* op[3] = {dout.p, dout.p + px, dout.p + 2 * px};
  launch_gaborish(ctx->stream, (int)width, (int)height, ip, op);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipMemcpyAsync(out, dout.p, 3 * px * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return JPXLG_OK;
}

// jpxlg_score_pair's buffers for a w x h pair, carved in this order from
// one 256-aligned region: the scale-0 linear RGB of both images (unless the
// inputs are read in place), two ping-pong sets of the next scales' linear
// RGB, this scale's positive XYB of both images, one channel's blur planes
// (both transposed XYB planes, three horizontal blurs, mu1 and s11; reused
// channel after channel and scale after scale), the pooling partials and
// sums, and the line-set table.
struct PairBuffers {
  float *in_r[3], *in_c[3];         // scale 0 linear RGB
  float *pp_r[2][3], *pp_c[2][3];   // scales >= 1 linear RGB (ping-pong)
  float *rx[3], *cx[3];             // positive XYB of the current scale
  float *x1T, *x2T, *h[3], *mu, *s11;
  double *partials, *sums;
  char* table;
};
static constexpr int kPairSets = 2 * 3 * kMaxScales;
static constexpr size_t kPairTable = kPairSets * (sizeof(LineSet) + 4 * sizeof(int));

static bool carve_pair(Arena& A, int w, int h, bool device_inputs, const float* reference, const float* candidate,
                       PairBuffers* B) {
  int n, sw[kMaxScales], sh[kMaxScales];
  scale_dims(w, h, &n, sw, sh);
  const size_t px = (size_t)w * h, p1 = n > 1 ? (size_t)sw[1] * sh[1] : 0;
  bool ok = true;
  for (int c = 0; c < 3; c++) {
    if (device_inputs) {
      B->in_r[c] = const_cast<float*>(reference) + c * px;
      B->in_c[c] = const_cast<float*>(candidate) + c * px;
    } else {
      B->in_r[c] = (float*)A.take(px * 4);
      B->in_c[c] = (float*)A.take(px * 4);
    }
    for (int k = 0; k < 2; k++) {
      B->pp_r[k][c] = p1 ? (float*)A.take(p1 * 4) : nullptr;
      B->pp_c[k][c] = p1 ? (float*)A.take(p1 * 4) : nullptr;
      ok = ok && (!p1 || (B->pp_r[k][c] && B->pp_c[k][c]));
    }
    B->rx[c] = (float*)A.take(px * 4);
    B->cx[c] = (float*)A.take(px * 4);
    ok = ok && B->in_r[c] && B->in_c[c] && B->rx[c] && B->cx[c];
  }
  const size_t tp = transposed_floats(w, h);  // the largest scale's transposed plane
  B->x1T = (float*)A.take(tp * 4);
  B->x2T = (float*)A.take(tp * 4);
  for (int m = 0; m < 3; m++) ok = ok && (B->h[m] = (float*)A.take(tp * 4)) != nullptr;
  B->mu = (float*)A.take(tp * 4);
  B->s11 = (float*)A.take(tp * 4);
  size_t units = 0;
  for (int s = 0; s < n; s++) units += (size_t)pool_units(sw[s], sh[s]);
  B->partials = (double*)A.take(units * 18 * 8);
  B->sums = (double*)A.take(kMaxScales * 18 * 8);
  B->table = (char*)A.take(kPairTable);
  return ok && B->x1T && B->x2T && B->mu && B->s11 && B->partials && B->sums && B->table;
}

// Device bytes jpxlg_score_pair's buffers take for a w x h pair.
static size_t score_pair_bytes(int w, int h, bool device_inputs) {
  Arena A{nullptr, ~(size_t)0 >> 1, 0};
  PairBuffers B;
  A.base = (char*)(uintptr_t)256;  // sizing only: nothing is written
  carve_pair(A, w, h, device_inputs, nullptr, nullptr, &B);
  return A.used + 256;
}

// ssimulacra2.rs:79-135 on a w x h pair whose buffers are carved out of
// [base, base + bytes) (256-aligned): the inputs are copied from the host,
// or read in place when device_inputs.  The per-scale sums come back into
// host_sums (pinned or pageable) on the context stream, synchronously.
static int score_pair_run(jpxlg_ctx* ctx, int w, int h, const float* reference, const float* candidate,
                          bool device_inputs, char* base, size_t bytes, double* host_sums, double* score,
                          double* terms, uint32_t* nterms) {
  int n, sw[kMaxScales], sh[kMaxScales];
  scale_dims(w, h, &n, sw, sh);
  Arena A{base, bytes, 0};
  PairBuffers B;
  if (!carve_pair(A, w, h, device_inputs, reference, candidate, &B)) return fail(ctx, JPXLG_E_OOM, "arena sizing");
  const size_t px = (size_t)w * h;
  hipStream_t st = ctx->stream;
  if (!device_inputs)
    for (int c = 0; c < 3; c++) {
      HIP_TRY(ctx, hipMemcpyAsync(B.in_r[c], reference + c * px, px * 4, hipMemcpyHostToDevice, st));
      HIP_TRY(ctx, hipMemcpyAsync(B.in_c[c], candidate + c * px, px * 4, hipMemcpyHostToDevice, st));
    }
  // line sets: per (scale, channel) the reference's, then the candidate's
  LineSet sets[kPairSets];
  int offs[kPairSets][4];  // band prefix {0, bands}, column prefix {0, columns}
  PoolLayout PL{};
  PL.nscales = n;
  int pu = 0, ns = 0;
  for (int s = 0; s < n; s++) {
    PL.off[s] = pu;
    PL.tiles[s] = pool_units(sw[s], sh[s]);
    for (int c = 0; c < 3; c++) {
      LineSet R{}, C{};
      R.w = C.w = sw[s];
      R.h = C.h = sh[s];
      R.pitch = C.pitch = line_pitch(sh[s]);
      R.x1 = C.x1 = B.rx[c];
      R.hT[0] = B.h[0];
      R.hT[1] = B.h[1];
      R.xT = B.x1T;
      R.out[0] = B.mu;
      R.out[1] = B.s11;
      C.x2 = B.cx[c];
      for (int m = 0; m < 3; m++) C.hT[m] = B.h[m];
      C.xT = B.x2T;
      C.x1T = B.x1T;
      C.x2T = B.x2T;
      C.mu1T = B.mu;
      C.s11T = B.s11;
      C.partials = B.partials + (size_t)pu * 18 + c * 6;
      for (const LineSet* L : {&R, &C}) {
        sets[ns] = *L;
        offs[ns][0] = 0;
        offs[ns][1] = blur_bands(L->w, L->h);
        offs[ns][2] = 0;
        offs[ns][3] = blur_columns(L->w, L->h);
        ns++;
      }
    }
    pu += PL.tiles[s];
  }
  char blob[kPairTable];
  memcpy(blob, sets, sizeof(LineSet) * ns);
  memcpy(blob + sizeof(LineSet) * kPairSets, offs, sizeof(int) * 4 * ns);
  int rc = upload_small(ctx, st, B.table, blob, kPairTable);
  if (rc) return rc;
  const LineSet* dsets = (const LineSet*)B.table;
  const int* doffs = (const int*)(B.table + sizeof(LineSet) * kPairSets);
  for (int s = 0, j = 0; s < n; s++) {
    float* const* rin = s == 0 ? B.in_r : B.pp_r[(s - 1) & 1];
    float* const* cin = s == 0 ? B.in_c : B.pp_c[(s - 1) & 1];
    float** rd = s + 1 < n ? B.pp_r[s & 1] : nullptr;
    float** cd = s + 1 < n ? B.pp_c[s & 1] : nullptr;
    launch_colour_metric(st, 1, sw[s], sh[s], rin, nullptr, nullptr, 0, B.rx, rd);
    launch_colour_metric(st, 1, sw[s], sh[s], cin, nullptr, nullptr, 0, B.cx, cd);
    for (int c = 0; c < 3; c++, j += 2) {
      launch_blur_h(st, 0, ctx->d_tab, dsets + j, doffs + 4 * j, 1, offs[j][1]);
      launch_blur_v(st, 0, ctx->d_tab, dsets + j, doffs + 4 * j + 2, 1, offs[j][3], 0);
      launch_blur_h(st, 1, ctx->d_tab, dsets + j + 1, doffs + 4 * (j + 1), 1, offs[j + 1][1]);
      launch_blur_v(st, 1, ctx->d_tab, dsets + j + 1, doffs + 4 * (j + 1) + 2, 1, offs[j + 1][3], 0);
    }
  }
  launch_reduce_scales(st, PL, B.partials, B.sums);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipMemcpyAsync(host_sums, B.sums, sizeof(double) * 18 * n, hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  MetricResult m = pool_terms(host_sums, n, sw, sh);
  if (score) *score = m.score;
  if (terms) memcpy(terms, m.terms, sizeof(double) * 18 * n);
  if (nterms) *nterms = (uint32_t)(18 * n);
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_score_pair(jpxlg_ctx* ctx, uint32_t width, uint32_t height,
                                    const float* reference, const float* candidate, double* score,
                                    double* terms, uint32_t* nterms) {
  if (!ctx || !reference || !candidate) return JPXLG_E_INVALID;
  if (width < 8 || height < 8) return fail(ctx, JPXLG_E_TOO_SMALL, "image is smaller than the metric's 8x8 floor");
  if (width > 65535 * 8 || height > 65535 * 8) return fail(ctx, JPXLG_E_UNSUPPORTED, "frame too large");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t bytes = score_pair_bytes((int)width, (int)height, false);
  DevBuf<char> mem;
  if (!device_alloc(ctx, (void**)&mem.p, bytes)) return fail(ctx, JPXLG_E_OOM, "device allocation failed");
  double sums[kMaxScales * 18];
  return score_pair_run(ctx, (int)width, (int)height, reference, candidate, false, mem.p, bytes, sums, score,
                        terms, nterms);
}

// The no-allocation score: every intermediate buffer lives in the
// file-scope array `scratch` (see its definition), through this context's
// device address of it.
extern "C" int32_t jpxlg_score_pair_noalloc(jpxlg_ctx* ctx, uint32_t width, uint32_t height,
                                            const float* d_reference, const float* d_candidate, double* score,
                                            double* terms, uint32_t* nterms) {
  if (!ctx || !d_reference || !d_candidate) return JPXLG_E_INVALID;
  if (width < 8 || height < 8) return fail(ctx, JPXLG_E_TOO_SMALL, "image is smaller than the metric's 8x8 floor");
  if (width > 65535 * 8 || height > 65535 * 8) return fail(ctx, JPXLG_E_UNSUPPORTED, "frame too large");
  if (!ctx->static_scratch) return fail(ctx, JPXLG_E_DEVICE, "the static scratch array has no device address");
  const size_t bytes = score_pair_bytes((int)width, (int)height, true);
  if (bytes > JPXLG_NOALLOC_SCRATCH_BYTES)
    return fail(ctx, JPXLG_E_UNSUPPORTED,
                "image too large for the no-allocation score: its pyramids need " + std::to_string(bytes) +
                    " bytes of the " + std::to_string((size_t)JPXLG_NOALLOC_SCRATCH_BYTES) + "-byte scratch array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  return score_pair_run(ctx, (int)width, (int)height, d_reference, d_candidate, true, ctx->static_scratch,
                        JPXLG_NOALLOC_SCRATCH_BYTES, ctx->static_sums, score, terms, nterms);
}

// ----------------------------------------------------------------------------
// Per-kernel timing
// ----------------------------------------------------------------------------

extern "C" int32_t jpxlg_profile_enable(jpxlg_ctx* ctx, int32_t on) {
  if (!ctx) return JPXLG_E_INVALID;
  ctx->prof = on != 0;
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_profile_only(jpxlg_ctx* ctx, const char* name) {
  if (!ctx) return JPXLG_E_INVALID;
  ctx->prof_only = name ? name : "";
  return JPXLG_OK;
}

extern "C" int32_t jpxlg_profile_read(jpxlg_ctx* ctx, char* names, double* total_ms,
                                      uint64_t* launches, int32_t cap) {
  if (!ctx) return JPXLG_E_INVALID;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->side));
  for (auto& r : ctx->recs) {
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, r.a, r.b));
    auto& t = ctx->totals[r.name];
    t.first += ms;
    t.second += 1;
    ctx->pool.push_back(r.a);
    ctx->pool.push_back(r.b);
  }
  ctx->recs.clear();
  int32_t n = 0;
  for (auto& kv : ctx->totals) {
    if (n >= cap) break;
    if (names) {
      strncpy(names + 48 * n, kv.first.c_str(), 47);
      names[48 * n + 47] = 0;
    }
    if (total_ms) total_ms[n] = kv.second.first;
    if (launches) launches[n] = kv.second.second;
    n++;
  }
  ctx->totals.clear();
  return n;
}

// ----------------------------------------------------------------------------
// Autotuning of launch shapes.  The candidates of a tunable launch compute
// the same sums (bit for bit), so the choice is speed only: each candidate
// runs the probe chain of a synthetic image of the requested size, the
// streaming pass is timed with events, and the fastest shape is kept for the
// size class (a chain's total pixels).
// ----------------------------------------------------------------------------

extern "C" int32_t jpxlg_autotune(jpxlg_ctx* ctx, uint32_t width, uint32_t height, int32_t* units_per_group,
                                  int32_t* swizzle) {
  if (!ctx) return JPXLG_E_INVALID;
  if (width < 8 || height < 8) return fail(ctx, JPXLG_E_TOO_SMALL, "autotune size below 8x8");
  // Deterministic content with flat, smooth and textured regions.
  std::vector<uint8_t> rgb((size_t)width * height * 3);
  uint32_t st = 0x9E3779B9u;
  for (uint32_t y = 0; y < height; y++)
    for (uint32_t x = 0; x < width; x++) {
      st = st * 1664525u + 1013904223u;
      const bool tex = ((x / 96) + (y / 64)) % 3 == 0;
      const int base = (int)(128 + 60 * sin(x * 0.01) * cos(y * 0.013));
      for (int c = 0; c < 3; c++) {
        int v = base + (tex ? (int)((st >> (8 * c)) & 63) - 32 : 0) + 20 * c;
        rgb[((size_t)y * width + x) * 3 + c] = (uint8_t)std::min(std::max(v, 0), 255);
      }
    }
  jpxlg_img* img = nullptr;
  int rc = jpxlg_image_upload_srgb(ctx, width, height, 8, rgb.data(), &img);
  if (rc) return rc;
  std::unique_ptr<jpxlg_img, void (*)(jpxlg_img*)> hold(img, jpxlg_image_destroy);
  const jpxlg_quantizer q{3000, 2, 4, 3, 3, 1, 4.0f, 1.0f};
  rc = jpxlg_plan_structure(img, &q, JPXLG_COVER_FIXED_DCT8, 0);
  if (rc) return rc;
  rc = jpxlg_quantize(img, &q);
  if (rc) return rc;
  const jpxlg_restoration r{1, 0, 7};
  const int cls = tune_class((double)width * height);
  const jpxlg_ctx::Shape keep = ctx->tune[cls];
  const bool prof = ctx->prof;
  const std::string only = ctx->prof_only;
  // the vertical pass's workgroups have four waves; the order is tuned
  static const jpxlg_ctx::Shape kCand[2] = {{4, 0}, {4, 1}};
  double best_ms = 1e30, ref_score = 0.0;
  jpxlg_ctx::Shape best = keep;
  bool first = true;
  // earlier timing records are folded away before the candidates run
  rc = jpxlg_profile_read(ctx, nullptr, nullptr, nullptr, 0);
  for (int a = 0; a < 2 && rc >= 0; a++) {
    ctx->tune[cls] = kCand[a];
    double score = 0.0;
    ctx->prof = false;
    for (int it = 0; it < 2 && rc >= 0; it++) rc = jpxlg_render_score(img, &r, &score, nullptr, nullptr, nullptr);
    if (rc < 0) break;
    if (first) ref_score = score;
    if (score != ref_score) {
      rc = fail(ctx, JPXLG_E_DEVICE, "a launch shape changed the metric sums");
      break;
    }
    ctx->prof = true;
    ctx->prof_only = "blur_pool";
    for (int it = 0; it < 5 && rc >= 0; it++) rc = jpxlg_render_score(img, &r, &score, nullptr, nullptr, nullptr);
    char names[48 * 4] = {0};
    double ms[4] = {0, 0, 0, 0};
    uint64_t launches[4] = {0, 0, 0, 0};
    const int nk = rc < 0 ? 0 : jpxlg_profile_read(ctx, names, ms, launches, 4);
    ctx->prof = false;
    ctx->prof_only.clear();
    if (nk <= 0) {
      if (rc >= 0) rc = nk < 0 ? nk : fail(ctx, JPXLG_E_DEVICE, "no timing records");
      break;
    }
    if (ms[0] < best_ms) {
      best_ms = ms[0];
      best = ctx->tune[cls];
    }
    first = false;
  }
  ctx->prof = prof;
  ctx->prof_only = only;
  ctx->tune[cls] = rc < 0 ? keep : best;
  if (rc < 0) return rc;
  if (units_per_group) *units_per_group = best.upw;
  if (swizzle) *swizzle = best.swizzle;
  return JPXLG_OK;
}
