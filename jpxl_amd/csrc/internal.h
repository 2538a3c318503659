// Internal types shared by the host runtime and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "dct_matrix.h"
#include "device_math.h"
#include "predictor.h"

namespace jpxlg {

// Square transform sizes on the path: DCT8x8, DCT16x16, DCT32x32.  Index 0,
// 1, 2; `cell_base` is each size's offset in the concatenated per-cell tables.
constexpr int kSides[3] = {8, 16, 32};
constexpr int kCellBase[3] = {0, 64, 64 + 256};
constexpr int kCellTotal = 64 + 256 + 1024;

__host__ __device__ constexpr int size_index(int side) { return side == 8 ? 0 : (side == 16 ? 1 : 2); }

// Frame geometry on the 8x8 block grid.
struct Geom {
  int w, h;        // pixels
  int bw, bh;      // 8x8 blocks
  int cstride;     // coefficient plane stride (8*bw)
  int crows;       // coefficient plane rows (8*bh)
  int tw, th;      // 64x64 chroma-from-luma tiles
};

// Constant tables uploaded once per context (device memory).
struct ConstTables {
  DctTables dct;
  float matrix[3][kCellTotal];    // dequantization matrices per channel, sizes 8/16/32 concatenated
  float donor[kCellTotal];        // quantizer-donor frequency weights
  uint16_t order[kCellTotal];     // natural coefficient order (position -> cell)
  float matrix_pos[3][kCellTotal];  // matrix and donor weights indexed by natural-order
  float donor_pos[kCellTotal];      // position (matrix[c][order[p]], donor[order[p]])
  float scale_f[3][4];            // ScaleF(c, n) for n = 1, 2, 4 (index by size), c < n
  double fir[5];                  // metric blur taps h[0..4] (symmetric), f64
};

// Per-probe quantizer parameters (device memory, written before each probe).
struct ProbeParams {
  float sc[3];          // Mul * qm[c]: the I.5.3 step scale
  double lambda[3];     // Lagrange multiplier per channel (bits per squared sample error)
  float mdc[3];         // I.2.1 LF multipliers
  float mul;            // I.5.3 Mul (hf_multiplier)
  float sigma;          // J.4.3 sigma for the uniform sharpness
  float zero_token_bits;
  int32_t truncate;     // trailing-truncation pass on/off
  // Adaptive quantization: with an AQ field (aq_adj non-null, one f32
  // adjustment per 8x8 atom) every varblock carries its own HfMul from the
  // field over its footprint, around the wire baseline; sc/sigma above are
  // the baseline's.  Null: every varblock at the baseline.
  const float* aq_adj;
  uint32_t gs;           // the wire global_scale
  uint32_t aq_baseline;  // the wire baseline HfMul
  int32_t aq_refine_only, aq_fine;
  float qm[3];           // per-channel qm multipliers (Y: 1)
  float sig_lut;         // the EPF sharpness LUT entry of sigma
  // The cover objective (CoverRateModel, request.rs:585-636; cell weights,
  // request.rs:650-690): residual-bit scale and fixed bits per DCT8/16/32,
  // and flat (1) or quantizer-donor (0) cell weights.
  double cover_rate[3];
  double cover_fixed[3];
  int32_t cover_flat;
  // A plan supplied by the caller: each varblock's HfMul at its top-left
  // block (bw x bh, u32), used instead of the AQ field when non-null.
  const uint32_t* vb_mul;
};

// field.rs:383-395 mul_of: baseline * 2^-adj rounded to the nearest integer,
// clamped to [1, 2 * baseline] (an unrepresentable result keeps the baseline).
__host__ __device__ inline uint32_t aq_mul_of(uint32_t base, float adj) {
  const double v = round((double)base * exp2((double)-adj));
  uint32_t mul = !(v > 0.0) ? 0u : (v >= 4294967295.0 ? 0xffffffffu : (uint32_t)v);
  const uint32_t hi = base > 0x7fffffffu ? 0xffffffffu : base * 2u;
  mul = mul < 1u ? 1u : (mul > hi ? hi : mul);
  return mul <= 0x7fffffffu ? mul : base;
}

// field.rs:338-378 mul_for_footprint over the n x n atoms at (bx, by): the
// f32 mean of the adjustments (rows, then columns; zero outside the grid),
// snapped to half octaves (legacy) or clamped to one octave (fine), capped
// at zero under refine-only.
// One sample of a per-block field in device memory (global loads on the
// device: the pointer comes out of a parameter block, where the compiler
// cannot see its address space).
__host__ __device__ inline float field_at(const float* p, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *(const __attribute__((address_space(1))) float*)((const __attribute__((address_space(1))) char*)p + 4u * i);
#else
  return p[i];
#endif
}

__host__ __device__ inline uint32_t u32_at(const uint32_t* p, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *(const __attribute__((address_space(1))) uint32_t*)((const __attribute__((address_space(1))) char*)p + 4u * i);
#else
  return p[i];
#endif
}

__host__ __device__ inline uint32_t aq_footprint_mul(const ProbeParams* pp, int bw, int bh, int bx, int by, int n) {
  if (pp->vb_mul) return u32_at(pp->vb_mul, (uint32_t)(by * bw + bx));
  const float* adj = pp->aq_adj;
  float sum = 0.0f;
  for (int dy = 0; dy < n; dy++)
    for (int dx = 0; dx < n; dx++) {
      const int x = bx + dx, y = by + dy;
      sum += (x < bw && y < bh) ? field_at(adj, (uint32_t)(y * bw + x)) : 0.0f;
    }
  float a = sum / (float)(n * n);
  if (pp->aq_fine) a = a < -1.0f ? -1.0f : (a > 1.0f ? 1.0f : a);
  else a = roundf(a * 2.0f) / 2.0f;
  if (pp->aq_refine_only) a = fminf(a, 0.0f);
  return aq_mul_of(pp->aq_baseline, a);
}

// The I.5.3 step scales (Mul * qm[c]) and the J.4.3 sigma of the varblock of
// side 8n at block (bx, by) (reconstruct.rs:27-36, restoration.rs:338-349).
__device__ inline void varblock_scales(const ProbeParams* pp, int bw, int bh, int bx, int by, int n, float* sc,
                                       float* sigma) {
  if (!pp->aq_adj && !pp->vb_mul) {
    sc[0] = pp->sc[0];
    sc[1] = pp->sc[1];
    sc[2] = pp->sc[2];
    if (sigma) *sigma = pp->sigma;
    return;
  }
  const uint32_t m = aq_footprint_mul(pp, bw, bh, bx, by, n);
  const double den = (double)pp->gs * (double)m;
  const float mul = den <= 0.0 ? 0.0f : (float)(65536.0 / den);
  sc[0] = mul * pp->qm[0];
  sc[1] = mul;
  sc[2] = mul * pp->qm[2];
  if (sigma) *sigma = mul * 0.46f * pp->sig_lut;
}

// lib.rs:4569-4577 / :1933-1938: the cover's metadata bits for a varblock's
// HfMul (8 for a non-baseline multiplier on the legacy lattice, a flat 2 on
// the fine one, none without a field).
__device__ inline double aq_signal_bits(const ProbeParams* pp, int bw, int bh, int bx, int by, int n) {
  if (!pp->aq_adj) return 0.0;
  if (pp->aq_fine) return 2.0;
  return aq_footprint_mul(pp, bw, bh, bx, by, n) == pp->aq_baseline ? 0.0 : 8.0;
}

// Adaptive-quantization modes (field.rs:52-91 AqMode) and the device
// scratch of the field's construction.
enum AqMode { kAqOff = 0, kAqMasking = 1, kAqUniform = 2, kAqFineMasking = 3, kAqFineUniform = 4, kAqEdgeRefine = 5 };
struct AqScratch {
  float mean;       // the frame mean of the (eroded) activities
  int32_t certified;
  int32_t active;   // some atom leaves the baseline (the field is not neutral)
};

// Counts of a structure built on the device (device memory): varblocks per
// size and the CfL LF / HF sample totals.
struct StructInfo {
  int count[3];
  uint32_t lf_total, hf_total;
};

// Chroma-from-luma structure of a plan (device memory).
struct CflParams {
  int32_t x_factor_lf, b_factor_lf;  // unbiased (-128..127)
};

// Device scratch of the frame-LF CfL refinement: the candidate factors of
// each channel and their residual-bit prices.
constexpr int kCflLfFastGroups = 256;
struct CflLfScratch {
  int32_t cands[2][4];
  int32_t ncand[2];
  unsigned long long cost[2][4];
  // Certified fast path: per-workgroup double-double partial sums (yy, yx,
  // xx, yb, bb, |yx|, |yb|), and whether their decision stands for the
  // sequential sums'.
  double2 part[kCflLfFastGroups][7];
  int32_t certified;
};

// One image of a probe batch: everything the per-probe chain (quantize,
// render, restoration, colour, blur + pooling, reduction) reads or writes.
// A probe of one image is a batch of one.
constexpr int kDescScales = 6;
struct ProbeDesc {
  Geom g;
  int nscales;
  int sw[kDescScales], sh[kDescScales];
  const uint32_t* list[3];  // varblock lists per size, entries (by << 16) | bx
  int count[3];
  const float* coeff[3];
  int32_t* quant[3];
  int32_t* lf[3];
  float* rx[3];             // render planes and their restoration ping-pong
  float* aux[3];
  float* sigma;
  const int8_t* cfl_x;
  const int8_t* cfl_b;
  const CflParams* cflp;
  const ProbeParams* pp;
  uint32_t* err;
  float* cand_x[kDescScales][3];    // the candidate's positive XYB pyramid
  double* partials;                 // 18 f64 per pooling unit (column), scale after scale
  int ptile_off[kDescScales];       // first unit of each scale in `partials`
  int punits[kDescScales];          // pooling units (columns) per scale
  double* sums;                     // 18 f64 per scale
  const float* thr;                 // depth quantizer thresholds and levels
  const float* lut;
  int levels;
};

// The descriptors of a batched launch are written by the host before the
// launch and never by a kernel.  Read through the constant address space,
// every pointer they hold is known to point to global memory, so the
// compiler emits global loads and stores for the planes rather than flat
// ones (which would also count against the LDS counter: every LDS wait
// would wait for them as well).
using DescSpace = const __attribute__((address_space(4))) ProbeDesc;
__device__ __forceinline__ DescSpace& desc_at(const ProbeDesc* D, int d) { return ((DescSpace*)D)[d]; }
__device__ __forceinline__ Geom geom_of(const __attribute__((address_space(4))) Geom& s) {
  Geom g;
  g.w = s.w, g.h = s.h, g.bw = s.bw, g.bh = s.bh, g.cstride = s.cstride, g.crows = s.crows, g.tw = s.tw, g.th = s.th;
  return g;
}

// Loads through a pointer the compiler cannot place, stated to be global
// (the 32-bit element index keeps a wave-uniform base in scalar registers).
template <class T>
__device__ __forceinline__ T ldg32(const T* p, uint32_t i) {
  return *(const __attribute__((address_space(1))) T*)((const __attribute__((address_space(1))) char*)p +
                                                         i * (uint32_t)sizeof(T));
}

template <class T>
__device__ __forceinline__ void stg32(T* p, uint32_t i, T v) {
  *(__attribute__((address_space(1))) T*)((__attribute__((address_space(1))) char*)p + i * (uint32_t)sizeof(T)) = v;
}

// Workgroup prefix sums of one batched launch: the workgroups of desc d are
// [off[d], off[d+1]).  The desc of a workgroup is found by binary search.
__device__ __forceinline__ int find_desc(const int* __restrict__ off, int nd, int b) {
  int lo = 0, hi = nd;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Error word bits raised by kernels (device memory).
enum : uint32_t { kErrQuantRange = 1u };

// ----------------------------------------------------------------------------
// Kernel launchers (defined in kernels_pixel.hip / kernels_metric.hip).
// ----------------------------------------------------------------------------

void launch_xyb_fwd(hipStream_t s, const Geom& g, int bits, const void* rgb, const float* lut,
                    float* xyb[3], float* lin[3], int* gray);
// `form` is a DctForm: the reference's butterflies (bit-exact) or one of the
// matrix-core contractions of dct_matrix.h.
// Forward DCT of a size's varblocks: count from the host, or (si non-null)
// from the device structure over an upper-bound grid.
void launch_dct_fwd(hipStream_t s, int side, int form, const Geom& g, const uint32_t* list, int count,
                    const StructInfo* si, const ConstTables* tab, const DctMats* mats, float* const xyb[3],
                    float* coeff[3]);
void launch_dct_blocks(hipStream_t s, int side, int inverse, int count, const ConstTables* tab,
                       const float* in, float* out);
// The per-probe chain over a batch of ProbeDescs (one image = a batch of
// one).  `off` (device) holds the workgroup prefix of each desc, `nwg` the
// total; src selects the restoration ping-pong input (0: rx, 1: aux).
// Quantization of every varblock size in one launch (DCT32 workgroups first):
// per-size prefixes off[0..2] (sides 8, 16, 32), nwg[0..2] workgroups.  With
// render, each workgroup also reconstructs its varblocks in the exact DCT
// form (render planes and EPF sigma, as launch_render_all).
void launch_quantize_all(hipStream_t s, const ProbeDesc* D, const int* const off[3], int nd, const int nwg[3],
                         const ConstTables* tab, bool render);
// Reconstruction of every varblock size in one launch (DCT32 workgroups
// first): per-size prefixes off[0..2] (sides 8, 16, 32), nwg[0..2] workgroups.
void launch_render_all(hipStream_t s, int form, const ProbeDesc* D, const int* const off[3], int nd,
                       const int nwg[3], const ConstTables* tab, const DctMats* mats);
void launch_gaborish_b(hipStream_t s, const ProbeDesc* D, const int* off, int nd, int nwg, int src);
void launch_epf_b(hipStream_t s, int step, const ProbeDesc* D, const int* off, int nd, int nwg, int src);
// The colour pass of every scale in one launch: one workgroup per 32x32 tile
// of scale 0 (pyramid_tiles per image, prefix `off`).
constexpr int kMaxPyr = 6;
int pyramid_tiles(int w, int h);
void launch_colour_pyr_b(hipStream_t s, const ProbeDesc* D, const int* off, int nd, int nwg, int src);
// The reference pyramid's positive XYB of one image from its linear RGB.
void launch_colour_pyr_ref(hipStream_t s, int ns, const int* sw, const int* sh, float* const in[3],
                           float* const (*out)[3]);
void launch_reduce_b(hipStream_t s, const ProbeDesc* D, int nd);
// Workgroup counts of the batched stages for one image.
int quant_groups(int side, int count);
int render_groups(int side, int count);
int epf_tiles(int w, int h);
int gaborish_blocks(int w, int h);
void launch_fill(hipStream_t s, float* p, float v, size_t n);
void launch_gaborish(hipStream_t s, int w, int h, float* const in[3], float* out[3]);
// One Jacobi step of the encoder's Gaborish inverse: out = x + (f - Kx)/centre.
void launch_gaborish_jacobi(hipStream_t s, int w, int h, float* const f[3], float* const x[3], float* out[3]);
void launch_epf(hipStream_t s, int step, int w, int h, float* const in[3], const float* sigma,
                float* out[3]);

// Colour + metric conversion at one scale.  mode 0: input is XYB (render
// output) -> opsin inverse -> depth quantizer -> linear; mode 1: input is
// linear RGB.  Writes positive XYB at this scale and, when `down` is non-null,
// the 2:1 box-filtered linear RGB of the next scale.
void launch_colour_metric(hipStream_t s, int mode, int w, int h, float* const in[3],
                          const float* thresholds, const float* depth_lut, int levels,
                          float* pxyb[3], float* down[3]);
// XYB -> linear RGB at the source depth, one plane set (test download).
void launch_linear_at_depth(hipStream_t s, int w, int h, float* const in[3], const float* thresholds,
                            const float* depth_lut, int levels, float* out[3]);
// Hierarchical cover (kernels_plan.hip): candidate costs for every square
// (f64, indexed by the candidate's top-left block) and the per-region
// quadtree decision written into the block map.
// Costs are double-double sums with a bound (rad*); regions the bounds
// cannot settle are queued (queue[0] = count, zeroed by the caller) and
// recomputed with the reference's sequential sums (every region when
// all_exact, the self-check mode).
void launch_cover(hipStream_t s, const Geom& g, const ConstTables* tab, const ProbeParams* pp,
                  float* const xyb[3], double* cost8, double* cost16, double* cost32, double* rad8,
                  double* rad16, double* rad32, uint32_t* queue, uint8_t* bmap, uint32_t* err, int all_exact);
// The structure's work lists (in CfL order), counts, and each varblock's CfL
// sample slots (LF offset per block; HF offset inside its tile, per block;
// tile bases), from the device block map.
// rows: 2 * structure_rows(g) uint4 of scratch (per-row counts and bases).
int structure_rows(const Geom& g);
void launch_structure_lists(hipStream_t s, const Geom& g, const uint8_t* bmap, uint32_t* const list[3],
                            uint32_t* lf_off, uint32_t* hf_local, uint32_t* tile_count, uint32_t* tile_base,
                            StructInfo* si, uint4* rows);
// One contiguous readback block of a structure: StructInfo at 0, CflParams at
// 32, the error word at 48, cfl_x and cfl_b (tiles bytes each) from 64, then
// the block map.
size_t structure_readback_bytes(int tiles, int bpx);
void launch_gather_structure(hipStream_t s, const StructInfo* si, const CflParams* cp, const uint32_t* err,
                             const int8_t* cx, const int8_t* cb, const uint8_t* bmap, int tiles, int bpx,
                             char* out);
// Chroma-from-luma samples of every listed varblock, in the reference's
// accumulation order (one workgroup per list entry up to each size's bound).
void launch_cfl_collect(hipStream_t s, const Geom& g, const ConstTables* tab, const ProbeParams* pp,
                        uint32_t* const list[3], const StructInfo* si, const uint32_t* lf_off,
                        const uint32_t* hf_local, const uint32_t* tile_base, float* const coeff[3], float4* lf_s,
                        float4* hf_s, uint16_t* hf_cell, uint32_t* err);
// The per-64x64-tile HF factors from each tile's sample range.
void launch_cfl_tiles(hipStream_t s, const Geom& g, const ConstTables* tab, const ProbeParams* pp,
                      const uint32_t* tile_base, const float4* hf_s, const uint16_t* hf_cell,
                      int8_t* cfl_x, int8_t* cfl_b, const int* gray);
// The frame-LF factors, on their own stream: sequential regression, then a
// grid-wide pricing of the candidates, then the selection.
void launch_cfl_lf(hipStream_t s, const ProbeParams* pp, const StructInfo* si, const float4* lf_s,
                   CflLfScratch* scratch, CflParams* cflp, const int* gray);
// The AQ field of `mode` from the source XYB planes into adj (bw x bh);
// act/ero are bw x bh scratch planes; scratch->active must be zero before.
void launch_aq_field(hipStream_t s, const Geom& g, int mode, float* const xyb[3], float* act, float* ero,
                     AqScratch* scratch, float* adj);
// Each listed varblock's HfMul at its top-left block.
// Exact order statistics of the per-atom variances: kSelectQueries ranks,
// each of set 0 (luma) or 1 (chroma), and the count of luma below a bound.
constexpr int kSelectQueries = 4;
struct SelectRanks {
  uint32_t rank[kSelectQueries];
  int set[kSelectQueries];
};
struct SelectScratch {
  uint32_t h1[2][2048];                // digit 1 (bits 31-21) of luma, chroma
  uint32_t h2[kSelectQueries][2048];   // digit 2 (bits 20-10) per query
  uint32_t h3[kSelectQueries][1024];   // digit 3 (bits 9-0) per query
  uint32_t flat;
  uint32_t pad_;
  uint32_t prefix[kSelectQueries];
  uint32_t rank[kSelectQueries];
  float value[kSelectQueries];
};
void launch_select(hipStream_t s, const float* luma, const float* chroma, int n, float flat_below,
                   const SelectRanks& r, SelectScratch* sc);
// Whole 32x32 preanalysis cells of a w x h frame, and their integer sums
// from the source samples (8-bit, or 9..16-bit reduced to their top 8 bits).
int preanalysis_cells(int w, int h);
void launch_preanalysis_cells(hipStream_t s, int w, int h, int bits, const void* rgb, CellStats* out);
// The per-block first-guess features in one pass: each atom's Y variance
// and X + B variance, the DCT8 summary's summary_groups(g) partials and its
// kSummaryBins-bin histogram (cleared first).
int summary_groups(const Geom& g);
void launch_block_features(hipStream_t s, const Geom& g, const ConstTables* tab, float* const xyb[3],
                           float* luma_var, float* chroma_var, SummaryPartial* parts, uint32_t* hist);
// The QuantizedFrameIr arena (every varblock's X, Y, B coefficients in LF
// group then block-raster order) and the per-group LF planes.
void launch_gather_ir(hipStream_t s, const Geom& g, uint32_t* const list[3], const int counts[3],
                      const uint32_t* lf_off, int32_t* const quant[3], int32_t* const lf[3], int32_t* arena,
                      int32_t* lf_out);
// Token-ready runs of the quantized plan (count, prefix sum, write): len_nz
// and start are indexed 3 * IR block offset + channel (start has one more
// entry, the total); arena receives the runs.
void launch_runs(hipStream_t s, const Geom& g, uint32_t* const list[3], const int counts[3], const uint32_t* lf_off,
                 int32_t* const quant[3], const ConstTables* tab, uint32_t* len_nz, uint32_t* start,
                 int32_t* arena);
void launch_aq_block_muls(hipStream_t s, const Geom& g, const ProbeParams* pp, uint32_t* const list[3],
                          const int counts[3], uint32_t* muls);
// One scale and channel of the metric's blur passes: the planes each pass
// reads and writes.  Transposed planes hold column x at x * pitch (pitch =
// h rounded up to even, so that pairs of samples load as one 8-byte word).
struct LineSet {
  int w, h, pitch;
  const float* x1;     // horizontal pass: the reference plane (row-major)
  const float* x2;     // horizontal pass (candidate): the candidate plane (row-major)
  float* hT[3];        // horizontal outputs per moment (transposed)
  float* xT;           // horizontal pass: its first moment's source, transposed
  const float* x1T;    // vertical pass (candidate): reference and candidate planes
  const float* x2T;    //   and the reference's mu1 and s11, transposed
  const float* mu1T;
  const float* s11T;
  float* out[2];       // vertical pass (reference): mu1 and s11 (transposed)
  double* partials;    // vertical pass (candidate): six sums per wave at unit * 18
};
// Workgroups of the horizontal pass over a w x h plane (bands of rows).
int blur_bands(int w, int h);
// Line pitch of the transposed planes of a plane h samples high, and their
// size in floats.
inline int line_pitch(int h) { return (h + 1) & ~1; }
inline size_t transposed_floats(int w, int h) { return (size_t)w * line_pitch(h); }
// Workgroups of the vertical pass over a w x h plane (groups of columns, or
// one long column each), and the partial rows of six sums per channel they
// leave (one per wave).
int blur_columns(int w, int h);
int pool_units(int w, int h);
// The blur passes over `nsets` line sets (device array) with workgroup
// prefixes `off` (nsets + 1 ints, device): the horizontal pass in bands
// (blur_bands per set), the vertical pass in column groups (blur_columns per
// set, with or without the XCD swizzle of the workgroup order).  cand 0: the
// reference's moments (x1, x1^2), 1: the candidate's (x2, x2^2, x1 x2) with
// the SSIM and edge pooling.
void launch_blur_h(hipStream_t s, int cand, const ConstTables* tab, const LineSet* sets, const int* off, int nsets,
                   int bands);
void launch_blur_v(hipStream_t s, int cand, const ConstTables* tab, const LineSet* sets, const int* off, int nsets,
                   int groups, int swizzle);
constexpr int kMaxPoolScales = 6;
// Where each scale's units start in the partials buffer.
struct PoolLayout {
  int nscales;
  int off[kMaxPoolScales];
  int tiles[kMaxPoolScales];
};
// Sums every scale's partials into 18 f64 sums per scale, in a fixed order.
void launch_reduce_scales(hipStream_t s, const PoolLayout& L, const double* partials, double* sums);

}  // namespace jpxlg
