/*
 * jpxlg.h -- C ABI of the MI355X (gfx950) VarDCT pixel path and SSIMULACRA2
 * quality controller.
 *
 * The library keeps each image's XYB planes, forward-DCT coefficients and
 * SSIMULACRA2 reference pyramid resident in HBM, and runs the
 * quantize -> reconstruct -> score chain of the quality search on the GPU.
 * Entropy coding and bitstream writing run on the host: the plan handed back
 * (quantized coefficients, LF planes, block map, CfL factors) is the input
 * of the reference's `attach_entropy`, and jpxlg_emit_codestream /
 * jpxlg_encode_emit write the JPEG XL codestream from it on the CPU.
 *
 * Conventions: plain C types, caller-owned host buffers, int32 status
 * (0 = OK, negative = JPXLG_E_*).  One context = one device + one HIP stream;
 * a context is not shared across threads, contexts are independent.
 *
 * Reference interfaces replaced (paths relative to JPXL/crates/):
 *   jpxlg_image_upload_srgb   PreparedFrame::from_srgb8_with / from_srgb16_with
 *                             (jpxl-encode-policy/src/source.rs:226, :355) and
 *                             PlanRenderEvaluator::from_srgb8 / from_srgb16
 *                             (jpxl-perceptual/src/evaluator.rs:87-122)
 *   jpxlg_plan_structure      the cover + CfL half of build_pixel_plan
 *                             (jpxl-encode-policy/src/lib.rs:555-845)
 *   jpxlg_quantize            quantize_group / quantize_square_varblock
 *                             (jpxl-encode-policy/src/lib.rs:4469, :3237)
 *   jpxlg_render_score        PerceptualEvaluator::evaluate =
 *                             PlanRenderer::render_linear_at_depth_with +
 *                             Ssimulacra2::score
 *                             (jpxl-encode-policy/src/quality.rs:87-157,
 *                              jpxl-plan-render/src/lib.rs:680,
 *                              jpxl-perceptual/src/ssimulacra2.rs:79)
 *   jpxlg_plan_download       the QuantizedFrameIr hand-off to attach_entropy
 *                             (jpxl-encode/src/vardct/plan.rs:614-636)
 *   jpxlg_emit_codestream     attach_entropy + emit_codestream_with_executor
 *                             (jpxl-encode-policy/src/lib.rs:1012,
 *                              jpxl-encode/src/vardct/write.rs:350)
 *   jpxlg_encode_bytes        rate::search_frame (byte target, exhaustive ladder)
 *                             (jpxl-encode-policy/src/rate.rs:1143, :387-921)
 *   jpxlg_encode_quality      search_frame_perceptual
 *                             (jpxl-encode-policy/src/quality.rs:2405)
 *   jpxlg_score_pair          jpxl_perceptual::score_pair
 *                             (jpxl-perceptual/src/lib.rs:195)
 *   jpxlg_dct_blocks          dct_2d_in_place / idct_2d_in_place
 *                             (jpxl-core/src/dct.rs:956, :982)
 *   jpxlg_epf_step            epf_step_rows (jpxl-core/src/restoration.rs:583)
 *   jpxlg_gaborish            gaborish_into (jpxl-core/src/restoration.rs:148)
 */
#ifndef JPXLG_H
#define JPXLG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JPXLG_OK 0
#define JPXLG_E_INVALID (-1)
#define JPXLG_E_UNSUPPORTED (-2)
#define JPXLG_E_OOM (-3)
#define JPXLG_E_DEVICE (-4)
#define JPXLG_E_TOO_SMALL (-5)
#define JPXLG_E_RANGE (-6)

typedef struct jpxlg_ctx jpxlg_ctx;
typedef struct jpxlg_img jpxlg_img;

/* I.5.3 / I.2.1 quantizer scalars of one probe plus the quantizer rule.
 * quantize.rs:301-366, lib.rs:2033-2172, request.rs:855-972. */
typedef struct {
  uint32_t global_scale;   /* 1..73728 */
  uint32_t hf_mul;         /* >= 1 */
  uint32_t quant_lf;       /* >= 1 */
  uint32_t x_qm_scale;     /* u(3), 2 = neutral */
  uint32_t b_qm_scale;     /* u(3), 2 = neutral */
  int32_t truncate;        /* 1 = QuantizerChoiceMode::TrailingTruncation */
  float lambda_scale;      /* request.lambda_scale (4.0 on the quality path) */
  float zero_token_bits;   /* request.zero_token_bits */
} jpxlg_quantizer;

/* J.1 restoration decisions of a plan (plan.rs:221-228, request.rs:905). */
typedef struct {
  int32_t epf_iters;   /* 0..3 */
  int32_t gaborish;    /* 0/1 */
  int32_t sharpness;   /* uniform EPF sharpness 0..7 */
} jpxlg_restoration;

/* Cover modes (request.rs CoverMode).  HIERARCHICAL prices candidates as the
 * quality and target-rate requests do (the calibrated rate model and the
 * quantizer-donor cell weights, request.rs:912-925); HIERARCHICAL_LEGACY as
 * the fixed-quantizer request does (EncodeRequest::defaults: the legacy rate
 * proxy and flat cell weights, request.rs:855-880). */
#define JPXLG_COVER_FIXED_DCT8 0
#define JPXLG_COVER_HIERARCHICAL 1
#define JPXLG_COVER_HIERARCHICAL_LEGACY 2

/* ---- context ------------------------------------------------------------ */
int32_t jpxlg_ctx_create(int32_t device, jpxlg_ctx** out);
/* Destroy every image created on a context before the context itself: an
 * image keeps a pointer to its context's stream and tables. */
void jpxlg_ctx_destroy(jpxlg_ctx* ctx);
const char* jpxlg_last_error(const jpxlg_ctx* ctx);
/* Bytes of device memory an image of this size keeps resident. */
uint64_t jpxlg_image_footprint(uint32_t width, uint32_t height);

/* ---- image: XYB + metric reference resident in HBM ------------------------ */
/* rgb: interleaved RGB, row-major; bits 8 (uint8) or 9..16 (uint16). */
int32_t jpxlg_image_upload_srgb(jpxlg_ctx* ctx, uint32_t width, uint32_t height, uint32_t bits,
                                const void* rgb, jpxlg_img** out);
/* Same, reading the interleaved samples from device memory (already in HBM). */
int32_t jpxlg_image_from_device(jpxlg_ctx* ctx, uint32_t width, uint32_t height, uint32_t bits,
                                const void* device_rgb, jpxlg_img** out);
void jpxlg_image_destroy(jpxlg_img* img);
int32_t jpxlg_image_info(const jpxlg_img* img, uint32_t* blocks_x, uint32_t* blocks_y,
                         int32_t* grayscale, uint32_t* metric_scales);

/* ---- plan structure: block map + CfL, then the forward DCT --------------- */
/* block_map: blocks_x * blocks_y bytes; 8/16/32 at a varblock's top-left 8x8
 * block, 0 on the blocks it covers.  cfl_x/cfl_b: one int8 per 64x64 tile
 * (tiles_x = ceil(blocks_x/8)); lf_factors: {x_factor_lf-128, b_factor_lf-128}.
 * NULL cfl pointers mean the neutral (all-zero) factors. */
int32_t jpxlg_set_structure(jpxlg_img* img, const uint8_t* block_map, const int8_t* cfl_x,
                            const int8_t* cfl_b, const int32_t* lf_factors);
/* Builds the structure on the device: the cover (fixed DCT8 or hierarchical
 * R-D quadtree under `q`) and, when enable_cfl, the CfL estimate. */
int32_t jpxlg_plan_structure(jpxlg_img* img, const jpxlg_quantizer* q, int32_t cover_mode,
                             int32_t enable_cfl);

/* prepare_gaborish_frame (jpxl-encode-policy/src/lib.rs:2256-2268,
 * jpxl-encode/src/vardct/gaborish.rs:218-315): replaces the image's source XYB
 * planes by their 12-step Jacobi Gaborish inverse (for plans rendered with
 * Gaborish on).  Drops any structure: plan or set it again afterwards. */
int32_t jpxlg_precondition_gaborish(jpxlg_img* img);

/* ---- adaptive quantization ------------------------------------------------ */
/* The AQ field of the image (AqMode, jpxl-encode-policy/src/field.rs:52-91;
 * DesiredQuantField::from_atlas_tuned at the default tuning, field.rs:231,
 * over AnalysisAtlas::analyze of the source planes, analysis.rs:111).  With
 * an active field (a mode other than OFF whose field is not neutral) every
 * later cover, CfL estimate, quantize and render under a quantizer q uses
 * AqSetup::build's wire quantizer (lib.rs:1859-1908: global_scale / 2,
 * quant_lf * 2, baseline HfMul hf_mul * 2; factor 16 on the fine lattice)
 * and each varblock's HfMul from the field over its footprint
 * (mul_for_footprint, field.rs:338).  OFF is the quality path's setting
 * (request.rs:903), MASKING the fixed-quantizer default (request.rs:1097).
 * Call before jpxlg_precondition_gaborish (the field reads the source
 * planes); drops any structure.  *active (optional) = the field is active. */
#define JPXLG_AQ_OFF 0
#define JPXLG_AQ_MASKING 1
#define JPXLG_AQ_UNIFORM 2
#define JPXLG_AQ_FINE_MASKING 3
#define JPXLG_AQ_FINE_UNIFORM 4
#define JPXLG_AQ_EDGE_REFINE 5
int32_t jpxlg_set_aq(jpxlg_img* img, int32_t mode, int32_t* active);
/* The AQ side of the last quantized plan, for the entropy coder: hf_mul
 * (blocks_x * blocks_y, optional) holds each varblock's HfMul at its
 * top-left block and 0 on covered blocks (VarblockDecision::hf_mul,
 * lib.rs:736-750); *wire the quantizer the frame header signals;
 * *refine_only AqSetup's fallback flag. */
int32_t jpxlg_plan_aq(jpxlg_img* img, uint32_t* hf_mul, jpxlg_quantizer* wire, int32_t* refine_only);

/* ---- probe ---------------------------------------------------------------- */
int32_t jpxlg_quantize(jpxlg_img* img, const jpxlg_quantizer* q);

/* Arithmetic of the square DCTs (forward DCT of the structure, IDCT of the
 * render).  EXACT is the reference's butterfly network, bit for bit
 * (jpxl-core/src/dct.rs:956-997; the default).  FP32 and BF16X3 are the
 * matrix-core contractions Out = L In^T L^T of the precision study: FP32
 * products, or bf16 hi/lo split operands (hi*hi + hi*lo + lo*hi) accumulated
 * in FP32.  They have no counterpart in the reference.  Changing the form
 * recomputes the forward coefficients of an existing structure. */
#define JPXLG_DCT_EXACT 0
#define JPXLG_DCT_FP32 1
#define JPXLG_DCT_BF16X3 2
int32_t jpxlg_set_dct_form(jpxlg_img* img, int32_t form);
/* Reconstructs the last quantized plan at the source depth and scores it.
 * terms (optional) receives 3 * scales * 6 pooled terms in the metric's
 * (component, scale, norm, map) order; *nterms their count. */
int32_t jpxlg_render_score(jpxlg_img* img, const jpxlg_restoration* r, double* score,
                           double* raw_error, double* terms, uint32_t* nterms);

/* ---- plan hand-off and debugging downloads ------------------------------- */
/* quant: 3 planes of (8*blocks_x) x (8*blocks_y) int32 in the footprint layout
 * (cell k of a side-s varblock at block (bx,by) sits at row 8*by + k/s,
 * column 8*bx + k%s); lf: 3 planes of blocks_x * blocks_y int32. */
int32_t jpxlg_plan_download(jpxlg_img* img, int32_t* quant, int32_t* lf, uint8_t* block_map,
                            int8_t* cfl_x, int8_t* cfl_b, int32_t* lf_factors);

/* The QuantizedFrameIr hand-off (jpxl-encode/src/vardct/plan.rs:519-636),
 * the layout attach_entropy consumes (jpxl-encode-policy/src/lib.rs:1012):
 * coeff_arena (192 * blocks_x * blocks_y int32) holds every varblock's X, Y
 * and B coefficient arrays back to back, each in the array's own raster
 * order (VarblockCoefficients; the arena of VarblockCoefficients::from_arena
 * with starts start, start + n, start + 2n for n = side * side);
 * lf_groups (3 * blocks_x * blocks_y int32) holds each LF group's three
 * LfQuantPlanes planes in raster order, groups in raster order (a group is
 * 256 x 256 blocks, clipped at the frame edge); varblocks (one per varblock,
 * at most blocks_x * blocks_y) lists them in the IR's order: LF groups in
 * raster order, inside a group by the block-raster order of their top-left
 * blocks (the BlockInfo order of QuantizedLfGroup::coefficients).  Any
 * output pointer may be NULL. */
typedef struct {
  uint32_t lf_group;  /* raster index of the LF group */
  uint32_t bx, by;    /* top-left 8x8 block in the frame */
  uint32_t side;      /* 8, 16 or 32 */
  uint32_t hf_mul;    /* the varblock's HfMul */
  uint32_t reserved;
  uint64_t start;     /* its X array's offset in coeff_arena */
} jpxlg_varblock;
int32_t jpxlg_plan_download_ir(jpxlg_img* img, int32_t* coeff_arena, int32_t* lf_groups, jpxlg_varblock* varblocks,
                               uint32_t* n_varblocks);

/* ---- codestream emission (CPU) -------------------------------------------- */
/* What the frame signals besides the plan: the fields of FrameDecision,
 * QuantizerDecision, LfCorrelationDecision and RestorationDecision
 * (jpxl-encode/src/vardct/plan.rs:409-470) that the headers and LfGlobal
 * write (jpxl-encode/src/vardct/headers.rs:104-233, write.rs:643-682). */
typedef struct {
  uint32_t width, height;           /* frame size in pixels */
  uint32_t bits_per_sample;         /* ImageMetadata bit depth, 1..16 */
  uint32_t global_scale, quant_lf;  /* I.2.1 Quantizer */
  uint32_t x_qm_scale, b_qm_scale;  /* frame header, 0..7 */
  int32_t x_factor_lf, b_factor_lf; /* LfChannelCorrelation factors minus 128 (both 0: the default) */
  uint32_t gaborish;                /* restoration filter: Gaborish on (1) or off (0) */
  uint32_t epf_iters;               /* 0..3 */
  uint32_t sharpness;               /* every block's EPF sharpness, 0..7 */
} jpxlg_frame_desc;

/* CodestreamSizing (write.rs:289-320): total bytes and where they went. */
typedef struct {
  uint64_t total, image_headers, frame_header_bits, toc_bits;
  uint64_t lf_global, lf_groups, hf_global, pass_groups; /* section bytes by kind (one section: all in pass_groups) */
} jpxlg_emit_sizes;

/* attach_entropy (jpxl-encode-policy/src/lib.rs:1012) +
 * emit_codestream_with_executor (jpxl-encode/src/vardct/write.rs:350): the
 * JPEG XL codestream of a quantized plan given in the jpxlg_plan_download_ir
 * layout (coeff_arena, lf_groups, varblocks) with the frame's CfL tiles
 * (cfl_x, cfl_b as jpxlg_plan_download; NULL = all zero).  Runs on the host
 * with `threads` workers (<= 0: every core).  out NULL prices the stream
 * (price_codestream_with, write.rs:302): *out_len and sizes are filled.  A
 * capacity below the stream's size returns JPXLG_E_TOO_SMALL with *out_len
 * the size needed.  The varblocks must tile the block grid exactly. */
int32_t jpxlg_emit_codestream(const jpxlg_frame_desc* frame, const int32_t* coeff_arena, const int32_t* lf_groups,
                              const jpxlg_varblock* varblocks, uint32_t n_varblocks, const int8_t* cfl_x,
                              const int8_t* cfl_b, int32_t threads, uint8_t* out, size_t capacity, size_t* out_len,
                              jpxlg_emit_sizes* sizes);

/* The plan's HF coefficients as token-ready runs, the hand-off that the
 * codestream stage reads fastest.  No reference counterpart: the same
 * QuantizedFrameIr content (jpxl-encode/src/vardct/plan.rs:519-636)
 * trimmed to what attach_entropy codes (hf_coeff.rs:480-575 writes, per
 * varblock and channel, the non-zero count and then the coefficients in
 * natural order up to the last non-zero one).  runs[3 i + c] belongs to
 * varblock i of the jpxlg_plan_download_ir order and channel c (0 X, 1 Y,
 * 2 B): its HF coefficients at natural-order positions nblocks ..
 * nblocks + count - 1 (nblocks = (side / 8)^2, the LLF cells carried by the
 * LF planes) are run_arena[start .. start + count - 1]; the last of them is
 * non-zero and `nonzeros` of them are (count 0 when the channel has no
 * non-zero HF coefficient). */
typedef struct {
  uint64_t start;
  uint32_t count, nonzeros;
} jpxlg_run;

/* Reads the image's last quantized plan back as runs: run_arena
 * (capacity entries; *arena_len receives the entries used, at most 189 per
 * 8x8 block of the frame), runs (3 per varblock), lf_groups and varblocks
 * as jpxlg_plan_download_ir (any output may be NULL except that run_arena
 * and runs go together).  The device computes the runs; only they cross
 * PCIe.  JPXLG_E_TOO_SMALL when the arena does not fit capacity. */
int32_t jpxlg_plan_download_runs(jpxlg_img* img, int32_t* run_arena, size_t capacity, size_t* arena_len,
                                 jpxlg_run* runs, int32_t* lf_groups, jpxlg_varblock* varblocks,
                                 uint32_t* n_varblocks);

/* jpxlg_emit_codestream over runs: the same stream, byte for byte, as over
 * the coefficient arena the runs were taken from.  Each varblock's `start`
 * field is ignored; runs that break their definition (a count past the
 * array, a run outside arena_len, a zero last coefficient, a non-zero count
 * that disagrees) are refused with JPXLG_E_INVALID. */
int32_t jpxlg_emit_codestream_runs(const jpxlg_frame_desc* frame, const int32_t* run_arena, size_t arena_len,
                                   const jpxlg_run* runs, const int32_t* lf_groups, const jpxlg_varblock* varblocks,
                                   uint32_t n_varblocks, const int8_t* cfl_x, const int8_t* cfl_b, int32_t threads,
                                   uint8_t* out, size_t capacity, size_t* out_len, jpxlg_emit_sizes* sizes);

/* The frame fields of an image's last quantized plan for
 * jpxlg_emit_codestream: its wire quantizer (jpxlg_plan_aq), LF factors,
 * source depth and `r` (NULL: the restoration of the image's last render).
 * With jpxlg_plan_download_ir and the CfL tiles of jpxlg_plan_download this
 * is everything the codestream stage reads, so the emit can run on another
 * host thread while the device works on the next image. */
int32_t jpxlg_plan_frame_desc(jpxlg_img* img, const jpxlg_restoration* r, jpxlg_frame_desc* frame);

/* The same over an image's last quantized plan (attach_entropy + emit after
 * search_frame_perceptual, jpxl-encode-policy/src/lib.rs:5501-5600): the IR
 * is read back into the context's pinned staging, the frame signals the
 * plan's wire quantizer, CfL tiles and LF factors, and `r` is the
 * restoration the plan was rendered with (NULL: that of the image's last
 * render, e.g. the quality search's). */
int32_t jpxlg_encode_emit(jpxlg_img* img, const jpxlg_restoration* r, int32_t threads, uint8_t* out,
                          size_t capacity, size_t* out_len, jpxlg_emit_sizes* sizes);

/* PerceptualEvaluator::evaluate(&ValidatedPixelPlan) (quality.rs:87-157):
 * a plan built elsewhere (the CPU planner) replaces the image's structure and
 * levels, and the next jpxlg_render_score renders and scores exactly it.
 * wire: the quantizer the frame signals (global_scale, quant_lf, qm scales;
 * hf_mul is every varblock's HfMul when hf_mul is NULL); block_map, cfl_x,
 * cfl_b, lf_factors as jpxlg_set_structure; hf_mul (optional, blocks_x *
 * blocks_y) each varblock's HfMul at its top-left block; quant and lf as
 * jpxlg_plan_download returns them (footprint layout).  A later quantize or
 * structure call replaces the plan. */
int32_t jpxlg_set_plan(jpxlg_img* img, const jpxlg_quantizer* wire, const uint8_t* block_map, const int8_t* cfl_x,
                       const int8_t* cfl_b, const int32_t* lf_factors, const uint32_t* hf_mul, const int32_t* quant,
                       const int32_t* lf);

#define JPXLG_BUF_XYB 0          /* 3 x W x H f32: source XYB */
#define JPXLG_BUF_COEFF 1        /* 3 x 8bw x 8bh f32: forward DCT (footprint) */
#define JPXLG_BUF_RENDER_XYB 2   /* 3 x W x H f32: reconstructed XYB after restoration */
#define JPXLG_BUF_LINEAR 3       /* 3 x W x H f32: reconstructed linear RGB at depth */
#define JPXLG_BUF_SIGMA 4        /* bw x bh f32: EPF sigma */
#define JPXLG_BUF_REF_XYB0 5     /* 3 x W x H f32: metric reference positive XYB, scale 0 */
#define JPXLG_BUF_AQ_FIELD 6     /* bw x bh f32: the active AQ field's per-atom adjustment */
int32_t jpxlg_image_download(jpxlg_img* img, int32_t which, void* host, size_t bytes);

/* ---- quality controller ---------------------------------------------------- */
#define JPXLG_EFFORT_FAST 0
#define JPXLG_EFFORT_BALANCED 1
typedef struct {
  double target;        /* minimum SSIMULACRA2 score */
  int32_t effort;       /* JPXLG_EFFORT_* */
  int32_t max_probes;   /* 0 = the effort's budget */
  int32_t mode;         /* 0 = JPXL navigator, 1 = bisection over the ladder, 2 = the navigator with
                           its finalists priced exactly by the codestream stage (the reference's
                           default build: the smallest feasible stream wins, quality.rs:1886-2010) */
} jpxlg_quality_request;

/* Why the search stopped where it did (QualityStatus, quality.rs:440-470). */
#define JPXLG_QUALITY_MET 0                     /* met inside the accept band */
#define JPXLG_QUALITY_MET_ADJACENT_RUNGS 1      /* met; the next coarser rung was verified not to */
#define JPXLG_QUALITY_MET_WORK_CAP 2            /* met; the budget ran out before tightening */
#define JPXLG_QUALITY_SATURATED_TOP 3           /* even the finest rung misses the target */
#define JPXLG_QUALITY_SATURATED_FLOOR 4         /* even the coarsest rung meets the target */
#define JPXLG_QUALITY_UNDER_TARGET_WORK_CAP 5   /* nothing met inside the budget */
#define JPXLG_QUALITY_RESCUED_FRESH_STRUCTURE 6 /* the chosen plan was rebuilt with its own structure */

typedef struct {
  double score;              /* achieved score of the chosen plan */
  uint32_t global_scale, hf_mul, quant_lf;
  uint32_t rung;
  int32_t met;               /* target met */
  int32_t probes;
  double probe_scores[16];
  uint32_t probe_rungs[16];
  double probe_ms[16];
  int32_t status;            /* JPXLG_QUALITY_* */
  uint32_t predicted_rung;   /* the first probe's rung */
  int32_t structural_builds; /* cover + CfL builds (the anchor and any rebuilt finalist) */
  int32_t structure_rebuilt; /* the chosen plan's structure was rebuilt at its own rung */
  uint8_t probe_fresh[16];   /* the probe built its own structure */
  int32_t exact_prices;      /* finalists priced by the codestream stage (mode 2) */
  uint64_t bytes;            /* the chosen finalist's codestream size when priced, else 0 */
} jpxlg_quality_result;

int32_t jpxlg_encode_quality(jpxlg_img* img, const jpxlg_quality_request* req,
                             jpxlg_quality_result* out);
/* ---- byte-rate controller ------------------------------------------------- */
/* search_frame over a byte target (jpxl-encode-policy/src/rate.rs:1143) as
 * the exhaustive ladder search does it (rate.rs:2016-2140 over search_ladder,
 * :387-921): a cold bracket doubling the effective scale from the request's
 * rung (global scale 32768), aimed bisection, a fill above the best; every
 * candidate rung is planned at its own quantizer on the device (cover, CfL,
 * quantize under EncodeRequest::for_target, request.rs:883-910) and priced
 * exactly by the codestream stage.  The largest stream at or under the
 * target wins and stays quantized on the image for jpxlg_encode_emit. */
typedef struct {
  uint64_t target_bytes;
  int32_t effort;       /* JPXLG_EFFORT_*: the request's chroma policy */
  uint32_t max_prices;  /* 0: RateSearchBudget::default (40), of which the ladder spends 24 */
} jpxlg_rate_request;

typedef struct {
  uint32_t rung, global_scale, hf_mul, quant_lf;
  uint64_t bytes;          /* the chosen plan's codestream size */
  int32_t saturated;       /* the finest rung fits the target */
  int32_t prices;          /* candidates priced */
  uint32_t trace_rungs[64];
  uint64_t trace_bytes[64];
  uint64_t floor_bytes;    /* JPXLG_E_RANGE for a target below the coarsest rung: that rung's size */
} jpxlg_rate_result;

int32_t jpxlg_encode_bytes(jpxlg_img* img, const jpxlg_rate_request* req, jpxlg_rate_result* out);

/* ---- first-guess features and the crossing predictor ---------------------- */
/* SourceFeatures (jpxl-encode-policy/src/quality_features.rs:21-47) with the
 * PreanalysisFeatures it carries (preanalysis.rs:99-137): the atlas's
 * per-atom variance quantiles, and the 32x32-cell structure statistics of
 * the source's 8-bit samples (has_preanalysis = 0 without a whole cell). */
typedef struct {
  uint32_t width, height;
  int32_t grayscale;
  float luma_variance_q10, luma_variance_q50, luma_variance_q90;
  float chroma_variance_q50, flat_fraction, edge_proxy;
  int32_t has_preanalysis;
  uint32_t cells;
  float share[7];  /* Flat, Gradient, ChromaCritical, Noisy, Texture, DirectionalEdge, TextLike */
  float log2var_mean, log2var_std, log2var_q10, log2var_q90;
  float edge_mean, edge_q90, flat_mean, noise_mean, noise_q90;
  float chroma_log2_mean, chroma_log2_q90;
  float orient_entropy_mean, dir_dominance_mean, axis_aligned_mean, hetero_frac;
} jpxlg_source_features;

/* TransformFeatureSummary (quality_features.rs:93-121) of the frame's
 * aligned DCT8x8 coefficients. */
typedef struct {
  uint64_t blocks;
  double ln_ac_y_mean, ln_ac_y_q50, ln_ac_y_q90, ln_ac_y_q99;
  double high_low_ratio, directional_asymmetry, chroma_ac_ratio;
  double near_zero_frac_1e3, near_zero_frac_1e2, dc_variance_y;
} jpxlg_transform_summary;

/* QualityPredictionV2 (quality_prediction.rs:48-77).  ood_flags: bit i for
 * feature i of the model's vector (the qpv2-stp/1 order of
 * quality_prediction.rs:373-385), bit 62 tiny_frame, bit 63 case_spread. */
#define JPXLG_OOD_TINY_FRAME (1ull << 62)
#define JPXLG_OOD_CASE_SPREAD (1ull << 63)
#define JPXLG_FALLBACK_NONE 0
#define JPXLG_FALLBACK_OOD_FEATURE 1
#define JPXLG_FALLBACK_WIDE_INTERVAL 2
#define JPXLG_FALLBACK_SATURATION_RISK 3
#define JPXLG_PREDICTOR_DISPATCH 0  /* quality_prediction::predict (case table, else knot model) */
#define JPXLG_PREDICTOR_CASES 1     /* predict_cases */
#define JPXLG_PREDICTOR_KNOTS 2     /* predict_v2 */
typedef struct {
  uint32_t median_rung, candidate_rung, interval_low, interval_high;
  double local_loss_exponent, saturation_risk;
  uint64_t ood_flags;
  int32_t fallback;             /* JPXLG_FALLBACK_* */
  int32_t large_frame_only_ood;
  int32_t model;                /* JPXLG_PREDICTOR_CASES or _KNOTS: the table that answered */
} jpxlg_crossing_prediction;

/* The image's source features and DCT8 summary, as the quality search reads
 * them (quality.rs:2449-2474; source_features quality_features.rs:173-213,
 * preanalysis_srgb8/16 preanalysis.rs:542-574, quality_transform_summary
 * jpxl-encode-policy/src/lib.rs:2769-2955).  Either pointer may be null. */
int32_t jpxlg_image_quality_features(jpxlg_img* img, jpxlg_source_features* source,
                                     jpxlg_transform_summary* transform);
/* quality_prediction::predict / predict_cases / predict_v2
 * (quality_prediction.rs:351-367, :453-598, :234-305) by `which`
 * (JPXLG_PREDICTOR_*), reserve = the effort's loss-relative reserve.
 * Returns 1 with *out filled, 0 when the chosen table gives no prediction
 * (predict_cases on a frame without preanalysis), or a negative status.
 * Host-only: needs no device. */
int32_t jpxlg_predict_crossing(const jpxlg_source_features* source, const jpxlg_transform_summary* transform,
                               double target, double reserve, int32_t which, jpxlg_crossing_prediction* out);

/* The quantizer a quality search of `effort` probes at ladder rung `rung`
 * (QuantizerChoice::at, jpxl-encode-policy/src/quantizer_ladder.rs:202-260,
 * over EncodeRequest::for_quality, request.rs:955-972). */
int32_t jpxlg_rung_quantizer(uint32_t rung, int32_t effort, jpxlg_quantizer* out);

/* ---- batches: many images of one context, one launch per stage ---------- */
/* One probe of n images: each quantized with its own quantizer, all
 * reconstructed with restoration r and scored, every stage one launch over
 * the batch's block-tile work list; scores[k] per image.  No counterpart in
 * the reference (which probes one image at a time through
 * PerceptualEvaluator::evaluate, quality.rs:87-157). */
int32_t jpxlg_batch_probe(uint32_t n, jpxlg_img* const* imgs, const jpxlg_quantizer* qs,
                          const jpxlg_restoration* r, double* scores);
/* search_frame_perceptual (quality.rs:2405) for n images at once: each
 * image's navigator runs as in jpxlg_encode_quality, and every round of
 * probes of the images still searching is one batched chain.  Each chosen
 * plan is left quantized on the device.  No counterpart in the reference. */
int32_t jpxlg_batch_encode_quality(uint32_t n, jpxlg_img* const* imgs, const jpxlg_quality_request* reqs,
                                   jpxlg_quality_result* out);

/* ---- captured probe chains (HIP graphs) -------------------------------- */
/* The quantize -> reconstruct -> restoration -> SSIMULACRA2 chain of n images
 * (one context, structures set, one DCT form) captured once as a HIP graph
 * with restoration r; each jpxlg_graph_probe replays it with one quantizer
 * per image and returns the scores (synchronous).  A graph is valid while
 * its images keep the structure they had at capture (a changed structure,
 * DCT form, AQ field or Gaborish precondition makes the probe fail with
 * JPXLG_E_INVALID); destroy it before its images.  No counterpart in the
 * reference, whose evaluator runs each probe eagerly
 * (PerceptualEvaluator::evaluate, quality.rs:87-157). */
typedef struct jpxlg_graph jpxlg_graph;
int32_t jpxlg_graph_create(uint32_t n, jpxlg_img* const* imgs, const jpxlg_restoration* r, jpxlg_graph** out);
int32_t jpxlg_graph_probe(jpxlg_graph* graph, const jpxlg_quantizer* qs, double* scores);
void jpxlg_graph_destroy(jpxlg_graph* graph);

/* ---- serving: continuous batching of quality requests ------------------- */
/* A server owns the navigators of the requests submitted to it (images of
 * its context, resident until released).  Each step() runs one round: the
 * pending probe of every active request, new requests' structures included,
 * as one batched chain, and the final quantize of requests whose search
 * ended.  Requests may be submitted between steps and join the next round.
 * One thread drives a server.  No counterpart in the reference (whose
 * encoder serves one request per call, jpxl/src/lib.rs:507-794). */
typedef struct jpxlg_server jpxlg_server;
int32_t jpxlg_server_create(jpxlg_ctx* ctx, jpxlg_server** out);
/* A request on an image the caller created (and keeps). */
int32_t jpxlg_server_submit(jpxlg_server* srv, jpxlg_img* img, const jpxlg_quality_request* req,
                            uint64_t* ticket);
/* A request as interleaved host samples (the jpxlg_image_upload_srgb
 * layout; rgb must stay valid until the request completes).  The server
 * admits such requests in arrival order while the projected device
 * footprint (jpxlg_image_footprint) of the admitted, unreleased ones fits
 * its budget (one is always admitted when none holds memory); admitting
 * creates the request's image, releasing it frees it.  Until admitted a
 * request is pending and polls 0.  When the device cannot hold an image the
 * budget allowed while others are admitted, the request goes back to the
 * head of the queue and the budget drops to the admitted footprint. */
int32_t jpxlg_server_submit_srgb(jpxlg_server* srv, uint32_t width, uint32_t height, uint32_t bits,
                                 const void* rgb, const jpxlg_quality_request* req, uint64_t* ticket);
/* The tickets of requests completed since the last call, in completion
 * order (up to cap; *n = how many were written).  A serving loop polls and
 * releases only these instead of every ticket it holds. */
int32_t jpxlg_server_completed(jpxlg_server* srv, uint64_t* tickets, uint32_t cap, uint32_t* n);
/* The admission budget in bytes (default: 90% of the device's free memory
 * when the server is created). */
int32_t jpxlg_server_set_budget(jpxlg_server* srv, uint64_t bytes);
/* The admitted requests' projected footprint, the budget, and the requests
 * waiting for admission (each pointer optional). */
int32_t jpxlg_server_admission(jpxlg_server* srv, uint64_t* used, uint64_t* budget, uint32_t* queued);
/* The image of a request (an admitted one, for a request given as samples),
 * e.g. to download its chosen plan before releasing it. */
int32_t jpxlg_server_image(jpxlg_server* srv, uint64_t ticket, jpxlg_img** out);
/* >= 0: requests completed by this round; < 0: JPXLG_E_*.  Admission runs
 * first. */
int32_t jpxlg_server_step(jpxlg_server* srv);
/* 1: done (result written), 0: queued or running, < 0: the request failed /
 * bad ticket. */
int32_t jpxlg_server_poll(jpxlg_server* srv, uint64_t ticket, jpxlg_quality_result* out);
/* Forgets a completed request; its image stays the caller's, or, for a
 * request given as samples, is freed (returning its footprint to the
 * budget). */
int32_t jpxlg_server_release(jpxlg_server* srv, uint64_t ticket);
/* Requests not yet completed (queued or running). */
int32_t jpxlg_server_pending(jpxlg_server* srv);
/* Ends queued requests unrun (JPXLG_E_INVALID), runs the admitted ones to
 * completion, then frees the server and the images it created. */
void jpxlg_server_destroy(jpxlg_server* srv);

/* ---- launch-shape autotuning ------------------------------------------- */
/* Times the pooling SSIMULACRA2 pass (the vertical blur, four waves per
 * workgroup) with and without an XCD-aware workgroup order on a synthetic
 * image of width x height and keeps the faster for chains of that many
 * pixels (size classes by log2 of a chain's total pixels); the outputs
 * (optional) report the choice (units_per_group: the waves per workgroup,
 * 4).  No candidate changes a result (a mismatch fails with
 * JPXLG_E_DEVICE).  Clears the per-kernel timing totals.  Untuned classes
 * use no swizzle.  No counterpart in the reference. */
int32_t jpxlg_autotune(jpxlg_ctx* ctx, uint32_t width, uint32_t height, int32_t* units_per_group,
                       int32_t* swizzle);

/* ---- per-kernel timing ----------------------------------------------------- */
/* When enabled, every launch the context issues is bracketed by HIP events on
 * its stream.  jpxlg_profile_read synchronises, folds the events into per-
 * kernel totals (name, milliseconds, launches) and returns the entry count;
 * names holds cap entries of 48 chars.  Reading also clears the totals. */
int32_t jpxlg_profile_enable(jpxlg_ctx* ctx, int32_t on);
/* Restricts the timing to launches of one kernel name (NULL or "" = all), so a
 * timed run carries two events per launch of that kernel only. */
int32_t jpxlg_profile_only(jpxlg_ctx* ctx, const char* name);
int32_t jpxlg_profile_read(jpxlg_ctx* ctx, char* names, double* total_ms, uint64_t* launches,
                           int32_t cap);

/* ---- stand-alone kernels (metric alone, transforms, filters) -------------- */
/* Planes are planar f32 linear sRGB, 3 x W x H, host memory. */
int32_t jpxlg_score_pair(jpxlg_ctx* ctx, uint32_t width, uint32_t height, const float* reference,
                         const float* candidate, double* score, double* terms, uint32_t* nterms);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): jpxlg_score_pair with the two images in DEVICE memory
 * (planar f32 linear sRGB, 3 x W x H, read only) and every intermediate
 * buffer in one file-scope device array of JPXLG_NOALLOC_SCRATCH_BYTES per
 * GPU, so the call allocates nothing.  A pair whose pyramids need more
 * (about 1.3 MP) fails with JPXLG_E_UNSUPPORTED.  Every call on one GPU, from
 * any context, uses the same array: calls that overlap on one GPU race.
 * Synchronous on the context's stream. */
#define JPXLG_NOALLOC_SCRATCH_BYTES ((size_t)96 << 20)
int32_t jpxlg_score_pair_noalloc(jpxlg_ctx* ctx, uint32_t width, uint32_t height, const float* d_reference,
                                 const float* d_candidate, double* score, double* terms, uint32_t* nterms);
/* count N x N blocks (N = 8, 16, 32), row-major; inverse = 0 forward DCT_2D. */
int32_t jpxlg_dct_blocks(jpxlg_ctx* ctx, int32_t side, int32_t inverse, uint32_t count,
                         const float* in, float* out);
/* One J.4 step over three W x H planes with a (ceil(W/8) x ceil(H/8)) sigma grid. */
int32_t jpxlg_epf_step(jpxlg_ctx* ctx, int32_t step, uint32_t width, uint32_t height,
                       const float* in, const float* sigma, float* out);
int32_t jpxlg_gaborish(jpxlg_ctx* ctx, uint32_t width, uint32_t height, const float* in,
                       float* out);

#ifdef __cplusplus
}
#endif
#endif
