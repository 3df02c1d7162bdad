/*
 * hpvcd.h — C ABI of the MI355X reconstruction back end.
 *
 * The reference (awxkee/hpvcd v0.3.1) is a Rust crate whose public API is
 * decode_heic(&[u8]) / Decoder::{decode, decode_rgb8, decode_yuv}
 * (src/lib.rs:1476-1497, src/decoder.rs:54-297).  Its reconstruction half
 * runs per TU through the ExecContext function table (src/exec.rs:33-130).
 * This library replaces that half: the host parse emits records
 * (hpvcd_records.h) and the entry points below reconstruct whole batches of
 * pictures on one GPU.  Every entry point returns an hpvcd_status; no
 * exception or abort crosses the ABI.
 */
#ifndef HPVCD_H
#define HPVCD_H

#include <stddef.h>
#include <stdint.h>
#include "hpvcd_records.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  1..13 follow DecodeError's variants in order
 * (src/error.rs:30-58); the last two are new to the GPU back end. */
typedef enum hpvcd_status {
    HPVCD_OK = 0,
    HPVCD_E_NOT_HEIF = 1,
    HPVCD_E_TRUNCATED_BOX = 2,
    HPVCD_E_MISSING_BOX = 3,
    HPVCD_E_UNSUPPORTED_ITEM = 4,
    HPVCD_E_BITSTREAM = 5,
    HPVCD_E_CABAC_DESYNC = 6,
    HPVCD_E_UNSUPPORTED_CHROMA = 7,
    HPVCD_E_UNSUPPORTED_BITDEPTH = 8,
    HPVCD_E_BAD_DIMENSIONS = 9,
    HPVCD_E_PARAMSET = 10,
    HPVCD_E_UNSUPPORTED = 11,
    HPVCD_E_LIMIT = 12,
    HPVCD_E_ALLOC = 13,
    HPVCD_E_DEVICE = 14,      /* HIP error or GPU/HBM admission failure */
    HPVCD_E_INVALID = 15      /* malformed records or arguments */
} hpvcd_status;

/* Thread-local message for the last non-OK status of this thread
 * (the Display text of DecodeError, src/error.rs:62-90). */
const char* hpvcd_last_error(void);
const char* hpvcd_version(void);

/* ------------------------------------------------------------------------ */
/* Seeded syntax-level generator: conformant reconstruction records for     */
/* synthetic pictures (and, below, the same pictures as coded HEVC / HEIC). */
/* ------------------------------------------------------------------------ */
typedef struct hpvcd_gen_config {
    int32_t  tile_width, tile_height;   /* coded size of each picture */
    int32_t  grid_cols, grid_rows;      /* pictures per output image */
    int32_t  out_width, out_height;     /* output image size (grid output) */
    uint8_t  chroma_idc, bit_depth, log2_ctb, log2_min_cb;
    uint8_t  log2_min_tb, log2_max_tb, max_transform_depth_intra, qp;
    uint8_t  log2_max_ts;               /* log2_max_transform_skip_block_size */
    uint8_t  transform_skip, implicit_rdpcm, cross_component, extended_precision;
    uint8_t  strong_intra_smoothing, sao, deblock, ts_rotation;
    uint8_t  matrix, full_range, out_kind, num_outputs;
    uint8_t  log2_res_scale_max;        /* CCP log2_res_scale drawn from [-max, max] */
    uint8_t  act;                       /* residual_adaptive_colour_transform_enabled_flag (4:4:4) */
    uint8_t  pcm;                       /* pcm_enabled_flag: 8x8..32x32 2Nx2N intra CUs may be I_PCM */
    uint8_t  pcm_bit_depth;             /* PcmBitDepth (luma and chroma), 1..bit_depth */
    uint8_t  pcm_lf_disabled;           /* pcm_loop_filter_disabled_flag */
    uint8_t  inter;                     /* pictures after the first are inter pictures of one sequence */
    uint8_t  num_refs;                  /* references per list (1..4) */
    uint8_t  weighted;                  /* explicit weighted prediction */
    uint8_t  amp;                       /* asymmetric motion partitions */
    uint8_t  gop;                       /* 0/1: each inter picture refers to the pictures just before it in decode
                                           order; n > 1: random access, hierarchical B GOPs of n pictures (pictures
                                           in decode order, outputs in display order) */
    uint8_t  _pad;
    float    p_split_cu, p_nxn, p_split_tu, p_cbf, p_sig, p_level_geom, p_ts, p_tqb;
    float    p_act;                     /* tu_residual_act_flag where the syntax is present */
    float    p_pcm;                     /* pcm_flag where the syntax is present */
    float    p_intra_cu;                /* intra CUs in inter pictures */
    float    p_skip;                    /* cu_skip_flag */
    float    p_bi;                      /* bi-prediction (0: P pictures, list 0 only) */
    float    mv_sigma;                  /* motion vector standard deviation in luma samples */
    uint32_t config_id;
    uint64_t base_seed;
    /* quantisation, loop-filter and picture-structure options */
    uint8_t  scaling_list;              /* 0 flat, 1 default lists (H.265 Tables 7-5/7-6), 2 random lists */
    uint8_t  cu_qp_delta;               /* cu_qp_delta_enabled_flag: per-CU QpY deltas */
    uint8_t  diff_cu_qp_delta_depth;    /* quantisation group = CTB >> depth */
    uint8_t  qp_delta_max;              /* CuQpDeltaVal drawn from [-max, max] */
    int8_t   pps_cb_qp_offset, pps_cr_qp_offset;
    int8_t   slice_cb_qp_offset, slice_cr_qp_offset;
    int8_t   beta_offset_div2, tc_offset_div2;   /* slice deblocking offsets (all slices) */
    uint8_t  num_slices;                /* slices per picture: CTB runs in tile-scan order (0/1 = one) */
    uint8_t  tile_cols, tile_rows;      /* uniform tile grid (0/1 = no tiles) */
    uint8_t  lf_across_tiles;           /* loop_filter_across_tiles_enabled_flag */
    uint8_t  lf_across_slices;          /* pps_loop_filter_across_slices_enabled_flag */
    uint8_t  random_slice_params;       /* each slice draws its deblocking offsets, disable and across flags */
    uint8_t  orientation;               /* EXIF orientation of the output images (0/1 = as decoded) */
    uint8_t  sign_hiding;               /* sign_data_hiding_enabled_flag of the coded stream */
    uint16_t conf_win[4];               /* conformance window: left, right, top, bottom (luma samples) */
    float    p_sao_merge;               /* sao_merge_left / sao_merge_up where the neighbour CTB is available */
    /* entropy-coding options of the coded stream (records unchanged) */
    uint8_t  persistent_rice;           /* persistent_rice_adaptation_enabled_flag */
    uint8_t  ts_context;                /* transform_skip_context_enabled_flag */
    uint8_t  bypass_alignment;          /* cabac_bypass_alignment_enabled_flag */
    uint8_t  wpp;                       /* entropy_coding_sync_enabled_flag: a substream per CTB row */
    /* inter syntax (shapes the records and the coded stream) */
    uint8_t  tmvp;                      /* temporal motion vector prediction (merge and AMVP candidates) */
    uint8_t  merge_cand;                /* MaxNumMergeCand 1..5 (0 = 5) */
    uint8_t  par_merge_level;           /* Log2ParMrgLevel 2..log2_ctb (0 = 2) */
    uint8_t  mvd_l1_zero;               /* mvd_l1_zero_flag of B slices */
    uint8_t  rps_in_sps;                /* short-term RPSs in the SPS (inter RPS prediction), slices name them */
    uint8_t  cabac_init;                /* cabac_init_present_flag: each slice draws cabac_init_flag */
    uint8_t  _pad2[2];
    float    p_merge;                   /* merge_flag of the PUs of CUs that are not skipped */
} hpvcd_gen_config;

typedef struct hpvcd_records hpvcd_records;   /* owned record set */

/* Fill *cfg with the distributions shared by every workload (SAO, deblock,
 * split/cbf/significance probabilities) for a tile of the given size. */
void hpvcd_gen_default_config(hpvcd_gen_config* cfg);
hpvcd_status hpvcd_gen_records(const hpvcd_gen_config* cfg, hpvcd_records** out);
/* Concatenate record sets (outputs renumbered); the inputs stay owned. */
hpvcd_status hpvcd_records_concat(hpvcd_records* const* sets, uint32_t count, hpvcd_records** out);
void hpvcd_records_view(const hpvcd_records* r, hpvcd_batch* out);
void hpvcd_records_free(hpvcd_records* r);

/* ------------------------------------------------------------------------ */
/* Coded HEVC: the record producer's bitstream side                        */
/* ------------------------------------------------------------------------ */
/* The generator's pictures as an HEVC elementary stream (Annex B start
 * codes): VPS, SPS, PPS, then each picture as one IDR access unit of CABAC
 * coded I slices.  Parsing it (hpvcd_parse_hevc) yields the records
 * hpvcd_gen_records makes for the same configuration.  Inter configurations
 * are refused.  *data is malloc'd; release it with hpvcd_bytes_free. */
hpvcd_status hpvcd_gen_hevc(const hpvcd_gen_config* cfg, uint8_t** data, size_t* size);
void hpvcd_bytes_free(uint8_t* data);
/* Parses an Annex B stream of intra pictures into records, one output image
 * per picture (its conformance window).  out_kind is an HPVCD_OUT_* kind, or
 * -1 for the kind decode_heic returns: luma for 4:0:0, else RGB, 16-bit
 * samples above 8-bit depth (src/lib.rs:1880-1916).  Replaces the syntax
 * half of the reference's FullDecoder (src/decode.rs:718-1435, :2495-3560,
 * :7617-8177; src/cabac; src/config.rs:566 parse_sps, :990 parse_pps). */
hpvcd_status hpvcd_parse_hevc(const uint8_t* data, size_t size, int32_t out_kind, hpvcd_records** out);
/* Introspection of the CABAC tables the parser and the writer share (the
 * context layout, initValues of initType 0..2, rangeTabLps, transIdxLps /
 * transIdxMps of H.265 9.3.2.2 and 9.3.4.3; the reference's
 * src/cabac/contexts.rs:139-400 and src/cabac/engine.rs:35-70).  The
 * element names are the reference's ContextSet field names. */
int32_t hpvcd_cabac_num_elements(void);
int32_t hpvcd_cabac_element(int32_t i, const char** name, int32_t* offset, int32_t* count);
int32_t hpvcd_cabac_tables(uint8_t* init /* 3 x count */, uint8_t* range_lps /* 64 x 4 */,
                           uint8_t* trans_lps /* 64 */, uint8_t* trans_mps /* 64 */);
int32_t hpvcd_cabac_init_states(int32_t init_type, int32_t qp, uint8_t* out /* count */);
/* Motion derivation probe (unit parity tests of the parser's merge / AMVP /
 * temporal derivations, src/motion.rs:155 derive_merge, :417 derive_amvp,
 * src/decode.rs:6180 DecoderNeighbors::temporal, :6626 derive_pu_merge).
 * The caller builds the current picture's motion cells (with each cell's
 * availability: decoded, same slice and tile) and optionally a collocated
 * picture's cells on the same grid. */
typedef struct hpvcd_mv_cell {
    uint8_t  inter;        /* 1: inter cell, 0: intra or not coded */
    uint8_t  pred;         /* bit 0 list 0, bit 1 list 1 */
    uint8_t  lt;           /* bit l: the list-l reference is long-term */
    uint8_t  avail;        /* current picture: available for motion */
    int8_t   ref_idx[2];
    uint8_t  _pad[2];
    int16_t  mv[2][2];     /* [list][x, y] quarter samples */
    int32_t  ref_poc[2];
} hpvcd_mv_cell;
typedef struct hpvcd_motion_case {
    int32_t width, height;             /* luma, multiples of 4 */
    int32_t log2_ctb, par_mrg_level, max_merge;
    int32_t is_b, tmvp, cur_poc, col_poc, col_from_l0;
    int32_t num_ref[2];
    int32_t ref_poc[2][16];
    uint8_t ref_lt[2][16];
    int32_t pu_x, pu_y, pu_w, pu_h;    /* the PU */
    int32_t cu_x, cu_y, cu_size;       /* its coding block */
    int32_t part_idx;
    const hpvcd_mv_cell* cur;          /* (width / 4) x (height / 4) cells */
    const hpvcd_mv_cell* col;          /* collocated picture, same grid, or NULL */
} hpvcd_motion_case;
typedef struct hpvcd_motion_result {
    int32_t       num_merge;           /* length of the merge candidate list */
    hpvcd_mv_cell merge[5];            /* the PU's merge motion for merge_idx 0..4 */
    int16_t       amvp[2][16][2][2];   /* [list][ref_idx][predictor][x, y] */
} hpvcd_motion_result;
hpvcd_status hpvcd_motion_probe(const hpvcd_motion_case* c, hpvcd_motion_result* r);

/* Checks every range, size and cross reference of a batch on the host (the
 * same checks hpvcd_job_create runs before touching the GPU). */
hpvcd_status hpvcd_validate_batch(const hpvcd_batch* b);
/* Bytes each output image needs (kind and size from the batch). */
size_t hpvcd_output_bytes(const hpvcd_output* o);

/* ------------------------------------------------------------------------ */
/* GPU reconstruction back end                                              */
/* ------------------------------------------------------------------------ */
typedef struct hpvcd_gpu hpvcd_gpu;           /* one device, one stream */
typedef struct hpvcd_job hpvcd_job;           /* a batch resident in HBM */

#define HPVCD_MODE_BATCHED 0   /* one launch per stage for the whole batch */
#define HPVCD_MODE_PLAIN   1   /* one launch per stage per picture */
#define HPVCD_MODE_GRAPH   2   /* batched stage chain replayed from a hipGraph */

#define HPVCD_STAGE_RECON   0  /* stop after intra reconstruction */
#define HPVCD_STAGE_DEBLOCK 1  /* stop after deblocking */
#define HPVCD_STAGE_SAO     2  /* stop after SAO */
#define HPVCD_STAGE_RGB     3  /* full pipeline (default) */

hpvcd_status hpvcd_gpu_open(int device, hpvcd_gpu** out);
void         hpvcd_gpu_close(hpvcd_gpu* g);
/* Free device memory on the GPU (admission control). */
hpvcd_status hpvcd_gpu_mem_info(hpvcd_gpu* g, size_t* free_bytes, size_t* total_bytes);

/* Upload a batch (host records) and prepare it for reconstruction.  The
 * records are copied; the caller may free them after return. */
hpvcd_status hpvcd_job_create(hpvcd_gpu* g, const hpvcd_batch* b, hpvcd_job** out);
void         hpvcd_job_free(hpvcd_job* j);
/* Device bytes the job holds. */
size_t       hpvcd_job_device_bytes(const hpvcd_job* j);
/* Host milliseconds hpvcd_job_create spent preparing the records
 * (validation, availability stamps, wavefront units, work lists) and
 * uploading them (pinned staging, one host-to-device copy, pool clears). */
hpvcd_status hpvcd_job_host_ms(const hpvcd_job* j, double* prep_ms, double* upload_ms);
/* Page-locked host memory (hipHostMalloc): output buffers that
 * hpvcd_job_download fills by DMA without a staging copy. */
hpvcd_status hpvcd_host_alloc(size_t bytes, void** out);
void         hpvcd_host_free(void* p);
/* Enqueue the reconstruction on the GPU's stream (asynchronous). */
hpvcd_status hpvcd_job_run(hpvcd_job* j, int mode, int stop_stage);
/* Wait for the stream. */
hpvcd_status hpvcd_gpu_sync(hpvcd_gpu* g);
/* Copy output images to host buffers (outputs[i] sized by hpvcd_output_bytes). */
hpvcd_status hpvcd_job_download(hpvcd_job* j, void** outputs);
/* Copy one picture's planes (u16 samples, coded size) at the stage the last
 * run stopped at; cb/cr may be NULL for 4:0:0.  A picture that a later
 * picture of the batch predicts from always completes SAO (its final samples
 * are the reference), so its planes are those after SAO whatever the stop. */
hpvcd_status hpvcd_job_download_planes(hpvcd_job* j, uint32_t picture, uint16_t* y, uint16_t* cb, uint16_t* cr);

/* Timed replay: `iters` runs of the given mode after `warmup` runs; writes
 * the mean milliseconds per run measured by HIP events on the job's stream
 * and, per kernel kind, the mean milliseconds per run
 * (kernel_ms[HPVCD_KERNEL_COUNT]). */
#define HPVCD_KERNEL_RESIDUAL 0
#define HPVCD_KERNEL_INTRA    1
#define HPVCD_KERNEL_DEBLOCK  2
#define HPVCD_KERNEL_SAO      3
#define HPVCD_KERNEL_RGB      4
#define HPVCD_KERNEL_PREP     5   /* static per-TB preparation for the intra wavefront */
#define HPVCD_KERNEL_MC       6   /* motion compensation of inter pictures */
#define HPVCD_KERNEL_COUNT    7
hpvcd_status hpvcd_job_time(hpvcd_job* j, int mode, int warmup, int iters, double* ms_per_run,
                            double* kernel_ms);

/* Launch shapes of the kernels that have a choice: the autotuner's knobs
 * (no counterpart in the reference, whose decode is scalar).  Every value
 * gives the same samples; only the speed differs.
 *   residual_block[k]: K1 threads per workgroup for TBs of 4 << k
 *                      (64, 128 or 256; block / n TBs per workgroup)
 *   deblock_block:     K3 threads (edge segments) per workgroup (64, 128, 256)
 *   rgb_rows:          K5 output rows per workgroup of 64 x 4 pixels (1, 2, 4)
 *   intra_per_cu:      K2 workgroups per CU of the wavefront (0: as many as
 *                      fit, else 1..8, capped by what fits)
 *   sao_parts:         K4 workgroups per luma CTB in launches of fewer CTBs
 *                      than 4 per CU (1, 2, 4) */
typedef struct hpvcd_tuning {
    uint16_t residual_block[4];
    uint16_t deblock_block;
    uint16_t rgb_rows;
    uint8_t  intra_per_cu;
    uint8_t  sao_parts;
    uint8_t  _pad[2];
} hpvcd_tuning;

void hpvcd_tuning_default(hpvcd_tuning* t);
/* HPVCD_E_INVALID (message names the field) for a value outside the lists. */
hpvcd_status hpvcd_tuning_check(const hpvcd_tuning* t);
/* The shapes a GPU's launches use; a job's captured graph is captured again
 * on its next replay after a change.  hpvcd_gpu_open starts from the
 * defaults, then from the entry for hpvcd_gpu_tuning_key() in the file that
 * the environment variable HPVCD_TUNING_FILE names, when it has one. */
hpvcd_status hpvcd_gpu_set_tuning(hpvcd_gpu* g, const hpvcd_tuning* t);
hpvcd_status hpvcd_gpu_get_tuning(hpvcd_gpu* g, hpvcd_tuning* t);
/* "<arch>:<CUs>cu" of the GPU, e.g. "gfx950:256cu" (at most 63 chars). */
hpvcd_status hpvcd_gpu_tuning_key(hpvcd_gpu* g, char* key, size_t key_size);
/* Autotuner: coordinate search over the knobs on a resident job.  Each knob
 * in turn takes each of its values with the others at their best so far;
 * a value is timed as the mean of `iters` batched runs (HIP events) after one
 * warm-up run, and kept when it is faster.  The GPU keeps the best shapes
 * (written to *best); best_ms is its time per run, default_ms that of the
 * shapes the search started from. */
hpvcd_status hpvcd_tune(hpvcd_job* j, int iters, hpvcd_tuning* best, double* best_ms, double* default_ms);
/* Tuning cache: a text file of lines "<key> r4 r8 r16 r32 deblock rgb intra
 * sao".  save replaces the key's line (or appends one); load fills *t from
 * it: HPVCD_E_INVALID when the file has no valid line for the key. */
hpvcd_status hpvcd_tuning_save(const char* path, const char* key, const hpvcd_tuning* t);
hpvcd_status hpvcd_tuning_load(const char* path, const char* key, hpvcd_tuning* t);

/* One-call convenience: upload, reconstruct (batched launches), download. */
hpvcd_status hpvcd_gpu_decode_batch(hpvcd_gpu* g, const hpvcd_batch* b, void** outputs);

/* Allocation-free form of hpvcd_gpu_decode_batch (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as tight
 * serving loops or code that forbids hipMalloc.  Same inputs (host records)
 * and outputs (host image buffers), same results.  Every device array of the
 * batch lives in a file-scope device array of hpvcd_noalloc_scratch_bytes()
 * bytes (64 MiB, one copy per GPU, held until the process ends); no
 * hipMalloc / hipFree / hipHostMalloc is called, records are staged from
 * ordinary host memory.  A batch that needs more is refused with
 * HPVCD_E_LIMIT before any device work.  Calls on one GPU share the array:
 * within this process they are serialised; overlapping calls from other
 * processes or library copies on the same GPU would race. */
hpvcd_status hpvcd_gpu_decode_batch_noalloc(hpvcd_gpu* g, const hpvcd_batch* b, void** outputs);
size_t hpvcd_noalloc_scratch_bytes(void);

/* Host half of hpvcd_job_create on its own (no GPU needed): validation and
 * preparation of a batch on `threads` host threads (0 = up to 16), the
 * milliseconds it took, and a digest of every prepared array (equal for
 * every thread count). */
hpvcd_status hpvcd_prep_digest(const hpvcd_batch* b, int32_t threads, uint64_t* digest, double* ms);

/* ------------------------------------------------------------------------ */
/* Drop-in decoder: the shape of the reference's public API                 */
/* (src/decoder.rs:54-234 Decoder::{from_settings, decode, decode_rgb8},    */
/*  src/lib.rs:1476 decode_heic, :1951 decode_heic_rgb8, src/settings.rs:59 */
/*  HeicSettings, src/limits.rs:85-102 ParseLimits).                        */
/* ------------------------------------------------------------------------ */
typedef struct hpvcd_decoder hpvcd_decoder;   /* one GPU, host worker threads; decode calls may overlap */

#define HPVCD_BACKEND_CPU_REF     0  /* the reference's own CPU decoder: not in this library (E_UNSUPPORTED) */
#define HPVCD_BACKEND_GPU_PLAIN   1  /* one launch per stage per picture */
#define HPVCD_BACKEND_GPU_BATCHED 2  /* one launch per stage for every picture of the image (default) */

typedef struct hpvcd_settings {
    int32_t  threads;            /* host worker threads; 0 = available parallelism (DecodeThreads::Available) */
    uint32_t max_dimension;      /* ParseLimits::DEFAULT_MAX_DIMENSION 16384 */
    uint64_t max_pixels;         /* 64 Mi pixels */
    uint64_t max_box_size;       /* 256 MiB */
    uint64_t max_item_size;      /* 256 MiB */
    uint32_t max_items;          /* 4096 */
    uint32_t max_tiles;          /* 16384 */
    uint8_t  decode_alpha, decode_gain_map;   /* HeicSettings flags; auxiliary images are not decoded here */
    uint8_t  _pad[2];
    int32_t  backend;            /* HPVCD_BACKEND_* */
    int32_t  device;             /* GPU ordinal */
} hpvcd_settings;

/* A decoded image (DecodedImage / Rgb8Image, src/lib.rs:430-455): pixels
 * are row-major, interleaved, in page-locked memory the library owns until
 * hpvcd_image_free. */
typedef struct hpvcd_image {
    uint32_t width, height;      /* displayed size (after orientation) */
    uint8_t  channels;           /* 3 (RGB) or 1 (luma) */
    uint8_t  bit_depth;          /* bit depth of the coded samples */
    uint8_t  sample_bytes;       /* 1 (8-bit output) or 2 */
    uint8_t  orientation;        /* HPVCD_ORIENT_* applied (0/1: none) */
    void*    pixels;
    size_t   pixels_len;         /* bytes */
    uint8_t  matrix;             /* CICP matrix_coefficients used for RGB */
    uint8_t  full_range;
    uint8_t  _pad[6];
} hpvcd_image;

/* HeicSettings::new() with ParseLimits::new() and backend GPU_BATCHED. */
void         hpvcd_settings_default(hpvcd_settings* s);
/* Decoder::from_settings: opens the GPU and the host worker threads. */
hpvcd_status hpvcd_decoder_new(const hpvcd_settings* s, hpvcd_decoder** out);
void         hpvcd_decoder_free(hpvcd_decoder* d);
/* Records in (what the host parse of one HEIC image emits), one image per
 * output of the batch out (images[b->num_outputs]), reconstructed by the
 * decoder's backend.  Each image is filled even when the call fails part
 * way (pixels NULL); free every one with hpvcd_image_free. */
hpvcd_status hpvcd_decode_records(hpvcd_decoder* d, const hpvcd_batch* b, hpvcd_image* images);
/* HEIC bytes in (borrowed for the call), one image out: Decoder::decode
 * (src/decoder.rs:224, src/lib.rs:1497 decode_heic_impl): RGB for colour,
 * luma for 4:0:0, 16-bit samples above 8-bit depth, orientation applied.
 * The container, parameter sets and CABAC are parsed on the host (grid tiles
 * on the decoder's threads); reconstruction runs on the decoder's GPU.
 * Grid tiles that fail to parse are left black, as the reference leaves
 * them.  Free the image with hpvcd_image_free. */
hpvcd_status hpvcd_decode_heic(hpvcd_decoder* d, const uint8_t* data, size_t size, hpvcd_image* img);
/* Decoder::decode_rgb8 (src/decoder.rs:234, src/lib.rs:1964): always 3 x
 * 8-bit, deeper samples >> (bit_depth - 8), 4:0:0 as gray triples. */
hpvcd_status hpvcd_decode_heic_rgb8(hpvcd_decoder* d, const uint8_t* data, size_t size, hpvcd_image* img);
/* Release an image's pixels (the Vec ownership hand-off of the reference). */
void         hpvcd_image_free(hpvcd_image* img);

/* Raw planes of a HEIC image (DecodedYuv, src/lib.rs:403): the coded sample
 * values of the visible window (a grid stitched), in coded orientation (the
 * orientation is reported, not applied), 8-bit samples as u8 and deeper ones
 * as u16.  The three planes share one page-locked block the library owns
 * until hpvcd_yuv_free. */
typedef struct hpvcd_yuv {
    uint32_t width, height;                 /* luma plane */
    uint32_t chroma_width, chroma_height;   /* each chroma plane; 0 for 4:0:0 */
    uint8_t  chroma_idc;                    /* 0, 1 (4:2:0) or 3 (4:4:4) */
    uint8_t  bit_depth;
    uint8_t  sample_bytes;                  /* 1 or 2 */
    uint8_t  orientation;                   /* HPVCD_ORIENT_* of the image, not applied */
    uint8_t  matrix, full_range;            /* colour signalling (as hpvcd_image) */
    uint8_t  _pad[2];
    void*    y;                             /* row-major planes, stride = plane width */
    void*    cb;                            /* NULL for 4:0:0 */
    void*    cr;
    size_t   bytes;                         /* of the whole block */
    uint64_t _reserved;
} hpvcd_yuv;

/* Decoder::decode_yuv (src/decoder.rs:229, src/lib.rs:762 decode_heic_yuv_with). */
hpvcd_status hpvcd_decode_heic_yuv(hpvcd_decoder* d, const uint8_t* data, size_t size, hpvcd_yuv* yuv);
void         hpvcd_yuv_free(hpvcd_yuv* yuv);
/* An HEVC elementary stream (Annex B) decoded to its frames in output order
 * (decode_hevc, src/video.rs:398; VideoDecoder::decode, :607): the stream's
 * pictures are parsed on the host (hpvcd_parse_hevc: POC, RPS, reference
 * lists, merge / AMVP, weights) and reconstructed on the GPU in dependency
 * waves.  out_kind is an HPVCD_OUT_* kind for every frame, or -1 for RGB /
 * luma by chroma format, 8 or 16 bit by depth.  *frames is a library-owned
 * array of *count images; release it with hpvcd_frames_free. */
hpvcd_status hpvcd_decode_hevc(hpvcd_decoder* d, const uint8_t* data, size_t size, int32_t out_kind,
                               hpvcd_image** frames, uint32_t* count);
void         hpvcd_frames_free(hpvcd_image* frames, uint32_t count);

/* The container half on its own: a HEIC file's primary image as records
 * (one output).  out_kind: -1 the kind hpvcd_decode_heic returns, -2 the
 * RGB8 kind of hpvcd_decode_heic_rgb8, else an HPVCD_OUT_* kind.  settings
 * may be NULL (ParseLimits defaults, all host threads).  Replaces
 * heif::parse (src/heif.rs:242) and the record-producing half of
 * decode_heic_impl / decode_grid (src/lib.rs:1497, :1724). */
hpvcd_status hpvcd_parse_heic(const uint8_t* data, size_t size, const hpvcd_settings* settings, int32_t out_kind,
                              hpvcd_records** out);
/* The generator's configuration as a HEIC file: one hvc1 item, or a grid
 * item over one hidden hvc1 item per picture (hvcC, ispe, colr nclx,
 * irot/imir; iloc v1 with the grid descriptor in idat).  Parsing it gives
 * the records hpvcd_gen_records makes.  Free with hpvcd_bytes_free. */
hpvcd_status hpvcd_gen_heic(const hpvcd_gen_config* cfg, uint8_t** data, size_t* size);

/* ------------------------------------------------------------------------ */
/* Serving loop (no counterpart in the reference; Decoder reuse semantics   */
/* src/decoder.rs:54-92): HEIC requests arrive at any time, are parsed on    */
/* host worker threads and are continuously batched onto the GPU.            */
/* ------------------------------------------------------------------------ */
typedef struct hpvcd_server hpvcd_server;

typedef struct hpvcd_server_config {
    uint32_t max_batch_pictures;   /* pictures (tiles) in one batch; a larger request still goes alone */
    uint32_t parse_workers;        /* requests parsed at once, each on threads / parse_workers host threads */
    uint64_t max_batch_bytes;      /* estimated device bytes of one batch */
    uint64_t hbm_reserve;          /* free HBM held back from admission */
    uint64_t batch_delay_us;     /* how long a free launch slot waits for requests still being parsed
                                      to join the batch, from the first ready request (0: no wait) */
} hpvcd_server_config;

/* Where one request's time went (milliseconds from submission). */
typedef struct hpvcd_request_stats {
    double   latency_ms;           /* submission to image in host memory */
    double   parse_ms;             /* submission to records (queueing for a parse worker included) */
    double   queue_ms;             /* records to launch (waiting for a slot and the batch) */
    double   device_ms;            /* launch to image (preparation, upload, kernels, download) */
    uint32_t batch_requests;       /* requests that shared its batch */
    uint32_t batch_pictures;       /* pictures in its batch */
} hpvcd_request_stats;

void         hpvcd_server_config_default(hpvcd_server_config* c);
/* settings: limits, host threads, backend and device as for the decoder.
 * The server opens two launch slots (streams) on the device, so that one
 * batch's host preparation and upload overlap the previous batch's kernels.
 * Admission: a batch takes requests in arrival order while its pictures,
 * its estimated bytes and the free HBM (less hbm_reserve) allow. */
hpvcd_status hpvcd_server_new(const hpvcd_settings* settings, const hpvcd_server_config* c, hpvcd_server** out);
/* Queue a HEIC decode (the bytes are copied); rgb8 != 0 asks for the
 * hpvcd_decode_heic_rgb8 kind.  Returns at once with a ticket. */
hpvcd_status hpvcd_server_submit(hpvcd_server* s, const uint8_t* data, size_t size, int32_t rgb8, uint64_t* ticket);
/* Block until the request is done; the image (the one hpvcd_decode_heic
 * gives for it alone) passes to the caller (hpvcd_image_free). */
hpvcd_status hpvcd_server_wait(hpvcd_server* s, uint64_t ticket, hpvcd_image* img, hpvcd_request_stats* stats);
/* Batches launched so far, and the requests and pictures they held. */
void         hpvcd_server_totals(hpvcd_server* s, uint64_t* batches, uint64_t* requests, uint64_t* pictures);
/* Finishes every queued request, then stops the workers and closes the GPU. */
void         hpvcd_server_free(hpvcd_server* s);

#ifdef __cplusplus
}
#endif
#endif /* HPVCD_H */
