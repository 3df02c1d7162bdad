/*
 * oracle.h — CPU oracle for the HEVC reconstruction back end.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the scalar
 * reconstruction path of the reference decoder (awxkee/hpvcd v0.3.1, Rust),
 * which cannot be built here.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker; the
 * product path (hpvcd_amd) never links or calls it.
 *
 * Every function names the reference file:line it restates.  Inputs are the
 * packed records of include/hpvcd_records.h, consumed in decode order exactly
 * as the reference consumes parse results.
 */
#ifndef HPVCD_ORACLE_H
#define HPVCD_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/hpvcd_records.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- kernel-level entry points (one ExecContext slot each) ------------- */

/* src/transform.rs:701 dequant_params + :764 / :794 dequantize_*_scalar.
 * levels/out: n*n row-major.  scaling: NULL = flat, else 64 coeffs + dc. */
void oracle_dequant(const int32_t* levels, int n, int qp_prime, int bit_depth, int extended,
                    const uint8_t* scaling, int32_t* out);
/* src/transform.rs:728 transform_skip_params + :779 / :810 */
void oracle_dequant_ts(const int32_t* levels, int n, int qp_prime, int bit_depth, int extended,
                       const uint8_t* scaling, int32_t* out);
/* src/transform.rs:489 inv_dct_n_into; i16 != 0 selects the 8-bit i16 storage
 * path (second stage wraps to i16). */
void oracle_inv_dct(const int32_t* coeff, int n, int bit_depth, int max_range, int nx, int i16,
                    int32_t* out);
/* src/transform.rs:535 inv_transform_n_into with DST4 (:30) */
void oracle_inv_dst4(const int32_t* coeff, int bit_depth, int max_range, int i16, int32_t* out);
/* dense-matrix form of the inverse DCT (src/transform.rs:535 with the DCT
 * matrix): used only to cross-check the partial butterfly. */
void oracle_inv_dct_dense(const int32_t* coeff, int n, int bit_depth, int max_range, int32_t* out);
/* HEVC N-point DCT matrix entry T_N[k][m] (k = basis index). */
int oracle_dct_coef(int n, int k, int m);
int oracle_max_range(int bit_depth, int extended); /* src/transform.rs:686 */

/* src/intra.rs:84 substitute_refs_into + :151 filter_refs_into + :257
 * predict_into_scalar.  raw_above/raw_left: 2n entries, -1 = unavailable;
 * corner -1 = unavailable.  out: n*n. */
void oracle_intra_predict(int corner, const int32_t* raw_above, const int32_t* raw_left, int n,
                          int mode, int filter_is_luma, int smoothing_enabled, int strong,
                          int edge_filters, int bit_depth, uint16_t* out);

/* src/intra.rs:151 filter_refs_into (above/left: 2n+1 entries each) */
void oracle_intra_filter_refs(const int32_t* above, const int32_t* left, int n, int mode, int is_luma, int enabled,
                              int strong, int bd, int32_t* fa, int32_t* fl);
/* src/sao.rs:179 apply_sao_plane_scalar on one CTB rectangle [x0,x1)x[y0,y1) */
void oracle_sao_ctb(uint16_t* dst, const uint16_t* src, int w, int h, int x0, int y0, int x1, int y1, int type,
                    const int16_t* offsets, int band_pos, int eo_class, int bd);

/* src/deblock.rs:241 deblock_luma_segment on a 4-line segment stored as
 * lines[4][8] = p3 p2 p1 p0 q0 q1 q2 q3 (in place). */
void oracle_deblock_luma_segment(int32_t lines[4][8], int beta, int tc, int max_val);
/* src/deblock.rs:216 luma_decision: 0 skip, 1 strong, 2 weak (+do_p1/do_q1) */
int oracle_luma_decision(int32_t lines[4][8], int beta, int tc, int* do_p1, int* do_q1);
int oracle_qpc(int qpi, int chroma_idc, int bit_depth); /* src/decode.rs:7599 */

/* src/mc.rs:631 / :682 weighted sample prediction on one sample (KATs) */
int oracle_weighted_uni(int pred, int w, int o, int log2wd, int bit_depth);
int oracle_weighted_bi(int p0, int p1, int w0, int w1, int o0, int o1, int log2wd, int bit_depth);

/* src/act.rs:43 inverse YCgCo-R on one residual triple (in place) */
void oracle_act_inverse(int32_t* y, int32_t* cg, int32_t* co);

/* ---- picture-level entry points ----------------------------------------- */

#define ORACLE_STAGE_RECON   0   /* reconstructed, before the loop filters */
#define ORACLE_STAGE_DEBLOCK 1   /* after deblocking */
#define ORACLE_STAGE_SAO     2   /* after SAO (decoder output planes) */

/* Decode picture p of the batch.  Planes are u16, coded size, row-major:
 * y: width*height, cb/cr: ceil(w/sub_w)*ceil(h/sub_h).  Returns 0 on success. */
int oracle_picture_planes(const hpvcd_batch* b, uint32_t p, int stage, uint16_t* y, uint16_t* cb,
                          uint16_t* cr);

/* YUV -> RGB of one picture's window (src/yuv.rs:533) into out (win_w*win_h*ch
 * samples of the picture's output kind). */
int oracle_picture_rgb(const hpvcd_batch* b, uint32_t p, const uint16_t* y, const uint16_t* cb,
                       const uint16_t* cr, void* out);

/* Full decode of a batch: every output image (tightly packed, pre-zeroed by
 * the caller) receives its pictures' windows (src/lib.rs:1610 stitch).
 * threads: worker threads (grid rows in parallel as src/lib.rs:1680 does when
 * rows_parallel != 0, else one job per picture).  Returns 0 on success. */
int oracle_decode_batch(const hpvcd_batch* b, void** outputs, int threads, int rows_parallel);

#ifdef __cplusplus
}
#endif
#endif
