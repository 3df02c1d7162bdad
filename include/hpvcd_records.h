/*
 * hpvcd_records.h — packed reconstruction records handed from the host
 * syntax parser (or the seeded syntax generator) to the reconstruction back
 * end.  Plain C, no device types: the same structs are read by the HIP
 * kernels, the C++ host runtime and the CPU oracle.
 *
 * One batch holds N pictures of mixed size.  Every picture owns a contiguous
 * range of each array; ranges are absolute indices into the batch arrays.
 *
 * What the records replace in the reference (awxkee/hpvcd v0.3.1): the
 * reference reconstructs each TU as soon as it is parsed
 * (src/decode.rs:3478 transform_unit -> :4724 reconstruct_luma /
 * :5280 reconstruct_chroma) and keeps per-4x4 side maps for the loop filters
 * (src/decode.rs:101-335: qp_y_map, edge_v/h, bs_v/h, tqb, pcm, slice_idx).
 * Here the parse emits those facts as records instead of calling the kernels.
 */
#ifndef HPVCD_RECORDS_H
#define HPVCD_RECORDS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- picture flags (hpvcd_picture.flags) -------------------------------- */
#define HPVCD_PF_STRONG_INTRA_SMOOTHING     (1u << 0)  /* sps.strong_intra_smoothing_enabled_flag */
#define HPVCD_PF_INTRA_SMOOTHING_DISABLED   (1u << 1)  /* RExt intra_smoothing_disabled_flag */
#define HPVCD_PF_INTRA_BOUNDARY_FILTER_OFF  (1u << 2)  /* SCC intra_boundary_filtering_disabled_flag */
#define HPVCD_PF_EXTENDED_PRECISION         (1u << 3)  /* extended_precision_processing_flag */
#define HPVCD_PF_IMPLICIT_RDPCM             (1u << 4)  /* implicit_rdpcm_enabled_flag */
#define HPVCD_PF_TS_ROTATION                (1u << 5)  /* transform_skip_rotation_enabled_flag */
#define HPVCD_PF_SAO_LUMA                   (1u << 6)  /* slice_sao_luma_flag */
#define HPVCD_PF_SAO_CHROMA                 (1u << 7)  /* slice_sao_chroma_flag */
#define HPVCD_PF_PCM_LF_DISABLED            (1u << 8)  /* pcm_loop_filter_disabled_flag */
#define HPVCD_PF_CONSTRAINED_INTRA          (1u << 9)  /* constrained_intra_pred_flag */
#define HPVCD_PF_LF_ACROSS_TILES            (1u << 10) /* loop_filter_across_tiles_enabled_flag */
#define HPVCD_PF_LF_ACROSS_SLICES           (1u << 11) /* pps_loop_filter_across_slices_enabled_flag */
#define HPVCD_PF_FULL_RANGE                 (1u << 12) /* colour: full-range YCbCr */
#define HPVCD_PF_TILES                      (1u << 13) /* more than one in-picture tile */
#define HPVCD_PF_DEBLOCK_DISABLED           (1u << 14) /* pps/slice default deblocking disabled */

/* ---- transform-block flags (hpvcd_tb.flags) ----------------------------- */
#define HPVCD_TB_CBF          (1u << 0)   /* coded residual present */
#define HPVCD_TB_TS           (1u << 1)   /* transform_skip_flag */
#define HPVCD_TB_TQB          (1u << 2)   /* cu_transquant_bypass_flag */
#define HPVCD_TB_INTER        (1u << 3)   /* CU is inter: no intra prediction, MC result is the predictor */
#define HPVCD_TB_RDPCM_H      (1u << 4)   /* explicit RDPCM, horizontal (inter, parsed) */
#define HPVCD_TB_RDPCM_V      (1u << 5)   /* explicit RDPCM, vertical (inter, parsed) */
/* tu_residual_act_flag (4:4:4 only, src/decode.rs:3619 transform_unit_act):
 * set on the luma TB and on the Cb and Cr TBs that follow it directly, all
 * three co-located and of one size.  Their residuals are computed in 32-bit
 * arithmetic, Cb/Cr get cross-component prediction from the luma residual,
 * then the inverse YCgCo-R (src/act.rs:43) mixes the triple before each is
 * added to its prediction.  qp carries the ACT-adjusted Qp'. */
#define HPVCD_TB_ACT          (1u << 6)
/* I_PCM block (src/decode.rs:2768 decode_pcm_cu): no prediction; the
 * coefficient words carry the raw pcm_sample values (position, value) and qp
 * carries the shift BitDepth - PcmBitDepth, so each sample is value << qp. */
#define HPVCD_TB_PCM          (1u << 7)

/* ---- prediction-unit flags (hpvcd_pu.flags) ----------------------------- */
#define HPVCD_PU_L0           (1u << 0)   /* predFlagL0 */
#define HPVCD_PU_L1           (1u << 1)   /* predFlagL1 */

/* hpvcd_ref.picture of a reference that is not available (the PU reads
 * nothing from that list; with no list left the prediction is mid-grey,
 * src/decode.rs:7206) */
#define HPVCD_REF_MISSING     0xffffffffu

/* ---- per-4x4 luma cell info word (cell_info[]) --------------------------
 * bits 0-1  bS of the vertical edge at the cell's left side (0 when the
 *           cell's left side is not a TU/PU edge: edge_v folded in)
 * bits 2-3  bS of the horizontal edge at the cell's top side
 * bit  4    transquant bypass CU
 * bit  5    I_PCM CU
 * bit  6    intra CU (constrained intra prediction)
 * bits 8-15 QpY (signed 8-bit)
 * bits 16-23 slice index (into hpvcd_picture.slices)
 * bits 24-31 in-picture tile id                                              */
#define HPVCD_CELL_BSV(w)    ((w) & 3u)
#define HPVCD_CELL_BSH(w)    (((w) >> 2) & 3u)
#define HPVCD_CELL_TQB       (1u << 4)
#define HPVCD_CELL_PCM       (1u << 5)
#define HPVCD_CELL_INTRA     (1u << 6)
#define HPVCD_CELL_QP(w)     ((int32_t)(int8_t)(((w) >> 8) & 0xffu))
#define HPVCD_CELL_SLICE(w)  (((w) >> 16) & 0xffu)
#define HPVCD_CELL_TILE(w)   (((w) >> 24) & 0xffu)
#define HPVCD_CELL_PACK(bsv, bsh, extra, qp, slice, tile)                                   \
    ((uint32_t)(bsv) | ((uint32_t)(bsh) << 2) | (uint32_t)(extra) |                         \
     (((uint32_t)(uint8_t)(int8_t)(qp)) << 8) | ((uint32_t)(slice) << 16) | ((uint32_t)(tile) << 24))

/* ---- coefficient word -----------------------------------------------------
 * bits 0-9   position y*n + x inside the n x n block (n <= 32)
 * bits 10-31 signed level (22 bits; covers the extended-precision range of
 *            bit depths up to 15)                                           */
#define HPVCD_COEFF_POS(c)    ((uint32_t)(c) & 0x3ffu)
#define HPVCD_COEFF_LEVEL(c)  (((int32_t)(c)) >> 10)
#define HPVCD_COEFF_PACK(pos, level) (((uint32_t)(pos) & 0x3ffu) | ((uint32_t)(level) << 10))
#define HPVCD_COEFF_LEVEL_MAX ((1 << 21) - 1)
#define HPVCD_COEFF_LEVEL_MIN (-(1 << 21))

#define HPVCD_MAX_SLICES 8

typedef struct hpvcd_slice_params {
    uint8_t disabled;          /* slice_deblocking_filter_disabled_flag */
    int8_t  beta_offset_div2;  /* slice_beta_offset_div2 */
    int8_t  tc_offset_div2;    /* slice_tc_offset_div2 */
    uint8_t lf_across;         /* slice_loop_filter_across_slices_enabled_flag */
} hpvcd_slice_params;

/* One coded picture (one HEIF tile or one stand-alone item). */
typedef struct hpvcd_picture {
    int32_t  width, height;            /* coded luma size */
    uint8_t  chroma_idc;               /* 0 mono, 1 4:2:0, 2 4:2:2, 3 4:4:4 */
    uint8_t  bit_depth_luma, bit_depth_chroma;
    uint8_t  log2_ctb;
    uint32_t flags;                    /* HPVCD_PF_* */
    int8_t   pps_cb_qp_offset, pps_cr_qp_offset;
    uint8_t  matrix;                   /* CICP matrix_coefficients used for RGB */
    uint8_t  num_slices;
    int8_t   log2_sao_offset_scale_luma, log2_sao_offset_scale_chroma; /* informational */
    uint8_t  out_kind;                 /* HPVCD_OUT_* of the picture's output image */
    uint8_t  _pad0;
    hpvcd_slice_params slices[HPVCD_MAX_SLICES];
    int32_t  slice_qp;                 /* QpY used where a cell has no map entry */
    /* RGB conversion window (decode_grid / decode_heic_impl): win_w x win_h
     * output pixels starting at coded sample (crop_left, crop_top). */
    int32_t  crop_left, crop_top, win_w, win_h;
    /* placement of the window's top-left copy_w x copy_h block in the output */
    int32_t  dst_x, dst_y, copy_w, copy_h;
    uint32_t out_index;                /* index into hpvcd_batch.outputs */
    int32_t  scaling_index;            /* index into hpvcd_batch.scaling, -1 = flat */
    /* absolute ranges into the batch arrays */
    uint32_t tb_begin, tb_count;
    uint32_t coeff_begin, coeff_count;
    uint32_t ctb_begin;                /* ceil(w/ctb) * ceil(h/ctb) SAO entries */
    uint32_t cell_begin;               /* ceil(w/4) * ceil(h/4) cell entries */
    /* inter prediction: the picture's prediction units and the reference
     * entries they name (all slices' lists, in any order) */
    uint32_t pu_begin, pu_count;
    uint32_t ref_begin, ref_count;
} hpvcd_picture;

/* One transform block of one colour component, in decode order. */
typedef struct hpvcd_tb {
    uint16_t x, y;        /* top-left, in samples of the component plane */
    uint8_t  log2n;       /* 2..5 */
    uint8_t  comp;        /* 0 Y, 1 Cb, 2 Cr */
    uint8_t  mode;        /* intra prediction mode 0..34 (chroma: derived mode) */
    uint8_t  qp;          /* Qp' of the component (QpBdOffset included) */
    uint16_t flags;       /* HPVCD_TB_* */
    uint8_t  nx;          /* 1 + largest column holding a nonzero level */
    int8_t   res_scale;   /* cross-component ResScaleVal (chroma), 0 = none */
    uint32_t coeff_off;   /* first level, relative to the picture's coeff_begin */
    uint32_t coeff_cnt;   /* number of nonzero levels */
    uint32_t order;       /* reference-availability stamp */
} hpvcd_tb;

/* One inter prediction unit (src/decode.rs:7000 motion_compensate): luma
 * position and size, the lists it predicts from and a quarter-sample luma
 * motion vector per list.  Its samples of every plane are predicted before
 * the TBs of its CU add their residuals (those TBs carry HPVCD_TB_INTER);
 * the TB records still tile every inter CU, skipped ones included. */
typedef struct hpvcd_pu {
    uint16_t x, y;        /* luma position in the picture */
    uint8_t  w, h;        /* luma size: multiples of 4, at most 64 */
    uint8_t  flags;       /* HPVCD_PU_L0 | HPVCD_PU_L1 */
    uint8_t  _pad;
    uint16_t ref[2];      /* reference entry per list, relative to the picture's ref_begin */
    int16_t  mv[2][2];    /* [list][x, y] in quarter luma samples */
} hpvcd_pu;

/* One reference-list entry (RefPicList[X][i] with its pred_weight_table
 * row, src/inter.rs:141).  The picture is an earlier picture of the same
 * batch with the same size and format: its final (deblocked, SAO) samples
 * are read in place on the device. */
typedef struct hpvcd_ref {
    uint32_t picture;      /* batch picture index, or HPVCD_REF_MISSING */
    uint8_t  weighted;     /* explicit weighted prediction (weighted_pred_flag / weighted_bipred_flag) */
    uint8_t  log2_denom[2];/* luma_log2_weight_denom, ChromaLog2WeightDenom */
    uint8_t  _pad;
    int16_t  weight[3];    /* Y, Cb, Cr weights */
    int16_t  offset[3];    /* Y, Cb, Cr offsets in reconstructed-sample units */
} hpvcd_ref;

/* SAO parameters of one CTB, components [Y, Cb, Cr]; offsets already carry
 * the edge-offset sign convention and log2_sao_offset_scale. */
typedef struct hpvcd_sao {
    uint8_t type_idx[3];  /* 0 off, 1 band, 2 edge */
    uint8_t band_pos[3];
    uint8_t eo_class[3];
    uint8_t _pad[3];
    int16_t offsets[3][4];
} hpvcd_sao;

/* Scaling lists: for sizeId 0..3 and matrixId 0..5 (Intra Y/Cb/Cr, Inter
 * Y/Cb/Cr): 64 coefficients (16 used for 4x4), then the DC value, then a
 * flat flag (all 16). */
#define HPVCD_SCALING_ENTRY_BYTES 66
#define HPVCD_SCALING_SET_BYTES   (4 * 6 * HPVCD_SCALING_ENTRY_BYTES)

#define HPVCD_OUT_RGB8   0   /* 3 x u8, 8-bit pictures */
#define HPVCD_OUT_RGB16  1   /* 3 x u16 holding bit_depth-bit values */
#define HPVCD_OUT_RGB8_FROM16 2 /* decode_rgb8 of a >8-bit picture: >> (bd-8) */
#define HPVCD_OUT_LUMA8  3
#define HPVCD_OUT_LUMA16 4
/* Raw planes (Decoder::decode_yuv, src/lib.rs:762-1207): Y (width x height),
 * then for colour pictures Cb and Cr (ceil(width / sub_w) x ceil(height /
 * sub_h) each), planar, the coded sample values (u8 for 8-bit pictures, u16
 * above).  The orientation is reported, not applied.  A picture copies
 * copy_w x copy_h luma samples from (crop_left, crop_top) of its coded plane
 * to (dst_x, dst_y); its chroma copy starts at (crop_left / sub_w, crop_top /
 * sub_h) and lands at (col * tcw, row * tch) with col = dst_x / win_w, row =
 * dst_y / win_h, tcw = ceil(max(win_w, copy_w) / sub_w) and tch likewise,
 * at most tcw x tch samples; source coordinates clamp to the coded plane
 * (src/lib.rs:143 plane_window, :882 stitch_native_yuv_band). */
#define HPVCD_OUT_YUV420_8  5
#define HPVCD_OUT_YUV420_16 6
#define HPVCD_OUT_YUV444_8  7
#define HPVCD_OUT_YUV444_16 8
#define HPVCD_OUT_YUV400_8  9
#define HPVCD_OUT_YUV400_16 10
#define HPVCD_OUT_IS_YUV(k) ((k) >= HPVCD_OUT_YUV420_8 && (k) <= HPVCD_OUT_YUV400_16)
/* chroma_format_idc of a YUV kind */
#define HPVCD_OUT_YUV_CHROMA(k) ((k) <= HPVCD_OUT_YUV420_16 ? 1 : ((k) <= HPVCD_OUT_YUV444_16 ? 3 : 0))
#define HPVCD_OUT_YUV_WIDE(k) (((k) - HPVCD_OUT_YUV420_8) & 1)

/* EXIF orientation of an output image (src/metadata.rs:31 Orientation,
 * applied after RGB conversion as src/lib.rs:2017 apply_orientation does).
 * Pictures are placed on the width x height canvas as decoded; the output
 * buffer holds the oriented image, whose size swaps for 5..8.  5 and 7
 * (transpose, transverse) swap the size but keep the buffer as decoded, as
 * the reference does. */
#define HPVCD_ORIENT_NORMAL     1
#define HPVCD_ORIENT_FLIP_H     2
#define HPVCD_ORIENT_ROTATE_180 3
#define HPVCD_ORIENT_FLIP_V     4
#define HPVCD_ORIENT_TRANSPOSE  5
#define HPVCD_ORIENT_ROTATE_90  6
#define HPVCD_ORIENT_TRANSVERSE 7
#define HPVCD_ORIENT_ROTATE_270 8

typedef struct hpvcd_output {
    int32_t  width, height;   /* output canvas size (before orientation) */
    uint8_t  kind;            /* HPVCD_OUT_* */
    uint8_t  channels;        /* 3 or 1 (YUV kinds: planes, 3 or 1) */
    uint8_t  bytes_per_sample;
    uint8_t  orientation;     /* HPVCD_ORIENT_* (0 = as decoded) */
} hpvcd_output;

/* A batch of pictures.  All pointers are host pointers for submission. */
typedef struct hpvcd_batch {
    uint32_t num_pictures;   const hpvcd_picture* pictures;
    uint32_t num_tbs;        const hpvcd_tb*      tbs;
    uint32_t num_coeffs;     const uint32_t*      coeffs;
    uint32_t num_ctbs;       const hpvcd_sao*     sao;
    uint32_t num_cells;      const uint32_t*      cell_info;
    uint32_t num_scaling;    const uint8_t*       scaling;   /* num_scaling * HPVCD_SCALING_SET_BYTES */
    uint32_t num_outputs;    const hpvcd_output*  outputs;
    uint32_t num_pus;        const hpvcd_pu*      pus;
    uint32_t num_refs;       const hpvcd_ref*     refs;
} hpvcd_batch;

#ifdef __cplusplus
}
#endif
#endif /* HPVCD_RECORDS_H */
