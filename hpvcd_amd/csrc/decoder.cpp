// decoder.cpp — the drop-in decoder object of include/hpvcd.h: settings and
// limits (src/settings.rs:59 HeicSettings, src/limits.rs:85-102
// ParseLimits), one GPU and its backend mode (src/decoder.rs:54
// Decoder::from_settings), HEIC bytes in (src/decoder.rs:224 decode, :234
// decode_rgb8: container and CABAC parse on the host, reconstruction on the
// GPU) or records in, into library-owned images (src/lib.rs:430-455
// DecodedImage / Rgb8Image; hpvcd_image_free is the Vec ownership hand-off).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/hpvcd.h"
#include "heif.h"
#include "internal.h"

using namespace hpvcd;

struct hpvcd_decoder {
    hpvcd_settings s;
    hpvcd_gpu* g = nullptr;
    int threads = 1;
};

namespace {

// oriented size of an output canvas (src/lib.rs:2007 oriented_dimensions)
void oriented_size(const hpvcd_output& o, uint32_t* w, uint32_t* h) {
    // raw planes keep the coded orientation
    const bool swap = !HPVCD_OUT_IS_YUV(o.kind) && o.orientation >= HPVCD_ORIENT_TRANSPOSE &&
                      o.orientation <= HPVCD_ORIENT_ROTATE_270;
    *w = (uint32_t)(swap ? o.height : o.width);
    *h = (uint32_t)(swap ? o.width : o.height);
}

}  // namespace

extern "C" {

void hpvcd_settings_default(hpvcd_settings* s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->threads = 0;
    s->max_dimension = 16384;
    s->max_pixels = 64ull * 1024 * 1024;
    s->max_box_size = 256ull * 1024 * 1024;
    s->max_item_size = 256ull * 1024 * 1024;
    s->max_items = 4096;
    s->max_tiles = 16384;
    s->decode_alpha = 1;
    s->decode_gain_map = 1;
    s->backend = HPVCD_BACKEND_GPU_BATCHED;
    s->device = 0;
}

hpvcd_status hpvcd_decoder_new(const hpvcd_settings* s, hpvcd_decoder** out) {
    if (!out) return set_error(HPVCD_E_INVALID, "null argument");
    *out = nullptr;
    hpvcd_settings def;
    hpvcd_settings_default(&def);
    const hpvcd_settings& S = s ? *s : def;
    if (S.backend == HPVCD_BACKEND_CPU_REF)
        return set_error(HPVCD_E_UNSUPPORTED,
                         "backend CPU_REF is the reference decoder itself; this library decodes on the GPU only");
    if (S.backend != HPVCD_BACKEND_GPU_PLAIN && S.backend != HPVCD_BACKEND_GPU_BATCHED)
        return set_error(HPVCD_E_INVALID, "unknown backend");
    if (S.max_dimension == 0 || S.max_pixels == 0) return set_error(HPVCD_E_INVALID, "zero limits");
    hpvcd_decoder* d = new (std::nothrow) hpvcd_decoder();
    if (!d) return set_error(HPVCD_E_ALLOC, "out of memory");
    d->s = S;
    // DecodeThreads: Available = available parallelism, Fixed(n) = max(n, 1)
    d->threads = S.threads > 0 ? S.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    hpvcd_status st = hpvcd_gpu_open(S.device, &d->g);
    if (st) {
        delete d;
        return st;
    }
    *out = d;
    return HPVCD_OK;
}

void hpvcd_decoder_free(hpvcd_decoder* d) {
    if (!d) return;
    hpvcd_gpu_close(d->g);
    delete d;
}

void hpvcd_image_free(hpvcd_image* img) {
    if (!img) return;
    hpvcd_host_free(img->pixels);
    img->pixels = nullptr;
    img->pixels_len = 0;
}

hpvcd_status hpvcd_decode_records(hpvcd_decoder* d, const hpvcd_batch* b, hpvcd_image* images) {
    if (!d || !b || (b->num_outputs && !images)) return set_error(HPVCD_E_INVALID, "null argument");
    for (uint32_t o = 0; o < b->num_outputs; o++) std::memset(&images[o], 0, sizeof(hpvcd_image));
    hpvcd_status st = hpvcd_validate_batch(b);
    if (st) return st;
    // ParseLimits::check_dims on every output image (src/decoder.rs check_dims)
    for (uint32_t o = 0; o < b->num_outputs; o++) {
        const hpvcd_output& O = b->outputs[o];
        if ((uint32_t)O.width > d->s.max_dimension || (uint32_t)O.height > d->s.max_dimension ||
            (uint64_t)O.width * (uint64_t)O.height > d->s.max_pixels)
            return set_error(HPVCD_E_LIMIT, "image dimensions exceed the decoder's limits");
    }
    for (uint32_t o = 0; o < b->num_outputs; o++) {
        const hpvcd_output& O = b->outputs[o];
        hpvcd_image& I = images[o];
        oriented_size(O, &I.width, &I.height);
        I.channels = O.channels;
        I.sample_bytes = O.bytes_per_sample;
        I.orientation = O.orientation;
        for (uint32_t p = 0; p < b->num_pictures; p++)
            if (b->pictures[p].out_index == o) {
                I.bit_depth = b->pictures[p].bit_depth_luma;
                I.matrix = b->pictures[p].matrix;
                I.full_range = (b->pictures[p].flags & HPVCD_PF_FULL_RANGE) ? 1 : 0;
                break;
            }
        I.pixels_len = hpvcd_output_bytes(&O);
        if ((st = hpvcd_host_alloc(I.pixels_len, &I.pixels))) {
            for (uint32_t k = 0; k <= o; k++) hpvcd_image_free(&images[k]);
            return st;
        }
    }
    if (b->num_pictures == 0) {  // nothing coded (every grid tile failed): black images
        for (uint32_t o = 0; o < b->num_outputs; o++) std::memset(images[o].pixels, 0, images[o].pixels_len);
        return HPVCD_OK;
    }
    hpvcd_job* j = nullptr;
    if ((st = job_create_validated(d->g, b, &j))) {
        for (uint32_t o = 0; o < b->num_outputs; o++) hpvcd_image_free(&images[o]);
        return st;
    }
    const int mode = d->s.backend == HPVCD_BACKEND_GPU_PLAIN ? HPVCD_MODE_PLAIN : HPVCD_MODE_BATCHED;
    st = hpvcd_job_run(j, mode, HPVCD_STAGE_RGB);
    if (!st) {
        std::vector<void*> outs(b->num_outputs);
        for (uint32_t o = 0; o < b->num_outputs; o++) outs[o] = images[o].pixels;
        st = hpvcd_job_download(j, outs.data());
    }
    hpvcd_job_free(j);
    if (st)
        for (uint32_t o = 0; o < b->num_outputs; o++) hpvcd_image_free(&images[o]);
    return st;
}

namespace {

// out_kind receives the output kind the container chose
hpvcd_status decode_heic_kind(hpvcd_decoder* d, const uint8_t* data, size_t size, int kind, hpvcd_image* img,
                              int* out_kind = nullptr) {
    if (!d || !data || !img) return set_error(HPVCD_E_INVALID, "null argument");
    std::memset(img, 0, sizeof(*img));
    hpvcd_records* r = nullptr;
    HeicInfo info;
    hpvcd_status st = heic_records(data, size, limits_of(&d->s), d->threads, kind, &r, &info);
    if (st) return st;
    hpvcd_batch b;
    hpvcd_records_view(r, &b);
    if (out_kind) *out_kind = b.num_outputs ? b.outputs[0].kind : -1;
    st = hpvcd_decode_records(d, &b, img);
    hpvcd_records_free(r);
    return st;
}

}  // namespace

hpvcd_status hpvcd_decode_heic(hpvcd_decoder* d, const uint8_t* data, size_t size, hpvcd_image* img) {
    return decode_heic_kind(d, data, size, HEIC_KIND_DEFAULT, img);
}

hpvcd_status hpvcd_decode_heic_rgb8(hpvcd_decoder* d, const uint8_t* data, size_t size, hpvcd_image* img) {
    return decode_heic_kind(d, data, size, HEIC_KIND_RGB8, img);
}

hpvcd_status hpvcd_decode_heic_yuv(hpvcd_decoder* d, const uint8_t* data, size_t size, hpvcd_yuv* yuv) {
    if (!yuv) return set_error(HPVCD_E_INVALID, "null argument");
    std::memset(yuv, 0, sizeof(*yuv));
    hpvcd_image img;
    int kind = -1;
    hpvcd_status st = decode_heic_kind(d, data, size, HEIC_KIND_YUV, &img, &kind);
    if (st) return st;
    const int chroma = HPVCD_OUT_IS_YUV(kind) ? HPVCD_OUT_YUV_CHROMA(kind) : 0;
    yuv->width = img.width;
    yuv->height = img.height;
    yuv->chroma_idc = (uint8_t)chroma;
    if (chroma) {
        const uint32_t sw = chroma == 1 ? 2 : 1, sh = chroma == 1 ? 2 : 1;
        yuv->chroma_width = (img.width + sw - 1) / sw;
        yuv->chroma_height = (img.height + sh - 1) / sh;
    }
    yuv->bit_depth = img.bit_depth;
    yuv->sample_bytes = img.sample_bytes;
    yuv->orientation = img.orientation;
    yuv->matrix = img.matrix;
    yuv->full_range = img.full_range;
    uint8_t* base = static_cast<uint8_t*>(img.pixels);
    yuv->y = base;
    if (chroma) {
        const size_t ly = (size_t)img.width * img.height * img.sample_bytes;
        const size_t lc = (size_t)yuv->chroma_width * yuv->chroma_height * img.sample_bytes;
        yuv->cb = base + ly;
        yuv->cr = base + ly + lc;
    }
    yuv->bytes = img.pixels_len;
    return HPVCD_OK;
}

void hpvcd_yuv_free(hpvcd_yuv* yuv) {
    if (!yuv) return;
    hpvcd_host_free(yuv->y);
    std::memset(yuv, 0, sizeof(*yuv));
}

hpvcd_status hpvcd_decode_hevc(hpvcd_decoder* d, const uint8_t* data, size_t size, int32_t out_kind,
                               hpvcd_image** frames, uint32_t* count) {
    if (!d || !data || !frames || !count) return set_error(HPVCD_E_INVALID, "null argument");
    *frames = nullptr;
    *count = 0;
    hpvcd_records* r = nullptr;
    hpvcd_status st = hpvcd_parse_hevc(data, size, out_kind, &r);
    if (st) return st;
    hpvcd_batch b;
    hpvcd_records_view(r, &b);
    hpvcd_image* imgs = nullptr;
    if (b.num_outputs) {
        imgs = static_cast<hpvcd_image*>(std::calloc(b.num_outputs, sizeof(hpvcd_image)));
        if (!imgs) {
            hpvcd_records_free(r);
            return set_error(HPVCD_E_ALLOC, "out of memory");
        }
    }
    st = hpvcd_decode_records(d, &b, imgs);
    const uint32_t n = b.num_outputs;
    hpvcd_records_free(r);
    if (st) {
        std::free(imgs);
        return st;
    }
    *frames = imgs;
    *count = n;
    return HPVCD_OK;
}

void hpvcd_frames_free(hpvcd_image* frames, uint32_t count) {
    if (!frames) return;
    for (uint32_t i = 0; i < count; i++) hpvcd_image_free(&frames[i]);
    std::free(frames);
}

}  // extern "C"

static_assert(sizeof(hpvcd_settings) == 56, "hpvcd_settings layout");
static_assert(sizeof(hpvcd_image) == 40, "hpvcd_image layout");
static_assert(sizeof(hpvcd_yuv) == 64, "hpvcd_yuv layout");
