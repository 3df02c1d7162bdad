// server.cpp — the serving loop: HEIC decode requests arrive at any time,
// are parsed on host worker threads, and are continuously batched onto the
// GPU: everything parsed when a launch slot frees up goes out in one batch
// (tiles of several requests share one launch per stage), as far as the
// batch caps and the free HBM admit.  Two launch slots, each with its own
// stream, overlap one batch's host preparation and upload with the previous
// batch's kernels.  Each request's image is the one hpvcd_decode_heic gives
// for it alone (the batch changes which launch a picture rides in, not its
// arithmetic).  The reference has no serving layer; its Decoder reuse
// semantics are src/decoder.rs:54-92.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hpvcd.h"
#include "heif.h"
#include "internal.h"

using namespace hpvcd;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    uint64_t id = 0;
    std::vector<uint8_t> file;
    int kind = HEIC_KIND_DEFAULT;
    hpvcd_records* rec = nullptr;
    size_t pictures = 0, est_bytes = 0;
    hpvcd_status st = HPVCD_OK;
    std::string err;
    hpvcd_image img;
    bool done = false, collected = false;
    Clock::time_point t_submit, t_parsed, t_launch, t_done;
    uint32_t batch_requests = 0, batch_pictures = 0;
};

struct Slot {
    hpvcd_gpu* g = nullptr;
    hpvcd_job* job = nullptr;
    hpvcd_records* rec = nullptr;
    std::vector<Request*> reqs;
};

// Device bytes a request's batch share needs, from its records: planes and
// residual pools, per-sample reference words, records and outputs (an
// upper-end estimate of what hpvcd_job_create lays out).
size_t estimate_bytes(const hpvcd_batch& b) {
    size_t n = 0;
    for (uint32_t p = 0; p < b.num_pictures; p++) {
        const hpvcd_picture& P = b.pictures[p];
        const size_t s = (size_t)P.width * P.height + 2 * (size_t)chroma_w(P) * chroma_h(P);
        n += s * 20;
    }
    n += (size_t)b.num_tbs * 96 + (size_t)b.num_coeffs * 8 + (size_t)b.num_ctbs * 64 + (size_t)b.num_cells * 8;
    for (uint32_t o = 0; o < b.num_outputs; o++) n += hpvcd_output_bytes(&b.outputs[o]);
    return n + ((size_t)4 << 20);
}

}  // namespace

struct hpvcd_server {
    hpvcd_settings s;
    hpvcd_server_config cfg;
    Slot slots[2];
    std::vector<std::thread> parsers;
    std::thread dispatcher;
    std::mutex mu;
    std::condition_variable cv_parse, cv_ready, cv_done;
    std::deque<Request*> to_parse, ready;
    std::map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    int parse_threads = 1;
    int parsing = 0;  // requests a parse worker holds
    // totals for hpvcd_server_stats
    uint64_t batches = 0, batched_requests = 0, batched_pictures = 0;

    void parse_loop() {
        for (;;) {
            Request* r;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv_parse.wait(lock, [&] { return stop || !to_parse.empty(); });
                if (stop && to_parse.empty()) return;
                r = to_parse.front();
                to_parse.pop_front();
                parsing++;
            }
            hpvcd_records* rec = nullptr;
            hpvcd_status st = heic_records(r->file.data(), r->file.size(), limits_of(&s), parse_threads, r->kind,
                                           &rec, nullptr);
            std::string err = st ? hpvcd_last_error() : "";
            if (!st) {
                hpvcd_batch b;
                hpvcd_records_view(rec, &b);
                st = hpvcd_validate_batch(&b);
                if (st) err = hpvcd_last_error();
                for (uint32_t o = 0; !st && o < b.num_outputs; o++) {
                    const hpvcd_output& O = b.outputs[o];
                    if ((uint32_t)O.width > s.max_dimension || (uint32_t)O.height > s.max_dimension ||
                        (uint64_t)O.width * (uint64_t)O.height > s.max_pixels) {
                        st = HPVCD_E_LIMIT;
                        err = "image dimensions exceed the server's limits";
                    }
                }
                if (!st) {
                    r->pictures = b.num_pictures;
                    r->est_bytes = estimate_bytes(b);
                }
            }
            std::lock_guard<std::mutex> lock(mu);
            parsing--;
            r->t_parsed = Clock::now();
            if (st) {
                if (rec) hpvcd_records_free(rec);
                r->st = st;
                r->err = err;
                r->done = true;
                r->t_done = r->t_parsed;
                cv_done.notify_all();
            } else {
                r->rec = rec;
                ready.push_back(r);
            }
            // a new ready request, or one parse fewer for a waiting slot
            cv_ready.notify_one();
        }
    }

    // download a finished slot's images and hand them to their requests
    void complete(Slot& sl) {
        std::vector<void*> outs;
        hpvcd_batch b;
        hpvcd_records_view(sl.rec, &b);
        hpvcd_status st = HPVCD_OK;
        std::vector<hpvcd_image> imgs(b.num_outputs);
        for (uint32_t o = 0; o < b.num_outputs; o++) {
            hpvcd_image& I = imgs[o];
            std::memset(&I, 0, sizeof(I));
            const hpvcd_output& O = b.outputs[o];
            const bool swap = O.orientation >= HPVCD_ORIENT_TRANSPOSE && O.orientation <= HPVCD_ORIENT_ROTATE_270;
            I.width = (uint32_t)(swap ? O.height : O.width);
            I.height = (uint32_t)(swap ? O.width : O.height);
            I.channels = (O.kind == HPVCD_OUT_LUMA8 || O.kind == HPVCD_OUT_LUMA16) ? 1 : 3;
            I.sample_bytes = (O.kind == HPVCD_OUT_RGB16 || O.kind == HPVCD_OUT_LUMA16) ? 2 : 1;
            I.orientation = O.orientation;
            for (uint32_t p = 0; p < b.num_pictures; p++)
                if (b.pictures[p].out_index == o) {
                    I.bit_depth = b.pictures[p].bit_depth_luma;
                    I.matrix = b.pictures[p].matrix;
                    I.full_range = (b.pictures[p].flags & HPVCD_PF_FULL_RANGE) ? 1 : 0;
                    break;
                }
            I.pixels_len = hpvcd_output_bytes(&O);
            if (!st) st = hpvcd_host_alloc(I.pixels_len, &I.pixels);
            outs.push_back(I.pixels);
        }
        std::string err;
        if (!st) st = hpvcd_job_download(sl.job, outs.data());
        if (st) err = hpvcd_last_error();
        hpvcd_job_free(sl.job);
        hpvcd_records_free(sl.rec);
        sl.job = nullptr;
        sl.rec = nullptr;
        const auto now = Clock::now();
        std::lock_guard<std::mutex> lock(mu);
        // outputs follow the requests in batch order: request k owns output k
        for (size_t k = 0; k < sl.reqs.size(); k++) {
            Request* r = sl.reqs[k];
            if (st) {
                r->st = st;
                r->err = err;
                if (k < imgs.size()) hpvcd_image_free(&imgs[k]);
            } else {
                r->img = imgs[k];
            }
            r->done = true;
            r->t_done = now;
        }
        sl.reqs.clear();
        cv_done.notify_all();
    }

    void dispatch_loop() {
        std::deque<int> busy;  // slots in launch order
        for (;;) {
            std::vector<Request*> batch;
            {
                std::unique_lock<std::mutex> lock(mu);
                if (busy.empty()) cv_ready.wait(lock, [&] { return stop || !ready.empty(); });
                if (stop && ready.empty() && busy.empty()) return;
                // a free slot waits a little for requests still in the parse
                if (cfg.batch_delay_us && (int)busy.size() < 2 && !ready.empty()) {
                    const auto until = ready.front()->t_parsed + std::chrono::microseconds(cfg.batch_delay_us);
                    cv_ready.wait_until(lock, until, [&] {
                        size_t pics = 0;
                        for (const Request* q : ready) pics += q->pictures;
                        return stop || (to_parse.empty() && parsing == 0) || pics >= cfg.max_batch_pictures;
                    });
                }
                // continuous batching: everything parsed now, within the caps
                size_t pics = 0, bytes = 0;
                size_t free_b = 0, total_b = 0;
                const int s0 = busy.empty() ? 0 : 1 - busy.front();
                if (hpvcd_gpu_mem_info(slots[s0].g, &free_b, &total_b)) free_b = 0;
                const size_t budget = free_b > cfg.hbm_reserve ? free_b - cfg.hbm_reserve : 0;
                while (!ready.empty() && (int)busy.size() < 2) {
                    Request* r = ready.front();
                    const bool first = batch.empty();
                    if (!first && (pics + r->pictures > cfg.max_batch_pictures ||
                                   bytes + r->est_bytes > cfg.max_batch_bytes || bytes + r->est_bytes > budget))
                        break;
                    // a request larger than the budget waits until in-flight work frees HBM
                    if (first && r->est_bytes > budget && !busy.empty()) break;
                    ready.pop_front();
                    batch.push_back(r);
                    pics += r->pictures;
                    bytes += r->est_bytes;
                }
            }
            if (!batch.empty()) {
                const int si = busy.empty() ? 0 : 1 - busy.front();
                Slot& sl = slots[si];
                std::vector<hpvcd_records*> sets;
                for (Request* r : batch) sets.push_back(r->rec);
                hpvcd_status st = hpvcd_records_concat(sets.data(), (uint32_t)sets.size(), &sl.rec);
                for (Request* r : batch) {
                    hpvcd_records_free(r->rec);
                    r->rec = nullptr;
                }
                hpvcd_batch b;
                if (!st) {
                    hpvcd_records_view(sl.rec, &b);
                    st = job_create_validated(sl.g, &b, &sl.job);
                }
                if (!st) st = hpvcd_job_run(sl.job, s.backend == HPVCD_BACKEND_GPU_PLAIN ? HPVCD_MODE_PLAIN
                                                                                        : HPVCD_MODE_BATCHED,
                                            HPVCD_STAGE_RGB);
                const auto now = Clock::now();
                if (st) {
                    std::string err = hpvcd_last_error();
                    if (sl.job) hpvcd_job_free(sl.job);
                    if (sl.rec) hpvcd_records_free(sl.rec);
                    sl.job = nullptr;
                    sl.rec = nullptr;
                    std::lock_guard<std::mutex> lock(mu);
                    for (Request* r : batch) {
                        r->st = st;
                        r->err = err;
                        r->done = true;
                        r->t_launch = r->t_done = now;
                    }
                    cv_done.notify_all();
                } else {
                    std::lock_guard<std::mutex> lock(mu);
                    for (Request* r : batch) {
                        r->t_launch = now;
                        r->batch_requests = (uint32_t)batch.size();
                        r->batch_pictures = b.num_pictures;
                    }
                    batches++;
                    batched_requests += batch.size();
                    batched_pictures += b.num_pictures;
                    sl.reqs = batch;
                    busy.push_back(si);
                }
            }
            // finish the oldest batch when both slots are busy or nothing
            // else is ready to prepare
            bool idle;
            {
                std::lock_guard<std::mutex> lock(mu);
                idle = ready.empty();
            }
            if (!busy.empty() && ((int)busy.size() == 2 || idle || batch.empty())) {
                const int si = busy.front();
                busy.pop_front();
                complete(slots[si]);
            }
        }
    }
};

extern "C" {

void hpvcd_server_config_default(hpvcd_server_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->max_batch_pictures = 256;
    c->max_batch_bytes = (uint64_t)16 << 30;
    c->hbm_reserve = (uint64_t)8 << 30;
    c->parse_workers = 4;
}

hpvcd_status hpvcd_server_new(const hpvcd_settings* s, const hpvcd_server_config* c, hpvcd_server** out) {
    if (!out) return set_error(HPVCD_E_INVALID, "null argument");
    *out = nullptr;
    hpvcd_settings def;
    hpvcd_settings_default(&def);
    hpvcd_server_config cdef;
    hpvcd_server_config_default(&cdef);
    const hpvcd_settings& S = s ? *s : def;
    const hpvcd_server_config& K = c ? *c : cdef;
    if (S.backend != HPVCD_BACKEND_GPU_PLAIN && S.backend != HPVCD_BACKEND_GPU_BATCHED)
        return set_error(HPVCD_E_UNSUPPORTED, "the server runs the GPU backends only");
    if (K.max_batch_pictures == 0 || K.parse_workers == 0) return set_error(HPVCD_E_INVALID, "zero server limits");
    hpvcd_server* v = new (std::nothrow) hpvcd_server();
    if (!v) return set_error(HPVCD_E_ALLOC, "out of memory");
    v->s = S;
    v->cfg = K;
    for (auto& sl : v->slots)
        if (hpvcd_status st = hpvcd_gpu_open(S.device, &sl.g)) {
            for (auto& x : v->slots)
                if (x.g) hpvcd_gpu_close(x.g);
            delete v;
            return st;
        }
    const int host = S.threads > 0 ? S.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    v->parse_threads = std::max(1, host / (int)K.parse_workers);
    for (uint32_t i = 0; i < K.parse_workers; i++) v->parsers.emplace_back([v] { v->parse_loop(); });
    v->dispatcher = std::thread([v] { v->dispatch_loop(); });
    *out = v;
    return HPVCD_OK;
}

hpvcd_status hpvcd_server_submit(hpvcd_server* v, const uint8_t* data, size_t size, int32_t rgb8, uint64_t* ticket) {
    if (!v || !data || !ticket) return set_error(HPVCD_E_INVALID, "null argument");
    Request* r = new (std::nothrow) Request();
    if (!r) return set_error(HPVCD_E_ALLOC, "out of memory");
    r->file.assign(data, data + size);
    r->kind = rgb8 ? HEIC_KIND_RGB8 : HEIC_KIND_DEFAULT;
    std::memset(&r->img, 0, sizeof(r->img));
    std::lock_guard<std::mutex> lock(v->mu);
    if (v->stop) {
        delete r;
        return set_error(HPVCD_E_INVALID, "server is shutting down");
    }
    r->id = v->next_id++;
    r->t_submit = Clock::now();
    v->reqs[r->id] = r;
    v->to_parse.push_back(r);
    v->cv_parse.notify_one();
    *ticket = r->id;
    return HPVCD_OK;
}

hpvcd_status hpvcd_server_wait(hpvcd_server* v, uint64_t ticket, hpvcd_image* img, hpvcd_request_stats* stats) {
    if (!v || !img) return set_error(HPVCD_E_INVALID, "null argument");
    std::memset(img, 0, sizeof(*img));
    std::unique_lock<std::mutex> lock(v->mu);
    auto it = v->reqs.find(ticket);
    if (it == v->reqs.end()) return set_error(HPVCD_E_INVALID, "unknown ticket");
    Request* r = it->second;
    v->cv_done.wait(lock, [&] { return r->done; });
    v->reqs.erase(it);
    lock.unlock();
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->latency_ms = ms_between(r->t_submit, r->t_done);
        stats->parse_ms = ms_between(r->t_submit, r->t_parsed);
        if (r->t_launch.time_since_epoch().count()) {
            stats->queue_ms = ms_between(r->t_parsed, r->t_launch);
            stats->device_ms = ms_between(r->t_launch, r->t_done);
        }
        stats->batch_requests = r->batch_requests;
        stats->batch_pictures = r->batch_pictures;
    }
    const hpvcd_status st = r->st;
    if (st) set_error(st, r->err.c_str());
    else *img = r->img;
    delete r;
    return st;
}

void hpvcd_server_totals(hpvcd_server* v, uint64_t* batches, uint64_t* requests, uint64_t* pictures) {
    if (!v) return;
    std::lock_guard<std::mutex> lock(v->mu);
    if (batches) *batches = v->batches;
    if (requests) *requests = v->batched_requests;
    if (pictures) *pictures = v->batched_pictures;
}

void hpvcd_server_free(hpvcd_server* v) {
    if (!v) return;
    {
        std::lock_guard<std::mutex> lock(v->mu);
        v->stop = true;
    }
    v->cv_parse.notify_all();
    for (auto& t : v->parsers) t.join();
    {
        std::lock_guard<std::mutex> lock(v->mu);
        v->cv_ready.notify_all();
    }
    v->dispatcher.join();
    for (auto& kv : v->reqs) {
        hpvcd_image_free(&kv.second->img);
        if (kv.second->rec) hpvcd_records_free(kv.second->rec);
        delete kv.second;
    }
    for (auto& sl : v->slots) hpvcd_gpu_close(sl.g);
    delete v;
}

}  // extern "C"

static_assert(sizeof(hpvcd_server_config) == 32, "hpvcd_server_config layout");
static_assert(sizeof(hpvcd_request_stats) == 40, "hpvcd_request_stats layout");
