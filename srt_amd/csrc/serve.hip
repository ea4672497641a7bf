// Render-serving loop: continuous batching of render requests into the shared
// wavefront path queues.
//
// The reference has no serving path; its renderer loops render one frame at a
// time, synchronously (src/renderer/renderer.cu:176-275).  Here a dispatcher
// thread owns the context's device buffers while serving runs:
//   1. admission: queued requests (FIFO) are admitted while their device film
//      (12 B accumulator + 16 B output per pixel) fits in the context's HBM
//      budget; a request that can never fit fails with SRT_E_OOM;
//   2. batching: each chunk takes the scene of the admitted request with the
//      smallest key (remaining paths at 2 Gpaths/s + w x submit time, w =
//      serve_age / 1000: 0 is shortest remaining first, 1 earliest due time)
//      and fills the chunk (2^24 paths by default) with the remaining paths
//      of the admitted requests on that scene in key order (small requests
//      share a chunk; a large one is cut into chunks, and newer small
//      requests overtake it until its waiting time outweighs their sizes),
//      one segment per request; requests admitted while a chunk runs join
//      the next chunk;
//   3. completion: a request whose last paths were in the chunk has its film
//      copied to the caller's buffer and its device memory released.
// A request's paths run in increasing path-id order with its own PCG streams,
// so its image is bit-identical alone or in any mix.
#include "srt_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>

namespace srt {

struct Request {
    uint64_t ticket;
    srt_request req;
    int32_t status = 0;               // 0 queued, 1 admitted/running, 2 done, < 0 error
    int64_t total = 0, next = 0;      // paths, next path id to schedule
    DevBuf acc, out;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int64_t chunks = 0;
    bool inflight = false;            // part of the chunk the dispatcher is running
};

// default paths per serving chunk: about 9 ms of GPU time at 1.9 Gpaths/s,
// the longest a newly admitted request waits for the device
constexpr int64_t kServeChunk = 1ll << 24;
// paths per millisecond that the scheduling due times assume (2 Gpaths/s)
constexpr double kServeRate = 2.0e6;

struct ServeState {
    std::thread th;
    std::mutex mu;
    std::condition_variable cv, done_cv;
    bool stop = false;
    int64_t pool = 0;
    std::deque<Request*> pending;
    std::vector<Request*> active;
    std::map<uint64_t, Request*> all;
    uint64_t next_ticket = 1;
    std::chrono::steady_clock::time_point t0;
    int64_t served_paths = 0, chunks = 0;
};

static double now_ms(ServeState* S) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - S->t0).count();
}

__global__ void k_zero_f(float* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

static size_t request_bytes(const srt_request& r) { return (size_t)r.width * r.height * (12 + 16); }

static void finish(ServeState* S, Request* q, int32_t status) {
    q->status = status;
    q->t_done = now_ms(S);
}

static void dispatcher(srt_ctx* c) {
    ServeState* S = (ServeState*)c->serve;
    (void)hipSetDevice(c->device);
    std::vector<Segment> segs;
    std::vector<int64_t> take;                // paths of each batch request in the chunk
    for (;;) {
        std::vector<Request*> batch;
        SceneRec* scene = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            // admission (FIFO) against the HBM budget
            while (!S->pending.empty()) {
                Request* q = S->pending.front();
                size_t need = request_bytes(q->req);
                if (c->used.load() + need > c->budget) {
                    if (S->active.empty() && release_build_scratch(c)) continue;   // kept BVH-build chunks first
                    if (S->active.empty()) {          // nothing will ever free memory for it
                        S->pending.pop_front();
                        set_err(c, SRT_E_OOM, "request %llu needs %zu bytes beyond the HBM budget",
                                (unsigned long long)q->ticket, need);
                        finish(S, q, SRT_E_OOM);
                        S->done_cv.notify_all();
                        continue;
                    }
                    break;
                }
                S->pending.pop_front();
                int r = dev_alloc(c, q->acc, (size_t)q->req.width * q->req.height * 12);
                if (!r) r = dev_alloc(c, q->out, (size_t)q->req.width * q->req.height * 16);
                if (r) {
                    dev_free(c, q->acc);
                    dev_free(c, q->out);
                    finish(S, q, r);
                    S->done_cv.notify_all();
                    continue;
                }
                size_t n3 = (size_t)q->req.width * q->req.height * 3;
                k_zero_f<<<(unsigned)((n3 + 255) / 256), 256, 0, c->stream>>>((float*)q->acc.p, n3);
                q->status = 1;
                q->t_admit = now_ms(S);
                S->active.push_back(q);
            }
            if (S->active.empty()) {
                if (S->stop) break;
                S->cv.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
            // chunk: the scene of the admitted request with the smallest
            // key, remaining paths / kServeRate (the time its remaining
            // work would take on an idle device) + w x submit time; the
            // chunk takes that scene's requests in key order, up to
            // S->pool paths.  Small requests thus overtake a large one
            // between its chunks until w x its waiting time outweighs
            // their sizes (ties: arrival order).
            const double age_w = c->serve_age_pm * 1e-3;
            auto due = [age_w](const Request* q) {
                return age_w * q->t_submit + (double)(q->total - q->next) / kServeRate;
            };
            Request* first = nullptr;
            for (Request* q : S->active)
                if (!first || due(q) < due(first)) first = q;
            uint64_t sh = first->req.scene;
            auto it = c->scenes.find(sh);
            scene = it == c->scenes.end() ? nullptr : it->second;
            std::vector<Request*> cand;
            for (Request* q : S->active)
                if (q->req.scene == sh) cand.push_back(q);
            std::stable_sort(cand.begin(), cand.end(), [&](const Request* a, const Request* b) { return due(a) < due(b); });
            int64_t left = S->pool;
            take.clear();
            for (Request* q : cand) {
                if (left == 0 || (int)batch.size() == kMaxSegments) break;
                q->inflight = true;
                batch.push_back(q);
                take.push_back(std::min(q->total - q->next, left));
                left -= take.back();
            }
        }
        if (!scene) {
            std::lock_guard<std::mutex> lk(S->mu);
            for (Request* q : batch) {
                finish(S, q, SRT_E_NOT_FOUND);
                dev_free(c, q->acc);
                dev_free(c, q->out);
                S->active.erase(std::find(S->active.begin(), S->active.end(), q));
            }
            S->done_cv.notify_all();
            continue;
        }
        segs.clear();
        int64_t off = 0;
        for (size_t k = 0; k < batch.size(); k++) {
            if (take[k] == 0) continue;
            Request* q = batch[k];
            const srt_request& r = q->req;
            Segment sg;
            sg.path_begin = q->next;
            sg.n = (int32_t)take[k];
            sg.slot_offset = (int32_t)off;
            sg.wh = r.width * r.height;
            sg.width = r.width;
            sg.height = r.height;
            sg.spp = r.params.samples_per_pixel;
            sg.seed = r.params.seed;
            sg.max_depth = r.params.max_depth;
            sg.rr_depth = r.params.rr_depth;
            sg.use_nee = (r.params.use_nee && scene->nlights > 0) ? 1 : 0;
            sg.use_mis = r.params.use_mis ? 1 : 0;
            sg.mlp_mode = r.params.mlp_mode;
            sg.cam = make_camera(r.camera.position, r.camera.target, r.camera.up, r.camera.fov_deg, r.width, r.height);
            sg.film_acc = (float*)q->acc.p;
            sg.film_out = (float4*)q->out.p;
            sg.last = (q->next + take[k] >= q->total) ? 1 : 0;
            segs.push_back(sg);
            off += take[k];
        }
        hipEvent_t ev;
        prof_begin(c, "serve_chunk", &ev);
        int rc = run_chunk(c, scene, segs.data(), (int)segs.size(), nullptr);
        prof_end(c, "serve_chunk", ev, off);
        prof_collect(c);
        std::vector<Request*> completed;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->chunks++;
            S->served_paths += off;
            for (size_t k = 0; k < batch.size(); k++) {
                Request* q = batch[k];
                q->inflight = false;
                if (take[k] == 0) continue;
                q->next += take[k];
                q->chunks++;
                if (rc != SRT_OK || q->next >= q->total) completed.push_back(q);
            }
        }
        for (Request* q : completed) {
            int r = rc;
            if (!r && q->req.rgba_out)
                r = hip_check(c, hipMemcpy(q->req.rgba_out, q->out.p, (size_t)q->req.width * q->req.height * 16,
                                           hipMemcpyDeviceToHost), "result download");
            std::lock_guard<std::mutex> lk(S->mu);
            dev_free(c, q->acc);
            dev_free(c, q->out);
            S->active.erase(std::find(S->active.begin(), S->active.end(), q));
            finish(S, q, r ? r : 2);
        }
        if (!completed.empty()) S->done_cv.notify_all();
    }
}

static void serve_free(srt_ctx* c) {
    ServeState* S = (ServeState*)c->serve;
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        // queued requests that were never admitted are cancelled
        for (Request* q : S->pending) finish(S, q, SRT_E_STATE);
        S->pending.clear();
    }
    S->cv.notify_all();
    if (S->th.joinable()) S->th.join();
    for (auto& kv : S->all) {
        dev_free(c, kv.second->acc);
        dev_free(c, kv.second->out);
        delete kv.second;
    }
    delete S;
    c->serve = nullptr;
}

}  // namespace srt

using namespace srt;

extern "C" {

int srt_serve_start(srt_ctx* c, int64_t path_pool) {
    if (!c) return SRT_E_INVALID;
    if (c->serve) return set_err(c, SRT_E_STATE, "serving already started");
    if (path_pool != 0 && !path_pool_ok(path_pool))
        return set_err(c, SRT_E_INVALID, "path_pool must be 0 (the context's) or in [%lld, %lld]",
                       (long long)kMinPathPool, (long long)kMaxPathPool);
    ServeState* S = new ServeState();
    S->pool = path_pool > 0 ? path_pool : std::min<int64_t>(c->pool_request, kServeChunk);
    S->t0 = std::chrono::steady_clock::now();
    c->serve = S;
    c->serve_free = serve_free;
    S->th = std::thread(dispatcher, c);
    return SRT_OK;
}

int srt_serve_stop(srt_ctx* c, int32_t drain) {
    if (!c) return SRT_E_INVALID;
    if (!c->serve) return set_err(c, SRT_E_STATE, "serving not started");
    ServeState* S = (ServeState*)c->serve;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        if (!drain) {
            // cancel what has not started: queued requests and admitted ones
            // with no path scheduled yet (their films are freed here)
            for (Request* q : S->pending) finish(S, q, SRT_E_STATE);
            S->pending.clear();
            std::vector<Request*> keep;
            for (Request* q : S->active) {
                if (q->next == 0 && !q->inflight) {
                    finish(S, q, SRT_E_STATE);
                    dev_free(c, q->acc);
                    dev_free(c, q->out);
                } else {
                    keep.push_back(q);
                }
            }
            S->active.swap(keep);
            S->done_cv.notify_all();
        }
        S->done_cv.wait(lk, [&] { return S->pending.empty() && S->active.empty(); });
    }
    serve_free(c);
    return SRT_OK;
}

int srt_submit(srt_ctx* c, const srt_request* r, uint64_t* ticket) {
    if (!c || !r || !ticket) return SRT_E_INVALID;
    ServeState* S = (ServeState*)c->serve;
    if (!S) return SRT_E_STATE;
    const srt_pt_params& p = r->params;
    if (r->width <= 0 || r->height <= 0 || (int64_t)r->width * r->height > (1ll << 26) || p.samples_per_pixel <= 0 ||
        p.max_depth < 0 || p.rr_depth < 0 || p.mlp_mode < 0 || p.mlp_mode > 1 ||
        !(r->camera.fov_deg > 0.0f && r->camera.fov_deg < 180.0f))
        return SRT_E_INVALID;
    if (p.mlp_mode == 1 && !mlp_for_scene(c, r->scene)) return SRT_E_STATE;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = c->scenes.find(r->scene);
    if (it == c->scenes.end() || !it->second->has_bvh) return SRT_E_NOT_FOUND;
    Request* q = new Request();
    q->ticket = S->next_ticket++;
    q->req = *r;
    q->total = (int64_t)r->width * r->height * p.samples_per_pixel;
    q->t_submit = now_ms(S);
    S->all[q->ticket] = q;
    S->pending.push_back(q);
    *ticket = q->ticket;
    S->cv.notify_all();
    return SRT_OK;
}

static void fill_result(const Request* q, srt_result* out) {
    out->status = q->status;
    out->submit_ms = q->t_submit;
    out->admit_ms = q->t_admit;
    out->done_ms = q->t_done;
    out->paths = q->total;
    out->chunks = q->chunks;
}

int srt_poll(srt_ctx* c, uint64_t ticket, srt_result* out) {
    if (!c || !out) return SRT_E_INVALID;
    ServeState* S = (ServeState*)c->serve;
    if (!S) return SRT_E_STATE;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return SRT_E_NOT_FOUND;
    fill_result(it->second, out);
    return SRT_OK;
}

int srt_wait(srt_ctx* c, uint64_t ticket, int32_t timeout_ms, srt_result* out) {
    if (!c || !out) return SRT_E_INVALID;
    ServeState* S = (ServeState*)c->serve;
    if (!S) return SRT_E_STATE;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return SRT_E_NOT_FOUND;
    Request* q = it->second;
    auto pred = [&] { return q->status == 2 || q->status < 0; };
    if (timeout_ms < 0) S->done_cv.wait(lk, pred);
    else S->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
    fill_result(q, out);
    return SRT_OK;
}

int srt_serve_stats(srt_ctx* c, int64_t* served_paths, int64_t* chunks, int32_t* queued, int32_t* active) {
    if (!c) return SRT_E_INVALID;
    ServeState* S = (ServeState*)c->serve;
    if (!S) return SRT_E_STATE;
    std::lock_guard<std::mutex> lk(S->mu);
    if (served_paths) *served_paths = S->served_paths;
    if (chunks) *chunks = S->chunks;
    if (queued) *queued = (int32_t)S->pending.size();
    if (active) *active = (int32_t)S->active.size();
    return SRT_OK;
}

}  // extern "C"
