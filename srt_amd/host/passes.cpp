#include "passes.h"

#include <cstring>

namespace srt {

void check_srt(srt_ctx* ctx, int code, const char* what) {
    if (code < 0) throw RenderError(std::string(what) + ": " + (ctx ? srt_last_error(ctx) : "no context"));
}

Film::Film(srt_ctx* c, int w, int h) : ctx(c), width(w), height(h) {
    void* p = nullptr;
    check_srt(ctx, srt_device_alloc(ctx, (size_t)w * h * 16, &p), "film allocation");
    pixels = (float*)p;
    std::vector<float> zero((size_t)w * h * 4, 0.0f);
    upload(zero.data());
}
Film::~Film() {
    if (pixels) srt_device_free(ctx, pixels);
}
void Film::download(float* host) const {
    check_srt(ctx, srt_memcpy(ctx, host, pixels, (size_t)width * height * 16, 1), "film download");
}
void Film::upload(const float* host) {
    check_srt(ctx, srt_memcpy(ctx, pixels, host, (size_t)width * height * 16, 0), "film upload");
}

// ---------------------------------------------------------------------------
// PathTracer / Wavefront (path_tracer.cu:5-33)
// ---------------------------------------------------------------------------
REGISTER_RENDER_PASS_CPP(PathTracer);
REGISTER_RENDER_PASS_CPP(Wavefront);

void PathTracer::init() {
    if (!scene || !scene->ctx || !scene->handle) throw RenderError("PathTracer needs a scene uploaded to the device");
}

void PathTracer::render(std::shared_ptr<Film> film) {
    if (!enable) return;
    srt_camera cam{};
    for (int i = 0; i < 3; i++) {
        cam.position[i] = scene->camera->position[i];
        cam.target[i] = scene->camera->target[i];
        cam.up[i] = scene->camera->up[i];
    }
    cam.fov_deg = scene->camera->fov;
    srt_pt_params p{};
    p.samples_per_pixel = samples_per_pixel;
    p.max_depth = max_depth;
    p.rr_depth = rr_depth;
    p.use_nee = use_nee ? 1 : 0;
    p.use_mis = use_mis ? 1 : 0;
    p.seed = seed + frame;
    p.mlp_mode = 0;
    frame++;
    check_srt(scene->ctx, srt_render(scene->ctx, scene->handle, &cam, &p, width, height, film->get_pixels(), 1), "PathTracer");
}

Json PathTracer::params() const {
    Json j = Json::object();
    j["samples_per_pixel"] = Json(samples_per_pixel);
    j["max_depth"] = Json(max_depth);
    j["rr_depth"] = Json(rr_depth);
    j["use_nee"] = Json(use_nee);
    j["use_mis"] = Json(use_mis);
    j["seed"] = Json((double)seed);
    return j;
}

// ---------------------------------------------------------------------------
// Accumulate (accumulate.cu:10-29)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// LightTracer (passes/integrator/light_tracer.cu:13-37)
// ---------------------------------------------------------------------------
REGISTER_RENDER_PASS_CPP(LightTracer);

void LightTracer::init() {
    if (!scene || !scene->ctx || !scene->handle) throw RenderError("LightTracer needs a scene uploaded to the device");
}

void LightTracer::render(std::shared_ptr<Film> film) {
    if (!enable) return;
    srt_camera cam{};
    for (int i = 0; i < 3; i++) {
        cam.position[i] = scene->camera->position[i];
        cam.target[i] = scene->camera->target[i];
        cam.up[i] = scene->camera->up[i];
    }
    cam.fov_deg = scene->camera->fov;
    srt_pt_params p{};
    p.samples_per_pixel = samples_per_pixel;
    p.max_depth = max_depth;
    p.rr_depth = rr_depth;
    p.seed = seed + frame;
    frame++;
    check_srt(scene->ctx, srt_render_light(scene->ctx, scene->handle, &cam, &p, width, height, film->get_pixels(), 1),
              "LightTracer");
}

Json LightTracer::params() const {
    Json j = Json::object();
    j["samples_per_pixel"] = Json(samples_per_pixel);
    j["max_depth"] = Json(max_depth);
    j["rr_depth"] = Json(rr_depth);
    j["seed"] = Json((double)seed);
    return j;
}

// ---------------------------------------------------------------------------
// Simple (passes/npr/simple.cu:14-33)
// ---------------------------------------------------------------------------
REGISTER_RENDER_PASS_CPP(Simple);

void Simple::init() {
    if (!scene || !scene->ctx || !scene->handle) throw RenderError("Simple needs a scene uploaded to the device");
}

void Simple::render(std::shared_ptr<Film> film) {
    if (!enable) return;
    srt_camera cam{};
    for (int i = 0; i < 3; i++) {
        cam.position[i] = scene->camera->position[i];
        cam.target[i] = scene->camera->target[i];
        cam.up[i] = scene->camera->up[i];
    }
    cam.fov_deg = scene->camera->fov;
    const uint32_t s = seed + frame;
    frame++;
    check_srt(scene->ctx,
              srt_render_simple(scene->ctx, scene->handle, &cam, (int32_t)type, samples_per_pixel, s, width, height,
                                film->get_pixels(), 1),
              "Simple");
}

Json Simple::params() const {
    static const char* names[] = {"Depth", "Normal", "BaseColor", "Ambient", "FaceOrientation"};
    Json j = Json::object();
    j["type"] = Json(std::string(names[(int)type]));
    j["samples_per_pixel"] = Json(samples_per_pixel);
    j["seed"] = Json((double)seed);
    return j;
}

REGISTER_RENDER_PASS_CPP(Accumulate);

Accumulate::~Accumulate() {
    if (accumulated && scene && scene->ctx) srt_device_free(scene->ctx, accumulated);
}

void Accumulate::init() {
    if (!scene || !scene->ctx) throw RenderError("Accumulate needs a scene context");
    void* p = nullptr;
    check_srt(scene->ctx, srt_device_alloc(scene->ctx, (size_t)width * height * 16, &p), "accumulate buffer");
    accumulated = (float*)p;
    std::vector<float> zero((size_t)width * height * 4, 0.0f);
    check_srt(scene->ctx, srt_memcpy(scene->ctx, accumulated, zero.data(), zero.size() * 4, 0), "accumulate init");
}

void Accumulate::render(std::shared_ptr<Film> film) {
    if (!enable || !scene->is_static()) {
        frame_count = 0;
        return;
    }
    check_srt(scene->ctx,
              srt_accumulate(scene->ctx, accumulated, film->get_pixels(), (int64_t)width * height, frame_count, 1),
              "Accumulate");
    frame_count++;
}

Json Accumulate::params() const { return Json::object(); }

// ---------------------------------------------------------------------------
// ToneMapping (tonemapping.cu:5-70)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Denoise (passes/postprocess/denoise.cu:54-114)
// ---------------------------------------------------------------------------
REGISTER_RENDER_PASS_CPP(Denoise);

void Denoise::render(std::shared_ptr<Film> film) {
    if (!enable) return;
    check_srt(scene->ctx, srt_denoise(scene->ctx, film->get_pixels(), width, height, 1), "Denoise");
}

REGISTER_RENDER_PASS_CPP(ToneMapping);

void ToneMapping::render(std::shared_ptr<Film> film) {
    if (!enable) return;
    check_srt(scene->ctx,
              srt_tonemap(scene->ctx, film->get_pixels(), (int64_t)width * height, (int32_t)type, exposure, use_gamma ? 1 : 0, 1),
              "ToneMapping");
}

Json ToneMapping::params() const {
    static const char* names[] = {"None", "Clamp", "Reinhard", "Uncharted2", "ACES"};
    Json j = Json::object();
    j["type"] = Json(names[(int)type]);
    j["exposure"] = Json((double)exposure);
    j["use_gamma"] = Json(use_gamma);
    return j;
}

}  // namespace srt
