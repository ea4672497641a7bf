// The path-tracing passes and the post-processing passes on the hot path.
//
//   PathTracer  src/passes/integrator/path_tracer.{h,cu}  (megakernel pass)
//   Wavefront   src/passes/wavefront/wavefront.{h,cu}     (stage-per-kernel pass)
//   LightTracer src/passes/integrator/light_tracer.{h,cu}  (particle tracing from the lights)
//   Simple      src/passes/npr/simple.{h,cu}                (first-hit NPR pass)
//   Accumulate  src/passes/postprocess/accumulate.{h,cu}
//   Denoise     src/passes/postprocess/denoise.{h,cu}       (a-trous filter in place of the OptiX denoiser)
//   ToneMapping src/passes/postprocess/tonemapping.{h,cu}
//
// PathTracer and Wavefront read the same JSON keys with the same defaults
// (samples_per_pixel 1, max_depth 16, rr_depth 4, use_nee true, use_mis true)
// and both run the wavefront HIP integrator: per path it reproduces the
// megakernel (the reference wavefront pass deviates from it and cannot run,
// its closest-hit module is commented out).  `seed` replaces the host random
// frame seed (path_tracer.cu:17): frame k renders with seed + k.
#pragma once
#include "renderpass.h"

namespace srt {

class PathTracer : public RenderPass {
protected:
    REGISTER_RENDER_PASS(PathTracer);

    int samples_per_pixel{1};
    int max_depth{16};
    int rr_depth{4};
    bool use_nee{true};
    bool use_mis{true};
    uint32_t seed{0};
    uint32_t frame{0};

public:
    PathTracer() {}
    void init() override;
    void render(std::shared_ptr<Film> film) override;
    void update() override {}          // the scene refits its shared BVH in Scene::update
    Json params() const override;

    friend void from_json(const Json& j, PathTracer& p) {
        if (j.is_null()) return;
        p.samples_per_pixel = j.value("samples_per_pixel", 1);
        p.max_depth = j.value("max_depth", 16);
        p.rr_depth = j.value("rr_depth", 4);
        p.use_nee = j.value("use_nee", true);
        p.use_mis = j.value("use_mis", true);
        p.seed = (uint32_t)j.value("seed", 0);
    }
};

class Wavefront : public PathTracer {
    REGISTER_RENDER_PASS(Wavefront);

public:
    friend void from_json(const Json& j, Wavefront& p) { from_json(j, (PathTracer&)p); }
};

// LightTracer: samples_per_pixel 1, max_depth 16, rr_depth 4
// (light_tracer.h:15-17, :28-35); frame k renders with seed + k.
class LightTracer : public RenderPass {
    REGISTER_RENDER_PASS(LightTracer);

    int samples_per_pixel{1};
    int max_depth{16};
    int rr_depth{4};
    uint32_t seed{0};
    uint32_t frame{0};

public:
    LightTracer() {}
    void init() override;
    void render(std::shared_ptr<Film> film) override;
    void update() override {}
    Json params() const override;

    friend void from_json(const Json& j, LightTracer& p) {
        if (j.is_null()) return;
        p.samples_per_pixel = j.value("samples_per_pixel", 1);
        p.max_depth = j.value("max_depth", 16);
        p.rr_depth = j.value("rr_depth", 4);
        p.seed = (uint32_t)j.value("seed", 0);
    }
};

// Simple: `type` Depth | Normal | BaseColor | Ambient | FaceOrientation
// (unknown names read as Depth, simple.h:24-31), samples_per_pixel 16; frame k
// renders with seed + k (the reference draws a random seed per frame).
class Simple : public RenderPass {
public:
    REGISTER_RENDER_PASS(Simple);
    enum class Type { Depth, Normal, BaseColor, Ambient, FaceOrientation };

private:
    Type type{Type::Depth};
    int samples_per_pixel{16};
    uint32_t seed{0};
    uint32_t frame{0};

public:
    Simple() {}
    void init() override;
    void render(std::shared_ptr<Film> film) override;
    void update() override {}          // the scene refits its shared BVH in Scene::update
    Json params() const override;

    static Type string_to_type(const std::string& str) {
        if (str == "Depth") return Type::Depth;
        if (str == "Normal") return Type::Normal;
        if (str == "BaseColor") return Type::BaseColor;
        if (str == "Ambient") return Type::Ambient;
        if (str == "FaceOrientation") return Type::FaceOrientation;
        return Type::Depth;
    }

    friend void from_json(const Json& j, Simple& p) {
        if (j.is_null()) return;
        p.type = Simple::string_to_type(j.value("type", "Depth"));
        p.samples_per_pixel = j.value("samples_per_pixel", 16);
        p.seed = (uint32_t)j.value("seed", 0);
    }
};

class Accumulate : public RenderPass {
    REGISTER_RENDER_PASS(Accumulate);

    int frame_count{0};
    float* accumulated{nullptr};     // device float4 per pixel

public:
    Accumulate() {}
    ~Accumulate() override;
    void init() override;
    void render(std::shared_ptr<Film> film) override;
    int get_frame_count() const { return frame_count; }
    Json params() const override;
    friend void from_json(const Json&, Accumulate&) {}
};

// Denoise: no parameters (denoise.h:31); filters the film in place.
class Denoise : public RenderPass {
    REGISTER_RENDER_PASS(Denoise);

public:
    Denoise() {}
    void render(std::shared_ptr<Film> film) override;
    friend void from_json(const Json&, Denoise&) {}
};

class ToneMapping : public RenderPass {
public:
    REGISTER_RENDER_PASS(ToneMapping);
    enum class Type { None, Clamp, Reinhard, Uncharted2, ACES };

private:
    Type type{Type::None};
    float exposure{0.0f};
    bool use_gamma{true};

public:
    ToneMapping() {}
    void render(std::shared_ptr<Film> film) override;
    Json params() const override;

    static Type string_to_type(const std::string& str) {
        if (str == "None") return Type::None;
        if (str == "Clamp") return Type::Clamp;
        if (str == "Reinhard") return Type::Reinhard;
        if (str == "Uncharted2") return Type::Uncharted2;
        if (str == "ACES") return Type::ACES;
        return Type::None;
    }

    friend void from_json(const Json& j, ToneMapping& p) {
        if (j.is_null()) return;
        p.type = ToneMapping::string_to_type(j.value("type", "None"));
        p.exposure = j.value("exposure", 0.0f);
        p.use_gamma = j.value("use_gamma", true);
    }
};

}  // namespace srt
