#include "renderer.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>

#include "passes.h"

namespace srt {

static std::string parent_dir(const std::string& p) {
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
}
static std::string join(const std::string& dir, const std::string& rel) {
    if (!rel.empty() && rel[0] == '/') return rel;
    return dir + "/" + rel;
}

Renderer::Renderer(const std::string& path) : config_path(path) {
    int r = srt_create(0, 0, &ctx);
    if (r < 0) throw RenderError("no HIP device available (srt_create returned " + std::to_string(r) + ")");
}

Renderer::~Renderer() {
    passes.clear();
    film.reset();
    scene.reset();
    if (ctx) srt_destroy(ctx);
}

// renderer.cu:3-18
void Renderer::load_passes(const Json& config) {
    for (auto& c : config.items()) {
        std::string name = c.at("name").get_string();
        std::shared_ptr<RenderPass> pass = RenderPassFactory::create_pass(name, c["params"]);
        pass->set_enable(c.value("enable", true));
        pass->set_online(online);
        pass->resize(width, height);
        pass->set_scene(scene);
        pass->init();
        passes.push_back(pass);
    }
}

static void vec3(const Json& j, float out[3]) {
    if (!j.is_array() || j.size() != 3) throw RenderError("expected a 3-vector in the scene config");
    for (int i = 0; i < 3; i++) out[i] = j.at((size_t)i).get_float();
}

// renderer.cu:42-174
void Renderer::load_scene(const Json& config) {
    std::string model_path = join(parent_dir(config_path), config.at("model").get_string());
    std::string ext = model_path.substr(model_path.find_last_of('.') + 1);
    scene = std::make_shared<Scene>();
    if (ext == "obj") load_obj(model_path, *scene);
    else if (ext == "srtscene") load_srtscene(model_path, *scene);
    else throw RenderError("unsupported model format '." + ext + "' (obj or srtscene)");
    if (width == 0 || height == 0) resize(1920, 1080);

    if (config.contains("camera")) {
        const Json& cc = config.at("camera");
        auto cam = std::make_shared<Camera>();
        cam->type = cc.at("type").get_string();
        if (cam->type != "perspective") throw RenderError("Unknown camera type: " + cam->type);
        vec3(cc.at("position"), cam->position);
        vec3(cc.at("target"), cam->target);
        vec3(cc.at("up"), cam->up);
        cam->fov = cc.value("fov", 60.0f);
        scene->camera = cam;
    } else {
        throw RenderError("No camera specified");
    }
    if (config.contains("environment")) {
        const Json& e = config.at("environment");
        std::string t = e.at("type").get_string();
        if (t == "constant") vec3(e.at("color"), scene->background);
        else throw RenderError("Unknown environment type: " + t);
    }
    if (config.contains("animation")) {
        // a list of keyframed node animations (renderer.cu:71-121), or a flag
        // that enables the skinning animation stored in an .srtscene
        const Json& a = config.at("animation");
        if (a.is_array()) {
            for (size_t i = 0; i < a.size(); i++) scene->node_anims.push_back(parse_node_animation(a.at(i)));
            scene->enable_animation = true;
        } else {
            scene->enable_animation = a.get_bool();
        }
    }
    scene->build_gscene(ctx);
    film = std::make_shared<Film>(ctx, width, height);
}

void Renderer::print_profile() const {
    char names[4096];
    double ms[64];
    int64_t launches[64], items[64];
    int k = srt_stats(ctx, names, sizeof(names), ms, launches, items, 64);
    const char* p = names;
    for (int i = 0; i < k; i++) {
        std::printf("  %-12s launches %8lld\n", p, (long long)launches[i]);
        p += std::strlen(p) + 1;
    }
}

void ImageRenderer::run() {
    auto t0 = std::chrono::steady_clock::now();
    for (auto& pass : passes) pass->render(film);
    auto t1 = std::chrono::steady_clock::now();
    std::printf("Render: %.3f ms\n", std::chrono::duration<double, std::milli>(t1 - t0).count());
    if (!filename.empty()) {
        std::vector<float> px((size_t)width * height * 4);
        film->download(px.data());
        save_image(join(parent_dir(config_path), filename), width, height, px.data());
    }
    print_profile();
}

void InteractiveRenderer::run() {
    float t = 0.0f;
    for (int i = 0; i < frames; i++) {
        scene->update(t);
        for (auto& pass : passes) {
            pass->update();
            pass->render(film);
        }
        scene->camera->set_moved(false);
        t += 1.0f / 30.0f;
    }
    if (!filename.empty()) {
        std::vector<float> px((size_t)width * height * 4);
        film->download(px.data());
        save_image(join(parent_dir(config_path), filename), width, height, px.data());
    }
}

void VideoRenderer::run() {
    if (frame <= 0) throw RenderError("video needs frame > 0");
    std::string out = join(parent_dir(config_path), filename.empty() ? std::string("video.pfm") : filename);
    std::string stem = out.substr(0, out.find_last_of('.'));
    float step = 1.0f / 30.0f, t = 0.0f;
    std::vector<float> px((size_t)width * height * 4);
    for (int i = 0; i < frame; i++) {
        scene->update(t);
        for (auto& pass : passes) {
            pass->update();
            pass->render(film);
        }
        film->download(px.data());
        char buf[32];
        std::snprintf(buf, sizeof(buf), "_%04d.pfm", i);
        save_image(stem + buf, width, height, px.data());
        t += step;
    }
}

std::shared_ptr<Renderer> create_renderer(const std::string& path, const Json& config) {
    std::string type = config.at("type").get_string();
    std::string output = config.value("output", "");
    int w = 0, h = 0;
    if (config.contains("resolution")) {
        w = config.at("resolution").at((size_t)0).get_int();
        h = config.at("resolution").at((size_t)1).get_int();
    }
    int frame = config.value("frame", 0);
    std::shared_ptr<Renderer> r;
    if (type == "image") r = std::make_shared<ImageRenderer>(path, output);
    else if (type == "interactive") r = std::make_shared<InteractiveRenderer>(path, output, frame > 0 ? frame : 16);
    else if (type == "video") r = std::make_shared<VideoRenderer>(path, output, frame);
    else throw RenderError("Unknown renderer type: " + type);
    r->resize(w, h);
    return r;
}

int run_cli(int argc, char** argv) {
    std::string config, render, passes_p, scene_p;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&](std::string& dst) {
            if (i + 1 >= argc) throw RenderError("missing value for " + a);
            dst = argv[++i];
        };
        if (a == "-c" || a == "--config") next(config);
        else if (a == "-r" || a == "--render") next(render);
        else if (a == "-p" || a == "--passes") next(passes_p);
        else if (a == "-s" || a == "--scene") next(scene_p);
        else if (a == "-h" || a == "--help") {
            std::printf("usage: srtracer -c config.json [-r render.json] [-p passes.json] [-s scene.json]\n");
            return 0;
        } else throw RenderError("unknown option " + a);
    }
    if (config.empty()) throw RenderError("no config file given (-c)");
    auto load = [](const std::string& p) {
        try {
            return Json::load(p);
        } catch (const json_error& e) {
            throw RenderError(e.what());
        }
    };
    Json cfg = load(config);
    if (!render.empty()) cfg["render"] = load(render);
    if (!passes_p.empty()) cfg["passes"] = load(passes_p);
    if (!scene_p.empty()) cfg["scene"] = load(scene_p);
    auto renderer = create_renderer(config, cfg.at("render"));
    renderer->load_scene(cfg.at("scene"));
    renderer->load_passes(cfg.at("passes"));
    renderer->run();
    return 0;
}

}  // namespace srt

// C entry points of the pass layer (tests and embedding hosts)
extern "C" {

// Effective parameters of a pass created by name from JSON params (no device
// needed).  Returns the length written, or -1 if the pass name is unknown /
// the JSON is invalid (message in out).
int srt_host_pass_params(const char* name, const char* params_json, char* out, int out_len) {
    std::string res;
    int rc = 0;
    try {
        srt::Json p = (params_json && *params_json) ? srt::Json::parse(params_json) : srt::Json();
        auto pass = srt::RenderPassFactory::create_pass(name, p);
        res = pass->params().dump();
    } catch (const std::exception& e) {
        res = e.what();
        rc = -1;
    }
    if (out && out_len > 0) {
        std::snprintf(out, (size_t)out_len, "%s", res.c_str());
    }
    return rc < 0 ? rc : (int)res.size();
}

// Write width*height RGBA floats (film layout) to path (.exr, .hdr, .png, .bmp,
// .jpg, .pfm or .ppm by extension).  Returns 0, or -1 with the message on stderr.
int srt_host_save_image(const char* path, int width, int height, const float* rgba) {
    try {
        srt::save_image(path, width, height, rgba);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return -1;
    }
}

// 3x4 row-major transform at time t of one node animation given as the JSON
// object of the scene config's "animation" list.  0, or -1 (message in err).
int srt_host_node_transform(const char* animation_json, float t, float* out12, char* err, int err_len) {
    try {
        srt::parse_node_animation(srt::Json::parse(animation_json)).transform(t, out12);
        return 0;
    } catch (const std::exception& e) {
        if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", e.what());
        return -1;
    }
}

// Host-side scene loading as the renderer does it (OBJ + MTL with the
// reference's assimp material mapping, or .srtscene), without a device: the
// scene's sizes in counts[0..3] = vertices, triangles, meshes, materials, and,
// for every non-null array, its contents (positions 3 per vertex, indices 3
// per triangle, mesh_first meshes + 1, mesh_material per mesh, materials, mesh
// names NUL-separated within names_len bytes).  0, or -1 (message in err).
int srt_host_load_model(const char* path, int32_t counts[4], float* positions, uint32_t* indices, int32_t* mesh_first,
                        int32_t* mesh_material, srt_material* materials, char* names, int names_len, char* err,
                        int err_len) {
    try {
        srt::Scene s;
        std::string p = path ? path : "";
        std::string ext = p.substr(p.find_last_of('.') + 1);
        if (ext == "obj") srt::load_obj(p, s);
        else if (ext == "srtscene") srt::load_srtscene(p, s);
        else throw srt::RenderError("unsupported model format '." + ext + "' (obj or srtscene)");
        counts[0] = (int32_t)(s.positions.size() / 3);
        counts[1] = (int32_t)(s.indices.size() / 3);
        counts[2] = (int32_t)s.mesh_material.size();
        counts[3] = (int32_t)s.materials.size();
        if (positions) std::memcpy(positions, s.positions.data(), s.positions.size() * sizeof(float));
        if (indices) std::memcpy(indices, s.indices.data(), s.indices.size() * sizeof(uint32_t));
        if (mesh_first) std::memcpy(mesh_first, s.mesh_first.data(), s.mesh_first.size() * sizeof(int32_t));
        if (mesh_material) std::memcpy(mesh_material, s.mesh_material.data(), s.mesh_material.size() * sizeof(int32_t));
        if (materials) std::memcpy(materials, s.materials.data(), s.materials.size() * sizeof(srt_material));
        if (names && names_len > 0) {
            int o = 0;
            for (const auto& n : s.mesh_names) {
                if (o + (int)n.size() + 1 > names_len) break;
                std::memcpy(names + o, n.c_str(), n.size() + 1);
                o += (int)n.size() + 1;
            }
        }
        return 0;
    } catch (const std::exception& e) {
        if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", e.what());
        return -1;
    }
}

int srt_host_run(int argc, char** argv) {
    try {
        return srt::run_cli(argc, argv);
    } catch (const std::exception&) {
        return 255;
    }
}
}
