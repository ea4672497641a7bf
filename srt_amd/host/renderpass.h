// Render-pass interface, factory and film: the drop-in boundary of the
// reference renderer (src/passes/renderpass.h:10-80, src/passes/film.h:7-20).
//
// Same names and call sequence as the reference: the renderer creates a pass
// by name from its JSON params, then calls set_enable / set_online / resize /
// set_scene / init(), and per frame update() + render(film).  The passes here
// call the HIP library through its C ABI (include/srt_hip.h).
#pragma once
#include <functional>
#include <iostream>
#include <memory>
#include <string>
#include <unordered_map>

#include "json.h"
#include "scene_io.h"
#include "srt_hip.h"

namespace srt {

// film.h:7-20: RGBA float film in device memory, row-major, idx = iy * W + ix
class Film {
public:
    Film(srt_ctx* ctx, int w, int h);
    ~Film();
    Film(const Film&) = delete;
    Film& operator=(const Film&) = delete;
    int get_width() const { return width; }
    int get_height() const { return height; }
    float* get_pixels() { return pixels; }        // device pointer, 4 floats per pixel
    void download(float* host_rgba) const;
    void upload(const float* host_rgba);

private:
    srt_ctx* ctx;
    int width, height;
    float* pixels = nullptr;
};

class RenderPass {
protected:
    bool enable = true;
    bool online = false;
    int width = 0, height = 0;
    std::shared_ptr<Scene> scene = nullptr;

public:
    RenderPass() {}
    virtual ~RenderPass() {}

    void set_enable(bool _enable) { enable = _enable; }
    void set_online(bool _online) { online = _online; }
    void resize(int _w, int _h) { width = _w; height = _h; }
    void set_scene(std::shared_ptr<Scene> _scene) { scene = _scene; }
    bool is_enabled() const { return enable; }

    virtual void init() {}
    virtual void render(std::shared_ptr<Film> film) = 0;
    virtual void update() {}
    virtual void render_ui() {}
    // effective parameters as JSON (for logs and tests)
    virtual Json params() const { return Json::object(); }
};

class RenderPassFactory {
private:
    using map_type = std::unordered_map<std::string, std::function<std::shared_ptr<RenderPass>(const Json&)>>;

public:
    static map_type& get_map() {
        static map_type map;
        return map;
    }

    template <class T>
    struct Register {
        Register(const std::string& name) {
            auto& map = get_map();
            map.emplace(name, [](const Json& params) {
                auto p = std::make_shared<T>();
                from_json(params, *p);
                return p;
            });
        }
    };

    static std::shared_ptr<RenderPass> create_pass(const std::string& name, const Json& params) {
        auto& map = get_map();
        auto it = map.find(name);
        if (it == map.end()) throw RenderError("Render pass " + name + " not found!");
        return it->second(params);
    }

    static void print_registered_pass() {
        auto& map = get_map();
        std::cout << "Registered Render Pass:" << std::endl;
        for (auto& it : map) std::cout << "    " << it.first << std::endl;
    }
};

#define REGISTER_RENDER_PASS(T) static RenderPassFactory::Register<T> reg
#define REGISTER_RENDER_PASS_CPP(T) RenderPassFactory::Register<T> T::reg(#T)

// throw RenderError with the library's message when a C-ABI call fails
void check_srt(srt_ctx* ctx, int code, const char* what);

}  // namespace srt
