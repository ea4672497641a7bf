// Renderer front end: config -> scene -> passes -> run
// (src/renderer/renderer.{h,cu}, src/main/main.cu).
//
//   image        one pass sequence, then the film is saved (renderer.cu:176-191)
//   interactive  the reference opens a window; here it is headless: `frame`
//                frames (default 16) of update + render with the camera held
//                still after the first, so Accumulate averages them, then the
//                film is saved to `output` if given
//   video        `frame` frames at t += 1/30 with update + render, each saved as
//                <output stem>_<i>.pfm (the reference pipes PNGs to ffmpeg)
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "json.h"
#include "renderpass.h"
#include "scene_io.h"

namespace srt {

class Renderer {
protected:
    std::string config_path;
    int width = 0, height = 0;
    bool online = false;
    srt_ctx* ctx = nullptr;
    std::shared_ptr<Scene> scene;
    std::shared_ptr<Film> film;
    std::vector<std::shared_ptr<RenderPass>> passes;

public:
    explicit Renderer(const std::string& path);
    virtual ~Renderer();
    void resize(int w, int h) { width = w; height = h; }
    virtual void load_scene(const Json& config);
    void load_passes(const Json& config);
    virtual void run() = 0;
    std::shared_ptr<Film> get_film() { return film; }
    std::shared_ptr<Scene> get_scene() { return scene; }
    const std::vector<std::shared_ptr<RenderPass>>& get_passes() const { return passes; }
    int get_width() const { return width; }
    int get_height() const { return height; }
    void print_profile() const;
};

class ImageRenderer : public Renderer {
    std::string filename;

public:
    ImageRenderer(const std::string& path, const std::string& out) : Renderer(path), filename(out) {}
    void run() override;
};

class InteractiveRenderer : public Renderer {
    std::string filename;
    int frames;

public:
    InteractiveRenderer(const std::string& path, const std::string& out, int f) : Renderer(path), filename(out), frames(f) {}
    void run() override;
};

class VideoRenderer : public Renderer {
    std::string filename;
    int frame;

public:
    VideoRenderer(const std::string& path, const std::string& out, int f) : Renderer(path), filename(out), frame(f) {}
    void run() override;
};

// main.cu:19-41
std::shared_ptr<Renderer> create_renderer(const std::string& path, const Json& config);
// main.cu:44-70: -c config [-r render] [-p passes] [-s scene]; returns exit status
int run_cli(int argc, char** argv);

}  // namespace srt
