// Host scene model and file I/O for the C++ render passes.
//
// Mirrors the parts of the reference's Scene (src/scene/scene.h:65-130) that
// the path-tracing passes read: meshes, materials, camera, background,
// skinning and keyframed node animation, and the upload to the device (Scene::build_gscene,
// src/scene/scene.cu:84-288), which here is srt_scene_upload +
// srt_scene_build_bvh: one BVH per scene shared by every pass.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "json.h"
#include "srt_hip.h"

namespace srt {

// Errors of the pass layer.  The reference prints "ERROR::..." and calls
// exit(-1); here the message is printed the same way and an exception carries
// it to the caller (srtracer's main turns it into exit status 255).
class RenderError : public std::runtime_error {
public:
    explicit RenderError(const std::string& m);
};

// camera.h:93-170 (perspective pinhole; vertical fov in degrees)
struct Camera {
    std::string type = "perspective";
    float position[3] = {0, 0, 0}, target[3] = {0, 0, -1}, up[3] = {0, 1, 0};
    float fov = 60.0f;
    bool moved = true;               // camera.h:104 (set by reset)
    void set_moved(bool m) { moved = m; }
};

// A keyframed rigid animation of the meshes of one name (animation.h:24-53,
// animation.cpp, the "animation" list of the scene config, renderer.cu:71-121):
// linear interpolation of translation and scale, slerp of rotation,
// repeat extrapolation over the last key's time, T * R * S.
struct NodeAnimation {
    std::string name;
    float duration = 0.0f;
    std::vector<float> t_keys, r_keys, s_keys;             // time stamps of each channel
    std::vector<float> translations, rotations, scales;    // 3, 4 (x, y, z, w), 3 floats per key
    void transform(float t, float out[12]) const;          // 3x4 row-major
};
NodeAnimation parse_node_animation(const Json& c);

class Scene {
public:
    std::shared_ptr<Camera> camera;
    float background[3] = {0, 0, 0};
    bool enable_animation = false;

    // world-space geometry, one identity instance per mesh
    std::vector<float> positions, normals, tangents;
    std::vector<uint32_t> indices;
    std::vector<int32_t> mesh_first{0};
    std::vector<int32_t> mesh_material;
    std::vector<srt_material> materials;
    std::vector<std::string> mesh_names;

    // linear blend skinning animation: per-vertex bones and precomputed
    // per-frame bone matrices (3x4 row-major)
    int32_t num_bones = 0, num_frames = 0;
    std::vector<int32_t> bone_ids;
    std::vector<float> bone_weights;
    std::vector<float> bone_frames;
    // rigid node animations: bone 0 is the identity, bone i + 1 moves the meshes named node_anims[i].name
    std::vector<NodeAnimation> node_anims;

    // device residency
    srt_ctx* ctx = nullptr;
    srt_scene handle = 0;
    int current_frame = -1;

    ~Scene();
    bool has_animation() const { return num_frames > 0 || !node_anims.empty(); }
    bool is_static() const { return (!has_animation() || !enable_animation) && !camera->moved; }   // scene.h:93
    int num_triangles() const { return (int)(indices.size() / 3); }
    void add_mesh(const std::string& name, const std::vector<float>& pos, const std::vector<uint32_t>& tris, int32_t material);
    // upload to the device and build the BVH (scene.cu:84-90)
    void build_gscene(srt_ctx* ctx);
    // animation time t in seconds, 30 frames per second (renderer.cu:240);
    // skins the mesh and refits the BVH when the frame changes (scene.cu:309-395)
    void update(float t);
};

// OBJ + MTL reader with the reference's assimp material mapping
// (src/importer/assimp.cu:17-79): Disney BxDF, base colour = Kd (or Ks if Kd
// is black), Ke -> intensity |Ke| and emission colour Ke/|Ke|, Ni -> ior,
// roughness 0.5, specular 0.5, clearcoatGloss 0.001; polygons fan-triangulated;
// one mesh per group/material.
void load_obj(const std::string& path, Scene& s);
// Binary scene written by srt_amd.scenes.save_srtscene (procedural workloads).
void load_srtscene(const std::string& path, Scene& s);
// Image::save_to_file (src/scene/texture.cu:219-230): flip x and y, then
// write .pfm (float RGB) or .ppm (8-bit RGB, clamped).
void save_image(const std::string& path, int width, int height, const float* rgba);

}  // namespace srt
