#include "scene_io.h"

#include "image_out.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <sstream>

namespace srt {

RenderError::RenderError(const std::string& m) : std::runtime_error(m) { std::cout << "ERROR::" << m << std::endl; }

static void check(srt_ctx* ctx, int r, const char* what) {
    if (r < 0) throw RenderError(std::string(what) + ": " + srt_last_error(ctx));
}

Scene::~Scene() {
    if (ctx && handle) srt_scene_release(ctx, handle);
}

void Scene::add_mesh(const std::string& name, const std::vector<float>& pos, const std::vector<uint32_t>& tris, int32_t material) {
    uint32_t base = (uint32_t)(positions.size() / 3);
    positions.insert(positions.end(), pos.begin(), pos.end());
    for (uint32_t t : tris) indices.push_back(base + t);
    mesh_first.push_back((int32_t)(indices.size() / 3));
    mesh_material.push_back(material);
    mesh_names.push_back(name);
}

void Scene::build_gscene(srt_ctx* c) {
    ctx = c;
    srt_scene_desc d{};
    d.num_vertices = (int32_t)(positions.size() / 3);
    d.positions = positions.data();
    d.normals = normals.empty() ? nullptr : normals.data();
    d.tangents = tangents.empty() ? nullptr : tangents.data();
    d.num_triangles = (int32_t)(indices.size() / 3);
    d.indices = indices.data();
    d.num_meshes = (int32_t)mesh_material.size();
    d.mesh_first_triangle = mesh_first.data();
    d.mesh_material = mesh_material.data();
    d.num_materials = (int32_t)materials.size();
    d.materials = materials.data();
    for (int i = 0; i < 3; i++) d.env_color[i] = background[i];
    check(ctx, srt_scene_upload(ctx, &d, &handle), "scene upload");
    check(ctx, srt_scene_build_bvh(ctx, handle, nullptr), "BVH build");
    if (num_bones > 0) check(ctx, srt_scene_set_skin(ctx, handle, bone_ids.data(), bone_weights.data()), "skin weights");
    if (!node_anims.empty()) {
        if (num_bones > 0) throw RenderError("a skinned scene cannot also have node animations");
        // every vertex follows the animation of its mesh's name (rigid: one bone, weight 1)
        const size_t nv = positions.size() / 3;
        std::vector<int32_t> ids(4 * nv, -1), owner(nv, -1);
        std::vector<float> w(4 * nv, 0.0f);
        for (size_t v = 0; v < nv; v++) { ids[4 * v] = 0; w[4 * v] = 1.0f; }
        for (size_t a = 0; a < node_anims.size(); a++) {
            bool found = false;
            for (size_t m = 0; m + 1 < mesh_first.size(); m++) {
                if (mesh_names[m] != node_anims[a].name) continue;
                found = true;
                for (int32_t f = mesh_first[m]; f < mesh_first[m + 1]; f++)
                    for (int k = 0; k < 3; k++) {
                        const uint32_t v = indices[3 * (size_t)f + k];
                        if (owner[v] >= 0 && owner[v] != (int32_t)a)
                            throw RenderError("a vertex is shared by two animated meshes: " + node_anims[a].name);
                        owner[v] = (int32_t)a;
                        ids[4 * v] = (int32_t)(a + 1);
                    }
            }
            if (!found) throw RenderError("animation of unknown mesh " + node_anims[a].name);
        }
        check(ctx, srt_scene_set_skin(ctx, handle, ids.data(), w.data()), "node animation weights");
    }
}

// ---------------------------------------------------------------------------
// node animation (animation.cpp, quaternion.h:104-117, transform.h:148-167)
// ---------------------------------------------------------------------------
namespace {
struct Quat { float x, y, z, w; };
Quat qadd(Quat a, Quat b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
Quat qscale(Quat a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
float qdot(Quat a, Quat b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
Quat qnormalize(Quat a) { return qscale(a, 1.0f / std::sqrt(qdot(a, a))); }
Quat slerp(Quat q1, Quat q2, float t) {
    const float cos_theta = std::min(std::max(qdot(q1, q2), -1.0f), 1.0f);
    if (std::fabs(cos_theta) > 1.0f - 1e-6f) return qadd(qscale(q1, 1.0f - t), qscale(q2, t));
    if (cos_theta < 0.0f) return slerp(q1, qscale(q2, -1.0f), t);
    const float theta = std::acos(cos_theta), theta_p = theta * t;
    const Quat qperp = qnormalize(qadd(q2, qscale(q1, -cos_theta)));
    return qadd(qscale(q1, std::cos(theta_p)), qscale(qperp, std::sin(theta_p)));
}
// key interval of time t over stamps ts: (index, factor); the scale channel
// compares with '<' where the others use '<=' (animation.cpp:37-81)
template <int N>
void interp(const std::vector<float>& ts, const std::vector<float>& v, float t, bool strict, const float* dflt, float* out) {
    const size_t n = ts.size();
    if (n == 0) { for (int k = 0; k < N; k++) out[k] = dflt[k]; return; }
    if (n == 1) { for (int k = 0; k < N; k++) out[k] = v[k]; return; }
    size_t idx = 0;
    for (; idx < n - 1; idx++)
        if (strict ? (t < ts[idx + 1]) : (t <= ts[idx + 1])) break;
    if (idx == n - 1) { for (int k = 0; k < N; k++) out[k] = v[N * idx + k]; return; }
    const float f = (t - ts[idx]) / (ts[idx + 1] - ts[idx]);
    if (N == 4) {
        const Quat a{v[4 * idx], v[4 * idx + 1], v[4 * idx + 2], v[4 * idx + 3]};
        const Quat b{v[4 * idx + 4], v[4 * idx + 5], v[4 * idx + 6], v[4 * idx + 7]};
        const Quat q = slerp(a, b, f);
        out[0] = q.x; out[1] = q.y; out[2] = q.z; out[3] = q.w;
        return;
    }
    for (int k = 0; k < N; k++) out[k] = v[N * idx + k] * (1.0f - f) + v[N * (idx + 1) + k] * f;
}
void mat4_mul(const float a[16], const float b[16], float o[16]) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float s = 0.0f;
            for (int k = 0; k < 4; k++) s += a[4 * i + k] * b[4 * k + j];
            o[4 * i + j] = s;
        }
}
}  // namespace

void NodeAnimation::transform(float t, float out[12]) const {
    t = std::fmod(t, duration);                                    // Extrapolation::Repeat
    static const float zero3[3] = {0, 0, 0}, one3[3] = {1, 1, 1}, idq[4] = {0, 0, 0, 1};
    float T[3], R[4], S[3];
    interp<3>(t_keys, translations, t, false, zero3, T);
    interp<4>(r_keys, rotations, t, false, idq, R);
    interp<3>(s_keys, scales, t, true, one3, S);
    const float x = R[0], y = R[1], z = R[2], w = R[3];
    const float x2 = x * x, y2 = y * y, z2 = z * z, xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y, wz = w * z;
    const float Mt[16] = {1, 0, 0, T[0], 0, 1, 0, T[1], 0, 0, 1, T[2], 0, 0, 0, 1};
    const float Mr[16] = {1.0f - 2.0f * (y2 + z2), 2.0f * (xy - wz), 2.0f * (xz + wy), 0.0f,
                          2.0f * (xy + wz), 1.0f - 2.0f * (x2 + z2), 2.0f * (yz - wx), 0.0f,
                          2.0f * (xz - wy), 2.0f * (yz + wx), 1.0f - 2.0f * (x2 + y2), 0.0f,
                          0.0f, 0.0f, 0.0f, 1.0f};
    const float Ms[16] = {S[0], 0, 0, 0, 0, S[1], 0, 0, 0, 0, S[2], 0, 0, 0, 0, 1};
    float TR[16], M[16];
    mat4_mul(Mt, Mr, TR);
    mat4_mul(TR, Ms, M);
    for (int i = 0; i < 12; i++) out[i] = M[i];
}

NodeAnimation parse_node_animation(const Json& c) {
    NodeAnimation a;
    a.name = c.at("name").get_string();
    const Json& keys = c.at("keys");
    if (!keys.is_array() || keys.size() == 0) throw RenderError("animation " + a.name + ": 'keys' must be a non-empty list");
    std::vector<float> ts;
    for (size_t i = 0; i < keys.size(); i++) ts.push_back(keys.at(i).get_float());
    auto channel = [&](const char* key, int n, std::vector<float>& stamps, std::vector<float>& vals) {
        if (!c.contains(key)) return;
        const Json& v = c.at(key);
        if (v.size() != ts.size()) throw RenderError("animation " + a.name + ": '" + key + "' needs one entry per key");
        for (size_t i = 0; i < v.size(); i++) {
            if ((int)v.at(i).size() != n) throw RenderError("animation " + a.name + ": bad '" + key + "' entry");
            for (int k = 0; k < n; k++) vals.push_back(v.at(i).at((size_t)k).get_float());
        }
        stamps = ts;
    };
    channel("translations", 3, a.t_keys, a.translations);
    channel("rotations", 4, a.r_keys, a.rotations);
    channel("scales", 3, a.s_keys, a.scales);
    a.duration = ts.back();
    return a;
}

void Scene::update(float t) {
    if (!has_animation() || !enable_animation || !handle) return;
    if (!node_anims.empty()) {
        // bone 0: the identity for meshes without an animation
        std::vector<float> bones(12 * (node_anims.size() + 1), 0.0f);
        bones[0] = bones[5] = bones[10] = 1.0f;
        for (size_t i = 0; i < node_anims.size(); i++) node_anims[i].transform(t, &bones[12 * (i + 1)]);
        check(ctx, srt_scene_skin_refit(ctx, handle, bones.data(), (int32_t)(node_anims.size() + 1)), "node animation");
        return;
    }
    int f = (int)std::floor(t * 30.0f + 1e-4f) % num_frames;
    if (f < 0) f += num_frames;
    if (f == current_frame) return;
    check(ctx, srt_scene_skin_refit(ctx, handle, bone_frames.data() + (size_t)f * num_bones * 12, num_bones), "skin + refit");
    current_frame = f;
}

// ---------------------------------------------------------------------------
// OBJ / MTL
// ---------------------------------------------------------------------------
namespace {
struct MtlRec {
    float Kd[3] = {0, 0, 0}, Ks[3] = {0, 0, 0}, Ke[3] = {0, 0, 0};
    float Ni = 1.5f;
};

std::string dir_of(const std::string& p) {
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

std::map<std::string, MtlRec> read_mtl(const std::string& path) {
    std::map<std::string, MtlRec> out;
    std::ifstream f(path);
    if (!f.is_open()) throw RenderError("Failed to open file: " + path);
    std::string line, cur;
    while (std::getline(f, line)) {
        size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag)) continue;
        if (tag == "newmtl") { ss >> cur; out[cur] = MtlRec(); continue; }
        if (cur.empty()) continue;
        MtlRec& m = out[cur];
        if (tag == "Kd") ss >> m.Kd[0] >> m.Kd[1] >> m.Kd[2];
        else if (tag == "Ks") ss >> m.Ks[0] >> m.Ks[1] >> m.Ks[2];
        else if (tag == "Ke") ss >> m.Ke[0] >> m.Ke[1] >> m.Ke[2];
        else if (tag == "Ni") ss >> m.Ni;
    }
    return out;
}

// assimp.cu:17-79
srt_material convert(const MtlRec& r) {
    srt_material m{};
    m.type = SRT_BXDF_DISNEY;
    float e[3] = {r.Ke[0] * 1.0f, r.Ke[1] * 1.0f, r.Ke[2] * 1.0f};      // emissiveStrength 1
    float inten = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    for (int i = 0; i < 3; i++) m.emission_color[i] = inten > 0 ? e[i] * (1.0f / inten) : e[i];
    m.intensity = inten;
    bool kd = (r.Kd[0] * r.Kd[0] + r.Kd[1] * r.Kd[1] + r.Kd[2] * r.Kd[2]) > 0;
    for (int i = 0; i < 3; i++) m.base_color[i] = kd ? r.Kd[i] : r.Ks[i];
    m.ior = r.Ni;
    m.metallic = 0.0f;
    m.subsurface = 0.0f;
    m.roughness = std::max(0.5f, 0.001f);
    m.specular = 0.5f;
    m.specular_tint = 0.0f;
    m.anisotropic = 0.0f;
    m.sheen = 0.0f;
    m.sheen_tint = 0.0f;
    m.clearcoat = 0.0f;
    m.clearcoat_gloss = std::max(0.0f, 0.001f);
    m.spec_trans = 0.0f;
    return m;
}
}  // namespace

void load_obj(const std::string& path, Scene& s) {
    std::ifstream f(path);
    if (!f.is_open()) throw RenderError("Failed to open file: " + path);
    std::vector<float> v;
    std::map<std::string, MtlRec> mtl;
    std::map<std::string, int> mat_index;
    std::string line, group = "default", usemtl;
    // faces grouped by (group, material) in first-appearance order
    std::vector<std::pair<std::string, std::string>> keys;
    std::map<std::pair<std::string, std::string>, std::vector<int>> faces;
    auto resolve = [&](const std::string& tok) {
        int k = std::atoi(tok.substr(0, tok.find('/')).c_str());
        int nv = (int)(v.size() / 3);
        int idx = k > 0 ? k - 1 : nv + k;
        if (idx < 0 || idx >= nv) throw RenderError("OBJ face index out of range in " + path);
        return idx;
    };
    while (std::getline(f, line)) {
        size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag)) continue;
        if (tag == "v") {
            float x, y, z;
            ss >> x >> y >> z;
            v.push_back(x); v.push_back(y); v.push_back(z);
        } else if (tag == "mtllib") {
            std::string name;
            ss >> name;
            auto m = read_mtl(dir_of(path) + "/" + name);
            mtl.insert(m.begin(), m.end());
        } else if (tag == "usemtl") {
            ss >> usemtl;
        } else if (tag == "g" || tag == "o") {
            ss >> group;
        } else if (tag == "f") {
            std::vector<int> poly;
            std::string tok;
            while (ss >> tok) poly.push_back(resolve(tok));
            if (poly.size() < 3) continue;
            auto key = std::make_pair(group, usemtl);
            if (!faces.count(key)) keys.push_back(key);
            auto& fl = faces[key];
            for (size_t k = 1; k + 1 < poly.size(); k++) {   // fan triangulation
                fl.push_back(poly[0]); fl.push_back(poly[k]); fl.push_back(poly[k + 1]);
            }
        }
    }
    if (keys.empty()) throw RenderError("OBJ has no faces: " + path);
    for (auto& key : keys) {
        const std::string& mname = key.second;
        int mi;
        auto it = mat_index.find(mname);
        if (it != mat_index.end()) mi = it->second;
        else {
            MtlRec r;
            auto m = mtl.find(mname);
            if (m != mtl.end()) r = m->second;
            else { r.Kd[0] = r.Kd[1] = r.Kd[2] = 0.6f; }
            mi = (int)s.materials.size();
            s.materials.push_back(convert(r));
            mat_index[mname] = mi;
        }
        // per-mesh vertex arrays (unshared, as assimp splits meshes)
        const auto& fl = faces[key];
        std::vector<float> pos;
        std::vector<uint32_t> tri;
        std::map<int, uint32_t> remap;
        for (int g : fl) {
            auto r = remap.find(g);
            if (r == remap.end()) {
                uint32_t nid = (uint32_t)(pos.size() / 3);
                remap[g] = nid;
                pos.push_back(v[3 * g]); pos.push_back(v[3 * g + 1]); pos.push_back(v[3 * g + 2]);
                tri.push_back(nid);
            } else tri.push_back(r->second);
        }
        s.add_mesh(key.first, pos, tri, mi);
    }
}

// ---------------------------------------------------------------------------
// .srtscene: "SRTSCN1\0", int32 nv nf nmesh nmat has_nrm has_tan, float env[3],
// pos, [nrm], [tan], idx, mesh_first, mesh_mat, materials (80 B each),
// int32 nbones nframes, [bone_ids 4nv, bone_w 4nv, frames nframes*nbones*12]
// ---------------------------------------------------------------------------
void load_srtscene(const std::string& path, Scene& s) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw RenderError("Failed to open file: " + path);
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "SRTSCN1", 8) != 0) throw RenderError("not an srtscene file: " + path);
    int32_t h[6];
    f.read((char*)h, sizeof(h));
    float env[3];
    f.read((char*)env, sizeof(env));
    const int32_t nv = h[0], nf = h[1], nm = h[2], nmat = h[3];
    if (!f || nv <= 0 || nf <= 0 || nm <= 0 || nmat <= 0) throw RenderError("corrupt srtscene header: " + path);
    auto rd = [&](void* p, size_t bytes) {
        f.read((char*)p, (std::streamsize)bytes);
        if (!f) throw RenderError("truncated srtscene: " + path);
    };
    s.positions.resize(3 * (size_t)nv);
    rd(s.positions.data(), 12 * (size_t)nv);
    if (h[4]) { s.normals.resize(3 * (size_t)nv); rd(s.normals.data(), 12 * (size_t)nv); }
    if (h[5]) { s.tangents.resize(3 * (size_t)nv); rd(s.tangents.data(), 12 * (size_t)nv); }
    s.indices.resize(3 * (size_t)nf);
    rd(s.indices.data(), 12 * (size_t)nf);
    s.mesh_first.resize((size_t)nm + 1);
    rd(s.mesh_first.data(), 4 * ((size_t)nm + 1));
    s.mesh_material.resize((size_t)nm);
    rd(s.mesh_material.data(), 4 * (size_t)nm);
    s.materials.resize((size_t)nmat);
    rd(s.materials.data(), sizeof(srt_material) * (size_t)nmat);
    for (int i = 0; i < 3; i++) s.background[i] = env[i];
    s.mesh_names.assign((size_t)nm, "mesh");
    int32_t sk[2] = {0, 0};
    f.read((char*)sk, sizeof(sk));
    if (f && sk[0] > 0 && sk[1] > 0) {
        s.num_bones = sk[0];
        s.num_frames = sk[1];
        s.bone_ids.resize(4 * (size_t)nv);
        s.bone_weights.resize(4 * (size_t)nv);
        s.bone_frames.resize((size_t)sk[1] * sk[0] * 12);
        rd(s.bone_ids.data(), 16 * (size_t)nv);
        rd(s.bone_weights.data(), 16 * (size_t)nv);
        rd(s.bone_frames.data(), 4 * s.bone_frames.size());
        s.enable_animation = true;
    }
}

// ---------------------------------------------------------------------------
// images (film row 0 is the bottom row, column 0 the right edge: flip both,
// as Image::save_to_file does, texture.cu:219-230)
// ---------------------------------------------------------------------------

void save_image(const std::string& path, int W, int H, const float* rgba) {
    std::string ext = path.substr(path.find_last_of('.') + 1);
    static const char* known[] = {"pfm", "ppm", "exr", "hdr", "png", "bmp", "jpg"};
    if (std::find(std::begin(known), std::end(known), ext) == std::end(known))
        throw RenderError("Unsupported image format: " + ext + " (use .exr, .hdr, .png, .bmp, .jpg, .pfm or .ppm)");
    if (ext == "jpg" && (W > 65535 || H > 65535)) throw RenderError("image too large for JPEG: " + path);
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw RenderError("Failed to open file for writing: " + path);
    auto px = [&](int row_from_top, int col) {
        int r = H - 1 - row_from_top, c = W - 1 - col;     // flip y and x
        return rgba + 4 * ((size_t)r * W + c);
    };
    if (ext == "pfm") {
        f << "PF\n" << W << " " << H << "\n-1.0\n";
        for (int row = H - 1; row >= 0; row--)               // PFM stores rows bottom-up
            for (int col = 0; col < W; col++) f.write((const char*)px(row, col), 12);
    } else if (ext == "ppm") {
        f << "P6\n" << W << " " << H << "\n255\n";
        for (int row = 0; row < H; row++)
            for (int col = 0; col < W; col++) {
                const float* p = px(row, col);
                unsigned char b[3];
                for (int k = 0; k < 3; k++) {
                    float x = std::min(std::max(p[k], 0.0f), 1.0f);
                    b[k] = (unsigned char)(x * 255.0f + 0.5f);
                }
                f.write((const char*)b, 3);
            }
    } else if (ext == "exr") {
        write_exr(f, W, H, px);
    } else if (ext == "hdr") {
        write_hdr(f, W, H, px);
    } else if (ext == "png") {
        write_png(f, W, H, px);
    } else if (ext == "bmp") {
        write_bmp(f, W, H, px);
    } else {
        write_jpg(f, W, H, px);
    }
}

}  // namespace srt
