#include "scene_io.h"

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
}

void Scene::update(float t) {
    if (!has_animation() || !enable_animation || !handle) return;
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

// float -> IEEE half, round to nearest even (overflow to infinity, NaN kept)
static uint16_t to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);               // rounds to >= 65520: infinity
    if (ax < 0x38800000u) {                                                   // below the smallest normal half
        if (ax < 0x33000000u) return (uint16_t)sign;                          // below half the smallest subnormal
        const uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u;
        const uint32_t shift = 126u - e;                                      // 14..24
        uint32_t h = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((ax - 0x38000000u) >> 13);
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

// OpenEXR 2.0 scanline file, no compression, HALF channels (the reference
// writes HALF, texture.cu:96-117); channel names in the file's required
// alphabetical order A, B, G, R; rows top to bottom.
static void write_exr(std::ofstream& f, int W, int H, const std::function<const float*(int, int)>& px) {
    std::string hdr;
    auto put32 = [&](std::string& s, int32_t v) { s.append((const char*)&v, 4); };
    auto attr = [&](const char* name, const char* type, const std::string& value) {
        hdr += name; hdr += '\0';
        hdr += type; hdr += '\0';
        put32(hdr, (int32_t)value.size());
        hdr += value;
    };
    std::string ch;
    for (const char* name : {"A", "B", "G", "R"}) {
        ch += name; ch += '\0';
        put32(ch, 1);                        // HALF
        ch += std::string(4, '\0');           // pLinear + reserved
        put32(ch, 1); put32(ch, 1);          // x / y sampling
    }
    ch += '\0';
    std::string box;
    put32(box, 0); put32(box, 0); put32(box, W - 1); put32(box, H - 1);
    const float one = 1.0f, zero = 0.0f;
    attr("channels", "chlist", ch);
    attr("compression", "compression", std::string(1, '\0'));
    attr("dataWindow", "box2i", box);
    attr("displayWindow", "box2i", box);
    attr("lineOrder", "lineOrder", std::string(1, '\0'));
    attr("pixelAspectRatio", "float", std::string((const char*)&one, 4));
    std::string center((const char*)&zero, 4);
    center.append((const char*)&zero, 4);
    attr("screenWindowCenter", "v2f", center);
    attr("screenWindowWidth", "float", std::string((const char*)&one, 4));
    hdr += '\0';
    const uint32_t magic = 20000630u, version = 2u;
    f.write((const char*)&magic, 4);
    f.write((const char*)&version, 4);
    f.write(hdr.data(), (std::streamsize)hdr.size());
    const uint64_t line_bytes = 8ull * W;    // 4 channels x 2 bytes
    const uint64_t start = 8 + hdr.size() + 8ull * H;
    for (int y = 0; y < H; y++) {
        const uint64_t off = start + (uint64_t)y * (8 + line_bytes);
        f.write((const char*)&off, 8);
    }
    std::vector<uint16_t> line(4 * (size_t)W);
    static const int chan[4] = {3, 2, 1, 0};    // A, B, G, R from RGBA
    for (int y = 0; y < H; y++) {
        for (int k = 0; k < 4; k++)
            for (int x = 0; x < W; x++) line[(size_t)k * W + x] = to_half(px(y, x)[chan[k]]);
        const int32_t yy = y, size = (int32_t)line_bytes;
        f.write((const char*)&yy, 4);
        f.write((const char*)&size, 4);
        f.write((const char*)line.data(), (std::streamsize)line_bytes);
    }
}

void save_image(const std::string& path, int W, int H, const float* rgba) {
    std::string ext = path.substr(path.find_last_of('.') + 1);
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
    } else {
        throw RenderError("unsupported image format '." + ext + "' (use .exr, .pfm or .ppm)");
    }
}

}  // namespace srt
