// C ABI of the SRT HIP library: context, HBM budget, scene residency, render
// entry points and statistics.  See include/srt_hip.h for the contract and
// the reference interfaces each entry point replaces.
#include "srt_internal.h"
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <algorithm>

namespace srt {

int set_err(srt_ctx* c, int code, const char* fmt, ...) {
    if (!c) return code;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    std::lock_guard<std::mutex> lk(c->mu);
    c->err = buf;
    return code;
}

int hip_check(srt_ctx* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return SRT_OK;
    if (e == hipErrorOutOfMemory) return set_err(c, SRT_E_OOM, "%s: %s", what, hipGetErrorString(e));
    return set_err(c, SRT_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

bool release_build_scratch(srt_ctx* c) {
    std::vector<DevBuf> drop;
    {
        std::lock_guard<std::mutex> g(c->scratch_mu);
        drop.swap(c->build_scratch);
    }
    for (auto& b : drop) dev_free(c, b);
    return !drop.empty();
}

int dev_alloc(srt_ctx* c, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (c->used + bytes > c->budget) release_build_scratch(c);
    if (c->used + bytes > c->budget)
        return set_err(c, SRT_E_OOM, "allocation of %zu bytes exceeds the HBM budget (%zu of %zu used)", bytes, c->used.load(), c->budget);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess && release_build_scratch(c)) {
        (void)hipGetLastError();          // the failed attempt's error is not this call's result
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) return hip_check(c, e, "hipMalloc");
    b.p = p;
    b.bytes = bytes;
    c->used += bytes;
    return SRT_OK;
}

void dev_free(srt_ctx* c, DevBuf& b) {
    if (b.p) {
        (void)hipFree(b.p);
        c->used -= b.bytes;
    }
    b.p = nullptr;
    b.bytes = 0;
}

int dev_realloc(srt_ctx* c, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.p && b.bytes == bytes) return SRT_OK;
    dev_free(c, b);
    return dev_alloc(c, b, bytes);
}

void prof_begin(srt_ctx* c, const char* name, hipEvent_t* ev) {
    (void)name;
    *ev = nullptr;
    if (!c->profile) return;
    hipEvent_t e;
    if (!c->event_pool.empty()) { e = c->event_pool.back(); c->event_pool.pop_back(); }
    else if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, c->stream);
    *ev = e;
}

void prof_end(srt_ctx* c, const char* name, hipEvent_t ev0, int64_t items) {
    auto& st = c->stats[name];
    st.launches += 1;
    st.items += items;
    if (!c->profile || !ev0) return;
    hipEvent_t e;
    if (!c->event_pool.empty()) { e = c->event_pool.back(); c->event_pool.pop_back(); }
    else if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, c->stream);
    c->pending.push_back({name, {ev0, e}});
}

void prof_collect(srt_ctx* c) {
    for (auto& p : c->pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.second.second) == hipSuccess && hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess)
            c->stats[p.first].ms += ms;
        c->event_pool.push_back(p.second.first);
        c->event_pool.push_back(p.second.second);
    }
    c->pending.clear();
}

DevScene dev_scene(const SceneRec* s) {
    DevScene S;
    S.tris = (const float4*)s->tris.p;
    S.tri_rec = (const float4*)s->tri_rec.p;
    S.rec_flags = (s->has_nrm ? REC_NORMALS : 0) | (s->has_tan ? REC_TANGENTS : 0);
    S.nodes = (const BvhNode*)s->nodes.p;
    S.root = s->root;
    S.nodes4 = (const Bvh4Node*)s->nodes4.p;
    S.root4 = s->root4;
    S.num_nodes4 = s->num_nodes4;
    S.num_tris = s->nf;
    S.idx = (const uint32_t*)s->idx.p;
    S.pos = (const float*)s->pos.p;
    S.nrm = s->has_nrm ? (const float*)s->nrm.p : nullptr;
    S.tan = s->has_tan ? (const float*)s->tan.p : nullptr;
    S.mesh_mat = (const int32_t*)s->mesh_mat.p;
    S.mesh_first = (const int32_t*)s->mesh_first.p;
    S.area_cdf = (const float*)s->area_cdf.p;
    S.mats = (const MatRec*)s->mats.p;
    S.light_cdf = (const float*)s->light_cdf.p;
    S.light_mesh = (const int32_t*)s->light_mesh.p;
    S.num_lights = s->nlights;
    S.light_wsum = s->light_wsum;
    S.env[0] = s->env[0]; S.env[1] = s->env[1]; S.env[2] = s->env[2];
    S.num_inst = s->ninst;
    S.inst_xf = (const float4*)s->inst_xf.p;
    S.inst_inv = (const float4*)s->inst_inv.p;
    S.inst_root = (const int32_t*)s->inst_root.p;
    S.light_inst = (const int32_t*)s->light_inst.p;
    return S;
}

static void free_scene(srt_ctx* c, SceneRec* s) {
    DevBuf* all[] = {&s->pos, &s->nrm, &s->tan, &s->idx, &s->mesh_mat, &s->mesh_first, &s->area_cdf, &s->mats,
                     &s->light_cdf, &s->light_mesh, &s->tri_mesh, &s->orig_pos, &s->orig_nrm, &s->orig_tan,
                     &s->bone_ids, &s->bone_w, &s->bones, &s->nodes, &s->tris, &s->order, &s->node_parent,
                     &s->level_nodes, &s->node_valid, &s->nodes4, &s->tri_rec, &s->inst_xf, &s->inst_inv,
                     &s->inst_root, &s->light_inst};
    for (DevBuf* b : all) dev_free(c, *b);
    delete s;
}

static SceneRec* find_scene(srt_ctx* c, srt_scene h) {
    auto it = c->scenes.find(h);
    return it == c->scenes.end() ? nullptr : it->second;
}

static int upload(srt_ctx* c, DevBuf& b, const void* src, size_t bytes) {
    if (int r = dev_alloc(c, b, bytes)) return r;
    return hip_check(c, hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice), "upload");
}

}  // namespace srt

using namespace srt;

extern "C" {

int srt_create(int device, size_t hbm_budget_bytes, srt_ctx** out) {
    if (!out) return SRT_E_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return SRT_E_NO_DEVICE;
    if (device < 0 || device >= count) return SRT_E_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return SRT_E_HIP;
    srt_ctx* c = new srt_ctx();
    c->device = device;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    c->budget = hbm_budget_bytes ? hbm_budget_bytes : (size_t)(free_b * 0.9);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return SRT_E_HIP; }
    if (dev_alloc(c, c->counters, kCountersBytes)) { (void)hipStreamDestroy(c->stream); delete c; return SRT_E_OOM; }
    (void)hipMemset(c->counters.p, 0, kCountersBytes);
    if (int r = noalloc_init(c)) {
        dev_free(c, c->counters);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return r;
    }
    *out = c;
    return SRT_OK;
}

void srt_destroy(srt_ctx* c) {
    if (!c) return;
    if (c->serve_free) c->serve_free(c);
    if (c->mlp_free) c->mlp_free(c);
    (void)hipStreamSynchronize(c->stream);
    prof_collect(c);
    for (auto& kv : c->scenes) free_scene(c, kv.second);
    c->scenes.clear();
    for (int k = 0; k < 2; k++) for (auto& b : c->q[k]) dev_free(c, b);
    for (auto& b : c->hits) dev_free(c, b);
    for (auto& b : c->shq) dev_free(c, b);
    for (auto& b : c->Lbuf) dev_free(c, b);
    dev_free(c, c->counters); dev_free(c, c->film_acc); dev_free(c, c->film_out); dev_free(c, c->gstack);
    for (auto& kv : c->user_bufs) { DevBuf b; b.p = kv.first; b.bytes = kv.second; dev_free(c, b); }
    c->user_bufs.clear();
    for (auto& b : c->mlpbuf) dev_free(c, b);
    dev_free(c, c->scan_tmp); dev_free(c, c->train_in9); dev_free(c, c->train_t); dev_free(c, c->segbuf);
    release_build_scratch(c);
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* srt_last_error(const srt_ctx* c) { return c ? c->err.c_str() : "null context"; }

int srt_set_option(srt_ctx* c, const char* key, int64_t value) {
    SRT_IDLE(c);
    if (!c || !key) return SRT_E_INVALID;
    if (!strcmp(key, "path_pool")) {
        if (!path_pool_ok(value))
            return set_err(c, SRT_E_INVALID, "path_pool must be in [%lld, %lld]", (long long)kMinPathPool,
                           (long long)kMaxPathPool);
        c->pool_request = value;
        return SRT_OK;
    }
    if (!strcmp(key, "profile")) { c->profile = value != 0; return SRT_OK; }
    if (!strcmp(key, "count_visits")) { c->count_visits = value != 0; return SRT_OK; }
    if (!strcmp(key, "trace_stack_lds")) {
        if (value < 4 || value > 64) return set_err(c, SRT_E_INVALID, "trace_stack_lds must be in [4, 64]");
        c->trace_lds_depth = (int32_t)value;
        return SRT_OK;
    }
    if (!strcmp(key, "trace_refill")) {
        if (value < 1 || value > 64) return set_err(c, SRT_E_INVALID, "trace_refill must be in [1, 64]");
        c->trace_refill = (int32_t)value;
        return SRT_OK;
    }
    if (!strcmp(key, "shadow_refill")) {
        if (value < 1 || value > 64) return set_err(c, SRT_E_INVALID, "shadow_refill must be in [1, 64]");
        c->shadow_refill = (int32_t)value;
        return SRT_OK;
    }
    if (!strcmp(key, "trace_inner_exit")) {
        if (value < 0 || value > 63) return set_err(c, SRT_E_INVALID, "trace_inner_exit must be in [0, 63]");
        c->trace_inner_exit = (int32_t)value;
        return SRT_OK;
    }
    if (!strcmp(key, "serve_age")) {
        if (value < 0 || value > 1000) return set_err(c, SRT_E_INVALID, "serve_age must be in [0, 1000] (per mille)");
        c->serve_age_pm = (int32_t)value;
        return SRT_OK;
    }
    if (!strcmp(key, "trace_bvh_width")) {
        if (value != 2 && value != 4) return set_err(c, SRT_E_INVALID, "trace_bvh_width must be 2 or 4");
        c->trace_bvh_width = (int32_t)value;
        return SRT_OK;
    }
    if (!strcmp(key, "max_wide_nodes")) {
        if (value < 0 || value > kMaxNodes4)
            return set_err(c, SRT_E_INVALID, "max_wide_nodes must be in [0, %lld]", (long long)kMaxNodes4);
        c->max_wide_nodes = value;
        return SRT_OK;
    }
    if (!strcmp(key, "trace_blocks_per_cu")) {
        if (value < 0 || value > 64) return set_err(c, SRT_E_INVALID, "trace_blocks_per_cu must be in [0, 64]");
        c->trace_blocks_per_cu = (int32_t)value;
        return SRT_OK;
    }
    if (!strcmp(key, "trace_packet")) {
        if (value < 0 || value > 64) return set_err(c, SRT_E_INVALID, "trace_packet must be in [0, 64]");
        c->trace_packet = (int32_t)value;
        return SRT_OK;
    }
    if (!strcmp(key, "keep_build_scratch")) {
        if (value != 0 && value != 1) return set_err(c, SRT_E_INVALID, "keep_build_scratch must be 0 or 1");
        c->keep_build_scratch = value == 1;
        if (!c->keep_build_scratch) release_build_scratch(c);
        return SRT_OK;
    }
    if (!strcmp(key, "hbm_budget")) {
        if (value <= 0 || (size_t)value < c->used.load())
            return set_err(c, SRT_E_INVALID, "hbm_budget must be positive and cover the %zu bytes in use", c->used.load());
        c->budget = (size_t)value;
        return SRT_OK;
    }
    return set_err(c, SRT_E_INVALID, "unknown option '%s'", key);
}

size_t srt_hbm_used(const srt_ctx* c) { return c ? c->used.load() : 0; }

int srt_scene_upload(srt_ctx* c, const srt_scene_desc* d, srt_scene* out) {
    SRT_IDLE(c);
    if (!c || !d || !out) return SRT_E_INVALID;
    if (d->num_vertices <= 0 || !d->positions) return set_err(c, SRT_E_INVALID, "scene needs vertices");
    if (d->num_triangles <= 0 || !d->indices) return set_err(c, SRT_E_INVALID, "scene needs triangles");
    if (d->num_meshes <= 0 || !d->mesh_first_triangle || !d->mesh_material) return set_err(c, SRT_E_INVALID, "scene needs meshes");
    if (d->num_materials <= 0 || !d->materials) return set_err(c, SRT_E_INVALID, "scene needs materials");
    const int nv = d->num_vertices, nf = d->num_triangles, nm = d->num_meshes;
    if (d->mesh_first_triangle[0] != 0 || d->mesh_first_triangle[nm] != nf)
        return set_err(c, SRT_E_INVALID, "mesh_first_triangle must start at 0 and end at num_triangles");
    for (int m = 0; m < nm; m++) {
        if (d->mesh_first_triangle[m + 1] <= d->mesh_first_triangle[m])
            return set_err(c, SRT_E_INVALID, "mesh %d has no triangles", m);
        if (d->mesh_material[m] < 0 || d->mesh_material[m] >= d->num_materials)
            return set_err(c, SRT_E_INVALID, "mesh %d references material %d", m, d->mesh_material[m]);
    }
    for (int64_t i = 0; i < 3 * (int64_t)nf; i++)
        if (d->indices[i] >= (uint32_t)nv) return set_err(c, SRT_E_INVALID, "index %u out of range", d->indices[i]);
    for (int i = 0; i < d->num_materials; i++) {
        int t = d->materials[i].type;
        if (t < 0 || t > 3) return set_err(c, SRT_E_INVALID, "material %d has unknown BxDF type %d", i, t);
    }
    // per-mesh area CDF (mesh.cpp:3-18) and light CDF over intensity * area
    // (scene.cu:240-264), in host float arithmetic as the reference does
    std::vector<float> area_cdf(nf), mesh_area(nm);
    std::vector<int32_t> tri_mesh(nf);
    for (int m = 0; m < nm; m++) {
        float area = 0.0f;
        for (int f = d->mesh_first_triangle[m]; f < d->mesh_first_triangle[m + 1]; f++) {
            tri_mesh[f] = m;
            const float* v0 = d->positions + 3 * (size_t)d->indices[3 * (size_t)f];
            const float* v1 = d->positions + 3 * (size_t)d->indices[3 * (size_t)f + 1];
            const float* v2 = d->positions + 3 * (size_t)d->indices[3 * (size_t)f + 2];
            float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
            float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
            float cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
            float A = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
            area += A;
            area_cdf[f] = area;
        }
        for (int f = d->mesh_first_triangle[m]; f < d->mesh_first_triangle[m + 1]; f++) area_cdf[f] /= area;
        mesh_area[m] = area;
    }
    std::vector<float> light_cdf;
    std::vector<int32_t> light_mesh;
    float wsum = 0.0f;
    for (int m = 0; m < nm; m++) {
        const srt_material& mat = d->materials[d->mesh_material[m]];
        if (mat.intensity <= 0.0f) continue;
        light_mesh.push_back(m);
        float w = mat.intensity * mesh_area[m];
        wsum += w;
        light_cdf.push_back(wsum);
    }
    for (auto& x : light_cdf) x /= wsum;
    SceneRec* s = new SceneRec();
    for (int a = 0; a < 3; a++) { s->bounds_lo[a] = d->positions[a]; s->bounds_hi[a] = d->positions[a]; }
    for (int64_t i = 0; i < (int64_t)nv; i++)
        for (int a = 0; a < 3; a++) {
            float x = d->positions[3 * i + a];
            s->bounds_lo[a] = std::min(s->bounds_lo[a], x);
            s->bounds_hi[a] = std::max(s->bounds_hi[a], x);
        }
    s->nv = nv; s->nf = nf; s->nmesh = nm; s->nmat = d->num_materials; s->nlights = (int32_t)light_mesh.size();
    s->has_nrm = d->normals != nullptr;
    s->has_tan = d->tangents != nullptr;
    for (int i = 0; i < 3; i++) s->env[i] = d->env_color[i];
    s->light_wsum = wsum;
    s->h_mesh_area = mesh_area;
    s->h_mesh_box.assign(6 * (size_t)nm, 0.0f);
    for (int m = 0; m < nm; m++) {
        float* b = &s->h_mesh_box[6 * (size_t)m];
        for (int a = 0; a < 3; a++) { b[a] = INFINITY; b[3 + a] = -INFINITY; }
        for (int f = d->mesh_first_triangle[m]; f < d->mesh_first_triangle[m + 1]; f++)
            for (int k = 0; k < 3; k++) {
                const float* v = d->positions + 3 * (size_t)d->indices[3 * (size_t)f + k];
                for (int a = 0; a < 3; a++) { b[a] = std::min(b[a], v[a]); b[3 + a] = std::max(b[3 + a], v[a]); }
            }
    }
    int r = SRT_OK;
    if (!r) r = upload(c, s->pos, d->positions, 12 * (size_t)nv);
    if (!r && s->has_nrm) r = upload(c, s->nrm, d->normals, 12 * (size_t)nv);
    if (!r && s->has_tan) r = upload(c, s->tan, d->tangents, 12 * (size_t)nv);
    if (!r) r = upload(c, s->idx, d->indices, 12 * (size_t)nf);
    if (!r) r = upload(c, s->mesh_mat, d->mesh_material, 4 * (size_t)nm);
    if (!r) r = upload(c, s->mesh_first, d->mesh_first_triangle, 4 * (size_t)(nm + 1));
    if (!r) r = upload(c, s->area_cdf, area_cdf.data(), 4 * (size_t)nf);
    if (!r) r = upload(c, s->tri_mesh, tri_mesh.data(), 4 * (size_t)nf);
    if (!r) r = upload(c, s->mats, d->materials, sizeof(srt_material) * (size_t)d->num_materials);
    if (!r && s->nlights > 0) r = upload(c, s->light_cdf, light_cdf.data(), 4 * light_cdf.size());
    if (!r && s->nlights > 0) r = upload(c, s->light_mesh, light_mesh.data(), 4 * light_mesh.size());
    if (r) { free_scene(c, s); return r; }
    uint64_t h = c->next_scene++;
    s->handle = h;
    c->scenes[h] = s;
    *out = h;
    return SRT_OK;
}

int srt_scene_release(srt_ctx* c, srt_scene h) {
    SRT_IDLE(c);
    if (!c) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene %llu", (unsigned long long)h);
    (void)hipStreamSynchronize(c->stream);
    mlp_release_scene(c, h);
    free_scene(c, s);
    c->scenes.erase(h);
    return SRT_OK;
}

int srt_scene_lights(srt_ctx* c, srt_scene h, float* light_cdf, int32_t* light_mesh, float* area_cdf, float* mesh_area,
                     float* weight_sum) {
    SRT_IDLE(c);
    if (!c) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (light_cdf && s->nlights) hip_check(c, hipMemcpy(light_cdf, s->light_cdf.p, 4 * (size_t)s->nlights, hipMemcpyDeviceToHost), "copy");
    if (light_mesh && s->nlights) hip_check(c, hipMemcpy(light_mesh, s->light_mesh.p, 4 * (size_t)s->nlights, hipMemcpyDeviceToHost), "copy");
    if (area_cdf) hip_check(c, hipMemcpy(area_cdf, s->area_cdf.p, 4 * (size_t)s->nf, hipMemcpyDeviceToHost), "copy");
    if (mesh_area) memcpy(mesh_area, s->h_mesh_area.data(), 4 * (size_t)s->nmesh);
    if (weight_sum) *weight_sum = s->light_wsum;
    return s->nlights;
}

// Render entries build a default tree on first use: the two-level form for
// instanced scenes, the single tree otherwise (leaves of up to 4 triangles).
static int ensure_bvh(srt_ctx* c, SceneRec* s) {
    if (s->has_bvh) return SRT_OK;
    return s->ninst > 0 ? inst_build(c, s, kDefaultMaxLeaf, 0) : bvh_build(c, s, kDefaultMaxLeaf, 0);
}

int srt_scene_build_bvh(srt_ctx* c, srt_scene h, const srt_bvh_opts* o) {
    SRT_IDLE(c);
    if (!c) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    int max_leaf = o ? o->max_leaf : kDefaultMaxLeaf;
    int cluster = o ? o->cluster_size : 0;
    if (max_leaf <= 0) max_leaf = kDefaultMaxLeaf;
    hipEvent_t ev;
    prof_begin(c, "bvh_build", &ev);
    int r = s->ninst > 0 ? inst_build(c, s, max_leaf, cluster) : bvh_build(c, s, max_leaf, cluster);
    prof_end(c, "bvh_build", ev, s->nf);
    prof_collect(c);
    return r;
}

// Instances (the IAS over per-mesh GAS of optix_ray_tracer.cu:337-359 and the
// instance transforms of scene.cu:225-288): the meshes become object space;
// lights become the emissive instances (weight intensity x the mesh's
// object-space area, scene.cu:240-264).  The BVH must be built afterwards.
int srt_scene_set_instances(srt_ctx* c, srt_scene h, int32_t n, const int32_t* inst_mesh, const float* xf) {
    SRT_IDLE(c);
    if (!c || n <= 0 || !inst_mesh || !xf) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (s->nbones > 0 || s->bone_ids.p) return set_err(c, SRT_E_STATE, "a skinned scene cannot also have instances");
    if (n >= 0x3fffffff) return set_err(c, SRT_E_LIMIT, "too many instances");
    if ((int64_t)s->nf > kMaxInstTriangles) return set_err(c, SRT_E_LIMIT, "an instanced scene holds < 2^27 triangles");
    std::vector<float> inv(12 * (size_t)n);
    for (int i = 0; i < n; i++) {
        if (inst_mesh[i] < 0 || inst_mesh[i] >= s->nmesh)
            return set_err(c, SRT_E_INVALID, "instance %d references mesh %d", i, inst_mesh[i]);
        for (int k = 0; k < 12; k++)
            if (!std::isfinite(xf[12 * (size_t)i + k])) return set_err(c, SRT_E_INVALID, "instance %d transform is not finite", i);
        if (!affine_inverse(xf + 12 * (size_t)i, &inv[12 * (size_t)i]))
            return set_err(c, SRT_E_INVALID, "instance %d transform is singular", i);
    }
    // lights over instances, host float arithmetic in instance order
    std::vector<float> light_cdf;
    std::vector<int32_t> light_mesh, light_inst;
    float wsum = 0.0f;
    std::vector<srt_material> mats(s->nmat);
    std::vector<int32_t> mesh_mat(s->nmesh);
    int r = hip_check(c, hipMemcpy(mats.data(), s->mats.p, sizeof(srt_material) * mats.size(), hipMemcpyDeviceToHost), "materials");
    if (!r) r = hip_check(c, hipMemcpy(mesh_mat.data(), s->mesh_mat.p, 4 * mesh_mat.size(), hipMemcpyDeviceToHost), "materials");
    if (r) return r;
    for (int i = 0; i < n; i++) {
        const int m = inst_mesh[i];
        const srt_material& mat = mats[mesh_mat[m]];
        if (mat.intensity <= 0.0f) continue;
        light_mesh.push_back(m);
        light_inst.push_back(i);
        float w = mat.intensity * s->h_mesh_area[m];
        wsum += w;
        light_cdf.push_back(wsum);
    }
    for (auto& x : light_cdf) x /= wsum;
    // device copies: transforms as three float4 rows each
    DevBuf* bufs[] = {&s->inst_xf, &s->inst_inv, &s->inst_root, &s->light_inst, &s->light_cdf, &s->light_mesh};
    for (DevBuf* b : bufs) dev_free(c, *b);
    r = upload(c, s->inst_xf, xf, 48 * (size_t)n);
    if (!r) r = upload(c, s->inst_inv, inv.data(), 48 * (size_t)n);
    if (!r && !light_cdf.empty()) r = upload(c, s->light_cdf, light_cdf.data(), 4 * light_cdf.size());
    if (!r && !light_cdf.empty()) r = upload(c, s->light_mesh, light_mesh.data(), 4 * light_mesh.size());
    if (!r && !light_cdf.empty()) r = upload(c, s->light_inst, light_inst.data(), 4 * light_inst.size());
    if (r) return r;
    s->ninst = n;
    s->h_inst_mesh.assign(inst_mesh, inst_mesh + n);
    s->h_inst_xf.assign(xf, xf + 12 * (size_t)n);
    s->h_inst_inv = inv;
    inst_scene_box(s);
    s->nlights = (int32_t)light_cdf.size();
    s->light_wsum = wsum;
    s->has_bvh = false;
    return SRT_OK;
}

// Node animation (scene.cu:309-334 updates instance transforms; update_as,
// optix_ray_tracer.cu:443-463, refits): new transforms for every instance,
// then the top-level tree is refit.
int srt_scene_update_instances(srt_ctx* c, srt_scene h, const float* xf) {
    SRT_IDLE(c);
    if (!c || !xf) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (s->ninst <= 0) return set_err(c, SRT_E_STATE, "the scene has no instances");
    if (!s->has_bvh) return set_err(c, SRT_E_STATE, "no BVH built");
    hipEvent_t ev;
    prof_begin(c, "inst_refit", &ev);
    int r = inst_update(c, s, xf);
    prof_end(c, "inst_refit", ev, s->ninst);
    prof_collect(c);
    return r;
}

int srt_scene_bvh_info(srt_ctx* c, srt_scene h, int32_t info[8]) {
    SRT_IDLE(c);
    if (!c || !info) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (!s->has_bvh) return set_err(c, SRT_E_STATE, "no BVH built");
    info[0] = s->num_nodes; info[1] = s->root; info[2] = s->max_depth; info[3] = s->num_clusters;
    info[4] = s->num_top; info[5] = s->cluster_size; info[6] = s->nf; info[7] = s->max_leaf;
    return SRT_OK;
}

int srt_scene_bvh_download(srt_ctx* c, srt_scene h, void* nodes, uint32_t* order, float* tris) {
    SRT_IDLE(c);
    if (!c) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (s->ninst > 0) return set_err(c, SRT_E_INVALID, "an instanced scene holds no single binary tree");
    if (!s->has_bvh) return set_err(c, SRT_E_STATE, "no BVH built");
    if (int r = hip_check(c, hipStreamSynchronize(c->stream), "sync")) return r;
    if (nodes) if (int r = hip_check(c, hipMemcpy(nodes, s->nodes.p, (size_t)s->num_nodes * 64, hipMemcpyDeviceToHost), "copy")) return r;
    if (order) if (int r = hip_check(c, hipMemcpy(order, s->order.p, (size_t)s->nf * 4, hipMemcpyDeviceToHost), "copy")) return r;
    if (tris) if (int r = hip_check(c, hipMemcpy(tris, s->tris.p, (size_t)s->nf * 48, hipMemcpyDeviceToHost), "copy")) return r;
    return SRT_OK;
}

int srt_scene_set_skin(srt_ctx* c, srt_scene h, const int32_t* bone_ids, const float* bone_weights) {
    SRT_IDLE(c);
    if (!c || !bone_ids || !bone_weights) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (s->ninst > 0) return set_err(c, SRT_E_INVALID, "instanced scenes are animated by instance transforms (srt_scene_update_instances)");
    for (int64_t i = 0; i < 4 * (int64_t)s->nv; i++)
        if (bone_ids[i] < -1) return set_err(c, SRT_E_INVALID, "bone id %d invalid", bone_ids[i]);
    dev_free(c, s->bone_ids); dev_free(c, s->bone_w);
    dev_free(c, s->orig_pos); dev_free(c, s->orig_nrm); dev_free(c, s->orig_tan);
    int r = upload(c, s->bone_ids, bone_ids, 16 * (size_t)s->nv);
    if (!r) r = upload(c, s->bone_w, bone_weights, 16 * (size_t)s->nv);
    if (!r) r = dev_alloc(c, s->orig_pos, 12 * (size_t)s->nv);
    if (!r) r = hip_check(c, hipMemcpy(s->orig_pos.p, s->pos.p, 12 * (size_t)s->nv, hipMemcpyDeviceToDevice), "copy");
    if (!r && s->has_nrm) {
        r = dev_alloc(c, s->orig_nrm, 12 * (size_t)s->nv);
        if (!r) r = hip_check(c, hipMemcpy(s->orig_nrm.p, s->nrm.p, 12 * (size_t)s->nv, hipMemcpyDeviceToDevice), "copy");
    }
    if (!r && s->has_tan) {
        r = dev_alloc(c, s->orig_tan, 12 * (size_t)s->nv);
        if (!r) r = hip_check(c, hipMemcpy(s->orig_tan.p, s->tan.p, 12 * (size_t)s->nv, hipMemcpyDeviceToDevice), "copy");
    }
    return r;
}

int srt_scene_skin_refit(srt_ctx* c, srt_scene h, const float* bones_3x4, int32_t n_bones) {
    SRT_IDLE(c);
    if (!c || (n_bones > 0 && !bones_3x4)) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (s->ninst > 0) return set_err(c, SRT_E_INVALID, "instanced scenes are animated by instance transforms (srt_scene_update_instances)");
    if (int r = skin_vertices(c, s, bones_3x4, n_bones)) return r;
    hipEvent_t ev;
    prof_begin(c, "refit", &ev);
    int r = bvh_refit(c, s);
    prof_end(c, "refit", ev, s->num_nodes);
    prof_collect(c);
    return r;
}

int srt_scene_download_vertices(srt_ctx* c, srt_scene h, float* positions, float* normals, float* tangents) {
    SRT_IDLE(c);
    if (!c) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (int r = hip_check(c, hipStreamSynchronize(c->stream), "sync")) return r;
    if (positions) if (int r = hip_check(c, hipMemcpy(positions, s->pos.p, 12 * (size_t)s->nv, hipMemcpyDeviceToHost), "copy")) return r;
    if (normals && s->has_nrm) if (int r = hip_check(c, hipMemcpy(normals, s->nrm.p, 12 * (size_t)s->nv, hipMemcpyDeviceToHost), "copy")) return r;
    if (tangents && s->has_tan) if (int r = hip_check(c, hipMemcpy(tangents, s->tan.p, 12 * (size_t)s->nv, hipMemcpyDeviceToHost), "copy")) return r;
    return SRT_OK;
}

int srt_trace_rays(srt_ctx* c, srt_scene h, int32_t n, const float* o, const float* d, float tmin, const float* tmax,
                   int32_t any, float* t, float* u, float* v, int32_t* prim) {
    SRT_IDLE(c);
    if (!c || n < 0 || !o || !d || !tmax || !prim || (!any && (!t || !u || !v))) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (!s->has_bvh) return set_err(c, SRT_E_STATE, "no BVH built");
    if (n == 0) return SRT_OK;
    DevBuf bo, bd, bt, tt, tu, tv, tp;
    int r = upload(c, bo, o, 12 * (size_t)n);
    if (!r) r = upload(c, bd, d, 12 * (size_t)n);
    if (!r) r = upload(c, bt, tmax, 4 * (size_t)n);
    if (!r) r = dev_alloc(c, tt, 4 * (size_t)n);
    if (!r) r = dev_alloc(c, tu, 4 * (size_t)n);
    if (!r) r = dev_alloc(c, tv, 4 * (size_t)n);
    if (!r) r = dev_alloc(c, tp, 4 * (size_t)n);
    if (!r) r = trace_rays(c, s, n, (const float*)bo.p, (const float*)bd.p, tmin, (const float*)bt.p, any, (float*)tt.p,
                           (float*)tu.p, (float*)tv.p, (int32_t*)tp.p);
    if (!r) r = hip_check(c, hipMemcpy(prim, tp.p, 4 * (size_t)n, hipMemcpyDeviceToHost), "copy");
    if (!r && !any) {
        r = hip_check(c, hipMemcpy(t, tt.p, 4 * (size_t)n, hipMemcpyDeviceToHost), "copy");
        if (!r) r = hip_check(c, hipMemcpy(u, tu.p, 4 * (size_t)n, hipMemcpyDeviceToHost), "copy");
        if (!r) r = hip_check(c, hipMemcpy(v, tv.p, 4 * (size_t)n, hipMemcpyDeviceToHost), "copy");
    }
    DevBuf* all[] = {&bo, &bd, &bt, &tt, &tu, &tv, &tp};
    for (DevBuf* b : all) dev_free(c, *b);
    return r;
}

int srt_trace_rays_inst(srt_ctx* c, srt_scene h, int32_t n, const float* o, const float* d, float tmin, const float* tmax,
                        int32_t any, float* t, float* u, float* v, int32_t* prim, int32_t* inst) {
    SRT_IDLE(c);
    if (!c || n < 0 || !o || !d || !tmax || !prim || (!any && (!t || !u || !v || !inst))) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (!s->has_bvh) return set_err(c, SRT_E_STATE, "no BVH built");
    if (n == 0) return SRT_OK;
    DevBuf bo, bd, bt, tt, tu, tv, tp, ti;
    int r = upload(c, bo, o, 12 * (size_t)n);
    if (!r) r = upload(c, bd, d, 12 * (size_t)n);
    if (!r) r = upload(c, bt, tmax, 4 * (size_t)n);
    if (!r) r = dev_alloc(c, tt, 4 * (size_t)n);
    if (!r) r = dev_alloc(c, tu, 4 * (size_t)n);
    if (!r) r = dev_alloc(c, tv, 4 * (size_t)n);
    if (!r) r = dev_alloc(c, tp, 4 * (size_t)n);
    if (!r) r = dev_alloc(c, ti, 4 * (size_t)n);
    if (!r) r = trace_rays(c, s, n, (const float*)bo.p, (const float*)bd.p, tmin, (const float*)bt.p, any, (float*)tt.p,
                           (float*)tu.p, (float*)tv.p, (int32_t*)tp.p, any ? nullptr : (int32_t*)ti.p);
    if (!r) r = hip_check(c, hipMemcpy(prim, tp.p, 4 * (size_t)n, hipMemcpyDeviceToHost), "copy");
    if (!r && !any) {
        r = hip_check(c, hipMemcpy(t, tt.p, 4 * (size_t)n, hipMemcpyDeviceToHost), "copy");
        if (!r) r = hip_check(c, hipMemcpy(u, tu.p, 4 * (size_t)n, hipMemcpyDeviceToHost), "copy");
        if (!r) r = hip_check(c, hipMemcpy(v, tv.p, 4 * (size_t)n, hipMemcpyDeviceToHost), "copy");
        if (!r) {
            if (s->ninst > 0) r = hip_check(c, hipMemcpy(inst, ti.p, 4 * (size_t)n, hipMemcpyDeviceToHost), "copy");
            else for (int32_t i = 0; i < n; i++) inst[i] = -1;
        }
    }
    DevBuf* all[] = {&bo, &bd, &bt, &tt, &tu, &tv, &tp, &ti};
    for (DevBuf* b : all) dev_free(c, *b);
    return r;
}

int srt_trace_rays_noalloc(srt_ctx* c, srt_scene h, int32_t n, const float* o, const float* d, float tmin,
                           const float* tmax, int32_t any, float* t, float* u, float* v, int32_t* prim) {
    SRT_IDLE(c);
    if (!c || n < 0 || !o || !d || !tmax || !prim || (!any && (!t || !u || !v))) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (!s->has_bvh) return set_err(c, SRT_E_STATE, "no BVH built");
    if (n == 0) return SRT_OK;
    return trace_rays_noalloc(c, s, n, o, d, tmin, tmax, any, t, u, v, prim);
}

int srt_render(srt_ctx* c, srt_scene h, const srt_camera* cam, const srt_pt_params* p, int32_t W, int32_t H, float* rgba_out,
               int32_t out_on_device) {
    SRT_IDLE(c);
    if (!c || !cam || !p) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (W <= 0 || H <= 0) return set_err(c, SRT_E_INVALID, "resolution must be positive");
    if (p->samples_per_pixel <= 0) return set_err(c, SRT_E_INVALID, "samples_per_pixel must be positive");
    if (p->max_depth < 0 || p->rr_depth < 0) return set_err(c, SRT_E_INVALID, "depths must be >= 0");
    if ((int64_t)W * H > (1ll << 28)) return set_err(c, SRT_E_LIMIT, "resolution too large");
    if ((int64_t)W * H * p->samples_per_pixel > (1ll << 40)) return set_err(c, SRT_E_LIMIT, "too many paths");
    if (!(cam->fov_deg > 0.0f && cam->fov_deg < 180.0f)) return set_err(c, SRT_E_INVALID, "fov must be in (0, 180)");
    if (p->mlp_mode < 0 || p->mlp_mode > 1) return set_err(c, SRT_E_INVALID, "mlp_mode must be 0 or 1");
    if (p->mlp_mode == 1 && !mlp_for_scene(c, h)) return set_err(c, SRT_E_STATE, "mlp_mode 1 needs srt_mlp_create");
    if (int r = ensure_bvh(c, s)) return r;
    if (p->use_nee && s->nlights == 0) {
        // no area light: next-event estimation never fires (light.h sample on
        // an empty CDF is undefined in the reference); render without it
    }
    DevCamera dc = make_camera(cam->position, cam->target, cam->up, cam->fov_deg, W, H);
    DevParams dp;
    dp.spp = p->samples_per_pixel; dp.max_depth = p->max_depth; dp.rr_depth = p->rr_depth;
    dp.use_nee = (p->use_nee && s->nlights > 0) ? 1 : 0; dp.use_mis = p->use_mis ? 1 : 0; dp.seed = p->seed;
    dp.width = W; dp.height = H;
    dp.mlp_mode = p->mlp_mode;
    const int64_t wh = (int64_t)W * H;
    float* film = nullptr;
    if (rgba_out && out_on_device) film = rgba_out;
    else {
        if (!c->film_out.p || c->film_out.bytes < (size_t)wh * 16) {
            dev_free(c, c->film_out);
            if (int r = dev_alloc(c, c->film_out, (size_t)wh * 16)) return r;
        }
        film = (float*)c->film_out.p;
    }
    if (int r = render_wavefront(c, s, dc, dp, film)) return r;
    if (rgba_out && !out_on_device)
        return hip_check(c, hipMemcpy(rgba_out, film, (size_t)wh * 16, hipMemcpyDeviceToHost), "film download");
    return SRT_OK;
}

int srt_render_simple(srt_ctx* c, srt_scene h, const srt_camera* cam, int32_t type, int32_t spp, uint32_t seed, int32_t W,
                      int32_t H, float* rgba_out, int32_t out_on_device) {
    SRT_IDLE(c);
    if (!c || !cam || !rgba_out) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (W <= 0 || H <= 0) return set_err(c, SRT_E_INVALID, "resolution must be positive");
    if ((int64_t)W * H > (1ll << 28)) return set_err(c, SRT_E_LIMIT, "resolution too large");
    if (spp <= 0) return set_err(c, SRT_E_INVALID, "samples_per_pixel must be positive");
    if (type < 0 || type > 4) return set_err(c, SRT_E_INVALID, "type must be 0..4 (Depth, Normal, BaseColor, Ambient, FaceOrientation)");
    if (!(cam->fov_deg > 0.0f && cam->fov_deg < 180.0f)) return set_err(c, SRT_E_INVALID, "fov must be in (0, 180)");
    if (int r = ensure_bvh(c, s)) return r;
    DevCamera dc = make_camera(cam->position, cam->target, cam->up, cam->fov_deg, W, H);
    const int64_t wh = (int64_t)W * H;
    float* film = nullptr;
    if (out_on_device) film = rgba_out;
    else {
        if (!c->film_out.p || c->film_out.bytes < (size_t)wh * 16) {
            dev_free(c, c->film_out);
            if (int r = dev_alloc(c, c->film_out, (size_t)wh * 16)) return r;
        }
        film = (float*)c->film_out.p;
    }
    if (int r = render_simple(c, s, dc, type, spp, seed, W, H, film)) return r;
    prof_collect(c);
    if (!out_on_device) return hip_check(c, hipMemcpy(rgba_out, film, (size_t)wh * 16, hipMemcpyDeviceToHost), "film download");
    return SRT_OK;
}

int srt_render_light(srt_ctx* c, srt_scene h, const srt_camera* cam, const srt_pt_params* p, int32_t W, int32_t H,
                     float* rgba_out, int32_t out_on_device) {
    SRT_IDLE(c);
    if (!c || !cam || !p || !rgba_out) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (W <= 0 || H <= 0) return set_err(c, SRT_E_INVALID, "resolution must be positive");
    if ((int64_t)W * H > (1ll << 28)) return set_err(c, SRT_E_LIMIT, "resolution too large");
    if (p->samples_per_pixel <= 0) return set_err(c, SRT_E_INVALID, "samples_per_pixel must be positive");
    if ((int64_t)W * H * p->samples_per_pixel > (1ll << 40)) return set_err(c, SRT_E_LIMIT, "too many paths");
    if (p->max_depth < 0 || p->rr_depth < 0) return set_err(c, SRT_E_INVALID, "depths must be >= 0");
    if (!(cam->fov_deg > 0.0f && cam->fov_deg < 180.0f)) return set_err(c, SRT_E_INVALID, "fov must be in (0, 180)");
    if (s->nlights == 0) return set_err(c, SRT_E_INVALID, "the light tracer needs an area light in the scene");
    if (int r = ensure_bvh(c, s)) return r;
    DevCamera dc = make_camera(cam->position, cam->target, cam->up, cam->fov_deg, W, H);
    const int64_t wh = (int64_t)W * H;
    float* film = nullptr;
    if (out_on_device) film = rgba_out;
    else {
        if (!c->film_out.p || c->film_out.bytes < (size_t)wh * 16) {
            dev_free(c, c->film_out);
            if (int r = dev_alloc(c, c->film_out, (size_t)wh * 16)) return r;
        }
        film = (float*)c->film_out.p;
    }
    if (int r = render_light(c, s, dc, p->samples_per_pixel, p->max_depth, p->rr_depth, p->seed, W, H, film)) return r;
    prof_collect(c);
    if (!out_on_device) return hip_check(c, hipMemcpy(rgba_out, film, (size_t)wh * 16, hipMemcpyDeviceToHost), "film download");
    return SRT_OK;
}

int srt_mlp_train_frame(srt_ctx* c, srt_scene h, const srt_camera* cam, const srt_pt_params* p, int32_t W, int32_t H,
                        int64_t max_samples, int32_t steps, float* loss_out, int64_t* samples_out) {
    SRT_IDLE(c);
    if (!c || !cam || !p || W <= 0 || H <= 0 || max_samples <= 0 || steps <= 0) return SRT_E_INVALID;
    SceneRec* s = find_scene(c, h);
    if (!s) return set_err(c, SRT_E_NOT_FOUND, "unknown scene");
    if (!mlp_for_scene(c, h)) return set_err(c, SRT_E_STATE, "no MLP: call srt_mlp_create first");
    const MlpBind bind(c, h);
    if (p->samples_per_pixel <= 0 || p->max_depth < 3) return set_err(c, SRT_E_INVALID, "training paths need max_depth >= 3");
    if (int r = ensure_bvh(c, s)) return r;
    if (c->train_cap < max_samples) {
        dev_free(c, c->train_in9);
        dev_free(c, c->train_t);
        if (int r = dev_alloc(c, c->train_in9, (size_t)max_samples * 36)) return r;
        if (int r = dev_alloc(c, c->train_t, (size_t)max_samples * 12)) return r;
        c->train_cap = max_samples;
    }
    const int64_t wh = (int64_t)W * H;
    if (!c->film_out.p || c->film_out.bytes < (size_t)wh * 16) {
        dev_free(c, c->film_out);
        if (int r = dev_alloc(c, c->film_out, (size_t)wh * 16)) return r;
    }
    DevCamera dc = make_camera(cam->position, cam->target, cam->up, cam->fov_deg, W, H);
    DevParams dp;
    dp.spp = p->samples_per_pixel; dp.max_depth = p->max_depth; dp.rr_depth = p->rr_depth;
    dp.use_nee = (p->use_nee && s->nlights > 0) ? 1 : 0; dp.use_mis = p->use_mis ? 1 : 0; dp.seed = p->seed;
    dp.width = W; dp.height = H; dp.mlp_mode = 2;
    TrainOut tr;
    tr.in9 = (float*)c->train_in9.p;
    tr.target = (float*)c->train_t.p;
    tr.cap = max_samples;
    tr.count = 0;
    if (int r = render_wavefront(c, s, dc, dp, (float*)c->film_out.p, &tr)) return r;
    if (samples_out) *samples_out = tr.count;
    const int64_t batch = tr.count / steps;
    double lsum = 0.0;
    int done = 0;
    for (int k = 0; k < steps && batch > 0; k++) {
        float l = 0.0f;
        if (int r = mlp_train_dev(c, tr.in9 + (size_t)k * batch * 9, tr.target + (size_t)k * batch * 3, (int)batch,
                                  loss_out ? &l : nullptr))
            return r;
        lsum += l;
        done++;
    }
    if (loss_out) *loss_out = done ? (float)(lsum / done) : 0.0f;
    if (int r = hip_check(c, hipStreamSynchronize(c->stream), "train frame")) return r;
    prof_collect(c);
    return SRT_OK;
}

int srt_film_download(srt_ctx* c, float* rgba, int64_t n) {
    SRT_IDLE(c);
    if (!c || !rgba || n < 0) return SRT_E_INVALID;
    if (!c->film_out.p || (size_t)n * 16 > c->film_out.bytes) return set_err(c, SRT_E_STATE, "no film of that size");
    return hip_check(c, hipMemcpy(rgba, c->film_out.p, (size_t)n * 16, hipMemcpyDeviceToHost), "film download");
}

static int post_buffer(srt_ctx* c, float* host, int64_t n, DevBuf& tmp) {
    if (int r = dev_alloc(c, tmp, (size_t)n * 16)) return r;
    return hip_check(c, hipMemcpy(tmp.p, host, (size_t)n * 16, hipMemcpyHostToDevice), "upload");
}

int srt_tonemap(srt_ctx* c, float* rgba, int64_t n, int32_t type, float exposure, int32_t use_gamma, int32_t on_device) {
    SRT_IDLE(c);
    if (!c || !rgba || n < 0 || n > 0x7fffffff) return SRT_E_INVALID;
    if (type < 0 || type > 4) return set_err(c, SRT_E_INVALID, "unknown tonemap type %d", type);
    if (n == 0) return SRT_OK;
    if (on_device) return tonemap_dev(c, rgba, (int)n, type, exposure, use_gamma);
    DevBuf tmp;
    int r = post_buffer(c, rgba, n, tmp);
    if (!r) r = tonemap_dev(c, (float*)tmp.p, (int)n, type, exposure, use_gamma);
    if (!r) r = hip_check(c, hipMemcpy(rgba, tmp.p, (size_t)n * 16, hipMemcpyDeviceToHost), "download");
    dev_free(c, tmp);
    return r;
}

int srt_accumulate(srt_ctx* c, float* acc, float* rgba, int64_t n, int32_t frame_count, int32_t on_device) {
    SRT_IDLE(c);
    if (!c || !acc || !rgba || n < 0 || n > 0x7fffffff || frame_count < 0) return SRT_E_INVALID;
    if (n == 0) return SRT_OK;
    if (on_device) return accumulate_dev(c, acc, rgba, (int)n, frame_count);
    DevBuf ta, tp;
    int r = post_buffer(c, acc, n, ta);
    if (!r) r = post_buffer(c, rgba, n, tp);
    if (!r) r = accumulate_dev(c, (float*)ta.p, (float*)tp.p, (int)n, frame_count);
    if (!r) r = hip_check(c, hipMemcpy(acc, ta.p, (size_t)n * 16, hipMemcpyDeviceToHost), "download");
    if (!r) r = hip_check(c, hipMemcpy(rgba, tp.p, (size_t)n * 16, hipMemcpyDeviceToHost), "download");
    dev_free(c, ta);
    dev_free(c, tp);
    return r;
}

int srt_denoise(srt_ctx* c, float* rgba, int32_t W, int32_t H, int32_t on_device) {
    SRT_IDLE(c);
    if (!c || !rgba || W <= 0 || H <= 0) return SRT_E_INVALID;
    const int64_t n = (int64_t)W * H;
    if (n > (1ll << 28)) return set_err(c, SRT_E_LIMIT, "resolution too large");
    DevBuf tmp, tp;
    int r = dev_alloc(c, tmp, (size_t)n * 32);
    if (r) return r;
    if (on_device) r = denoise_dev(c, rgba, (float*)tmp.p, W, H);
    else {
        r = post_buffer(c, rgba, n, tp);
        if (!r) r = denoise_dev(c, (float*)tp.p, (float*)tmp.p, W, H);
        if (!r) r = hip_check(c, hipMemcpy(rgba, tp.p, (size_t)n * 16, hipMemcpyDeviceToHost), "download");
        dev_free(c, tp);
    }
    dev_free(c, tmp);
    prof_collect(c);
    return r;
}

int srt_device_alloc(srt_ctx* c, size_t bytes, void** out) {
    SRT_IDLE(c);
    if (!c || !out) return SRT_E_INVALID;
    DevBuf b;
    if (int r = dev_alloc(c, b, bytes)) return r;
    c->user_bufs[b.p] = b.bytes;
    *out = b.p;
    return SRT_OK;
}

int srt_device_free(srt_ctx* c, void* p) {
    SRT_IDLE(c);
    if (!c) return SRT_E_INVALID;
    auto it = c->user_bufs.find(p);
    if (it == c->user_bufs.end()) return set_err(c, SRT_E_NOT_FOUND, "pointer not allocated by srt_device_alloc");
    (void)hipStreamSynchronize(c->stream);
    DevBuf b;
    b.p = p;
    b.bytes = it->second;
    dev_free(c, b);
    c->user_bufs.erase(it);
    return SRT_OK;
}

int srt_memcpy(srt_ctx* c, void* dst, const void* src, size_t bytes, int32_t kind) {
    SRT_IDLE(c);
    if (!c || (!dst && bytes) || (!src && bytes) || kind < 0 || kind > 2) return SRT_E_INVALID;
    if (bytes == 0) return SRT_OK;
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    if (int r = hip_check(c, hipStreamSynchronize(c->stream), "sync")) return r;
    return hip_check(c, hipMemcpy(dst, src, bytes, k), "srt_memcpy");
}

int srt_stats(srt_ctx* c, char* names, int32_t names_len, double* ms, int64_t* launches, int64_t* items, int32_t max) {
    SRT_IDLE(c);
    if (!c) return SRT_E_INVALID;
    (void)hipStreamSynchronize(c->stream);
    prof_collect(c);
    int k = 0, off = 0;
    for (auto& kv : c->stats) {
        if (k >= max) break;
        int len = (int)kv.first.size() + 1;
        if (names) {
            if (off + len > names_len) break;
            memcpy(names + off, kv.first.c_str(), len);
        }
        off += len;
        if (ms) ms[k] = kv.second.ms;
        if (launches) launches[k] = kv.second.launches;
        if (items) items[k] = kv.second.items;
        k++;
    }
    return k;
}

void srt_stats_reset(srt_ctx* c) {
    if (c && c->serve) return;
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    prof_collect(c);
    c->stats.clear();
}

int srt_last_counts(srt_ctx* c, int64_t* segments, int64_t* shadow) {
    SRT_IDLE(c);
    if (!c) return SRT_E_INVALID;
    if (segments) *segments = c->last_segments;
    if (shadow) *shadow = c->last_shadow_rays;
    return SRT_OK;
}

int srt_last_visits(srt_ctx* c, int64_t out[4]) {
    SRT_IDLE(c);
    if (!c || !out) return SRT_E_INVALID;
    for (int k = 0; k < 4; k++) out[k] = c->last_visits[k];
    return SRT_OK;
}

}  // extern "C"
