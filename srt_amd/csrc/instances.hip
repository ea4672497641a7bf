// Two-level acceleration structure of instanced scenes: one tree per mesh
// over its object-space triangles (the GPU LBVH + binned-SAH build of
// bvh.hip, run mesh by mesh) and a 4-wide top-level tree over the instances'
// world boxes.  Replaces the reference's per-mesh GAS under one IAS with 3x4
// instance transforms (src/optix/optix_ray_tracer.cu:213-359) and its
// per-frame update (optix_ray_tracer.cu:443-463, fed by the instance
// transforms of src/scene/scene.cu:309-334).
//
// Layout after the build, all in the scene's arrays:
//   nodes4   = [mesh 0's 4-wide nodes | mesh 1's | ... | top-level nodes]
//   tris     = sorted triangles, mesh by mesh (mesh m's at its first
//              triangle), carrying scene primitive ids
//   tri_rec  = their shading records, in the same order
// Node refs inside a mesh's tree are relocated to these global positions;
// top-level children are top-level nodes or instance leaves (kInstLeaf | i).
// The top-level tree is small (one leaf per instance), so it is built and
// refit on the host and copied over.
#include "srt_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace srt {

// 3x4 affine inverse: double cofactors / determinant, translation -A^-1 t,
// rounded to float (the oracle computes it the same way).
bool affine_inverse(const float m[12], float out[12]) {
    const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], k = m[10];
    const double A = e * k - f * h, B = f * g - d * k, Cc = d * h - e * g;
    const double det = a * A + b * B + c * Cc;
    if (!(det != 0.0) || !std::isfinite(det)) return false;
    const double inv[9] = {A / det, (c * h - b * k) / det, (b * f - c * e) / det,
                           B / det, (a * k - c * g) / det, (c * d - a * f) / det,
                           Cc / det, (b * g - a * h) / det, (a * e - b * d) / det};
    const double t[3] = {m[3], m[7], m[11]};
    for (int r = 0; r < 3; r++) {
        out[4 * r + 0] = (float)inv[3 * r + 0];
        out[4 * r + 1] = (float)inv[3 * r + 1];
        out[4 * r + 2] = (float)inv[3 * r + 2];
        out[4 * r + 3] = (float)(-(inv[3 * r + 0] * t[0] + inv[3 * r + 1] * t[1] + inv[3 * r + 2] * t[2]));
    }
    return true;
}

namespace {

// Relocates one mesh's 4-wide nodes into the scene array: inner refs move by
// the mesh's first node, leaf refs by its first sorted triangle.
__global__ void __launch_bounds__(256) k_relocate4(const Bvh4Node* src, Bvh4Node* dst, int n, int32_t node_base,
                                                   uint32_t tri_base) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Bvh4Node x = src[i];
    int32_t r[4] = {x.ref.x, x.ref.y, x.ref.z, x.ref.w};
    for (int k = 0; k < 4; k++) {
        if (r[k] >= 0) r[k] += node_base;
        else r[k] = leaf_ref(leaf_first(r[k]) + tri_base, leaf_count(r[k]));
    }
    x.ref = make_int4(r[0], r[1], r[2], r[3]);
    x.src = make_int4(-1, -1, -1, -1);
    dst[i] = x;
}

// Mesh-local primitive ids of sorted triangles -> scene primitive ids.
__global__ void __launch_bounds__(256) k_prim_offset(float4* tris, int n, uint32_t base) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float4 a = tris[3 * (size_t)k];
    a.w = __uint_as_float(__float_as_uint(a.w) + base);
    tris[3 * (size_t)k] = a;
}

inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

// World box of an instance: points of its mesh through the transform in
// double, padded by 2^-14 of (coordinate magnitude + extent + the scene
// scale) and rounded outward.  The padding covers the difference between a
// world ray and the object-space ray the traversal computes from it (one
// float rounding per coordinate of M^-1 o and M^-1 d), so culling with these
// boxes never drops an instance whose object-space test would hit.  The
// points are the mesh's vertices (the tight box of the placed mesh) or, for
// meshes of more than kHullVerts vertices, the corners of its object box.
constexpr int64_t kHullVerts = 4096;
void inst_box(const float* M, const float* pts, int64_t npts, double scene_scale, float lo[3], float hi[3]) {
    double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t k = 0; k < npts; k++) {
        const double p[3] = {pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]};
        for (int r = 0; r < 3; r++) {
            const double w = (double)M[4 * r] * p[0] + (double)M[4 * r + 1] * p[1] + (double)M[4 * r + 2] * p[2] +
                             (double)M[4 * r + 3];
            l[r] = std::min(l[r], w);
            h[r] = std::max(h[r], w);
        }
    }
    for (int a = 0; a < 3; a++) {
        const double pad = (std::max(std::fabs(l[a]), std::fabs(h[a])) + (h[a] - l[a]) + scene_scale) * 0x1p-14;
        lo[a] = std::nextafter((float)(l[a] - pad), -INFINITY);
        hi[a] = std::nextafter((float)(h[a] + pad), INFINITY);
    }
}

struct TItem {
    float lo[3], hi[3], c[3];
    int32_t ref;
};

void grow(float lo[3], float hi[3], const TItem& t) {
    for (int a = 0; a < 3; a++) {
        lo[a] = std::min(lo[a], t.lo[a]);
        hi[a] = std::max(hi[a], t.hi[a]);
    }
}

// Top-down build: each node splits its instances at the centroid median of
// the widest axis, twice, into up to four children; one instance is a leaf.
// Nodes are appended in pre-order; `post` receives them children first.
int32_t tlas_node(std::vector<TItem>& it, int b, int e, int32_t first, std::vector<Bvh4Node>& nodes,
                  std::vector<int32_t>& post, int depth, int* max_depth) {
    *max_depth = std::max(*max_depth, depth);
    if (e - b == 1 && depth > 0) return it[b].ref;   // the root is always a node (its boxes cull every instance)
    std::vector<std::pair<int, int>> g{{b, e}};
    while (g.size() < 4) {
        int pick = -1;
        for (int k = 0; k < (int)g.size(); k++)
            if (g[k].second - g[k].first > 1 && (pick < 0 || g[k].second - g[k].first > g[pick].second - g[pick].first))
                pick = k;
        if (pick < 0) break;
        const int gb = g[pick].first, ge = g[pick].second;
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = gb; i < ge; i++)
            for (int a = 0; a < 3; a++) { clo[a] = std::min(clo[a], it[i].c[a]); chi[a] = std::max(chi[a], it[i].c[a]); }
        int ax = 0;
        for (int a = 1; a < 3; a++)
            if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
        const int mid = gb + (ge - gb) / 2;
        std::nth_element(it.begin() + gb, it.begin() + mid, it.begin() + ge, [ax](const TItem& x, const TItem& y) {
            return x.c[ax] < y.c[ax] || (x.c[ax] == y.c[ax] && x.ref < y.ref);
        });
        g[pick] = {gb, mid};
        g.insert(g.begin() + pick + 1, {mid, ge});
    }
    const int32_t me = (int32_t)nodes.size();
    nodes.push_back(Bvh4Node{});
    float lo[4][3], hi[4][3];
    int32_t ref[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; k++)
        for (int a = 0; a < 3; a++) { lo[k][a] = INFINITY; hi[k][a] = -INFINITY; }
    for (int k = 0; k < (int)g.size(); k++) {
        ref[k] = tlas_node(it, g[k].first, g[k].second, first, nodes, post, depth + 1, max_depth);
        for (int i = g[k].first; i < g[k].second; i++) grow(lo[k], hi[k], it[i]);
    }
    Bvh4Node& n = nodes[me];
    n.lox = make_float4(lo[0][0], lo[1][0], lo[2][0], lo[3][0]);
    n.hix = make_float4(hi[0][0], hi[1][0], hi[2][0], hi[3][0]);
    n.loy = make_float4(lo[0][1], lo[1][1], lo[2][1], lo[3][1]);
    n.hiy = make_float4(hi[0][1], hi[1][1], hi[2][1], hi[3][1]);
    n.loz = make_float4(lo[0][2], lo[1][2], lo[2][2], lo[3][2]);
    n.hiz = make_float4(hi[0][2], hi[1][2], hi[2][2], hi[3][2]);
    n.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
    n.src = make_int4(-1, -1, -1, -1);
    post.push_back(me);
    return first + me;
}

bool box_empty(const float* ob) { return !(ob[0] <= ob[3]); }

double scene_scale(const SceneRec* s) {
    double m = 0.0;
    for (int i = 0; i < s->ninst; i++) {
        const float* M = &s->h_inst_xf[12 * (size_t)i];
        const float* ob = &s->h_mesh_box[6 * (size_t)s->h_inst_mesh[i]];
        if (box_empty(ob)) continue;
        for (int k = 0; k < 8; k++) {
            const double p[3] = {(k & 1) ? ob[3] : ob[0], (k & 2) ? ob[4] : ob[1], (k & 4) ? ob[5] : ob[2]};
            for (int r = 0; r < 3; r++)
                m = std::max(m, std::fabs((double)M[4 * r] * p[0] + (double)M[4 * r + 1] * p[1] +
                                          (double)M[4 * r + 2] * p[2] + (double)M[4 * r + 3]));
        }
    }
    return m;
}

std::vector<TItem> inst_items(SceneRec* s) {
    const double scale = scene_scale(s);
    std::vector<TItem> it(s->ninst);
    for (int i = 0; i < s->ninst; i++) {
        TItem& t = it[i];
        t.ref = kInstLeaf | i;
        const float* ob = &s->h_mesh_box[6 * (size_t)s->h_inst_mesh[i]];
        if (box_empty(ob)) {   // a mesh without triangles: an empty box no ray enters
            for (int a = 0; a < 3; a++) { t.lo[a] = INFINITY; t.hi[a] = -INFINITY; t.c[a] = 0.0f; }
            continue;
        }
        const int m = s->h_inst_mesh[i];
        const int64_t p0 = s->h_mesh_pts_first[m], np = s->h_mesh_pts_first[m + 1] - p0;
        if (np <= kHullVerts) {
            inst_box(&s->h_inst_xf[12 * (size_t)i], &s->h_mesh_pts[3 * p0], np, scale, t.lo, t.hi);
        } else {
            float corners[24];
            for (int k = 0; k < 8; k++) {
                corners[3 * k] = (k & 1) ? ob[3] : ob[0];
                corners[3 * k + 1] = (k & 2) ? ob[4] : ob[1];
                corners[3 * k + 2] = (k & 4) ? ob[5] : ob[2];
            }
            inst_box(&s->h_inst_xf[12 * (size_t)i], corners, 8, scale, t.lo, t.hi);
        }
        for (int a = 0; a < 3; a++) t.c[a] = 0.5f * (t.lo[a] + t.hi[a]);
    }
    return it;
}

// child boxes of the top-level nodes from instance boxes, children first
void tlas_refit(SceneRec* s, const std::vector<TItem>& it) {
    for (int32_t k : s->h_tlas_order) {
        Bvh4Node& n = s->h_tlas[k];
        const int32_t refs[4] = {n.ref.x, n.ref.y, n.ref.z, n.ref.w};
        float* lox = &n.lox.x; float* hix = &n.hix.x; float* loy = &n.loy.x;
        float* hiy = &n.hiy.x; float* loz = &n.loz.x; float* hiz = &n.hiz.x;
        for (int c = 0; c < 4; c++) {
            if (!(lox[c] <= hix[c])) continue;                  // empty child
            float lo[3], hi[3];
            if (is_inst_ref(refs[c])) {
                const TItem& t = it[refs[c] & 0x3fffffff];
                for (int a = 0; a < 3; a++) { lo[a] = t.lo[a]; hi[a] = t.hi[a]; }
            } else {
                const Bvh4Node& ch = s->h_tlas[refs[c] - s->tlas_first];
                const float* cl[3] = {&ch.lox.x, &ch.loy.x, &ch.loz.x};
                const float* chh[3] = {&ch.hix.x, &ch.hiy.x, &ch.hiz.x};
                for (int a = 0; a < 3; a++) {
                    lo[a] = INFINITY; hi[a] = -INFINITY;
                    for (int q = 0; q < 4; q++)
                        if (cl[0][q] <= chh[0][q]) { lo[a] = std::min(lo[a], cl[a][q]); hi[a] = std::max(hi[a], chh[a][q]); }
                }
            }
            lox[c] = lo[0]; hix[c] = hi[0]; loy[c] = lo[1]; hiy[c] = hi[1]; loz[c] = lo[2]; hiz[c] = hi[2];
        }
    }
}

}  // namespace

// Scene box of an instanced scene (Scene::compute_aabb, scene.cu:299-307, with
// Transform::apply_aabb, transform.h:58-70): the union of each instance's
// mesh-box corners through apply_point, in float.
void inst_scene_box(SceneRec* s) {
    for (int a = 0; a < 3; a++) { s->bounds_lo[a] = INFINITY; s->bounds_hi[a] = -INFINITY; }
    for (int i = 0; i < s->ninst; i++) {
        const float* M = &s->h_inst_xf[12 * (size_t)i];
        const float* ob = &s->h_mesh_box[6 * (size_t)s->h_inst_mesh[i]];
        if (box_empty(ob)) continue;
        for (int k = 0; k < 8; k++) {
            const float x = (k & 4) ? ob[3] : ob[0], y = (k & 2) ? ob[4] : ob[1], z = (k & 1) ? ob[5] : ob[2];
            for (int r = 0; r < 3; r++) {
                const float w = M[4 * r] * x + M[4 * r + 1] * y + M[4 * r + 2] * z + M[4 * r + 3];
                s->bounds_lo[r] = std::fmin(s->bounds_lo[r], w);
                s->bounds_hi[r] = std::fmax(s->bounds_hi[r], w);
            }
        }
    }
}

// Each mesh's distinct vertices (object space), for the instance boxes.
static int mesh_points(srt_ctx* c, SceneRec* s, const std::vector<int32_t>& mesh_first) {
    if (!s->h_mesh_pts_first.empty()) return SRT_OK;
    std::vector<float> pos(3 * (size_t)s->nv);
    std::vector<uint32_t> idx(3 * (size_t)s->nf);
    int r = hip_check(c, hipMemcpy(pos.data(), s->pos.p, 12 * (size_t)s->nv, hipMemcpyDeviceToHost), "vertices");
    if (!r) r = hip_check(c, hipMemcpy(idx.data(), s->idx.p, 12 * (size_t)s->nf, hipMemcpyDeviceToHost), "indices");
    if (r) return r;
    std::vector<int32_t> seen(s->nv, -1);
    s->h_mesh_pts.clear();
    s->h_mesh_pts_first.assign(1, 0);
    for (int m = 0; m < s->nmesh; m++) {
        for (int64_t k = 3 * (int64_t)mesh_first[m]; k < 3 * (int64_t)mesh_first[m + 1]; k++) {
            const uint32_t v = idx[k];
            if (seen[v] == m) continue;
            seen[v] = m;
            s->h_mesh_pts.insert(s->h_mesh_pts.end(), &pos[3 * (size_t)v], &pos[3 * (size_t)v] + 3);
        }
        s->h_mesh_pts_first.push_back((int64_t)(s->h_mesh_pts.size() / 3));
    }
    return SRT_OK;
}

int inst_build(srt_ctx* c, SceneRec* s, int max_leaf, int cluster_size) {
    hipStream_t st = c->stream;
    const int nm = s->nmesh;
    std::vector<int32_t> mesh_first(nm + 1);
    if (int r = hip_check(c, hipMemcpy(mesh_first.data(), s->mesh_first.p, 4 * (size_t)(nm + 1), hipMemcpyDeviceToHost), "meshes"))
        return r;
    if (int r = mesh_points(c, s, mesh_first)) return r;
    // 1. one tree per mesh: the bvh.hip build on a view of the mesh's triangles
    struct Blas {
        DevBuf nodes4, tris, tri_rec;
        int32_t root4 = 0, nn4 = 0, depth4 = 0;
    };
    std::vector<Blas> bl(nm);
    auto free_blas = [&]() {
        for (Blas& b : bl) { dev_free(c, b.nodes4); dev_free(c, b.tris); dev_free(c, b.tri_rec); }
    };
    int64_t total4 = 0;
    int max_depth4 = 0;
    for (int m = 0; m < nm; m++) {
        SceneRec v;
        v.nv = s->nv;
        v.nf = mesh_first[m + 1] - mesh_first[m];
        v.nmesh = nm;
        v.nmat = s->nmat;
        v.has_nrm = s->has_nrm;
        v.has_tan = s->has_tan;
        v.pos.p = s->pos.p;
        v.nrm.p = s->nrm.p;
        v.tan.p = s->tan.p;
        v.idx.p = (char*)s->idx.p + 12 * (size_t)mesh_first[m];
        v.tri_mesh.p = (char*)s->tri_mesh.p + 4 * (size_t)mesh_first[m];
        v.mesh_mat.p = s->mesh_mat.p;
        v.mats.p = s->mats.p;
        if (v.nf == 0) continue;          // no tree; its instances hold empty boxes
        int r = bvh_build(c, &v, max_leaf, cluster_size);
        if (!r && v.root >= 0 && v.num_nodes4 == 0)
            r = set_err(c, SRT_E_LIMIT, "mesh %d's tree exceeds the 4-wide node limit", m);
        // keep the 4-wide tree, the sorted triangles and their records; drop the rest
        DevBuf* drop[] = {&v.nodes, &v.order, &v.node_parent, &v.level_nodes, &v.node_valid};
        for (DevBuf* b : drop) dev_free(c, *b);
        bl[m].nodes4 = v.nodes4;
        bl[m].tris = v.tris;
        bl[m].tri_rec = v.tri_rec;
        bl[m].root4 = v.root4;
        bl[m].nn4 = v.num_nodes4;
        bl[m].depth4 = v.depth4;
        if (r) { free_blas(); return r; }
        total4 += v.num_nodes4;
        max_depth4 = std::max(max_depth4, v.depth4);
    }
    // 2. top-level tree on the host
    std::vector<TItem> it = inst_items(s);
    std::vector<Bvh4Node> tl;
    std::vector<int32_t> post;
    int tdepth = 0;
    const int32_t tlas_first = (int32_t)total4;
    int32_t root4 = tlas_node(it, 0, s->ninst, tlas_first, tl, post, 0, &tdepth);
    const int64_t all4 = total4 + (int64_t)tl.size();
    if (!nodes4_fit(all4, c->max_wide_nodes)) {
        free_blas();
        return set_err(c, SRT_E_LIMIT, "%lld 4-wide nodes exceed the limit of an instanced scene", (long long)all4);
    }
    // 3. scene arrays
    DevBuf nodes4, tris, tri_rec;
    int r = dev_alloc(c, nodes4, (size_t)std::max<int64_t>(all4, 1) * sizeof(Bvh4Node));
    if (!r) r = dev_alloc(c, tris, (size_t)s->nf * 48);
    if (!r) r = dev_alloc(c, tri_rec, (size_t)s->nf * 8 * sizeof(float4));
    if (r) {
        dev_free(c, nodes4); dev_free(c, tris); dev_free(c, tri_rec);
        free_blas();
        return r;
    }
    std::vector<int32_t> mesh_root(nm);
    int64_t base4 = 0;
    for (int m = 0; m < nm; m++) {
        const uint32_t f0 = (uint32_t)mesh_first[m];
        const int n = mesh_first[m + 1] - mesh_first[m];
        if (bl[m].nn4 > 0)
            k_relocate4<<<nblk(bl[m].nn4), 256, 0, st>>>((const Bvh4Node*)bl[m].nodes4.p, (Bvh4Node*)nodes4.p + base4,
                                                        bl[m].nn4, (int32_t)base4, f0);
        if (n > 0) {
            (void)hipMemcpyAsync((char*)tris.p + 48 * (size_t)f0, bl[m].tris.p, 48 * (size_t)n, hipMemcpyDeviceToDevice, st);
            (void)hipMemcpyAsync((char*)tri_rec.p + 128 * (size_t)f0, bl[m].tri_rec.p, 128 * (size_t)n,
                                 hipMemcpyDeviceToDevice, st);
            k_prim_offset<<<nblk(n), 256, 0, st>>>((float4*)tris.p + 3 * (size_t)f0, n, f0);
        }
        const int32_t br = bl[m].root4;
        if (n == 0) mesh_root[m] = kInstReturn;   // never entered: the instance boxes are empty
        else mesh_root[m] = br >= 0 ? (int32_t)(br + base4) : leaf_ref(leaf_first(br) + f0, leaf_count(br));
        base4 += bl[m].nn4;
    }
    if (!tl.empty())
        r = hip_check(c, hipMemcpyAsync((Bvh4Node*)nodes4.p + tlas_first, tl.data(), tl.size() * sizeof(Bvh4Node),
                                        hipMemcpyHostToDevice, st), "top-level nodes");
    std::vector<int32_t> inst_root(s->ninst);
    for (int i = 0; i < s->ninst; i++) inst_root[i] = mesh_root[s->h_inst_mesh[i]];
    dev_free(c, s->inst_root);
    if (!r) r = dev_alloc(c, s->inst_root, 4 * (size_t)s->ninst);
    if (!r) r = hip_check(c, hipMemcpyAsync(s->inst_root.p, inst_root.data(), 4 * (size_t)s->ninst, hipMemcpyHostToDevice, st),
                          "instance roots");
    if (!r) r = hip_check(c, hipGetLastError(), "instanced build launch");
    if (!r) r = hip_check(c, hipStreamSynchronize(st), "instanced build");
    free_blas();
    if (r) {
        dev_free(c, nodes4); dev_free(c, tris); dev_free(c, tri_rec);
        return r;
    }
    // 4. the scene now holds the two-level tree (4-wide only; no binary tree)
    DevBuf* old[] = {&s->nodes, &s->order, &s->node_parent, &s->level_nodes, &s->node_valid, &s->nodes4, &s->tris,
                     &s->tri_rec};
    for (DevBuf* b : old) dev_free(c, *b);
    s->nodes4 = nodes4;
    s->tris = tris;
    s->tri_rec = tri_rec;
    s->num_nodes4 = (int32_t)all4;
    s->root4 = root4;
    s->tlas_first = tlas_first;
    s->h_tlas = tl;
    s->h_tlas_order = post;
    s->h_mesh_root4 = mesh_root;
    // stack: the top-level levels, the return marker, the deepest mesh tree
    s->depth4 = tdepth + 1 + max_depth4 + 1;
    s->root = -1;
    s->num_nodes = 0;
    s->max_depth = 0;
    s->level_first.clear();
    s->num_clusters = 0;
    s->num_top = (int32_t)tl.size();
    s->cluster_size = cluster_size;
    s->max_leaf = max_leaf;
    s->has_bvh = true;
    return SRT_OK;
}

int inst_update(srt_ctx* c, SceneRec* s, const float* xf) {
    std::vector<float> inv(12 * (size_t)s->ninst);
    for (int i = 0; i < s->ninst; i++) {
        for (int k = 0; k < 12; k++)
            if (!std::isfinite(xf[12 * (size_t)i + k])) return set_err(c, SRT_E_INVALID, "instance %d transform is not finite", i);
        if (!affine_inverse(xf + 12 * (size_t)i, &inv[12 * (size_t)i]))
            return set_err(c, SRT_E_INVALID, "instance %d transform is singular", i);
    }
    s->h_inst_xf.assign(xf, xf + 12 * (size_t)s->ninst);
    s->h_inst_inv = inv;
    // the scene box stays the one of the transforms given to
    // srt_scene_set_instances (scene.h:72: computed once, without animation)
    std::vector<TItem> it = inst_items(s);
    tlas_refit(s, it);
    hipStream_t st = c->stream;
    int r = hip_check(c, hipMemcpyAsync(s->inst_xf.p, xf, 48 * (size_t)s->ninst, hipMemcpyHostToDevice, st), "transforms");
    if (!r) r = hip_check(c, hipMemcpyAsync(s->inst_inv.p, inv.data(), 48 * (size_t)s->ninst, hipMemcpyHostToDevice, st),
                          "transforms");
    if (!r && !s->h_tlas.empty())
        r = hip_check(c, hipMemcpyAsync((Bvh4Node*)s->nodes4.p + s->tlas_first, s->h_tlas.data(),
                                        s->h_tlas.size() * sizeof(Bvh4Node), hipMemcpyHostToDevice, st),
                      "top-level nodes");
    if (!r) r = hip_check(c, hipStreamSynchronize(st), "instance update");
    return r;
}

}  // namespace srt
