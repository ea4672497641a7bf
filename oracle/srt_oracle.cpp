// SRT path-tracing hot path: CPU oracle implementation (test infrastructure).
// Only the test suite, the smoke check and the benchmark's CPU-baseline leg
// load the library built from this file.  See srt_oracle.hpp for the list of
// reference files restated here and for the deviations D1-D4.
#include "srt_oracle.hpp"

#include <atomic>
#include <chrono>
#include <thread>
#include <cstdio>

namespace orc {

// ---------------------------------------------------------------------------
// BVH build
// ---------------------------------------------------------------------------
namespace {

struct KNode {            // Karras node data after the bottom-up pass
    Box box;
    int count = 0;
    int first = 0;
    float cost = 0.0f;
    bool collapse = false;
};

inline int kchild_is_leaf(int32_t c) { return c < 0; }
inline int kchild_leaf(int32_t c) { return -c - 1; }

}  // namespace

void bvh_build(const Scene& s, const BvhParams& prm, Bvh& out) {
    const int n = s.nf;
    out = Bvh();
    if (n <= 0) { out.root = leaf_ref(0, 1); return; }
    // 1. triangle boxes, box centroids, centroid bounds
    std::vector<Box> tb(n);
    std::vector<float> cen(3 * (size_t)n);
    Box cb;
    for (int i = 0; i < n; i++) {
        for (int v = 0; v < 3; v++) tb[i].grow(&s.pos[3 * (size_t)s.idx[3 * (size_t)i + v]]);
        for (int a = 0; a < 3; a++) cen[3 * (size_t)i + a] = (tb[i].lo[a] + tb[i].hi[a]) * 0.5f;
        cb.grow(&cen[3 * (size_t)i]);
    }
    float ext[3];
    for (int a = 0; a < 3; a++) ext[a] = cb.hi[a] - cb.lo[a];
    // 2. 63-bit Morton codes, sorted by (code, primitive id)
    std::vector<uint64_t> code(n);
    for (int i = 0; i < n; i++) {
        uint32_t q[3];
        for (int a = 0; a < 3; a++) q[a] = morton_axis(cen[3 * (size_t)i + a], cb.lo[a], ext[a]);
        code[i] = morton3(q[0], q[1], q[2]);
    }
    out.order.resize(n);
    for (int i = 0; i < n; i++) out.order[i] = (uint32_t)i;
    std::sort(out.order.begin(), out.order.end(), [&](uint32_t a, uint32_t b) {
        return code[a] < code[b] || (code[a] == code[b] && a < b);
    });
    out.codes.resize(n);
    for (int k = 0; k < n; k++) out.codes[k] = code[out.order[k]];
    const std::vector<uint64_t>& sc = out.codes;

    // 3. Karras binary radix tree: internal nodes 0..n-2, children refs
    //    >= 0 internal, < 0 leaf (-k-1 for sorted position k)
    std::vector<int32_t> kl(std::max(n - 1, 0)), kr(std::max(n - 1, 0));
    std::vector<int> kfirst(std::max(n - 1, 0)), klast(std::max(n - 1, 0));
    for (int i = 0; i < n - 1; i++) {
        int d = (kdelta(sc, n, i, i + 1) - kdelta(sc, n, i, i - 1)) >= 0 ? 1 : -1;
        int dmin = kdelta(sc, n, i, i - d);
        int lmax = 2;
        while (kdelta(sc, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (kdelta(sc, n, i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        int dnode = kdelta(sc, n, i, j);
        int sp = 0;
        int t = l;
        do {
            t = (t + 1) / 2;
            if (kdelta(sc, n, i, i + (sp + t) * d) > dnode) sp += t;
        } while (t > 1);
        int gamma = i + sp * d + std::min(d, 0);
        int lo = std::min(i, j), hi = std::max(i, j);
        kl[i] = (lo == gamma) ? -(gamma + 1) : gamma;
        kr[i] = (hi == gamma + 1) ? -(gamma + 1 + 1) : gamma + 1;
        kfirst[i] = lo;
        klast[i] = hi;
    }
    out.karras_left = kl;
    out.karras_right = kr;

    // 4. bottom-up: boxes, counts, SAH costs, collapse decisions
    std::vector<KNode> kn(std::max(n - 1, 0));
    auto leaf_cost = [&](int k) { return prm.c_isect * 1.0f * box_area(tb[out.order[k]]); };
    if (n >= 2) {
        // post-order over internal nodes from root 0
        std::vector<std::pair<int, int>> st;
        st.push_back({0, 0});
        while (!st.empty()) {
            auto& top = st.back();
            int i = top.first;
            if (top.second == 0) {
                top.second = 1;
                if (!kchild_is_leaf(kl[i])) st.push_back({kl[i], 0});
                continue;
            }
            if (top.second == 1) {
                top.second = 2;
                if (!kchild_is_leaf(kr[i])) st.push_back({kr[i], 0});
                continue;
            }
            st.pop_back();
            KNode& N = kn[i];
            Box bl, br; int cl, cr; float costl, costr;
            if (kchild_is_leaf(kl[i])) { int k = kchild_leaf(kl[i]); bl = tb[out.order[k]]; cl = 1; costl = leaf_cost(k); }
            else { bl = kn[kl[i]].box; cl = kn[kl[i]].count; costl = kn[kl[i]].cost; }
            if (kchild_is_leaf(kr[i])) { int k = kchild_leaf(kr[i]); br = tb[out.order[k]]; cr = 1; costr = leaf_cost(k); }
            else { br = kn[kr[i]].box; cr = kn[kr[i]].count; costr = kn[kr[i]].cost; }
            N.box = bl; N.box.grow(br);
            N.count = cl + cr;
            N.first = kfirst[i];
            float sa = box_area(N.box);
            float ci = prm.c_trav * sa + costl + costr;
            float cf = prm.c_isect * (float)N.count * sa;
            N.collapse = (N.count <= prm.max_leaf) && (cf <= ci);
            N.cost = N.collapse ? cf : ci;
        }
    }

    // 5. cluster frontier: highest nodes with count <= T, left to right
    // the library's automatic size; a cluster holds at least a leaf's triangles
    const int T = std::max(prm.max_leaf, prm.cluster_size > 0 ? prm.cluster_size : std::max(4, n / 262144));
    out.cluster_size = T;
    std::vector<int32_t> clusters;     // Karras refs (>=0 internal, <0 leaf)
    if (n == 1) clusters.push_back(-1);
    else if (n <= T) clusters.push_back(0);
    else {
        std::vector<int32_t> st;
        st.push_back(0);
        while (!st.empty()) {
            int32_t c = st.back(); st.pop_back();
            int cnt = kchild_is_leaf(c) ? 1 : kn[c].count;
            if (cnt <= T) { clusters.push_back(c); continue; }
            st.push_back(kr[c]);   // right pushed first so left is visited first
            st.push_back(kl[c]);
        }
    }
    const int m = (int)clusters.size();
    out.num_clusters = m;
    const int top_nodes = m - 1;
    out.num_top_nodes = top_nodes;
    const int total_nodes = top_nodes + std::max(n - 1, 0);
    out.nodes.assign(std::max(total_nodes, 1), BvhNode());
    std::memset(out.nodes.data(), 0, sizeof(BvhNode) * out.nodes.size());

    auto final_ref_of = [&](int32_t c) -> int32_t {
        if (kchild_is_leaf(c)) return leaf_ref((uint32_t)kchild_leaf(c), 1u);
        if (kn[c].collapse) return leaf_ref((uint32_t)kn[c].first, (uint32_t)kn[c].count);
        return top_nodes + c;
    };
    auto box_of = [&](int32_t c) -> Box { return kchild_is_leaf(c) ? tb[out.order[kchild_leaf(c)]] : kn[c].box; };
    auto count_of = [&](int32_t c) -> int { return kchild_is_leaf(c) ? 1 : kn[c].count; };
    auto set_child = [&](BvhNode& nd, int side, const Box& b, int32_t ref) {
        if (side == 0) {
            nd.lx[0] = b.lo[0]; nd.lx[1] = b.hi[0]; nd.ly[0] = b.lo[1]; nd.ly[1] = b.hi[1];
            nd.lz[0] = b.lo[2]; nd.lz[1] = b.hi[2]; nd.left = ref;
        } else {
            nd.rx[0] = b.lo[0]; nd.rx[1] = b.hi[0]; nd.ry[0] = b.lo[1]; nd.ry[1] = b.hi[1];
            nd.rz[0] = b.lo[2]; nd.rz[1] = b.hi[2]; nd.right = ref;
        }
    };

    // 6. top-level binned SAH over the clusters (breadth-first node numbering)
    if (m > 1) {
        struct Item { Box box; float c[3]; int count; int32_t ref; };
        std::vector<Item> items(m);
        for (int q = 0; q < m; q++) {
            items[q].box = box_of(clusters[q]);
            for (int a = 0; a < 3; a++) items[q].c[a] = (items[q].box.lo[a] + items[q].box.hi[a]) * 0.5f;
            items[q].count = count_of(clusters[q]);
            items[q].ref = final_ref_of(clusters[q]);
        }
        const int B = prm.sah_bins;
        struct Task { int b, e, node; };
        std::vector<Task> queue;
        queue.push_back({0, m, 0});
        int next = 1;
        std::vector<Item> tmp(m);
        for (size_t qi = 0; qi < queue.size(); qi++) {
            Task tk = queue[qi];
            Box cbx;
            for (int i = tk.b; i < tk.e; i++) cbx.grow(items[i].c);
            float best = INFINITY; int best_axis = -1, best_split = -1;
            for (int a = 0; a < 3; a++) {
                float ex = cbx.hi[a] - cbx.lo[a];
                if (!(ex > 0.0f)) continue;
                float scale = (float)B / ex;
                std::vector<Box> bb(B); std::vector<int> bc(B, 0);
                for (int i = tk.b; i < tk.e; i++) {
                    int bi = (int)((items[i].c[a] - cbx.lo[a]) * scale);
                    if (bi > B - 1) bi = B - 1;
                    if (bi < 0) bi = 0;
                    bb[bi].grow(items[i].box);
                    bc[bi] += items[i].count;
                }
                std::vector<float> la(B), ra(B); std::vector<int> lc(B), rc(B);
                Box acc; int cnt = 0;
                for (int k = 0; k < B; k++) { acc.grow(bb[k]); cnt += bc[k]; la[k] = box_area(acc); lc[k] = cnt; }
                acc = Box(); cnt = 0;
                for (int k = B - 1; k >= 0; k--) { acc.grow(bb[k]); cnt += bc[k]; ra[k] = box_area(acc); rc[k] = cnt; }
                for (int k = 0; k < B - 1; k++) {
                    if (lc[k] == 0 || rc[k + 1] == 0) continue;
                    float cost = la[k] * (float)lc[k] + ra[k + 1] * (float)rc[k + 1];
                    if (cost < best) { best = cost; best_axis = a; best_split = k; }
                }
            }
            int mid;
            if (best_axis < 0) {
                mid = (tk.b + tk.e) / 2;
            } else {
                float ex = cbx.hi[best_axis] - cbx.lo[best_axis];
                float scale = (float)B / ex;
                int w = tk.b;
                for (int pass = 0; pass < 2; pass++)
                    for (int i = tk.b; i < tk.e; i++) {
                        int bi = (int)((items[i].c[best_axis] - cbx.lo[best_axis]) * scale);
                        if (bi > B - 1) bi = B - 1;
                        if (bi < 0) bi = 0;
                        bool left = bi <= best_split;
                        if ((pass == 0) == left) tmp[w++] = items[i];
                    }
                int nleft = 0;
                for (int i = tk.b; i < tk.e; i++) {
                    int bi = (int)((items[i].c[best_axis] - cbx.lo[best_axis]) * scale);
                    if (bi > B - 1) bi = B - 1;
                    if (bi < 0) bi = 0;
                    if (bi <= best_split) nleft++;
                }
                for (int i = tk.b; i < tk.e; i++) items[i] = tmp[i];
                mid = tk.b + nleft;
            }
            BvhNode& nd = out.nodes[tk.node];
            int ranges[2][2] = {{tk.b, mid}, {mid, tk.e}};
            for (int side = 0; side < 2; side++) {
                int b = ranges[side][0], e = ranges[side][1];
                Box bx;
                for (int i = b; i < e; i++) bx.grow(items[i].box);
                int32_t ref;
                if (e - b == 1) ref = items[b].ref;
                else { ref = next++; queue.push_back({b, e, ref}); }
                set_child(nd, side, bx, ref);
            }
        }
        out.root = 0;
    } else {
        out.root = final_ref_of(clusters[0]);
    }

    // 7. Karras nodes that stay in the final tree: inside a cluster, not
    //    collapsed, and no collapsed ancestor (only ancestors with
    //    count <= max_leaf can be collapsed, so the walk up is short)
    {
        std::vector<int> parent(std::max(n - 1, 0), -1);
        for (int i = 0; i < n - 1; i++) {
            if (!kchild_is_leaf(kl[i])) parent[kl[i]] = i;
            if (!kchild_is_leaf(kr[i])) parent[kr[i]] = i;
        }
        for (int i = 0; i < n - 1; i++) {
            if (kn[i].count > T || kn[i].collapse) continue;
            bool hidden = false;
            for (int p = parent[i]; p >= 0 && kn[p].count <= prm.max_leaf; p = parent[p])
                if (kn[p].collapse) { hidden = true; break; }
            if (hidden) continue;
            BvhNode& nd = out.nodes[top_nodes + i];
            set_child(nd, 0, box_of(kl[i]), final_ref_of(kl[i]));
            set_child(nd, 1, box_of(kr[i]), final_ref_of(kr[i]));
        }
    }

    // 8. triangles in sorted order
    out.tris.assign(12 * (size_t)n, 0.0f);
    for (int k = 0; k < n; k++) {
        uint32_t p = out.order[k];
        for (int v = 0; v < 3; v++)
            for (int a = 0; a < 3; a++) out.tris[12 * (size_t)k + 4 * v + a] = s.pos[3 * (size_t)s.idx[3 * (size_t)p + v] + a];
        out.tris[12 * (size_t)k + 3] = u2f(p);
        out.tris[12 * (size_t)k + 7] = u2f((uint32_t)s.tri_mesh[p]);
        out.tris[12 * (size_t)k + 11] = u2f((uint32_t)s.mesh_mat[s.tri_mesh[p]]);   // material of the triangle
    }

    // 9. depth (internal nodes on the longest root-to-leaf path)
    int maxd = 0;
    if (out.root >= 0) {
        std::vector<std::pair<int32_t, int>> st;
        st.push_back({out.root, 1});
        while (!st.empty()) {
            auto e = st.back(); st.pop_back();
            maxd = std::max(maxd, e.second);
            const BvhNode& nd = out.nodes[e.first];
            if (nd.left >= 0) st.push_back({nd.left, e.second + 1});
            if (nd.right >= 0) st.push_back({nd.right, e.second + 1});
        }
    }
    out.max_depth = maxd;
}

// Refit: new triangle positions in sorted order, then child boxes bottom-up
// over the unchanged topology.
void bvh_refit(const Scene& s, Bvh& bvh) {
    const int n = s.nf;
    for (int k = 0; k < n; k++) {
        uint32_t p = bvh.order[k];
        for (int v = 0; v < 3; v++)
            for (int a = 0; a < 3; a++) bvh.tris[12 * (size_t)k + 4 * v + a] = s.pos[3 * (size_t)s.idx[3 * (size_t)p + v] + a];
    }
    auto leaf_box = [&](int32_t ref) {
        Box b;
        uint32_t f = leaf_first(ref), c = leaf_count(ref);
        for (uint32_t k = f; k < f + c; k++)
            for (int v = 0; v < 3; v++) b.grow(&bvh.tris[12 * (size_t)k + 4 * v]);
        return b;
    };
    if (bvh.root < 0) return;
    // post-order over final nodes
    std::vector<Box> nb(bvh.nodes.size());
    std::vector<std::pair<int32_t, int>> st;
    st.push_back({bvh.root, 0});
    while (!st.empty()) {
        auto& top = st.back();
        int32_t i = top.first;
        BvhNode& nd = bvh.nodes[i];
        if (top.second == 0) { top.second = 1; if (nd.left >= 0) { st.push_back({nd.left, 0}); } continue; }
        if (top.second == 1) { top.second = 2; if (nd.right >= 0) { st.push_back({nd.right, 0}); } continue; }
        st.pop_back();
        Box bl = nd.left >= 0 ? nb[nd.left] : leaf_box(nd.left);
        Box br = nd.right >= 0 ? nb[nd.right] : leaf_box(nd.right);
        nd.lx[0] = bl.lo[0]; nd.lx[1] = bl.hi[0]; nd.ly[0] = bl.lo[1]; nd.ly[1] = bl.hi[1]; nd.lz[0] = bl.lo[2]; nd.lz[1] = bl.hi[2];
        nd.rx[0] = br.lo[0]; nd.rx[1] = br.hi[0]; nd.ry[0] = br.lo[1]; nd.ry[1] = br.hi[1]; nd.rz[0] = br.lo[2]; nd.rz[1] = br.hi[2];
        Box u = bl; u.grow(br);
        nb[i] = u;
    }
}

// One BVH per mesh of an instanced scene, over the mesh's object-space
// triangles; sorted triangles and the sorted order carry scene primitive ids.
void blas_build(const Scene& s, const BvhParams& prm, std::vector<Bvh>& out) {
    out.assign(s.nmesh, Bvh());
    for (int m = 0; m < s.nmesh; m++) {
        const int f0 = s.mesh_first[m], n = s.mesh_first[m + 1] - f0;
        Scene sub;
        sub.nv = s.nv;
        sub.pos = s.pos;
        sub.nf = n;
        sub.idx.assign(s.idx.begin() + 3 * (size_t)f0, s.idx.begin() + 3 * (size_t)(f0 + n));
        sub.nmesh = 1;
        sub.mesh_first = {0, n};
        sub.mesh_mat = {s.mesh_mat[m]};
        sub.mats = s.mats;
        sub.tri_mesh.assign(n, 0);
        bvh_build(sub, prm, out[m]);
        for (auto& o : out[m].order) o += (uint32_t)f0;
        for (int k = 0; k < n; k++) {
            float* tr = &out[m].tris[12 * (size_t)k];
            tr[3] = u2f(f2u(tr[3]) + (uint32_t)f0);
            tr[7] = u2f((uint32_t)m);
        }
    }
}

// ---------------------------------------------------------------------------
// Traversal
// ---------------------------------------------------------------------------
namespace {
const float kRobust = 1.0000004f;   // 1 + 2 gamma_3 (Ize 2013, "Robust BVH Ray Traversal")
// Closest-hit traversals cull boxes at the best distance so far times
// 1 + 2^-12, as the kernels do: a triangle's computed t can lie below its own
// box's computed entry distance by more than the slab rounding, and the margin
// keeps every triangle with t <= best distance in the test, so the
// (t, primitive) minimum does not depend on the visiting order (coplanar
// faces).  Any-hit traversals cull at tmax (their answer is order-independent).
const float kCullTol = 1.000244140625f;

// Slab test with the near/far plane of each axis chosen by the sign of the
// inverse direction (Williams et al. 2005).  A ray lying exactly in a slab
// plane with a zero direction component gives 0 * inf = NaN, which fmax/fmin
// ignore, so that axis adds no constraint (conservative).
inline bool slab(const float lo[3], const float hi[3], const float o[3], const float inv[3], float tmin, float tmax,
                 float* tnear, bool closest = true) {
    float tn = tmin, tf = INFINITY;
    for (int a = 0; a < 3; a++) {
        float nb = (inv[a] < 0.0f) ? hi[a] : lo[a];
        float fb = (inv[a] < 0.0f) ? lo[a] : hi[a];
        tn = std::fmax(tn, (nb - o[a]) * inv[a]);
        tf = std::fmin(tf, (fb - o[a]) * inv[a]);
    }
    tf = std::fmin(tf * kRobust, tmax * (closest ? kCullTol : kRobust));
    *tnear = tn;
    return tn <= tf;
}
}  // namespace

template <bool ANY>
static bool trace_impl(const Bvh& bvh, const float o[3], const float d[3], float tmin, float tmax, Hit* hit, int* nv, int* nt) {
    RayPre rp = ray_pre(o, d);
    float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    float best_t = tmax; int32_t best_prim = 0x7fffffff; int32_t best_pos = -1; float bu = 0, bv = 0;
    int32_t stack[256]; int sp = 0;
    stack[sp++] = bvh.root;
    int visits = 0, tests = 0;
    while (sp > 0) {
        int32_t ref = stack[--sp];
        if (ref < 0) {
            uint32_t f = leaf_first(ref), c = leaf_count(ref);
            for (uint32_t k = f; k < f + c; k++) {
                const float* tr = &bvh.tris[12 * (size_t)k];
                tests++;
                float t, u, v;
                if (!intersect_tri(rp, tr, tr + 4, tr + 8, tmin, best_t, &t, &u, &v)) continue;
                int32_t prim = (int32_t)f2u(tr[3]);
                if (ANY) { if (nv) *nv = visits; if (nt) *nt = tests; return true; }
                if (t < best_t || (t == best_t && prim < best_prim)) {
                    best_t = t; best_prim = prim; best_pos = (int32_t)k; bu = u; bv = v;
                }
            }
            continue;
        }
        visits++;
        const BvhNode& nd = bvh.nodes[ref];
        float llo[3] = {nd.lx[0], nd.ly[0], nd.lz[0]}, lhi[3] = {nd.lx[1], nd.ly[1], nd.lz[1]};
        float rlo[3] = {nd.rx[0], nd.ry[0], nd.rz[0]}, rhi[3] = {nd.rx[1], nd.ry[1], nd.rz[1]};
        float tl, tr_;
        bool hl = slab(llo, lhi, o, inv, tmin, best_t, &tl, !ANY);
        bool hr = slab(rlo, rhi, o, inv, tmin, best_t, &tr_, !ANY);
        if (hl && hr) {
            if (tr_ < tl) { stack[sp++] = nd.left; stack[sp++] = nd.right; }
            else { stack[sp++] = nd.right; stack[sp++] = nd.left; }
        } else if (hl) stack[sp++] = nd.left;
        else if (hr) stack[sp++] = nd.right;
    }
    if (nv) *nv = visits;
    if (nt) *nt = tests;
    if (ANY) return false;
    if (best_pos < 0) return false;
    hit->t = best_t; hit->u = bu; hit->v = bv; hit->prim = best_prim; hit->pos = best_pos;
    return true;
}

bool trace_closest(const Bvh& bvh, const float o[3], const float d[3], float tmin, float tmax, Hit* hit, int* nv, int* nt) {
    return trace_impl<false>(bvh, o, d, tmin, tmax, hit, nv, nt);
}
bool trace_any(const Bvh& bvh, const float o[3], const float d[3], float tmin, float tmax) {
    return trace_impl<true>(bvh, o, d, tmin, tmax, nullptr, nullptr, nullptr);
}

bool brute_closest(const Scene& s, const float o[3], const float d[3], float tmin, float tmax, Hit* hit) {
    RayPre rp = ray_pre(o, d);
    float best_t = tmax; int32_t best = -1; float bu = 0, bv = 0;
    for (int f = 0; f < s.nf; f++) {
        float t, u, v;
        const float* v0 = &s.pos[3 * (size_t)s.idx[3 * (size_t)f]];
        const float* v1 = &s.pos[3 * (size_t)s.idx[3 * (size_t)f + 1]];
        const float* v2 = &s.pos[3 * (size_t)s.idx[3 * (size_t)f + 2]];
        if (!intersect_tri(rp, v0, v1, v2, tmin, best_t, &t, &u, &v)) continue;
        if (best < 0 || t < best_t || (t == best_t && f < best)) { best_t = t; best = f; bu = u; bv = v; }
    }
    if (best < 0) return false;
    hit->t = best_t; hit->u = bu; hit->v = bv; hit->prim = best; hit->pos = -1;
    return true;
}
bool brute_any(const Scene& s, const float o[3], const float d[3], float tmin, float tmax) {
    RayPre rp = ray_pre(o, d);
    for (int f = 0; f < s.nf; f++) {
        float t, u, v;
        if (intersect_tri(rp, &s.pos[3 * (size_t)s.idx[3 * (size_t)f]], &s.pos[3 * (size_t)s.idx[3 * (size_t)f + 1]],
                          &s.pos[3 * (size_t)s.idx[3 * (size_t)f + 2]], tmin, tmax, &t, &u, &v))
            return true;
    }
    return false;
}

// Instanced scenes: every instance's ray in object space (inverse transform,
// apply_point / apply_vector in float), the instance's mesh BVH (or all its
// triangles), closest hit = lexicographic minimum of (t, primitive id,
// instance) over all instances, so no top-level tree is needed here.
static bool trace_inst(const Scene& s, const std::vector<Bvh>* blas, const float o[3], const float d[3], float tmin,
                       float tmax, bool any, Hit* hit) {
    bool found = false;
    Hit best;
    best.t = tmax;
    for (int i = 0; i < s.ninst(); i++) {
        const float* Mi = &s.inst_inv[12 * (size_t)i];
        float oo[3], dd[3];
        xf_point(Mi, o, oo);
        xf_vector(Mi, d, dd);
        const int m = s.inst_mesh[i];
        const float tm = found ? best.t : tmax;
        Hit h;
        bool ok = false;
        if (blas) {
            ok = any ? trace_any((*blas)[m], oo, dd, tmin, tmax) : trace_closest((*blas)[m], oo, dd, tmin, tm, &h);
        } else {
            RayPre rp = ray_pre(oo, dd);
            float bt = tm;
            int32_t bp = -1;
            float bu = 0, bv = 0;
            for (int f = s.mesh_first[m]; f < s.mesh_first[m + 1]; f++) {
                float t, u, v;
                if (!intersect_tri(rp, &s.pos[3 * (size_t)s.idx[3 * (size_t)f]], &s.pos[3 * (size_t)s.idx[3 * (size_t)f + 1]],
                                   &s.pos[3 * (size_t)s.idx[3 * (size_t)f + 2]], tmin, any ? tmax : bt, &t, &u, &v))
                    continue;
                if (any) { ok = true; break; }
                if (bp < 0 || t < bt || (t == bt && f < bp)) { bt = t; bp = f; bu = u; bv = v; }
            }
            if (!any && bp >= 0) { ok = true; h.t = bt; h.prim = bp; h.u = bu; h.v = bv; h.pos = -1; }
        }
        if (any) {
            if (ok) return true;
            continue;
        }
        if (ok && (!found || h.t < best.t || (h.t == best.t && (h.prim < best.prim || (h.prim == best.prim && i < best.inst))))) {
            best = h;
            best.inst = i;
            found = true;
        }
    }
    if (any || !found) return false;
    *hit = best;
    return true;
}

// ---------------------------------------------------------------------------
// Integrator (path_tracer.cu:50-169)
// ---------------------------------------------------------------------------
template <class R> struct HitInfo {
    V3<R> pos, normal, color;
    Onb<R> onb;
    int mat;
    int mesh;
};

// hit_info.h:27-65: object-space reconstruction, then the instance transform
// (position by apply_point, geometric normal by apply_vector without
// renormalising, the shading frame axis by axis, transform.h:37-52, 72-75);
// inst < 0: world-space mesh (identity instance)
template <class R>
static void get_hitinfo(const Scene& s, int prim, R u, R v, HitInfo<R>& info, int inst = -1) {
    const uint32_t* ix = &s.idx[3 * (size_t)prim];
    V3<R> v0 = cvt<R>(&s.pos[3 * (size_t)ix[0]]), v1 = cvt<R>(&s.pos[3 * (size_t)ix[1]]), v2 = cvt<R>(&s.pos[3 * (size_t)ix[2]]);
    R w = R(1) - u - v;
    info.pos = v0 * w + v1 * u + v2 * v;
    V3<R> normal = normalize(cross(v1 - v0, v2 - v0));
    V3<R> sn = normal;
    if (!s.nrm.empty()) {
        V3<R> n0 = cvt<R>(&s.nrm[3 * (size_t)ix[0]]), n1 = cvt<R>(&s.nrm[3 * (size_t)ix[1]]), n2 = cvt<R>(&s.nrm[3 * (size_t)ix[2]]);
        sn = normalize(n0 * w + n1 * u + n2 * v);
    }
    if (dot(normal, sn) < R(0)) normal = -normal;
    V3<R> tg;
    if (!s.tan.empty()) {
        V3<R> t0 = cvt<R>(&s.tan[3 * (size_t)ix[0]]), t1 = cvt<R>(&s.tan[3 * (size_t)ix[1]]), t2 = cvt<R>(&s.tan[3 * (size_t)ix[2]]);
        tg = t0 * w + t1 * u + t2 * v;
    } else {
        tg = (M<R>::abs(sn.x) > M<R>::abs(sn.y)) ? V3<R>(R(0), R(1), R(0)) : V3<R>(R(1), R(0), R(0));
    }
    info.normal = normal;
    info.mesh = s.tri_mesh[prim];
    info.mat = s.mesh_mat[info.mesh];
    const MaterialRec& m = s.mats[info.mat];
    info.color = V3<R>((R)m.base_color[0], (R)m.base_color[1], (R)m.base_color[2]);
    info.onb = Onb<R>(sn, tg);
    if (inst >= 0) {
        const float* T = &s.inst_xf[12 * (size_t)inst];
        info.pos = xf_point<R>(T, info.pos);
        info.normal = xf_vector<R>(T, info.normal);
        info.onb = Onb<R>(xf_vector<R>(T, info.onb.z), xf_vector<R>(T, info.onb.x), xf_vector<R>(T, info.onb.y));
    }
}

template <class R> struct LightSample { V3<R> pos, normal, emission; R pdf; };

// light.h:61-76 and gmesh.h:45-80 (float lower bound, D2)
template <class R>
static LightSample<R> light_sample(const Scene& s, V2<R> rnd) {
    const int nl = s.nlights();
    std::vector<float> const& lc = s.light_cdf;
    int idx;
    R lower;
    {
        // binary search on float cdf values
        int lf = 0, rg = nl - 1;
        while (lf <= rg) { int mid = (lf + rg) / 2; if ((R)lc[mid] <= rnd.x) lf = mid + 1; else rg = mid - 1; }
        idx = std::min(rg + 1, nl - 1);
    }
    lower = (idx == 0) ? R(0) : (R)lc[idx - 1];
    rnd.x = (rnd.x - lower) / ((R)lc[idx] - lower);
    int mesh = s.light_mesh[idx];
    int f0 = s.mesh_first[mesh], fn = s.mesh_first[mesh + 1] - f0;
    const float* ac = &s.area_cdf[f0];
    int f;
    {
        int lf = 0, rg = fn - 1;
        while (lf <= rg) { int mid = (lf + rg) / 2; if ((R)ac[mid] <= rnd.x) lf = mid + 1; else rg = mid - 1; }
        f = std::min(rg + 1, fn - 1);
    }
    R lo2 = (f == 0) ? R(0) : (R)ac[f - 1];
    rnd.x = (rnd.x - lo2) / ((R)ac[f] - lo2);
    V2<R> uv = uniform_sample_triangle(rnd);
    int prim = f0 + f;
    const uint32_t* ix = &s.idx[3 * (size_t)prim];
    V3<R> v0 = cvt<R>(&s.pos[3 * (size_t)ix[0]]), v1 = cvt<R>(&s.pos[3 * (size_t)ix[1]]), v2 = cvt<R>(&s.pos[3 * (size_t)ix[2]]);
    R w = R(1) - uv.x - uv.y;
    LightSample<R> ls;
    ls.pos = v0 * w + v1 * uv.x + v2 * uv.y;
    V3<R> normal = normalize(cross(v1 - v0, v2 - v0));
    if (!s.nrm.empty()) {
        V3<R> n0 = cvt<R>(&s.nrm[3 * (size_t)ix[0]]), n1 = cvt<R>(&s.nrm[3 * (size_t)ix[1]]), n2 = cvt<R>(&s.nrm[3 * (size_t)ix[2]]);
        V3<R> sn = n0 * w + n1 * uv.x + n2 * uv.y;
        if (dot(normal, sn) < R(0)) normal = -normal;
    }
    ls.normal = normal;
    if (s.instanced()) {
        // GInstance::sample (gmesh.h:77-78): apply_point / apply_vector of the instance transform
        const float* T = &s.inst_xf[12 * (size_t)s.light_inst[idx]];
        ls.pos = xf_point<R>(T, ls.pos);
        ls.normal = xf_vector<R>(T, ls.normal);
    }
    const MaterialRec& m = s.mats[s.mesh_mat[mesh]];
    ls.emission = V3<R>((R)m.emission_color[0], (R)m.emission_color[1], (R)m.emission_color[2]) * (R)m.intensity;
    ls.pdf = (R)m.intensity / (R)s.light_weight_sum;
    return ls;
}

template <class R> static inline bool valid(R x) { return !std::isnan(x) && !std::isinf(x); }

struct Tracer {
    const Scene* s;
    const Bvh* bvh;    // null: brute force
    const std::vector<Bvh>* blas = nullptr;   // instanced scenes: per-mesh trees (null: brute force)
    bool closest(const float o[3], const float d[3], float tmin, float tmax, Hit* h) const {
        if (s->instanced()) return trace_inst(*s, blas, o, d, tmin, tmax, false, h);
        return bvh ? trace_closest(*bvh, o, d, tmin, tmax, h) : brute_closest(*s, o, d, tmin, tmax, h);
    }
    bool any(const float o[3], const float d[3], float tmin, float tmax) const {
        if (s->instanced()) return trace_inst(*s, blas, o, d, tmin, tmax, true, nullptr);
        return bvh ? trace_any(*bvh, o, d, tmin, tmax) : brute_any(*s, o, d, tmin, tmax);
    }
};

// ntrace / nshadow count the closest-hit traces and the shadow rays issued
// (the wavefront integrator's queue entries for the same paths).
template <class R>
static V3<R> render_pixel(const Scene& s, const Tracer& tr, const Camera& cam, const PtParams& p, int W, int H, int pixel,
                          int sample_begin, int sample_end, int64_t* ntrace, int64_t* nshadow) {
    const int ix = pixel % W, iy = pixel / W;
    V3<R> result(R(0));
    V3<R> X = cvt<R>(cam.X), Y = cvt<R>(cam.Y), Z = cvt<R>(cam.Z), CP = cvt<R>(cam.pos);
    V3<R> env((R)s.env[0], (R)s.env[1], (R)s.env[2]);
    // a scene without area lights has no light CDF to sample: NEE is off
    // (the reference would index an empty CDF in light.h:63)
    const bool use_nee = p.use_nee && s.nlights() > 0;
    for (int smp = sample_begin; smp < sample_end; smp++) {
        Pcg32 rng;
        path_rng(rng, (uint32_t)pixel, p.seed, (uint32_t)smp);
        // path_tracer.cu:66-68, camera.h:107-117, camera.h:27-32
        R r0 = (R)rng.next_float();
        R xx = ((R)ix + r0) / (R)W;
        R r1 = (R)rng.next_float();
        R yy = ((R)iy + r1) / (R)H;
        R px = (xx - R(0.5)) * (R)cam.nw, py = (yy - R(0.5)) * (R)cam.nh;
        V3<R> dir = normalize(V3<R>(px, py, R(1)));
        V3<R> zero(R(0));
        V3<R> ro = zero.x * X + zero.y * Y + zero.z * Z + CP;
        V3<R> rd = dir.x * X + dir.y * Y + dir.z * Z;

        bool specular = true;
        R scatter_pdf = R(1);
        V3<R> L(R(0)), beta(R(1));
        for (int depth = 0; depth < p.max_depth; depth++) {
            float of[3] = {(float)ro.x, (float)ro.y, (float)ro.z}, df[3] = {(float)rd.x, (float)rd.y, (float)rd.z};
            Hit h;
            (*ntrace)++;
            bool hit = tr.closest(of, df, 1e-3f, 1e16f, &h);
            if (!hit) { L = L + beta * env; break; }
            HitInfo<R> info;
            get_hitinfo<R>(s, h.prim, (R)h.u, (R)h.v, info, h.inst);
            const MaterialRec& m = s.mats[info.mat];
            BxDF<R> bx(m);
            if (mat_is_emissive(m)) {
                R cos_i = dot(info.normal, -rd);
                if (cos_i <= (R)1e-4f) break;
                R mis = R(1);
                if (use_nee && !specular) {
                    mis = R(0);
                    if (p.use_mis) {
                        R t2 = dot(info.pos - ro, info.pos - ro);
                        R lpdf = ((R)m.intensity / (R)s.light_weight_sum) * t2 / cos_i;
                        mis = scatter_pdf / (lpdf + scatter_pdf);
                    }
                }
                V3<R> Le = V3<R>((R)m.emission_color[0], (R)m.emission_color[1], (R)m.emission_color[2]) * (R)m.intensity;
                L = L + beta * Le * mis;
                break;
            }
            R cos_i = dot(info.normal, -rd);
            if (cos_i < R(0)) info.normal = -info.normal;
            if (use_nee && !mat_is_specular(m)) {
                V2<R> lr;
                lr.x = (R)rng.next_float();
                lr.y = (R)rng.next_float();
                LightSample<R> ls = light_sample<R>(s, lr);
                V3<R> sdir = normalize(ls.pos - info.pos);
                R cos_o = dot(info.normal, sdir);
                R cos_light = dot(ls.normal, -sdir);
                R t = length(ls.pos - info.pos);
                if ((cos_light > (R)1e-4f) && (ls.pdf > (R)1e-4f) && (t > (R)1e-3f) && (cos_o > R(0) || mat_is_transmissive(m))) {
                    float so[3] = {(float)info.pos.x, (float)info.pos.y, (float)info.pos.z};
                    float sd[3] = {(float)sdir.x, (float)sdir.y, (float)sdir.z};
                    float tmaxs = (float)(t - (R)1e-3f);
                    (*nshadow)++;
                    bool visible = !tr.any(so, sd, 1e-3f, tmaxs);
                    if (visible) {
                        R mis = R(1);
                        R lpdf = ls.pdf * t * t / cos_light;
                        if (p.use_mis) {
                            R mpdf = mat_pdf(bx, sdir, -rd, info.onb, info.color);
                            mis = lpdf / (lpdf + mpdf);
                        }
                        V3<R> f = mat_eval(bx, sdir, -rd, info.onb, info.color);
                        L = L + beta * f * M<R>::abs(cos_o) * ls.emission * mis / lpdf;
                    }
                }
            }
            V2<R> br;
            br.x = (R)rng.next_float();
            br.y = (R)rng.next_float();
            BxDFSample<R> ms = mat_sample(bx, -rd, br, info.onb, info.color);
            if (ms.pdf <= (R)1e-4f) break;
            if (dot(ms.wi, info.normal) < R(0) && !mat_is_transmissive(m)) break;
            beta = beta * (ms.f / ms.pdf);
            if (!ms.delta) beta = beta * M<R>::abs(dot(ms.wi, info.normal));
            specular = mat_is_specular(m);
            ro = info.pos;
            rd = ms.wi;
            scatter_pdf = ms.pdf;
            if (depth >= p.rr_depth) {
                R q = M<R>::fmax(M<R>::fmax(beta.x, M<R>::fmax(beta.y, beta.z)), (R)0.05f);
                if ((R)rng.next_float() > q) break;
                beta = beta / q;
            }
        }
        if (valid(L.x) && valid(L.y) && valid(L.z)) result = result + L / (R)p.spp;
    }
    return result;
}

}  // namespace orc

// ===========================================================================
// C entry points (loaded by the tests through ctypes)
// ===========================================================================
using namespace orc;

struct OrcScene { Scene s; };
static int64_t g_last_counts[2] = {0, 0};
struct OrcBvh { Bvh b; std::vector<Bvh> blas; };   // instanced scenes: blas, one per mesh

extern "C" {

void orc_pcg32(uint64_t initstate, uint64_t initseq, int n, uint32_t* out) {
    Pcg32 r; r.seed(initstate, initseq);
    for (int i = 0; i < n; i++) out[i] = r.next();
}
uint32_t orc_seed_hash(uint32_t pixel, uint32_t seed) { return seed_hash(pixel, seed); }
void orc_path_floats(uint32_t pixel, uint32_t seed, uint32_t sample, int n, float* out) {
    Pcg32 r; path_rng(r, pixel, seed, sample);
    for (int i = 0; i < n; i++) out[i] = r.next_float();
}

// fn: 0 sin 1 cos 2 log 3 exp 4 pow
void orc_mathf(int fn, int n, const float* x, const float* y, float* out) {
    for (int i = 0; i < n; i++) {
        switch (fn) {
        case 0: out[i] = pt_sinf(x[i]); break;
        case 1: out[i] = pt_cosf(x[i]); break;
        case 2: out[i] = pt_logf(x[i]); break;
        case 3: out[i] = pt_expf(x[i]); break;
        case 4: out[i] = pt_powf(x[i], y[i]); break;
        default: out[i] = 0.0f;
        }
    }
}

// BxDF probes.  op 0: eval(wi, wo) -> f[3]; op 1: pdf(wi, wo) -> pdf; op 2:
// sample(wo, rnd) -> f[3], wi[3], pdf, delta.  Directions in the local frame.
// out stride 8.
void orc_bxdf(int fp64, const MaterialRec* m, int op, int inner, int n, const double* wi, const double* wo,
              const double* color, const double* rnd, double* out) {
    for (int i = 0; i < n; i++) {
        double* o = out + 8 * (size_t)i;
        if (fp64) {
            BxDF<double> b(*m);
            V3<double> c(color[0], color[1], color[2]);
            V3<double> vo(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
            if (op == 0 || op == 1) {
                V3<double> vi(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
                if (op == 0) { V3<double> f = b.eval(vi, vo, c, inner != 0); o[0] = f.x; o[1] = f.y; o[2] = f.z; }
                else o[0] = b.sample_pdf(vi, vo, c, inner != 0);
            } else {
                BxDFSample<double> sm = b.sample(vo, V2<double>{rnd[2 * i], rnd[2 * i + 1]}, c, inner != 0);
                o[0] = sm.f.x; o[1] = sm.f.y; o[2] = sm.f.z; o[3] = sm.wi.x; o[4] = sm.wi.y; o[5] = sm.wi.z; o[6] = sm.pdf; o[7] = sm.delta;
            }
        } else {
            BxDF<float> b(*m);
            V3<float> c((float)color[0], (float)color[1], (float)color[2]);
            V3<float> vo((float)wo[3 * i], (float)wo[3 * i + 1], (float)wo[3 * i + 2]);
            if (op == 0 || op == 1) {
                V3<float> vi((float)wi[3 * i], (float)wi[3 * i + 1], (float)wi[3 * i + 2]);
                if (op == 0) { V3<float> f = b.eval(vi, vo, c, inner != 0); o[0] = f.x; o[1] = f.y; o[2] = f.z; }
                else o[0] = b.sample_pdf(vi, vo, c, inner != 0);
            } else {
                BxDFSample<float> sm = b.sample(vo, V2<float>{(float)rnd[2 * i], (float)rnd[2 * i + 1]}, c, inner != 0);
                o[0] = sm.f.x; o[1] = sm.f.y; o[2] = sm.f.z; o[3] = sm.wi.x; o[4] = sm.wi.y; o[5] = sm.wi.z; o[6] = sm.pdf; o[7] = sm.delta;
            }
        }
    }
}

void* orc_scene_new(int nv, const float* pos, const float* nrm, const float* tan, int nf, const uint32_t* idx, int nmesh,
                    const int32_t* mesh_first, const int32_t* mesh_mat, int nmat, const void* mats, const float* env) {
    OrcScene* o = new OrcScene();
    Scene& s = o->s;
    s.nv = nv; s.nf = nf; s.nmesh = nmesh; s.nmat = nmat;
    s.pos.assign(pos, pos + 3 * (size_t)nv);
    if (nrm) s.nrm.assign(nrm, nrm + 3 * (size_t)nv);
    if (tan) s.tan.assign(tan, tan + 3 * (size_t)nv);
    s.idx.assign(idx, idx + 3 * (size_t)nf);
    s.mesh_first.assign(mesh_first, mesh_first + nmesh + 1);
    s.mesh_mat.assign(mesh_mat, mesh_mat + nmesh);
    const MaterialRec* mr = (const MaterialRec*)mats;
    s.mats.assign(mr, mr + nmat);
    if (env) { s.env[0] = env[0]; s.env[1] = env[1]; s.env[2] = env[2]; }
    scene_finalize(s);
    return o;
}
void orc_scene_free(void* p) { delete (OrcScene*)p; }
void orc_scene_set_vertices(void* p, const float* pos, const float* nrm, const float* tan) {
    Scene& s = ((OrcScene*)p)->s;
    std::memcpy(s.pos.data(), pos, sizeof(float) * 3 * (size_t)s.nv);
    if (nrm && !s.nrm.empty()) std::memcpy(s.nrm.data(), nrm, sizeof(float) * 3 * (size_t)s.nv);
    if (tan && !s.tan.empty()) std::memcpy(s.tan.data(), tan, sizeof(float) * 3 * (size_t)s.nv);
}
// lights: out_cdf[nl], out_mesh[nl], area_cdf[nf], mesh_area[nmesh]; returns nl, weight sum in *wsum
int orc_scene_lights(void* p, float* out_cdf, int32_t* out_mesh, float* area_cdf, float* mesh_area, float* wsum) {
    Scene& s = ((OrcScene*)p)->s;
    int nl = s.nlights();
    if (out_cdf) std::memcpy(out_cdf, s.light_cdf.data(), sizeof(float) * nl);
    if (out_mesh) std::memcpy(out_mesh, s.light_mesh.data(), sizeof(int32_t) * nl);
    if (area_cdf) std::memcpy(area_cdf, s.area_cdf.data(), sizeof(float) * s.nf);
    if (mesh_area) std::memcpy(mesh_area, s.mesh_area.data(), sizeof(float) * s.nmesh);
    if (wsum) *wsum = s.light_weight_sum;
    return nl;
}
// light sample probe (float path): rnd[2n] -> out[8n] = pos3, normal3, pdf, 0
void orc_light_sample(void* p, int n, const float* rnd, float* out) {
    Scene& s = ((OrcScene*)p)->s;
    for (int i = 0; i < n; i++) {
        LightSample<float> ls = light_sample<float>(s, V2<float>{rnd[2 * i], rnd[2 * i + 1]});
        float* o = out + 8 * (size_t)i;
        o[0] = ls.pos.x; o[1] = ls.pos.y; o[2] = ls.pos.z; o[3] = ls.normal.x; o[4] = ls.normal.y; o[5] = ls.normal.z; o[6] = ls.pdf; o[7] = 0;
    }
}

void* orc_bvh_new(void* scene, int max_leaf, int cluster_size) {
    OrcBvh* b = new OrcBvh();
    BvhParams prm;
    prm.max_leaf = max_leaf;
    prm.cluster_size = cluster_size;
    const Scene& s = ((OrcScene*)scene)->s;
    if (s.instanced()) blas_build(s, prm, b->blas);
    else bvh_build(s, prm, b->b);
    return b;
}

// Instances: mesh per instance and 3x4 row-major transforms (12 floats each);
// the meshes become object space.  Returns 0, or -1 for a bad mesh index or
// a singular transform.
int orc_scene_set_instances(void* p, int n, const int32_t* inst_mesh, const float* xf) {
    Scene& s = ((OrcScene*)p)->s;
    std::vector<float> inv(12 * (size_t)n);
    for (int i = 0; i < n; i++) {
        if (inst_mesh[i] < 0 || inst_mesh[i] >= s.nmesh) return -1;
        if (!affine_inverse(xf + 12 * (size_t)i, &inv[12 * (size_t)i])) return -1;
    }
    s.inst_mesh.assign(inst_mesh, inst_mesh + n);
    s.inst_xf.assign(xf, xf + 12 * (size_t)n);
    s.inst_inv = inv;
    scene_finalize(s);
    return 0;
}
// the inverse transforms as computed (12 floats per instance)
void orc_scene_instance_inverses(void* p, float* out) {
    Scene& s = ((OrcScene*)p)->s;
    std::memcpy(out, s.inst_inv.data(), sizeof(float) * s.inst_inv.size());
}
// lights of an instanced scene: the instance of each light
void orc_scene_light_instances(void* p, int32_t* out) {
    Scene& s = ((OrcScene*)p)->s;
    std::memcpy(out, s.light_inst.data(), sizeof(int32_t) * s.light_inst.size());
}

// closest or any hit in an instanced scene: mode 0 closest over the per-mesh
// BVHs, 1 any over them, 2 closest brute force, 3 any brute force; prim -1
// (closest, miss) or 0/1 (any); inst: the instance hit (closest), else -1
void orc_trace_inst(void* scene, void* bvh, int mode, int n, const float* o, const float* d, float tmin, const float* tmax,
                    float* t, float* u, float* v, int32_t* prim, int32_t* inst) {
    Scene& s = ((OrcScene*)scene)->s;
    const std::vector<Bvh>* blas = (mode < 2 && bvh) ? &((OrcBvh*)bvh)->blas : nullptr;
    for (int i = 0; i < n; i++) {
        Hit h;
        const bool any = (mode & 1) != 0;
        bool r = trace_inst(s, blas, o + 3 * (size_t)i, d + 3 * (size_t)i, tmin, tmax[i], any, &h);
        if (any) { prim[i] = r ? 1 : 0; if (inst) inst[i] = -1; continue; }
        t[i] = r ? h.t : 0.0f; u[i] = r ? h.u : 0.0f; v[i] = r ? h.v : 0.0f;
        prim[i] = r ? h.prim : -1;
        if (inst) inst[i] = r ? h.inst : -1;
    }
}
void orc_bvh_free(void* p) { delete (OrcBvh*)p; }
// info[0]=num_nodes, [1]=root, [2]=max_depth, [3]=num_clusters, [4]=top nodes, [5]=cluster size
void orc_bvh_info(void* p, int32_t* info) {
    Bvh& b = ((OrcBvh*)p)->b;
    info[0] = (int32_t)b.nodes.size(); info[1] = b.root; info[2] = b.max_depth; info[3] = b.num_clusters;
    info[4] = b.num_top_nodes; info[5] = b.cluster_size;
}
void orc_bvh_copy(void* p, void* nodes, uint32_t* order, float* tris, uint64_t* codes) {
    Bvh& b = ((OrcBvh*)p)->b;
    if (nodes) std::memcpy(nodes, b.nodes.data(), sizeof(BvhNode) * b.nodes.size());
    if (order) std::memcpy(order, b.order.data(), sizeof(uint32_t) * b.order.size());
    if (tris) std::memcpy(tris, b.tris.data(), sizeof(float) * b.tris.size());
    if (codes) std::memcpy(codes, b.codes.data(), sizeof(uint64_t) * b.codes.size());
}
void orc_bvh_karras(void* p, int32_t* left, int32_t* right) {
    Bvh& b = ((OrcBvh*)p)->b;
    std::memcpy(left, b.karras_left.data(), sizeof(int32_t) * b.karras_left.size());
    std::memcpy(right, b.karras_right.data(), sizeof(int32_t) * b.karras_right.size());
}
void orc_bvh_refit(void* p, void* scene) { bvh_refit(((OrcScene*)scene)->s, ((OrcBvh*)p)->b); }

// mode 0: closest over BVH, 1: any over BVH, 2: closest brute force, 3: any brute force.
// out: t, u, v (float), prim (int, -1 miss), visits/tests counters optional.
void orc_trace(void* scene, void* bvh, int mode, int n, const float* o, const float* d, float tmin, const float* tmax,
               float* t, float* u, float* v, int32_t* prim, int32_t* visits, int32_t* tests) {
    Scene& s = ((OrcScene*)scene)->s;
    for (int i = 0; i < n; i++) {
        const float* oi = o + 3 * (size_t)i; const float* di = d + 3 * (size_t)i;
        Hit h; bool r;
        int nv = 0, nt = 0;
        switch (mode) {
        case 0: r = trace_closest(((OrcBvh*)bvh)->b, oi, di, tmin, tmax[i], &h, &nv, &nt); break;
        case 1: r = trace_any(((OrcBvh*)bvh)->b, oi, di, tmin, tmax[i]); break;
        case 2: r = brute_closest(s, oi, di, tmin, tmax[i], &h); break;
        default: r = brute_any(s, oi, di, tmin, tmax[i]); break;
        }
        if (mode == 1 || mode == 3) { prim[i] = r ? 1 : 0; continue; }
        if (r) { t[i] = h.t; u[i] = h.u; v[i] = h.v; prim[i] = h.prim; }
        else { t[i] = 0; u[i] = 0; v[i] = 0; prim[i] = -1; }
        if (visits) visits[i] = nv;
        if (tests) tests[i] = nt;
    }
}

// Camera rays for every pixel of sample 0 (debug / first-hit parity).
void orc_camera_basis(const float* pos, const float* target, const float* up, float fov, int W, int H, float* out14) {
    Camera c = make_camera(pos, target, up, fov, W, H);
    std::memcpy(out14, c.pos, 3 * sizeof(float)); std::memcpy(out14 + 3, c.X, 12); std::memcpy(out14 + 6, c.Y, 12);
    std::memcpy(out14 + 9, c.Z, 12); out14[12] = c.nw; out14[13] = c.nh;
}
void orc_primary_rays(const float* pos, const float* target, const float* up, float fov, int W, int H, uint32_t seed,
                      uint32_t sample, float* o, float* d) {
    Camera cam = make_camera(pos, target, up, fov, W, H);
    V3<float> X = cvt<float>(cam.X), Y = cvt<float>(cam.Y), Z = cvt<float>(cam.Z), CP = cvt<float>(cam.pos);
    for (int pixel = 0; pixel < W * H; pixel++) {
        Pcg32 rng; path_rng(rng, (uint32_t)pixel, seed, sample);
        float r0 = rng.next_float();
        float xx = ((float)(pixel % W) + r0) / (float)W;
        float r1 = rng.next_float();
        float yy = ((float)(pixel / W) + r1) / (float)H;
        float px = (xx - 0.5f) * cam.nw, py = (yy - 0.5f) * cam.nh;
        V3<float> dir = normalize(V3<float>(px, py, 1.0f));
        V3<float> zero(0.0f);
        V3<float> ro = zero.x * X + zero.y * Y + zero.z * Z + CP;
        V3<float> rd = dir.x * X + dir.y * Y + dir.z * Z;
        o[3 * pixel] = ro.x; o[3 * pixel + 1] = ro.y; o[3 * pixel + 2] = ro.z;
        d[3 * pixel] = rd.x; d[3 * pixel + 1] = rd.y; d[3 * pixel + 2] = rd.z;
    }
}

// Render rows [row_begin, H) with stride row_step; out is float4 (W*H) or, for
// fp64, double4.  Returns wall seconds of the render (threads over rows).
double orc_render(void* scene, void* bvh, const float* cam_pos, const float* cam_target, const float* cam_up, float fov,
                  int W, int H, int spp, int max_depth, int rr_depth, int use_nee, int use_mis, uint32_t seed, int fp64,
                  int threads, int row_begin, int row_step, int sample_begin, int sample_end, float* out_f, double* out_d) {
    Scene& s = ((OrcScene*)scene)->s;
    Tracer tr{&s, bvh ? &((OrcBvh*)bvh)->b : nullptr, bvh ? &((OrcBvh*)bvh)->blas : nullptr};
    Camera cam = make_camera(cam_pos, cam_target, cam_up, fov, W, H);
    PtParams p;
    p.spp = spp; p.max_depth = max_depth; p.rr_depth = rr_depth; p.use_nee = use_nee; p.use_mis = use_mis; p.seed = seed;
    if (sample_end <= 0) sample_end = spp;
    if (row_step <= 0) row_step = 1;
    std::atomic<int> next_row(row_begin);
    std::atomic<int64_t> traces(0), shadows(0);
    auto worker = [&]() {
        int64_t nt = 0, ns = 0;
        for (;;) {
            int row = next_row.fetch_add(row_step);
            if (row >= H) break;
            for (int x = 0; x < W; x++) {
                int pixel = row * W + x;
                if (fp64) {
                    V3<double> r = render_pixel<double>(s, tr, cam, p, W, H, pixel, sample_begin, sample_end, &nt, &ns);
                    out_d[4 * (size_t)pixel] = r.x; out_d[4 * (size_t)pixel + 1] = r.y; out_d[4 * (size_t)pixel + 2] = r.z;
                    out_d[4 * (size_t)pixel + 3] = 1.0;
                } else {
                    V3<float> r = render_pixel<float>(s, tr, cam, p, W, H, pixel, sample_begin, sample_end, &nt, &ns);
                    out_f[4 * (size_t)pixel] = r.x; out_f[4 * (size_t)pixel + 1] = r.y; out_f[4 * (size_t)pixel + 2] = r.z;
                    out_f[4 * (size_t)pixel + 3] = 1.0f;
                }
            }
        }
        traces += nt;
        shadows += ns;
    };
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int i = 1; i < threads; i++) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    auto t1 = std::chrono::steady_clock::now();
    g_last_counts[0] = traces.load();
    g_last_counts[1] = shadows.load();
    return std::chrono::duration<double>(t1 - t0).count();
}

// The scene box (aabb.h): the vertex box of a world-space scene; for an
// instanced one, Scene::compute_aabb (scene.cu:299-307): the union of each
// instance's mesh vertex box, its eight corners through apply_point
// (Transform::apply_aabb, transform.h:58-70), in float.
static void scene_box(const Scene& s, float lo[3], float hi[3]) {
    if (!s.instanced()) {
        for (int a = 0; a < 3; a++) { lo[a] = s.pos[a]; hi[a] = s.pos[a]; }
        for (int i = 0; i < s.nv; i++)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], s.pos[3 * (size_t)i + a]);
                hi[a] = std::max(hi[a], s.pos[3 * (size_t)i + a]);
            }
        return;
    }
    for (int a = 0; a < 3; a++) { lo[a] = INFINITY; hi[a] = -INFINITY; }
    for (int i = 0; i < s.ninst(); i++) {
        const int m = s.inst_mesh[i];
        float bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int f = s.mesh_first[m]; f < s.mesh_first[m + 1]; f++)
            for (int k = 0; k < 3; k++) {
                const float* v = &s.pos[3 * (size_t)s.idx[3 * (size_t)f + k]];
                for (int a = 0; a < 3; a++) { bl[a] = std::min(bl[a], v[a]); bh[a] = std::max(bh[a], v[a]); }
            }
        if (!(bl[0] <= bh[0])) continue;                  // a mesh without triangles
        const float* T = &s.inst_xf[12 * (size_t)i];
        for (int c = 0; c < 8; c++) {
            const float p[3] = {(c & 4) ? bh[0] : bl[0], (c & 2) ? bh[1] : bl[1], (c & 1) ? bh[2] : bl[2]};
            float w[3];
            xf_point(T, p, w);
            for (int a = 0; a < 3; a++) { lo[a] = std::fmin(lo[a], w[a]); hi[a] = std::fmax(hi[a], w[a]); }
        }
    }
}

// Simple NPR pass, float (shader/npr/simple.cu:42-97, passes/npr/simple.cu:14-33):
// samples 0..spp-1 of each pixel take the path tracer's camera ray (same
// stream), the closest hit, and one first-hit colouring (type 0 Depth, 1
// Normal, 2 BaseColor, 3 Ambient, 4 FaceOrientation; misses white); the pixel
// is the running sum of L / spp in sample order, alpha 1.  Depth scales the
// camera distance between the vertex box's min_distance and max_distance
// (aabb.h:24-41), returned in range2.
void orc_render_simple(void* scene, void* bvh, const float* cam_pos, const float* cam_target, const float* cam_up,
                       float fov, int W, int H, int spp, uint32_t seed, int type, float* out, float* range2) {
    Scene& s = ((OrcScene*)scene)->s;
    Tracer tr{&s, bvh ? &((OrcBvh*)bvh)->b : nullptr, bvh ? &((OrcBvh*)bvh)->blas : nullptr};
    Camera cam = make_camera(cam_pos, cam_target, cam_up, fov, W, H);
    float lo[3], hi[3];
    scene_box(s, lo, hi);
    bool inside = true;
    V3<float> far3, near3;
    float fa[3], na[3];
    for (int a = 0; a < 3; a++) {
        const float dlo = std::fabs(cam.pos[a] - lo[a]), dhi = std::fabs(cam.pos[a] - hi[a]);
        fa[a] = std::fmax(dlo, dhi);
        na[a] = std::fmin(dlo, dhi);
        inside = inside && cam.pos[a] >= lo[a] && cam.pos[a] <= hi[a];
    }
    far3 = V3<float>(fa[0], fa[1], fa[2]);
    near3 = V3<float>(na[0], na[1], na[2]);
    const float max_depth = length(far3);
    const float min_depth = inside ? 0.0f : length(near3);
    if (range2) { range2[0] = min_depth; range2[1] = max_depth; }
    V3<float> X = cvt<float>(cam.X), Y = cvt<float>(cam.Y), Z = cvt<float>(cam.Z), CP = cvt<float>(cam.pos);
    for (int pixel = 0; pixel < W * H; pixel++) {
        V3<float> result(0.0f);
        for (int smp = 0; smp < spp; smp++) {
            Pcg32 rng;
            path_rng(rng, (uint32_t)pixel, seed, (uint32_t)smp);
            float r0 = rng.next_float();
            float xx = ((float)(pixel % W) + r0) / (float)W;
            float r1 = rng.next_float();
            float yy = ((float)(pixel / W) + r1) / (float)H;
            float px = (xx - 0.5f) * cam.nw, py = (yy - 0.5f) * cam.nh;
            V3<float> dir = normalize(V3<float>(px, py, 1.0f));
            V3<float> zero(0.0f);
            V3<float> ro = zero.x * X + zero.y * Y + zero.z * Z + CP;
            V3<float> rd = dir.x * X + dir.y * Y + dir.z * Z;
            float of[3] = {ro.x, ro.y, ro.z}, df[3] = {rd.x, rd.y, rd.z};
            Hit h;
            V3<float> L(1.0f);
            if (tr.closest(of, df, 1e-3f, 1e16f, &h)) {
                HitInfo<float> info;
                get_hitinfo<float>(s, h.prim, h.u, h.v, info, h.inst);
                if (type == 0) {
                    const float d = length(info.pos - CP);
                    const float ratio = (d - min_depth) / (max_depth - min_depth);
                    L = V3<float>(ratio * 0.8f + 0.2f);
                } else if (type == 1) {
                    L = (info.normal + V3<float>(1.0f)) * 0.5f;
                } else if (type == 2) {
                    L = info.color;
                } else if (type == 3) {
                    L = info.color * (std::fabs(dot(info.normal, rd)) * 0.5f + 0.5f);
                } else {
                    L = dot(info.normal, rd) > 0.0f ? V3<float>(1.0f, 0.85f, 0.0f) : V3<float>(0.34f, 0.73f, 0.76f);
                }
            }
            const float fs = (float)spp;
            result = V3<float>(result.x + L.x / fs, result.y + L.y / fs, result.z + L.z / fs);
        }
        out[4 * (size_t)pixel] = result.x;
        out[4 * (size_t)pixel + 1] = result.y;
        out[4 * (size_t)pixel + 2] = result.z;
        out[4 * (size_t)pixel + 3] = 1.0f;
    }
}

// Denoise pass of this build (the reference's is the closed OptiX denoiser,
// passes/postprocess/denoise.cu:54-114): edge-avoiding a-trous wavelet filter
// (Dammertz et al., HPG 2010) on the colour alone, in place: 5 passes of the
// B3-spline kernel (1, 4, 6, 4, 1) / 16 at tap spacing 2^i, tap weight
// exp(-|g_p - g_q|^2 / sigma_i^2) on g = log(1 + c) per channel of the pass's
// input, sigma_i = 0.2 * 2^-i; taps outside the image skipped; alpha kept.
void orc_denoise(int W, int H, float* rgba) {
    const float h[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    std::vector<float> a(rgba, rgba + (size_t)W * H * 4), b((size_t)W * H * 4);
    std::vector<V3<float>> gbuf((size_t)W * H);
    for (int it = 0; it < 5; it++) {
        for (size_t p = 0; p < (size_t)W * H; p++)
            gbuf[p] = V3<float>(pt_logf(1.0f + a[4 * p]), pt_logf(1.0f + a[4 * p + 1]), pt_logf(1.0f + a[4 * p + 2]));
        auto guide = [&](const float* c) { return gbuf[(size_t)(c - a.data()) / 4]; };
        const float sigma = 0.2f * std::ldexp(1.0f, -it);
        const float inv_s2 = 1.0f / (sigma * sigma);
        const int step = 1 << it;
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const float* cp = &a[4 * ((size_t)y * W + x)];
                const V3<float> gp = guide(cp);
                V3<float> sum(0.0f);
                float wsum = 0.0f;
                for (int j = -2; j <= 2; j++) {
                    const int yy = y + j * step;
                    if (yy < 0 || yy >= H) continue;
                    for (int i = -2; i <= 2; i++) {
                        const int xx = x + i * step;
                        if (xx < 0 || xx >= W) continue;
                        const float* cq = &a[4 * ((size_t)yy * W + xx)];
                        const V3<float> d = gp - guide(cq);
                        const float w = (h[j + 2] * h[i + 2]) * pt_expf(-(dot(d, d) * inv_s2));
                        sum = sum + V3<float>(cq[0], cq[1], cq[2]) * w;
                        wsum = wsum + w;
                    }
                }
                float* o = &b[4 * ((size_t)y * W + x)];
                o[0] = sum.x / wsum; o[1] = sum.y / wsum; o[2] = sum.z / wsum; o[3] = cp[3];
            }
        a.swap(b);
    }
    std::copy(a.begin(), a.end(), rgba);
}

// camera.h:143-148 get_xy and the pixel index of light_tracer.cu:100-101
// (x < 1 that rounds x * W up to W goes to the last column / row)
static bool camera_pixel(const Camera& cam, V3<float> dir, int W, int H, int* idx) {
    const V3<float> d = -dir;
    const V3<float> l(dot(d, cvt<float>(cam.X)), dot(d, cvt<float>(cam.Y)), dot(d, cvt<float>(cam.Z)));
    const float px = l.x / l.z, py = l.y / l.z;
    const float x = px / cam.nw + 0.5f, y = py / cam.nh + 0.5f;
    if (!(x >= 0.0f && x < 1.0f && y >= 0.0f && y < 1.0f)) return false;
    const int ix = std::min((int)(x * (float)W), W - 1), iy = std::min((int)(y * (float)H), H - 1);
    *idx = ix + iy * W;
    return true;
}

// pixel index of each direction (-1 outside the image): the inverse of the camera rays
void orc_camera_pixel(const float* pos, const float* target, const float* up, float fov, int W, int H, int n,
                      const float* dirs, int32_t* out) {
    Camera cam = make_camera(pos, target, up, fov, W, H);
    for (int i = 0; i < n; i++) {
        int idx = -1;
        if (!camera_pixel(cam, V3<float>(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), W, H, &idx)) idx = -1;
        out[i] = idx;
    }
}

// LightTracer, float (shader/integrator/light_tracer.cu:57-176): particle
// (launch index idx, sample) uses the (idx, seed, sample) stream; splats are
// added in (idx, sample, depth) order into out (float4, cleared to (0,0,0,1)).
void orc_render_light(void* scene, void* bvh, const float* cam_pos, const float* cam_target, const float* cam_up,
                      float fov, int W, int H, int spp, int max_depth, int rr_depth, uint32_t seed, float* out) {
    Scene& s = ((OrcScene*)scene)->s;
    Tracer tr{&s, bvh ? &((OrcBvh*)bvh)->b : nullptr, bvh ? &((OrcBvh*)bvh)->blas : nullptr};
    Camera cam = make_camera(cam_pos, cam_target, cam_up, fov, W, H);
    const V3<float> CP = cvt<float>(cam.pos);
    const float aspect = (float)W / (float)H;
    for (int i = 0; i < W * H; i++) { out[4 * i] = 0.0f; out[4 * i + 1] = 0.0f; out[4 * i + 2] = 0.0f; out[4 * i + 3] = 1.0f; }
    auto splat = [&](int idx, V3<float> L) {
        out[4 * idx] = out[4 * idx] + L.x;
        out[4 * idx + 1] = out[4 * idx + 1] + L.y;
        out[4 * idx + 2] = out[4 * idx + 2] + L.z;
    };
    for (int idx = 0; idx < W * H; idx++) {
        for (int smp = 0; smp < spp; smp++) {
            Pcg32 rng;
            path_rng(rng, (uint32_t)idx, seed, (uint32_t)smp);
            V2<float> lr;
            lr.x = rng.next_float();
            lr.y = rng.next_float();
            LightSample<float> ls = light_sample<float>(s, lr);
            V2<float> sr;
            sr.x = rng.next_float();
            sr.y = rng.next_float();
            V3<float> wi = normalize(ls.normal + uniform_sample_sphere(sr));
            float cos_theta = dot(wi, ls.normal);
            float pw = cosine_hemisphere_pdf(cos_theta);
            V3<float> beta = ls.emission * cos_theta / ls.pdf;
            bool specular = true;
            V3<float> ro = ls.pos, rd = wi;
            for (int depth = 0; depth < max_depth; depth++) {
                float of[3] = {ro.x, ro.y, ro.z}, df[3] = {rd.x, rd.y, rd.z};
                Hit h;
                bool hit = tr.closest(of, df, 1e-3f, 1e16f, &h);
                HitInfo<float> info;
                if (hit) get_hitinfo<float>(s, h.prim, h.u, h.v, info, h.inst);
                if (specular) {
                    V3<float> to_camera = CP - ro;
                    const float t = length(to_camera);
                    to_camera = V3<float>(to_camera.x / t, to_camera.y / t, to_camera.z / t);
                    if ((!hit || t < length(info.pos - ro)) && dot(to_camera, rd) > 1.0f - 1e-6f) {
                        int px;
                        if (camera_pixel(cam, rd, W, H, &px)) {
                            V3<float> b = V3<float>(beta.x / pw, beta.y / pw, beta.z / pw);
                            splat(px, V3<float>(b.x / (float)spp, b.y / (float)spp, b.z / (float)spp));
                        }
                    }
                }
                if (!hit) break;
                const MaterialRec& m = s.mats[info.mat];
                if (mat_is_emissive(m)) break;
                BxDF<float> bx(m);
                if (dot(-rd, info.normal) < 0.0f) info.normal = -info.normal;
                if (!mat_is_specular(m)) {
                    const V3<float> sdir = normalize(CP - info.pos);
                    int px;
                    if (camera_pixel(cam, sdir, W, H, &px)) {
                        const float t = length(CP - info.pos);
                        float so[3] = {info.pos.x, info.pos.y, info.pos.z}, sd[3] = {sdir.x, sdir.y, sdir.z};
                        if (!tr.any(so, sd, 1e-3f, t - 1e-3f)) {
                            const V3<float> f = mat_eval(bx, sdir, -rd, info.onb, info.color);
                            V3<float> L = aspect * beta * f * dot(sdir, info.normal);
                            L = V3<float>(L.x / t, L.y / t, L.z / t);
                            L = V3<float>(L.x / t, L.y / t, L.z / t);
                            L = V3<float>(L.x / pw, L.y / pw, L.z / pw);
                            splat(px, V3<float>(L.x / (float)spp, L.y / (float)spp, L.z / (float)spp));
                        }
                    }
                }
                V2<float> br;
                br.x = rng.next_float();
                br.y = rng.next_float();
                BxDFSample<float> ms = mat_sample(bx, -rd, br, info.onb, info.color);
                if (ms.pdf <= 1e-5f) break;
                const V3<float> f = mat_eval(bx, ms.wi, -rd, info.onb, info.color);
                const V3<float> g = f * std::fabs(dot(ms.wi, info.normal));
                beta = beta * V3<float>(g.x / pw, g.y / pw, g.z / pw);
                pw = ms.pdf;
                specular = mat_is_specular(m);
                ro = info.pos;
                rd = ms.wi;
                if (depth >= rr_depth) {
                    const float p = std::fmin(std::fmax(beta.x, std::fmax(beta.y, beta.z)), 0.95f);
                    if (rng.next_float() > p) break;
                    beta = V3<float>(beta.x / p, beta.y / p, beta.z / p);
                }
            }
        }
    }
}

// closest-hit traces and shadow rays of the last orc_render
void orc_last_counts(int64_t* out2) {
    out2[0] = g_last_counts[0];
    out2[1] = g_last_counts[1];
}

// tonemapping.cu:14-68; exposure scale computed on the host as powf(2, exposure)
void orc_tonemap(int type, float exposure, int use_gamma, int n, float* rgba) {
    float ex = std::pow(2.0f, exposure);
    for (int i = 0; i < n; i++) {
        float* px = rgba + 4 * (size_t)i;
        V3<float> c = V3<float>(px[0], px[1], px[2]) * ex;
        switch (type) {
        case 1: c = V3<float>(clampf(c.x, 0.0f, 1.0f), clampf(c.y, 0.0f, 1.0f), clampf(c.z, 0.0f, 1.0f)); break;
        case 2: { float L = luminance(c); c = c / (1.0f + L); break; }
        case 3: {
            const float A = 0.15f, B = 0.50f, Cc = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
            auto op = [&](float x) { return ((x * (A * x + Cc * B) + D * E) / (x * (A * x + B) + D * F)) - E / F; };
            c = V3<float>(op(c.x), op(c.y), op(c.z));
            float Wt = 11.2f;
            float white = ((Wt * (A * Wt + Cc * B) + D * E) / (Wt * (A * Wt + B) + D * F)) - E / F;
            c = c / white;
            break;
        }
        case 4: {
            c = c * 0.6f;
            const float a = 2.51f, b = 0.03f, cc = 2.43f, d = 0.59f, e = 0.14f;
            auto op = [&](float x) { return (x * (a * x + b)) / (x * (cc * x + d) + e); };
            c = V3<float>(op(c.x), op(c.y), op(c.z));
            c = V3<float>(clampf(c.x, 0.0f, 1.0f), clampf(c.y, 0.0f, 1.0f), clampf(c.z, 0.0f, 1.0f));
            break;
        }
        default: break;
        }
        if (use_gamma) {
            float g = 1.0f / 2.2f;
            c = V3<float>(pt_powf(c.x, g), pt_powf(c.y, g), pt_powf(c.z, g));
        }
        px[0] = c.x; px[1] = c.y; px[2] = c.z; px[3] = 1.0f;
    }
}

// accumulate.cu:22-26
void orc_accumulate(int n, int frame_count, float* acc, float* pixels) {
    float w = 1.0f / (float)(frame_count + 1);
    for (int i = 0; i < 4 * n; i++) {
        acc[i] = acc[i] * (1.0f - w) + pixels[i] * w;
        pixels[i] = acc[i];
    }
}

// scene.cu:362-392, bones as 3x4 row-major matrices
void orc_skin(int nv, const float* orig_pos, const float* orig_nrm, const float* orig_tan, const int32_t* bone_ids,
              const float* bone_w, int nb, const float* bones, float* pos, float* nrm, float* tan) {
    for (int i = 0; i < nv; i++) {
        V3<float> v(0.0f), nn(0.0f), tt(0.0f);
        V3<float> ov = cvt<float>(orig_pos + 3 * (size_t)i);
        V3<float> on = orig_nrm ? cvt<float>(orig_nrm + 3 * (size_t)i) : V3<float>(0.0f);
        V3<float> ot = orig_tan ? cvt<float>(orig_tan + 3 * (size_t)i) : V3<float>(0.0f);
        for (int j = 0; j < 4; j++) {
            int id = bone_ids[4 * (size_t)i + j];
            float w = bone_w[4 * (size_t)i + j];
            if (id == -1) {
                if (j == 0) { v = ov; nn = on; tt = ot; }
                break;
            }
            if (id >= nb) continue;
            const float* m = bones + 12 * (size_t)id;
            V3<float> pp(m[0] * ov.x + m[1] * ov.y + m[2] * ov.z + m[3], m[4] * ov.x + m[5] * ov.y + m[6] * ov.z + m[7],
                         m[8] * ov.x + m[9] * ov.y + m[10] * ov.z + m[11]);
            v = v + pp * w;
            V3<float> pn(m[0] * on.x + m[1] * on.y + m[2] * on.z, m[4] * on.x + m[5] * on.y + m[6] * on.z,
                         m[8] * on.x + m[9] * on.y + m[10] * on.z);
            nn = nn + pn * w;
            V3<float> pt(m[0] * ot.x + m[1] * ot.y + m[2] * ot.z, m[4] * ot.x + m[5] * ot.y + m[6] * ot.z,
                         m[8] * ot.x + m[9] * ot.y + m[10] * ot.z);
            tt = tt + pt * w;
        }
        pos[3 * (size_t)i] = v.x; pos[3 * (size_t)i + 1] = v.y; pos[3 * (size_t)i + 2] = v.z;
        if (nrm) { nrm[3 * (size_t)i] = nn.x; nrm[3 * (size_t)i + 1] = nn.y; nrm[3 * (size_t)i + 2] = nn.z; }
        if (tan) { tan[3 * (size_t)i] = tt.x; tan[3 * (size_t)i + 1] = tt.y; tan[3 * (size_t)i + 2] = tt.z; }
    }
}

}  // extern "C"
