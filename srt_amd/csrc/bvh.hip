// GPU BVH construction, refit and linear blend skinning for gfx950.
//
// Replaces OptiX's acceleration-structure build and update
// (src/optix/optix_ray_tracer.cu:213-359 build, :443-463 update) and the
// skinning lambda of Scene::update_gscene (src/scene/scene.cu:336-395).
//
// Build (deterministic: every decision uses integers, an exact total-order
// min/max on floats, or float formulas evaluated in a fixed order, so the
// bytes are reproducible by a host execution of the same steps):
//   1. triangle boxes, box centroids, centroid bounds (two-level reduction)
//   2. 63-bit Morton codes of centroids, stable radix sort of (code, id)
//   3. Karras (HPG 2012) binary radix tree, index tie-break for equal codes
//   4. bottom-up pass, one launch per radix-tree depth (deepest first, nodes
//      sorted by depth): boxes, primitive counts, SAH costs, collapse-to-leaf
//      decisions (leaves of <= max_leaf triangles)
//   5. frontier of the highest subtrees with <= T primitives ("clusters")
//   6. top of the tree rebuilt with a 16-bin SAH sweep over the clusters, in one
//      workgroup, breadth-first node numbering
//   7. emit 64-byte nodes (two child boxes per node) and the triangles in
//      Morton order (48 bytes each, primitive and mesh id in the w lanes)
// Refit keeps the topology and recomputes child boxes bottom-up, one launch per
// tree level.  Kernel boundaries order the levels: per-node device-scope fences
// (an L2 write-back and invalidate each on gfx950) made the arrival-counter
// form of these passes several times slower.
#include "srt_internal.h"
#include <hipcub/hipcub.hpp>
#include <cstdio>
#include <algorithm>

namespace srt {

// total order on floats (-0 < +0): key, inverse, min/max
SRT_D uint32_t fkey(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
SRT_D float fromkey(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }
SRT_D float kmin(float a, float b) { return (fkey(b) < fkey(a)) ? b : a; }
SRT_D float kmax(float a, float b) { return (fkey(b) > fkey(a)) ? b : a; }

struct DBox { float lo[3], hi[3]; };
SRT_D DBox empty_box() {
    DBox b;
    for (int i = 0; i < 3; i++) { b.lo[i] = __builtin_huge_valf(); b.hi[i] = -__builtin_huge_valf(); }
    return b;
}
SRT_D void grow(DBox& b, const DBox& o) {
    for (int i = 0; i < 3; i++) { b.lo[i] = kmin(b.lo[i], o.lo[i]); b.hi[i] = kmax(b.hi[i], o.hi[i]); }
}
SRT_D float box_area(const DBox& b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    if (!(dx >= 0.0f)) return 0.0f;
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

// SoA box storage
struct BoxArr { float *lox, *loy, *loz, *hix, *hiy, *hiz; };
SRT_D DBox ldbox(const BoxArr& a, size_t i) {
    DBox b;
    b.lo[0] = a.lox[i]; b.lo[1] = a.loy[i]; b.lo[2] = a.loz[i];
    b.hi[0] = a.hix[i]; b.hi[1] = a.hiy[i]; b.hi[2] = a.hiz[i];
    return b;
}
SRT_D void stbox(const BoxArr& a, size_t i, const DBox& b) {
    a.lox[i] = b.lo[0]; a.loy[i] = b.lo[1]; a.loz[i] = b.lo[2];
    a.hix[i] = b.hi[0]; a.hiy[i] = b.hi[1]; a.hiz[i] = b.hi[2];
}

// ---------------------------------------------------------------------------
// 1. triangle boxes + centroid bounds
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_tri_boxes(const float* pos, const uint32_t* idx, int n, BoxArr tb, float* cen,
                                                   uint32_t* partial) {
    __shared__ uint32_t red[6][256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    DBox cb = empty_box();
    if (i < n) {
        DBox b = empty_box();
        for (int v = 0; v < 3; v++) {
            const float* p = pos + 3 * (size_t)idx[3 * (size_t)i + v];
            for (int a = 0; a < 3; a++) { b.lo[a] = kmin(b.lo[a], p[a]); b.hi[a] = kmax(b.hi[a], p[a]); }
        }
        stbox(tb, i, b);
        for (int a = 0; a < 3; a++) {
            float c = (b.lo[a] + b.hi[a]) * 0.5f;
            cen[3 * (size_t)i + a] = c;
            cb.lo[a] = c; cb.hi[a] = c;
        }
    }
    for (int a = 0; a < 3; a++) { red[a][threadIdx.x] = fkey(cb.lo[a]); red[3 + a][threadIdx.x] = fkey(cb.hi[a]); }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int a = 0; a < 3; a++) {
                red[a][threadIdx.x] = min(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = max(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * (size_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) k_reduce_bounds(const uint32_t* partial, int nblocks, uint32_t* out6) {
    __shared__ uint32_t red[6][256];
    uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int a = 0; a < 3; a++) { lo[a] = min(lo[a], partial[6 * (size_t)b + a]); hi[a] = max(hi[a], partial[6 * (size_t)b + 3 + a]); }
    for (int a = 0; a < 3; a++) { red[a][threadIdx.x] = lo[a]; red[3 + a][threadIdx.x] = hi[a]; }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 3; a++) {
                red[a][threadIdx.x] = min(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = max(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out6[threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// 2. Morton codes (21 bits per axis, x in the highest bit of each triple)
// ---------------------------------------------------------------------------
SRT_D uint64_t expand_bits21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffULL;
    x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}
SRT_D uint32_t morton_axis(float c, float lo, float ext) {
    float t = (ext > 0.0f) ? (c - lo) / ext : 0.0f;
    float q = t * 2097152.0f;
    return (q >= 2097151.0f) ? 2097151u : (q > 0.0f ? (uint32_t)q : 0u);
}
__global__ void __launch_bounds__(256) k_morton(const float* cen, const uint32_t* cb6, int n, uint64_t* keys, uint32_t* vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        float lo = fromkey(cb6[a]), hi = fromkey(cb6[3 + a]);
        q[a] = morton_axis(cen[3 * (size_t)i + a], lo, hi - lo);
    }
    keys[i] = (expand_bits21(q[0]) << 2) | (expand_bits21(q[1]) << 1) | expand_bits21(q[2]);
    vals[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// 3. Karras radix tree
// ---------------------------------------------------------------------------
SRT_D int kdelta(const uint64_t* c, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = c[i], b = c[j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}
// child refs: >= 0 internal node, < 0 leaf (-k-1)
__global__ void __launch_bounds__(256) k_karras(const uint64_t* sc, int n, int32_t* kl, int32_t* kr, int32_t* kfirst,
                                                int32_t* par_int, int32_t* par_leaf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (kdelta(sc, n, i, i + 1) - kdelta(sc, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = kdelta(sc, n, i, i - d);
    int lmax = 2;
    while (kdelta(sc, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(sc, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = kdelta(sc, n, i, j);
    int sp = 0, t = l;
    do {
        t = (t + 1) / 2;
        if (kdelta(sc, n, i, i + (sp + t) * d) > dnode) sp += t;
    } while (t > 1);
    int gamma = i + sp * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    int32_t left = (lo == gamma) ? -(gamma + 1) : gamma;
    int32_t right = (hi == gamma + 1) ? -(gamma + 2) : gamma + 1;
    kl[i] = left;
    kr[i] = right;
    kfirst[i] = lo;
    if (left < 0) par_leaf[-left - 1] = i; else par_int[left] = i;
    if (right < 0) par_leaf[-right - 1] = i; else par_int[right] = i;
}

// ---------------------------------------------------------------------------
// 4. bottom-up boxes / counts / SAH costs / collapse
// ---------------------------------------------------------------------------
constexpr int kMaxRadixDepth = 4096;   // depth of the per-triangle radix tree (16-bit sort keys)
struct KArr {
    BoxArr box;
    int32_t* count;
    float* cost;
    int32_t* collapse;
};
// Depth of radix-tree node i below the root (node 0), keyed for the sort that
// orders the nodes by depth
__global__ void __launch_bounds__(256) k_radix_depth(int m, const int32_t* par_int, uint32_t* key, uint32_t* idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint32_t depth = 0;
    for (int p = par_int[i]; p >= 0; p = par_int[p]) depth++;
    key[i] = depth;
    idx[i] = (uint32_t)i;
}
// One depth of the bottom-up pass (deepest first): box, triangle count, SAH
// cost and collapse decision of each node from its two children, which the
// previous launch (or the triangle boxes) provides.
__global__ void __launch_bounds__(256) k_bottom_up_level(const uint32_t* ids, int count, const uint32_t* order, BoxArr tb,
                                                         const int32_t* kl, const int32_t* kr, KArr K, int max_leaf,
                                                         float c_trav, float c_isect) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int p = (int)ids[k];
    DBox bl, br;
    int cl, cr;
    float costl, costr;
    int32_t L = kl[p], R = kr[p];
    if (L < 0) { int q = -L - 1; bl = ldbox(tb, order[q]); cl = 1; costl = c_isect * 1.0f * box_area(bl); }
    else { bl = ldbox(K.box, L); cl = K.count[L]; costl = K.cost[L]; }
    if (R < 0) { int q = -R - 1; br = ldbox(tb, order[q]); cr = 1; costr = c_isect * 1.0f * box_area(br); }
    else { br = ldbox(K.box, R); cr = K.count[R]; costr = K.cost[R]; }
    DBox b = bl;
    grow(b, br);
    int cnt = cl + cr;
    float sa = box_area(b);
    float ci = c_trav * sa + costl + costr;
    float cf = c_isect * (float)cnt * sa;
    int col = (cnt <= max_leaf) && (cf <= ci);
    stbox(K.box, p, b);
    K.count[p] = cnt;
    K.cost[p] = col ? cf : ci;
    K.collapse[p] = col;
}

// ---------------------------------------------------------------------------
// 5. cluster frontier (marks the first sorted position of each cluster)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_frontier(int n, int T, const int32_t* kfirst, const int32_t* kcount, const int32_t* par_int,
                                                  const int32_t* par_leaf, int32_t* mark) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    // x < n-1: internal node x; n-1 <= x < 2n-1: leaf x-(n-1)
    if (x >= 2 * n - 1) return;
    int32_t ref, first, cnt, parent;
    if (x < n - 1) { ref = x; first = kfirst[x]; cnt = kcount[x]; parent = par_int[x]; }
    else { int k = x - (n - 1); ref = -(k + 1); first = k; cnt = 1; parent = par_leaf[k]; }
    if (cnt > T) return;
    if (parent >= 0 && kcount[parent] <= T) return;
    mark[first] = ref;
}

// cluster items for the top-level SAH
struct Items {
    BoxArr box;
    float *cx, *cy, *cz;
    int32_t* count;
    int32_t* ref;
};
__global__ void __launch_bounds__(256) k_items(int m, const int32_t* cl, const uint32_t* order, BoxArr tb, KArr K,
                                               const int32_t* kfirst, int top, Items it) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    int32_t c = cl[q];
    DBox b;
    int cnt;
    int32_t ref;
    if (c < 0) {
        int k = -c - 1;
        b = ldbox(tb, order[k]); cnt = 1; ref = leaf_ref((uint32_t)k, 1u);
    } else {
        b = ldbox(K.box, c); cnt = K.count[c];
        ref = K.collapse[c] ? leaf_ref((uint32_t)kfirst[c], (uint32_t)cnt) : top + c;
    }
    stbox(it.box, q, b);
    it.cx[q] = (b.lo[0] + b.hi[0]) * 0.5f;
    it.cy[q] = (b.lo[1] + b.hi[1]) * 0.5f;
    it.cz[q] = (b.lo[2] + b.hi[2]) * 0.5f;
    it.count[q] = cnt;
    it.ref[q] = ref;
}

SRT_D void set_child(BvhNode* nd, int side, const DBox& b, int32_t ref) {
    if (side == 0) {
        nd->a.x = b.lo[0]; nd->a.y = b.hi[0]; nd->a.z = b.lo[1]; nd->a.w = b.hi[1];
        nd->c.x = b.lo[2]; nd->c.y = b.hi[2];
        nd->d.x = ref;
    } else {
        nd->b.x = b.lo[0]; nd->b.y = b.hi[0]; nd->b.z = b.lo[1]; nd->b.w = b.hi[1];
        nd->c.z = b.lo[2]; nd->c.w = b.hi[2];
        nd->d.y = ref;
    }
}

// ---------------------------------------------------------------------------
// 6. top-level binned SAH, one workgroup, breadth-first
// ---------------------------------------------------------------------------
constexpr int kTopThreads = 1024;
constexpr int kBins = 16;

SRT_D int bin_of(float c, float lo, float scale) {
    int bi = (int)((c - lo) * scale);
    if (bi > kBins - 1) bi = kBins - 1;
    if (bi < 0) bi = 0;
    return bi;
}
SRT_D float item_c(const Items& it, int a, int i) { return a == 0 ? it.cx[i] : (a == 1 ? it.cy[i] : it.cz[i]); }
SRT_D void copy_item(const Items& s, int i, const Items& d, int j) {
    d.box.lox[j] = s.box.lox[i]; d.box.loy[j] = s.box.loy[i]; d.box.loz[j] = s.box.loz[i];
    d.box.hix[j] = s.box.hix[i]; d.box.hiy[j] = s.box.hiy[i]; d.box.hiz[j] = s.box.hiz[i];
    d.cx[j] = s.cx[i]; d.cy[j] = s.cy[i]; d.cz[j] = s.cz[i];
    d.count[j] = s.count[i]; d.ref[j] = s.ref[i];
}

// Level-parallel form: one workgroup per node of a level finds its split and
// partitions its item range in place (stable: left items, then right items,
// each in their previous order); k_top_link then numbers the inner children of
// the level in order (breadth-first numbering, left before right) and writes
// the parents' child boxes and refs.  Each node's split depends only on its own
// items, so the tree is the one a sequential breadth-first build produces.
struct TopSplit {
    int32_t mid;
    float box[2][6];        // child boxes: lo xyz, hi xyz
};
constexpr int kSplitThreads = 256;
static_assert(kSplitThreads == 256, "k_top_split ranks items four waves at a time");
// min / max of ordered float keys over the 64 lanes of a wave (every lane gets them)
SRT_D uint32_t wave_min_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}
SRT_D uint32_t wave_max_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}
constexpr int kSplitWaves = 4;     // kSplitThreads / 64
// SAH over the 15 bin boundaries of each axis (cost = area x count per side of
// the boundary) from the bins' key boxes and counts: the first minimum in
// (axis, boundary) order; axis -1 when no boundary separates items (every
// extent zero or every item in one bin).
// The same choice by all 64 lanes of a wave (every lane gets the result):
// lane 16a + k holds bin k of axis a; inclusive prefix and suffix unions of
// the bins' key boxes and counts come from shuffles within each group of 16
// lanes (unions of keys are exact, so each prefix box equals the sequential
// one); the cost of boundary k uses the suffix of lane k + 1; the minimum over
// (cost, 16a + k) picks the first minimum in (axis, boundary) order, as the
// sequential loop does with its strict comparison.
SRT_D void sah_choose_wave(const uint32_t (*bin)[kBins][6], const int32_t (*cnt)[kBins], const float* ext, int32_t* axis,
                           int32_t* split) {
    static_assert(kBins == 16, "one group of 16 lanes per axis");
    const int lane = __lane_id();
    const int a = lane >> 4, k = lane & 15;
    const bool on = a < 3;
    const uint32_t kinf = fkey(__builtin_huge_valf()), kninf = fkey(-__builtin_huge_valf());
    uint32_t pl[6], sl[6];
    int32_t pc = 0, sc = 0;
    for (int q = 0; q < 6; q++) pl[q] = on ? bin[a][k][q] : (q < 3 ? kinf : kninf);
    if (on) pc = cnt[a][k];
    for (int q = 0; q < 6; q++) sl[q] = pl[q];
    sc = pc;
    for (int d = 1; d < 16; d <<= 1) {
        int32_t u = __shfl_up(pc, d, 16), w = __shfl_down(sc, d, 16);
        if (k >= d) pc += u;
        if (k + d < 16) sc += w;
        for (int q = 0; q < 6; q++) {
            const uint32_t x = (uint32_t)__shfl_up((int)pl[q], d, 16), y = (uint32_t)__shfl_down((int)sl[q], d, 16);
            if (k >= d) pl[q] = (q < 3) ? min(pl[q], x) : max(pl[q], x);
            if (k + d < 16) sl[q] = (q < 3) ? min(sl[q], y) : max(sl[q], y);
        }
    }
    DBox lb, rb;
    for (int q = 0; q < 3; q++) {
        lb.lo[q] = fromkey(pl[q]); lb.hi[q] = fromkey(pl[3 + q]);
        rb.lo[q] = fromkey(sl[q]); rb.hi[q] = fromkey(sl[3 + q]);
    }
    const float la = box_area(lb), ra = box_area(rb);
    // the right side of boundary k: suffix of bin k + 1
    const float ra1 = __shfl_down(ra, 1, 16);
    const int32_t rc1 = __shfl_down(sc, 1, 16);
    unsigned long long key = ~0ull;
    if (on && k < kBins - 1 && ext[a] > 0.0f && pc != 0 && rc1 != 0) {
        const float cost = la * (float)pc + ra1 * (float)rc1;
        if (cost < __builtin_huge_valf()) key = ((unsigned long long)__float_as_uint(cost) << 32) | (unsigned)lane;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(key, off);
        key = o < key ? o : key;
    }
    if (key == ~0ull) {
        *axis = -1;
        *split = -1;
    } else {
        *axis = (int32_t)((key & 0xffffffffu) >> 4);
        *split = (int32_t)(key & 15u);
    }
}
__global__ void __launch_bounds__(kSplitThreads) k_top_split(Items it, Items tmp, const int32_t* task_b, const int32_t* task_e,
                                                             TopSplit* out) {
    __shared__ uint32_t s_cb[6];
    // bins per wave (fewer lanes contend for one LDS word), merged after binning
    __shared__ uint32_t s_bin[kSplitWaves][3][kBins][6];
    __shared__ int32_t s_cnt[kSplitWaves][3][kBins];
    __shared__ int32_t s_scan[kSplitThreads];
    __shared__ uint32_t s_box[2][6];
    __shared__ int32_t s_axis, s_split, s_run;
    __shared__ float s_lo[3], s_ext[3];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int b = task_b[blockIdx.x], e = task_e[blockIdx.x];
    const uint32_t kinf = fkey(__builtin_huge_valf()), kninf = fkey(-__builtin_huge_valf());
    if (tid < 6) s_cb[tid] = (tid < 3) ? kinf : kninf;
    for (int j = tid; j < kSplitWaves * 3 * kBins * 6; j += kSplitThreads)
        (&s_bin[0][0][0][0])[j] = ((j % 6) < 3) ? kinf : kninf;
    for (int j = tid; j < kSplitWaves * 3 * kBins; j += kSplitThreads) (&s_cnt[0][0][0])[j] = 0;
    if (tid < 12) (&s_box[0][0])[tid] = ((tid % 6) < 3) ? kinf : kninf;
    __syncthreads();
    // centroid bounds: per-lane keys, a wave reduction, one LDS atomic per wave
    {
        uint32_t mn[3] = {kinf, kinf, kinf}, mx[3] = {kninf, kninf, kninf};
        for (int i = b + tid; i < e; i += kSplitThreads)
            for (int a = 0; a < 3; a++) {
                const uint32_t k = fkey(item_c(it, a, i));
                mn[a] = min(mn[a], k);
                mx[a] = max(mx[a], k);
            }
        for (int a = 0; a < 3; a++) {
            const uint32_t wmn = wave_min_u32(mn[a]), wmx = wave_max_u32(mx[a]);
            if (lane == 0) {
                atomicMin(&s_cb[a], wmn);
                atomicMax(&s_cb[3 + a], wmx);
            }
        }
    }
    __syncthreads();
    if (tid < 3) {
        float lo = fromkey(s_cb[tid]), hi = fromkey(s_cb[3 + tid]);
        s_lo[tid] = lo;
        s_ext[tid] = hi - lo;
    }
    __syncthreads();
    // bins for the three axes
    for (int i = b + tid; i < e; i += kSplitThreads) {
        DBox ib = ldbox(it.box, i);
        int cnt = it.count[i];
        for (int a = 0; a < 3; a++) {
            float ex = s_ext[a];
            if (!(ex > 0.0f)) continue;
            float scale = (float)kBins / ex;
            int bi = bin_of(item_c(it, a, i), s_lo[a], scale);
            for (int q = 0; q < 3; q++) {
                atomicMin(&s_bin[wv][a][bi][q], fkey(ib.lo[q]));
                atomicMax(&s_bin[wv][a][bi][3 + q], fkey(ib.hi[q]));
            }
            atomicAdd(&s_cnt[wv][a][bi], cnt);
        }
    }
    __syncthreads();
    // merge the waves' bins into copy 0 (min / max of keys and integer sums: order-free)
    for (int j = tid; j < 3 * kBins * 7; j += kSplitThreads) {
        const int w = j % 7, ab = j / 7;
        const int a = ab / kBins, k = ab % kBins;
        if (w < 6) {
            uint32_t v = s_bin[0][a][k][w];
            for (int q = 1; q < kSplitWaves; q++) v = (w < 3) ? min(v, s_bin[q][a][k][w]) : max(v, s_bin[q][a][k][w]);
            s_bin[0][a][k][w] = v;
        } else {
            int32_t v = s_cnt[0][a][k];
            for (int q = 1; q < kSplitWaves; q++) v += s_cnt[q][a][k];
            s_cnt[0][a][k] = v;
        }
    }
    __syncthreads();
    if (tid < 64) {
        int32_t ax, sp;
        sah_choose_wave(s_bin[0], s_cnt[0], s_ext, &ax, &sp);
        if (tid == 0) {
            s_axis = ax;
            s_split = sp;
            s_run = 0;
        }
    }
    __syncthreads();
    const int axis = s_axis, split = s_split;
    int mid;
    if (axis < 0) {
        mid = (b + e) / 2;
    } else {
        const float scale = (float)kBins / s_ext[axis];
        const float lo = s_lo[axis];
        // stable partition, 256 items per step: each wave ranks its items with
        // a ballot, the four wave counts go through LDS
        const unsigned long long below = (1ull << lane) - 1ull;
        for (int pass = 0; pass < 2; pass++) {
            const int base = (pass == 0) ? b : b + s_run;   // s_run holds nleft after pass 0
            __syncthreads();
            int run = 0;
            for (int start = b; start < e; start += kSplitThreads) {
                int i = start + tid;
                bool flag = false;
                if (i < e) {
                    bool left = bin_of(item_c(it, axis, i), lo, scale) <= split;
                    flag = (pass == 0) ? left : !left;
                }
                const unsigned long long m = __ballot(flag);
                if (lane == 0) s_scan[wv] = __popcll(m);
                __syncthreads();
                int wbase = 0, total = 0;
                for (int q = 0; q < kSplitWaves; q++) {
                    const int c = s_scan[q];
                    wbase += (q < wv) ? c : 0;
                    total += c;
                }
                if (flag) copy_item(it, i, tmp, base + run + wbase + __popcll(m & below));
                run += total;
                __syncthreads();
            }
            if (pass == 0 && tid == 0) s_run = run;
            __syncthreads();
        }
        mid = b + s_run;
        __syncthreads();
        for (int i = b + tid; i < e; i += kSplitThreads) copy_item(tmp, i, it, i);
        __syncthreads();
    }
    // child boxes: per-lane keys of each side, wave reductions, one LDS atomic per wave
    {
        uint32_t mn[2][3], mx[2][3];
        for (int sd = 0; sd < 2; sd++)
            for (int q = 0; q < 3; q++) { mn[sd][q] = kinf; mx[sd][q] = kninf; }
        for (int i = b + tid; i < e; i += kSplitThreads) {
            const int side = (i < mid) ? 0 : 1;
            const DBox ib = ldbox(it.box, i);
            for (int q = 0; q < 3; q++) {
                mn[side][q] = min(mn[side][q], fkey(ib.lo[q]));
                mx[side][q] = max(mx[side][q], fkey(ib.hi[q]));
            }
        }
        for (int sd = 0; sd < 2; sd++)
            for (int q = 0; q < 3; q++) {
                const uint32_t wmn = wave_min_u32(mn[sd][q]), wmx = wave_max_u32(mx[sd][q]);
                if (lane == 0) {
                    atomicMin(&s_box[sd][q], wmn);
                    atomicMax(&s_box[sd][3 + q], wmx);
                }
            }
    }
    __syncthreads();
    if (tid < 12) out[blockIdx.x].box[tid / 6][tid % 6] = fromkey(s_box[tid / 6][tid % 6]);
    if (tid == 0) out[blockIdx.x].mid = mid;
}


// Wide levels: a level of at most kWideTasks nodes (the first levels, where
// one workgroup per node would leave most CUs idle) is split by a grid of
// workgroups over chunks of kWideChunk items; chunk g of the level belongs to
// the node whose chunk range holds it.  Per-node accumulators live in global
// memory: centroid bounds, bins and child boxes are min / max of ordered float
// keys and integer sums (independent of the order of the atomics), the split
// is chosen by sah_choose_wave as above, and the partition places each chunk's
// items after the left / right items of the node's earlier chunks, ranked
// within the chunk by wave ballots: the same split, the same stable partition
// and the same child boxes as k_top_split.
constexpr int kWideTasks = 64;
constexpr int kWideChunk = 1024;      // 4 items per thread
constexpr int kWideMaxGrid = 1024;
struct WideAcc {
    uint32_t cb[kWideTasks][6];                 // centroid bounds (keys): lo xyz, hi xyz
    uint32_t bin[kWideTasks][3][kBins][6];      // bin boxes (keys)
    int32_t cnt[kWideTasks][3][kBins];          // bin triangle counts
    int32_t axis[kWideTasks], split[kWideTasks], nleft[kWideTasks];
    float lo[kWideTasks][3], ext[kWideTasks][3];
    uint32_t box[kWideTasks][2][6];             // child boxes (keys)
};
// Chunk ranges of the level's nodes (wave 0 scans the chunk counts; ntask <=
// 64) into s_first[0..ntask]; returns the number of chunks.
SRT_D int wide_chunks(int ntask, const int32_t* tb, const int32_t* te, int32_t* s_first) {
    if (threadIdx.x < 64u) {
        const int t = threadIdx.x;
        int v = t < ntask ? (te[t] - tb[t] + kWideChunk - 1) / kWideChunk : 0;
        for (int d = 1; d < 64; d <<= 1) {
            const int u = __shfl_up(v, d);
            if (t >= d) v += u;
        }
        s_first[t + 1] = v;
        if (t == 0) s_first[0] = 0;
    }
    __syncthreads();
    return s_first[ntask];
}
SRT_D int wide_task(const int32_t* s_first, int g) {
    int t = 0;
    while (g >= s_first[t + 1]) t++;
    return t;
}
__global__ void __launch_bounds__(256) k_wide_init(int ntask, WideAcc* A) {
    const uint32_t kinf = fkey(__builtin_huge_valf()), kninf = fkey(-__builtin_huge_valf());
    const int j0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (int j = j0; j < ntask * 6; j += stride) (&A->cb[0][0])[j] = (j % 6 < 3) ? kinf : kninf;
    for (int j = j0; j < ntask * 3 * kBins * 6; j += stride) (&A->bin[0][0][0][0])[j] = (j % 6 < 3) ? kinf : kninf;
    for (int j = j0; j < ntask * 3 * kBins; j += stride) (&A->cnt[0][0][0])[j] = 0;
    for (int j = j0; j < ntask * 12; j += stride) (&A->box[0][0][0])[j] = (j % 6 < 3) ? kinf : kninf;
}
__global__ void __launch_bounds__(256) k_wide_bounds(int ntask, const int32_t* tb, const int32_t* te, Items it, WideAcc* A) {
    __shared__ int32_t s_first[kWideTasks + 1];
    __shared__ uint32_t s_cb[6];
    const int tid = threadIdx.x, lane = tid & 63;
    const uint32_t kinf = fkey(__builtin_huge_valf()), kninf = fkey(-__builtin_huge_valf());
    const int total = wide_chunks(ntask, tb, te, s_first);
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const int t = wide_task(s_first, g);
        const int b = tb[t] + (g - s_first[t]) * kWideChunk, e = min(b + kWideChunk, te[t]);
        if (tid < 6) s_cb[tid] = (tid < 3) ? kinf : kninf;
        __syncthreads();
        uint32_t mn[3] = {kinf, kinf, kinf}, mx[3] = {kninf, kninf, kninf};
        for (int i = b + tid; i < e; i += 256)
            for (int a = 0; a < 3; a++) {
                const uint32_t k = fkey(item_c(it, a, i));
                mn[a] = min(mn[a], k);
                mx[a] = max(mx[a], k);
            }
        for (int a = 0; a < 3; a++) {
            const uint32_t wmn = wave_min_u32(mn[a]), wmx = wave_max_u32(mx[a]);
            if (lane == 0) {
                atomicMin(&s_cb[a], wmn);
                atomicMax(&s_cb[3 + a], wmx);
            }
        }
        __syncthreads();
        if (tid < 3) atomicMin(&A->cb[t][tid], s_cb[tid]);
        else if (tid < 6) atomicMax(&A->cb[t][tid], s_cb[tid]);
        __syncthreads();
    }
}
__global__ void __launch_bounds__(256) k_wide_bins(int ntask, const int32_t* tb, const int32_t* te, Items it, WideAcc* A) {
    __shared__ int32_t s_first[kWideTasks + 1];
    __shared__ uint32_t s_bin[kSplitWaves][3][kBins][6];
    __shared__ int32_t s_cnt[kSplitWaves][3][kBins];
    __shared__ float s_lo[3], s_ext[3];
    const int tid = threadIdx.x, wv = tid >> 6;
    const uint32_t kinf = fkey(__builtin_huge_valf()), kninf = fkey(-__builtin_huge_valf());
    const int total = wide_chunks(ntask, tb, te, s_first);
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const int t = wide_task(s_first, g);
        const int b = tb[t] + (g - s_first[t]) * kWideChunk, e = min(b + kWideChunk, te[t]);
        for (int j = tid; j < kSplitWaves * 3 * kBins * 6; j += 256) (&s_bin[0][0][0][0])[j] = ((j % 6) < 3) ? kinf : kninf;
        for (int j = tid; j < kSplitWaves * 3 * kBins; j += 256) (&s_cnt[0][0][0])[j] = 0;
        if (tid < 3) {
            const float lo = fromkey(A->cb[t][tid]), hi = fromkey(A->cb[t][3 + tid]);
            s_lo[tid] = lo;
            s_ext[tid] = hi - lo;
        }
        __syncthreads();
        for (int i = b + tid; i < e; i += 256) {
            const DBox ib = ldbox(it.box, i);
            const int cnt = it.count[i];
            for (int a = 0; a < 3; a++) {
                const float ex = s_ext[a];
                if (!(ex > 0.0f)) continue;
                const float scale = (float)kBins / ex;
                const int bi = bin_of(item_c(it, a, i), s_lo[a], scale);
                for (int q = 0; q < 3; q++) {
                    atomicMin(&s_bin[wv][a][bi][q], fkey(ib.lo[q]));
                    atomicMax(&s_bin[wv][a][bi][3 + q], fkey(ib.hi[q]));
                }
                atomicAdd(&s_cnt[wv][a][bi], cnt);
            }
        }
        __syncthreads();
        // merge the waves' bins and add the non-empty ones to the node's bins
        for (int j = tid; j < 3 * kBins * 7; j += 256) {
            const int w = j % 7, ab = j / 7;
            const int a = ab / kBins, k = ab % kBins;
            int32_t n = 0;
            for (int q = 0; q < kSplitWaves; q++) n += s_cnt[q][a][k];
            if (n == 0) continue;
            if (w < 6) {
                uint32_t v = s_bin[0][a][k][w];
                for (int q = 1; q < kSplitWaves; q++) v = (w < 3) ? min(v, s_bin[q][a][k][w]) : max(v, s_bin[q][a][k][w]);
                if (w < 3) atomicMin(&A->bin[t][a][k][w], v);
                else atomicMax(&A->bin[t][a][k][w], v);
            } else {
                atomicAdd(&A->cnt[t][a][k], n);
            }
        }
        __syncthreads();
    }
}
__global__ void __launch_bounds__(64) k_wide_sah(WideAcc* A) {
    const int t = blockIdx.x;
    float ext[3];
    for (int a = 0; a < 3; a++) {
        const float lo = fromkey(A->cb[t][a]), hi = fromkey(A->cb[t][3 + a]);
        ext[a] = hi - lo;
        if (threadIdx.x == 0) {
            A->lo[t][a] = lo;
            A->ext[t][a] = ext[a];
        }
    }
    int32_t ax, sp;
    sah_choose_wave(A->bin[t], A->cnt[t], ext, &ax, &sp);
    if (threadIdx.x == 0) {
        A->axis[t] = ax;
        A->split[t] = sp;
    }
}
// left items of each chunk (a node without a split keeps its order)
__global__ void __launch_bounds__(256) k_wide_count(int ntask, const int32_t* tb, const int32_t* te, Items it,
                                                    const WideAcc* A, int32_t* chunk_left) {
    __shared__ int32_t s_first[kWideTasks + 1];
    __shared__ int32_t s_w[kSplitWaves];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int total = wide_chunks(ntask, tb, te, s_first);
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const int t = wide_task(s_first, g);
        const int axis = A->axis[t];
        if (axis < 0) continue;          // uniform over the workgroup
        const int b = tb[t] + (g - s_first[t]) * kWideChunk, e = min(b + kWideChunk, te[t]);
        const float scale = (float)kBins / A->ext[t][axis], lo = A->lo[t][axis];
        const int split = A->split[t];
        int n = 0;
        for (int i = b + tid; i < e; i += 256) n += (bin_of(item_c(it, axis, i), lo, scale) <= split) ? 1 : 0;
        for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
        if (lane == 0) s_w[wv] = n;
        __syncthreads();
        if (tid == 0) chunk_left[g] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        __syncthreads();
    }
}
// stable partition of each chunk into tmp: its left items after the left
// items of the node's earlier chunks, its right items after all left items
// and the right items of the earlier chunks
__global__ void __launch_bounds__(256) k_wide_scatter(int ntask, const int32_t* tb, const int32_t* te, Items it, Items tmp,
                                                      WideAcc* A, const int32_t* chunk_left) {
    __shared__ int32_t s_first[kWideTasks + 1];
    __shared__ int32_t s_w[2][kSplitWaves];
    __shared__ int32_t s_sum[2][kSplitWaves];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int total = wide_chunks(ntask, tb, te, s_first);
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const int t = wide_task(s_first, g);
        const int axis = A->axis[t];
        if (axis < 0) continue;          // uniform over the workgroup
        const int k = g - s_first[t];
        const int tb0 = tb[t];
        const int b = tb0 + k * kWideChunk, e = min(b + kWideChunk, te[t]);
        // left items before this chunk and in the whole node
        int before = 0, all = 0;
        for (int q = s_first[t] + tid; q < s_first[t + 1]; q += 256) {
            const int c = chunk_left[q];
            all += c;
            before += (q < g) ? c : 0;
        }
        for (int off = 32; off > 0; off >>= 1) {
            before += __shfl_xor(before, off);
            all += __shfl_xor(all, off);
        }
        if (lane == 0) { s_sum[0][wv] = before; s_sum[1][wv] = all; }
        __syncthreads();
        const int lbefore = s_sum[0][0] + s_sum[0][1] + s_sum[0][2] + s_sum[0][3];
        const int nleft = s_sum[1][0] + s_sum[1][1] + s_sum[1][2] + s_sum[1][3];
        if (k == 0 && tid == 0) A->nleft[t] = nleft;
        const float scale = (float)kBins / A->ext[t][axis], lo = A->lo[t][axis];
        const int split = A->split[t];
        int lrun = tb0 + lbefore;                          // next left slot
        int rrun = tb0 + nleft + (b - tb0 - lbefore);      // next right slot
        for (int start = b; start < e; start += 256) {
            const int i = start + tid;
            bool has = i < e, left = false;
            if (has) left = bin_of(item_c(it, axis, i), lo, scale) <= split;
            const unsigned long long ml = __ballot(has && left), mr = __ballot(has && !left);
            if (lane == 0) { s_w[0][wv] = __popcll(ml); s_w[1][wv] = __popcll(mr); }
            __syncthreads();
            int lb = 0, rb = 0, lt = 0, rt = 0;
            for (int q = 0; q < kSplitWaves; q++) {
                lb += (q < wv) ? s_w[0][q] : 0;
                rb += (q < wv) ? s_w[1][q] : 0;
                lt += s_w[0][q];
                rt += s_w[1][q];
            }
            if (has) {
                const int dst = left ? lrun + lb + __popcll(ml & below) : rrun + rb + __popcll(mr & below);
                copy_item(it, i, tmp, dst);
            }
            lrun += lt;
            rrun += rt;
            __syncthreads();
        }
    }
}
// copy the partitioned chunks back and reduce the child boxes
__global__ void __launch_bounds__(256) k_wide_boxes(int ntask, const int32_t* tb, const int32_t* te, Items it, Items tmp,
                                                    WideAcc* A) {
    __shared__ int32_t s_first[kWideTasks + 1];
    __shared__ uint32_t s_box[2][6];
    const int tid = threadIdx.x, lane = tid & 63;
    const uint32_t kinf = fkey(__builtin_huge_valf()), kninf = fkey(-__builtin_huge_valf());
    const int total = wide_chunks(ntask, tb, te, s_first);
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const int t = wide_task(s_first, g);
        const int axis = A->axis[t];
        const int b = tb[t] + (g - s_first[t]) * kWideChunk, e = min(b + kWideChunk, te[t]);
        const int mid = axis < 0 ? (tb[t] + te[t]) / 2 : tb[t] + A->nleft[t];
        if (tid < 12) (&s_box[0][0])[tid] = ((tid % 6) < 3) ? kinf : kninf;
        __syncthreads();
        uint32_t mn[2][3], mx[2][3];
        for (int sd = 0; sd < 2; sd++)
            for (int q = 0; q < 3; q++) { mn[sd][q] = kinf; mx[sd][q] = kninf; }
        for (int i = b + tid; i < e; i += 256) {
            DBox ib;
            if (axis >= 0) {
                ib = ldbox(tmp.box, i);
                copy_item(tmp, i, it, i);
            } else {
                ib = ldbox(it.box, i);
            }
            const int side = (i < mid) ? 0 : 1;
            for (int q = 0; q < 3; q++) {
                mn[side][q] = min(mn[side][q], fkey(ib.lo[q]));
                mx[side][q] = max(mx[side][q], fkey(ib.hi[q]));
            }
        }
        for (int sd = 0; sd < 2; sd++)
            for (int q = 0; q < 3; q++) {
                const uint32_t wmn = wave_min_u32(mn[sd][q]), wmx = wave_max_u32(mx[sd][q]);
                if (lane == 0) {
                    atomicMin(&s_box[sd][q], wmn);
                    atomicMax(&s_box[sd][3 + q], wmx);
                }
            }
        __syncthreads();
        if (tid < 12) {
            const int sd = tid / 6, q = tid % 6;
            if (q < 3) atomicMin(&A->box[t][sd][q], s_box[sd][q]);
            else atomicMax(&A->box[t][sd][q], s_box[sd][q]);
        }
        __syncthreads();
    }
}
__global__ void __launch_bounds__(64) k_wide_final(int ntask, const int32_t* tb, const int32_t* te, const WideAcc* A,
                                                   TopSplit* out) {
    const int t = blockIdx.x, tid = threadIdx.x;
    if (t >= ntask) return;
    if (tid < 12) out[t].box[tid / 6][tid % 6] = fromkey(A->box[t][tid / 6][tid % 6]);
    if (tid == 0) out[t].mid = A->axis[t] < 0 ? (tb[t] + te[t]) / 2 : tb[t] + A->nleft[t];
}

// Levels of many small nodes: one wave per node (four per workgroup), each
// with its own LDS bins; the same bounds, bins, split, stable partition and
// child boxes as k_top_split, with wave reductions and ballots in place of the
// workgroup's.  A node of any size is handled (in rounds of 64 items).
constexpr int kWaveSplitMin = 1024;      // levels of more nodes than this use the wave form
__global__ void __launch_bounds__(256) k_top_split_wave(int ntask, Items it, Items tmp, const int32_t* task_b,
                                                        const int32_t* task_e, TopSplit* out) {
    __shared__ uint32_t s_bin[kSplitWaves][3][kBins][6];
    __shared__ int32_t s_cnt[kSplitWaves][3][kBins];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int t = blockIdx.x * kSplitWaves + wv;
    const bool active = t < ntask;       // uniform over the wave
    const int b = active ? task_b[t] : 0, e = active ? task_e[t] : 0;
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint32_t kinf = fkey(__builtin_huge_valf()), kninf = fkey(-__builtin_huge_valf());
    for (int j = lane; j < 3 * kBins * 6; j += 64) (&s_bin[wv][0][0][0])[j] = ((j % 6) < 3) ? kinf : kninf;
    for (int j = lane; j < 3 * kBins; j += 64) (&s_cnt[wv][0][0])[j] = 0;
    // centroid bounds
    float lo[3], ext[3];
    {
        uint32_t mn[3] = {kinf, kinf, kinf}, mx[3] = {kninf, kninf, kninf};
        for (int i = b + lane; i < e; i += 64)
            for (int a = 0; a < 3; a++) {
                const uint32_t k = fkey(item_c(it, a, i));
                mn[a] = min(mn[a], k);
                mx[a] = max(mx[a], k);
            }
        for (int a = 0; a < 3; a++) {
            lo[a] = fromkey(wave_min_u32(mn[a]));
            ext[a] = fromkey(wave_max_u32(mx[a])) - lo[a];
        }
    }
    __syncthreads();
    for (int i = b + lane; i < e; i += 64) {
        const DBox ib = ldbox(it.box, i);
        const int cnt = it.count[i];
        for (int a = 0; a < 3; a++) {
            if (!(ext[a] > 0.0f)) continue;
            const float scale = (float)kBins / ext[a];
            const int bi = bin_of(item_c(it, a, i), lo[a], scale);
            for (int q = 0; q < 3; q++) {
                atomicMin(&s_bin[wv][a][bi][q], fkey(ib.lo[q]));
                atomicMax(&s_bin[wv][a][bi][3 + q], fkey(ib.hi[q]));
            }
            atomicAdd(&s_cnt[wv][a][bi], cnt);
        }
    }
    __syncthreads();
    int32_t axis = -1, split = -1;
    if (active) sah_choose_wave(s_bin[wv], s_cnt[wv], ext, &axis, &split);
    int mid = (b + e) / 2;
    if (axis >= 0) {
        const float scale = (float)kBins / ext[axis], l0 = lo[axis];
        int nleft = 0;
        for (int start = b; start < e; start += 64) {
            const int i = start + lane;
            const bool left = i < e && bin_of(item_c(it, axis, i), l0, scale) <= split;
            nleft += __popcll(__ballot(left));
        }
        int lrun = b, rrun = b + nleft;
        for (int start = b; start < e; start += 64) {
            const int i = start + lane;
            const bool has = i < e;
            const bool left = has && bin_of(item_c(it, axis, i), l0, scale) <= split;
            const unsigned long long ml = __ballot(left), mr = __ballot(has && !left);
            if (has) copy_item(it, i, tmp, left ? lrun + __popcll(ml & below) : rrun + __popcll(mr & below));
            lrun += __popcll(ml);
            rrun += __popcll(mr);
        }
        mid = b + nleft;
    }
    __syncthreads();      // the partition's global writes are visible to the wave's other lanes
    uint32_t mn[2][3], mx[2][3];
    for (int sd = 0; sd < 2; sd++)
        for (int q = 0; q < 3; q++) { mn[sd][q] = kinf; mx[sd][q] = kninf; }
    for (int i = b + lane; i < e; i += 64) {
        DBox ib;
        if (axis >= 0) {
            ib = ldbox(tmp.box, i);
            copy_item(tmp, i, it, i);
        } else {
            ib = ldbox(it.box, i);
        }
        const int side = (i < mid) ? 0 : 1;
        for (int q = 0; q < 3; q++) {
            mn[side][q] = min(mn[side][q], fkey(ib.lo[q]));
            mx[side][q] = max(mx[side][q], fkey(ib.hi[q]));
        }
    }
    float bx[12];
    for (int sd = 0; sd < 2; sd++)
        for (int q = 0; q < 3; q++) {
            bx[sd * 6 + q] = fromkey(wave_min_u32(mn[sd][q]));
            bx[sd * 6 + 3 + q] = fromkey(wave_max_u32(mx[sd][q]));
        }
    if (active && lane < 12) {
        float v = bx[0];
        for (int j = 1; j < 12; j++) v = (lane == j) ? bx[j] : v;
        out[t].box[lane / 6][lane % 6] = v;
    }
    if (active && lane == 0) out[t].mid = mid;
}

// Numbering of a level's inner children as an exclusive scan over the nodes:
// k_link_count gives each node its number of inner children (child ranges of
// more than one item), the scan their first number, k_link_write the children
// (breadth-first numbering, left before right, as k_top_link), k_link_finish
// the next level's size and the next free node id.
__global__ void __launch_bounds__(256) k_link_count(int ntask, const int32_t* task_b, const int32_t* task_e,
                                                    const TopSplit* sp, int32_t* inner) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntask) return;
    const int b = task_b[t], e = task_e[t], mid = sp[t].mid;
    inner[t] = (mid - b > 1 ? 1 : 0) + (e - mid > 1 ? 1 : 0);
}
__global__ void __launch_bounds__(256) k_link_write(int ntask, const int32_t* task_b, const int32_t* task_e,
                                                    const int32_t* task_node, const TopSplit* sp, const int32_t* item_ref,
                                                    const int32_t* first, BvhNode* nodes, int32_t* next_b, int32_t* next_e,
                                                    int32_t* next_node, const int32_t* next_id) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntask) return;
    const int b = task_b[t], e = task_e[t], mid = sp[t].mid;
    const int rb[2] = {b, mid}, re[2] = {mid, e};
    const int32_t id0 = *next_id;
    int q = first[t];
    BvhNode* nd = nodes + task_node[t];
    for (int side = 0; side < 2; side++) {
        DBox bx;
        for (int k = 0; k < 3; k++) { bx.lo[k] = sp[t].box[side][k]; bx.hi[k] = sp[t].box[side][3 + k]; }
        int32_t ref;
        if (re[side] - rb[side] == 1) {
            ref = item_ref[rb[side]];
        } else {
            ref = id0 + q;
            next_b[q] = rb[side]; next_e[q] = re[side]; next_node[q] = ref;
            q++;
        }
        set_child(nd, side, bx, ref);
    }
}
__global__ void k_link_finish(int ntask, const int32_t* inner, const int32_t* first, int32_t* next_id, int32_t* next_count) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int32_t n = first[ntask - 1] + inner[ntask - 1];
    *next_count = n;
    *next_id += n;
}

// Numbers the inner children of one level (one workgroup; `next_id` holds the
// next free top-level node id) and writes the parents' children.
__global__ void __launch_bounds__(kTopThreads) k_top_link(int ntask, const int32_t* task_b, const int32_t* task_e,
                                                          const int32_t* task_node, const TopSplit* sp, const int32_t* item_ref,
                                                          BvhNode* nodes, int32_t* next_b, int32_t* next_e,
                                                          int32_t* next_node, int32_t* next_id, int32_t* next_count) {
    __shared__ int32_t s_scan[kTopThreads];
    __shared__ int32_t s_base, s_id;
    const int tid = threadIdx.x;
    if (tid == 0) { s_base = 0; s_id = *next_id; }
    __syncthreads();
    for (int start = 0; start < ntask; start += kTopThreads) {
        const int t = start + tid;
        int rb[2] = {0, 0}, re[2] = {0, 0}, inner = 0;
        if (t < ntask) {
            const int b = task_b[t], e = task_e[t], mid = sp[t].mid;
            rb[0] = b; re[0] = mid; rb[1] = mid; re[1] = e;
            inner = (re[0] - rb[0] > 1 ? 1 : 0) + (re[1] - rb[1] > 1 ? 1 : 0);
        }
        s_scan[tid] = inner;
        __syncthreads();
        for (int off = 1; off < kTopThreads; off <<= 1) {
            int v = (tid >= off) ? s_scan[tid - off] : 0;
            __syncthreads();
            s_scan[tid] += v;
            __syncthreads();
        }
        if (t < ntask) {
            int q = s_base + s_scan[tid] - inner;
            BvhNode* nd = nodes + task_node[t];
            for (int side = 0; side < 2; side++) {
                DBox bx;
                for (int k = 0; k < 3; k++) { bx.lo[k] = sp[t].box[side][k]; bx.hi[k] = sp[t].box[side][3 + k]; }
                int32_t ref;
                if (re[side] - rb[side] == 1) {
                    ref = item_ref[rb[side]];
                } else {
                    ref = s_id + q;
                    next_b[q] = rb[side]; next_e[q] = re[side]; next_node[q] = ref;
                    q++;
                }
                set_child(nd, side, bx, ref);
            }
        }
        __syncthreads();
        if (tid == 0) s_base += s_scan[kTopThreads - 1];
        __syncthreads();
    }
    if (tid == 0) {
        *next_count = s_base;
        *next_id = s_id + s_base;
    }
}

// ---------------------------------------------------------------------------
// 7. emit the Karras nodes that remain, triangles, parents, depth
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_emit(int n, int T, int top, int max_leaf, const int32_t* kl, const int32_t* kr,
                                              const int32_t* kfirst, const int32_t* par_int, const uint32_t* order, BoxArr tb,
                                              KArr K, BvhNode* nodes, int32_t* valid) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    if (K.count[i] > T || K.collapse[i]) return;
    for (int p = par_int[i]; p >= 0 && K.count[p] <= max_leaf; p = par_int[p])
        if (K.collapse[p]) return;
    BvhNode* nd = nodes + top + i;
    int32_t ch[2] = {kl[i], kr[i]};
    for (int side = 0; side < 2; side++) {
        int32_t c = ch[side];
        DBox b;
        int32_t ref;
        if (c < 0) { int k = -c - 1; b = ldbox(tb, order[k]); ref = leaf_ref((uint32_t)k, 1u); }
        else {
            b = ldbox(K.box, c);
            ref = K.collapse[c] ? leaf_ref((uint32_t)kfirst[c], (uint32_t)K.count[c]) : top + c;
        }
        set_child(nd, side, b, ref);
    }
    valid[top + i] = 1;
}

// Sorted triangle records for the traversal and the 128-byte shading records
// of the same triangles (positions, material, vertex normals and tangents),
// rewritten after every skinning step.
struct TriSrc {
    const float *pos, *nrm, *tan;
    const uint32_t* idx;
    const int32_t *tri_mesh, *mesh_mat;
};
__global__ void __launch_bounds__(256) k_tris(int n, const uint32_t* order, TriSrc src, float4* tris, float4* rec) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t p = order[k];
    const uint32_t* ix = src.idx + 3 * (size_t)p;
    const float* a = src.pos + 3 * (size_t)ix[0];
    const float* b = src.pos + 3 * (size_t)ix[1];
    const float* c = src.pos + 3 * (size_t)ix[2];
    const int32_t mesh = src.tri_mesh[p];
    const int32_t mat = src.mesh_mat[mesh];
    tris[3 * (size_t)k + 0] = make_float4(a[0], a[1], a[2], __uint_as_float(p));
    tris[3 * (size_t)k + 1] = make_float4(b[0], b[1], b[2], __int_as_float(mesh));
    tris[3 * (size_t)k + 2] = make_float4(c[0], c[1], c[2], __int_as_float(mat));
    // shading record (layout at DevScene::tri_rec)
    float nv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, t[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int v = 0; v < 3; v++) {
        for (int j = 0; j < 3; j++) {
            if (src.nrm) nv[3 * v + j] = src.nrm[3 * (size_t)ix[v] + j];
            if (src.tan) t[3 * v + j] = src.tan[3 * (size_t)ix[v] + j];
        }
    }
    float4* r = rec + 8 * (size_t)k;
    r[0] = make_float4(a[0], a[1], a[2], __int_as_float(mat));
    r[1] = make_float4(b[0], b[1], b[2], nv[0]);
    r[2] = make_float4(c[0], c[1], c[2], nv[1]);
    r[3] = make_float4(nv[2], nv[3], nv[4], nv[5]);
    r[4] = make_float4(nv[6], nv[7], nv[8], t[0]);
    r[5] = make_float4(t[1], t[2], t[3], t[4]);
    r[6] = make_float4(t[5], t[6], t[7], t[8]);
    r[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
static void launch_tris(SceneRec* s, hipStream_t st) {
    TriSrc src;
    src.pos = (const float*)s->pos.p;
    src.nrm = s->has_nrm ? (const float*)s->nrm.p : nullptr;
    src.tan = s->has_tan ? (const float*)s->tan.p : nullptr;
    src.idx = (const uint32_t*)s->idx.p;
    src.tri_mesh = (const int32_t*)s->tri_mesh.p;
    src.mesh_mat = (const int32_t*)s->mesh_mat.p;
    const int n = s->nf;
    k_tris<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, (const uint32_t*)s->order.p, src, (float4*)s->tris.p,
                                                        (float4*)s->tri_rec.p);
}

// parents in the final tree: parent index * 2 + side, -1 for the root
__global__ void __launch_bounds__(256) k_final_parents(int num_nodes, const BvhNode* nodes, const int32_t* valid, int32_t* parent) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_nodes || !valid[i]) return;
    int4 d = nodes[i].d;
    if (d.x >= 0) parent[d.x] = 2 * i;
    if (d.y >= 0) parent[d.y] = 2 * i + 1;
}

__global__ void __launch_bounds__(256) k_depth(int num_nodes, const BvhNode* nodes, const int32_t* valid, const int32_t* parent,
                                               int32_t* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_nodes || !valid[i]) return;
    int4 d = nodes[i].d;
    if (d.x >= 0 && d.y >= 0) return;      // only nodes holding a leaf end a path
    int depth = 1;
    for (int p = parent[i]; p >= 0; p = parent[p >> 1]) depth++;
    atomicMax(out, depth);
}

// ---------------------------------------------------------------------------
// refit: leaf boxes from the current triangles, then boxes bottom-up
// ---------------------------------------------------------------------------
SRT_D DBox leaf_box(const float4* tris, int32_t ref) {
    DBox b = empty_box();
    uint32_t f = leaf_first(ref), c = leaf_count(ref);
    for (uint32_t k = f; k < f + c; k++)
        for (int v = 0; v < 3; v++) {
            float4 p = tris[3 * (size_t)k + v];
            b.lo[0] = kmin(b.lo[0], p.x); b.hi[0] = kmax(b.hi[0], p.x);
            b.lo[1] = kmin(b.lo[1], p.y); b.hi[1] = kmax(b.hi[1], p.y);
            b.lo[2] = kmin(b.lo[2], p.z); b.hi[2] = kmax(b.hi[2], p.z);
        }
    return b;
}
SRT_D DBox node_union(const BvhNode* nd) {
    DBox l, r;
    float4 a = nd->a, b = nd->b, c = nd->c;
    l.lo[0] = a.x; l.hi[0] = a.y; l.lo[1] = a.z; l.hi[1] = a.w; l.lo[2] = c.x; l.hi[2] = c.y;
    r.lo[0] = b.x; r.hi[0] = b.y; r.lo[1] = b.z; r.hi[1] = b.w; r.lo[2] = c.z; r.hi[2] = c.w;
    grow(l, r);
    return l;
}
// Level of every binary-tree node (root = 0; 0xff for unused slots), keyed for
// the sort that orders the nodes by level
__global__ void __launch_bounds__(256) k_node_level(int num_nodes, const int32_t* valid, const int32_t* parent,
                                                    uint32_t* key, uint32_t* idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_nodes) return;
    uint32_t level = 0xffu;
    if (valid[i]) {
        level = 0;
        for (int p = parent[i]; p >= 0; p = parent[p >> 1]) level++;
    }
    key[i] = level;
    idx[i] = (uint32_t)i;
}
// first position of every key in the sorted keys
__global__ void __launch_bounds__(256) k_level_first(int n, const uint32_t* key, int32_t* first) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == 0 || key[i] != key[i - 1]) first[key[i]] = i;
}
// One level of the bottom-up refit (deepest level first): each node takes its
// leaf children's boxes from the current triangles and its inner children's
// boxes as the union of their two child boxes, written by the previous launch.
// Kernel boundaries order the levels, so no device-wide fences are needed.
__global__ void __launch_bounds__(256) k_refit_level(const uint32_t* ids, int count, BvhNode* nodes,
                                                     const float4* tris) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    BvhNode* nd = nodes + ids[k];
    const int4 d = nd->d;
    const DBox l = d.x < 0 ? leaf_box(tris, d.x) : node_union(nodes + d.x);
    const DBox r = d.y < 0 ? leaf_box(tris, d.y) : node_union(nodes + d.y);
    set_child(nd, 0, l, d.x);
    set_child(nd, 1, r, d.y);
}

// ---------------------------------------------------------------------------
// 4-wide tree collapsed from the binary tree.  The binary leaves stay leaves;
// which binary inner nodes become 4-wide nodes is chosen by a bottom-up
// dynamic program that minimises the summed surface area of the 4-wide nodes
// (the surface-area estimate of the node steps of a traversal; the leaves'
// triangle tests do not depend on the choice).  Per binary inner node x:
//   W(x)    = area(x) + G(x, 4)    x as a 4-wide node
//   G(x, k) = min over j in 1..k-1 of H(l, j) + H(r, k - j)
//   H(c, j) = 0 for a leaf, W(c) for j = 1, min(W(c), G(c, j)) for j > 1
// (H(c, j): subtree c in at most j child slots, c kept as one slot or opened).
// The 4-wide nodes are then emitted top-down by levels following the stored
// choices (ties keep a child whole, and give the left child the fewest
// slots).  Nodes are numbered level by level in order (deterministic); boxes
// are copies of binary-tree boxes, so both trees give every ray the same box
// tests.
// ---------------------------------------------------------------------------
SRT_D DBox child_box(const BvhNode& nd, int side) {
    DBox b;
    if (side == 0) {
        b.lo[0] = nd.a.x; b.hi[0] = nd.a.y; b.lo[1] = nd.a.z; b.hi[1] = nd.a.w; b.lo[2] = nd.c.x; b.hi[2] = nd.c.y;
    } else {
        b.lo[0] = nd.b.x; b.hi[0] = nd.b.y; b.lo[1] = nd.b.z; b.hi[1] = nd.b.w; b.lo[2] = nd.c.z; b.hi[2] = nd.c.w;
    }
    return b;
}
struct Kids4 {
    int32_t ref[4], src[4];
    int n;
};
// dynamic-program tables per binary inner node: W, G(k) for k = 2, 3, 4, and
// the slots given to the left child for each k (2 bits each)
struct C4Dp {
    float* W;
    float* G;            // 3 per node
    uint32_t* J;         // bits 2(k-2) .. 2(k-2)+1: left slots for k
};
SRT_D float c4_H(const C4Dp& dp, int32_t c, int j) {
    if (c < 0) return 0.0f;
    const float w = dp.W[c];
    if (j == 1) return w;
    const float g = dp.G[3 * (size_t)c + (j - 2)];
    return g < w ? g : w;
}
__global__ void __launch_bounds__(256) k_c4_dp(const uint32_t* ids, int count, const BvhNode* nodes, C4Dp dp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int32_t x = (int32_t)ids[i];
    const BvhNode nd = nodes[x];
    DBox b = child_box(nd, 0);
    grow(b, child_box(nd, 1));
    const int32_t l = nd.d.x, r = nd.d.y;
    uint32_t jbits = 0;
    float g4 = 0.0f;
    for (int k = 2; k <= 4; k++) {
        float best = __builtin_huge_valf();
        int bj = 1;
        for (int j = 1; j < k; j++) {
            const float v = c4_H(dp, l, j) + c4_H(dp, r, k - j);
            if (v < best) { best = v; bj = j; }
        }
        dp.G[3 * (size_t)x + (k - 2)] = best;
        jbits |= (uint32_t)bj << (2 * (k - 2));
        if (k == 4) g4 = best;
    }
    dp.J[x] = jbits;
    dp.W[x] = box_area(b) + g4;
}
// children of the 4-wide node made from binary node x: subtree (c, j) is kept
// as one slot when c is a leaf, j == 1 or opening it is not cheaper, else
// opened into its children with the stored split of j slots
SRT_D Kids4 collapse_children(const BvhNode* nodes, int32_t x, const C4Dp& dp) {
    Kids4 k;
    k.n = 0;
    for (int j = 0; j < 4; j++) { k.ref[j] = 0; k.src[j] = -1; }
    int32_t st_ref[4], st_src[4], st_j[4];
    int sp = 0;
    const BvhNode nd = nodes[x];
    const int jl = min(max((int)(dp.J[x] >> 4) & 3, 1), 3);      // k = 4; every split gives each side a slot
    st_ref[sp] = nd.d.y; st_src[sp] = (x << 1) | 1; st_j[sp] = 4 - jl; sp++;
    st_ref[sp] = nd.d.x; st_src[sp] = x << 1; st_j[sp] = jl; sp++;
    while (sp > 0) {
        sp--;
        const int32_t c = st_ref[sp], src = st_src[sp];
        const int j = st_j[sp];
        const bool open = c >= 0 && j > 1 && dp.G[3 * (size_t)c + (j - 2)] < dp.W[c];
        if (!open) {
            k.ref[k.n] = c;
            k.src[k.n] = src;
            k.n++;
            continue;
        }
        const BvhNode nc = nodes[c];
        const int cl = min(max((int)(dp.J[c] >> (2 * (j - 2))) & 3, 1), j - 1);
        st_ref[sp] = nc.d.y; st_src[sp] = (c << 1) | 1; st_j[sp] = j - cl; sp++;
        st_ref[sp] = nc.d.x; st_src[sp] = c << 1; st_j[sp] = cl; sp++;
    }
    // slots by decreasing box area (stable): any-hit traversals visit the hit
    // children in slot order, the likeliest to hold an occluder first
    float area[4];
    for (int j = 0; j < k.n; j++) area[j] = box_area(child_box(nodes[k.src[j] >> 1], k.src[j] & 1));
    for (int a = 1; a < k.n; a++)
        for (int b = a; b > 0 && area[b] > area[b - 1]; b--) {
            const float ta = area[b]; area[b] = area[b - 1]; area[b - 1] = ta;
            const int32_t tr = k.ref[b]; k.ref[b] = k.ref[b - 1]; k.ref[b - 1] = tr;
            const int32_t ts = k.src[b]; k.src[b] = k.src[b - 1]; k.src[b - 1] = ts;
        }
    return k;
}
__global__ void __launch_bounds__(256) k_c4_count(int nf, const int32_t* level, const BvhNode* nodes, C4Dp dp,
                                                  uint32_t* ninner) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const Kids4 k = collapse_children(nodes, level[i], dp);
    uint32_t m = 0;
    for (int j = 0; j < k.n; j++) m += (k.ref[j] >= 0) ? 1u : 0u;
    ninner[i] = m;
}
SRT_D void set4(Bvh4Node& o, int j, const DBox& b) {
    float* lx = &o.lox.x; float* hx = &o.hix.x; float* ly = &o.loy.x; float* hy = &o.hiy.x;
    float* lz = &o.loz.x; float* hz = &o.hiz.x;
    lx[j] = b.lo[0]; hx[j] = b.hi[0]; ly[j] = b.lo[1]; hy[j] = b.hi[1]; lz[j] = b.lo[2]; hz[j] = b.hi[2];
}
__global__ void __launch_bounds__(256) k_c4_emit(int nf, const int32_t* level, const BvhNode* nodes, C4Dp dp,
                                                 const uint32_t* off, int base_this, int base_next, Bvh4Node* out,
                                                 int32_t* next_level) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const Kids4 k = collapse_children(nodes, level[i], dp);
    Bvh4Node o;
    int32_t* ref = &o.ref.x;
    int32_t* src = &o.src.x;
    uint32_t q = off[i];
    DBox empty;
    empty.lo[0] = empty.lo[1] = empty.lo[2] = __builtin_huge_valf();
    empty.hi[0] = empty.hi[1] = empty.hi[2] = -__builtin_huge_valf();
    for (int j = 0; j < 4; j++) {
        if (j < k.n) {
            const int32_t sx = k.src[j];
            set4(o, j, child_box(nodes[sx >> 1], sx & 1));
            src[j] = sx;
            if (k.ref[j] >= 0) {
                next_level[q] = k.ref[j];
                ref[j] = base_next + (int32_t)q;
                q++;
            } else {
                ref[j] = k.ref[j];
            }
        } else {
            set4(o, j, empty);
            src[j] = -1;
            ref[j] = 0;
        }
    }
    out[base_this + i] = o;
}
// after a binary-tree refit: each wide child box is copied from its binary slot again
__global__ void __launch_bounds__(256) k_refit4(int n4, const BvhNode* nodes, Bvh4Node* nodes4) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n4) return;
    Bvh4Node o = nodes4[i];
    const int32_t* src = &o.src.x;
    for (int j = 0; j < 4; j++)
        if (src[j] >= 0) set4(o, j, child_box(nodes[src[j] >> 1], src[j] & 1));
    nodes4[i] = o;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// scene.cu:362-392: skinned position / normal / tangent from up to 4 bones
__global__ void __launch_bounds__(256) k_skin(int nv, const float* op, const float* on, const float* ot, const int32_t* bid,
                                              const float* bw, int nb, const float* bones, float* pos, float* nrm, float* tan) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    f3 v = mk(0.0f), nn = mk(0.0f), tt = mk(0.0f);
    f3 ov = ld3(op + 3 * (size_t)i);
    f3 onn = on ? ld3(on + 3 * (size_t)i) : mk(0.0f);
    f3 ott = ot ? ld3(ot + 3 * (size_t)i) : mk(0.0f);
    for (int j = 0; j < 4; j++) {
        int id = bid[4 * (size_t)i + j];
        float w = bw[4 * (size_t)i + j];
        if (id == -1) {
            if (j == 0) { v = ov; nn = onn; tt = ott; }
            break;
        }
        if (id >= nb) continue;
        const float* m = bones + 12 * (size_t)id;
        f3 pp = mk(m[0] * ov.x + m[1] * ov.y + m[2] * ov.z + m[3], m[4] * ov.x + m[5] * ov.y + m[6] * ov.z + m[7],
                   m[8] * ov.x + m[9] * ov.y + m[10] * ov.z + m[11]);
        v = v + pp * w;
        f3 pn = mk(m[0] * onn.x + m[1] * onn.y + m[2] * onn.z, m[4] * onn.x + m[5] * onn.y + m[6] * onn.z,
                   m[8] * onn.x + m[9] * onn.y + m[10] * onn.z);
        nn = nn + pn * w;
        f3 pt = mk(m[0] * ott.x + m[1] * ott.y + m[2] * ott.z, m[4] * ott.x + m[5] * ott.y + m[6] * ott.z,
                   m[8] * ott.x + m[9] * ott.y + m[10] * ott.z);
        tt = tt + pt * w;
    }
    pos[3 * (size_t)i] = v.x; pos[3 * (size_t)i + 1] = v.y; pos[3 * (size_t)i + 2] = v.z;
    if (nrm) { nrm[3 * (size_t)i] = nn.x; nrm[3 * (size_t)i + 1] = nn.y; nrm[3 * (size_t)i + 2] = nn.z; }
    if (tan) { tan[3 * (size_t)i] = tt.x; tan[3 * (size_t)i + 1] = tt.y; tan[3 * (size_t)i + 2] = tt.z; }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {
// Build scratch: 256-byte aligned slices bumped out of 16 MiB (or larger)
// device chunks.  The chunks go back to the context when the build ends and
// the next build takes them again (the smallest kept chunk that fits), so a
// rebuild of the same scene calls neither hipMalloc nor hipFree: each hipFree
// of a chunk cost about 0.25 ms of host time, a third of a 1M-triangle
// rebuild.  Kernels of a build still queued when its chunks go back are
// ordered before the next build's by the context's stream.
struct Tmp {
    static constexpr size_t kChunk = (size_t)16 << 20;
    srt_ctx* c;
    std::vector<DevBuf> chunks;
    size_t used = 0, cap = 0;      // of the last chunk
    bool failed = false;           // some get() returned nullptr
    explicit Tmp(srt_ctx* cc) : c(cc) {}
    ~Tmp() {
        if (!c->keep_build_scratch) {
            for (auto& b : chunks) dev_free(c, b);
            return;
        }
        std::lock_guard<std::mutex> g(c->scratch_mu);
        for (auto& b : chunks) c->build_scratch.push_back(b);
    }
    bool take_kept(size_t size, DevBuf& out) {
        std::lock_guard<std::mutex> g(c->scratch_mu);
        auto& kept = c->build_scratch;
        int best = -1;
        for (int i = 0; i < (int)kept.size(); i++)
            if (kept[i].bytes >= size && (best < 0 || kept[i].bytes < kept[best].bytes)) best = i;
        if (best < 0) return false;
        out = kept[best];
        kept.erase(kept.begin() + best);
        return true;
    }
    template <class T> T* get(size_t count) {
        const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
        if (chunks.empty() || used + bytes > cap) {
            DevBuf b;
            const size_t size = std::max(bytes, kChunk);
            if (!take_kept(size, b) && dev_alloc(c, b, size)) {
                failed = true;
                return nullptr;
            }
            chunks.push_back(b);
            used = 0;
            cap = b.bytes;
        }
        T* p = (T*)((char*)chunks.back().p + used);
        used += bytes;
        return p;
    }
};
inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)std::max<int64_t>(1, (n + b - 1) / b); }
}  // namespace

int bvh_build(srt_ctx* c, SceneRec* s, int max_leaf, int cluster_size) {
    if (max_leaf < 1 || max_leaf > kMaxLeaf) return set_err(c, SRT_E_LIMIT, "max_leaf must be in [1, %d]", kMaxLeaf);
    const int n = s->nf;
    if (n < 1) return set_err(c, SRT_E_INVALID, "scene has no triangles");
    if ((int64_t)n > kMaxTriangles) return set_err(c, SRT_E_LIMIT, "too many triangles for the leaf encoding");
    hipStream_t st = c->stream;
    // default: clusters of at most 4 triangles under the binned-SAH top (at
    // most 2^18 clusters): 3.5 % faster traversal than 64 on the 1M-triangle
    // field, for a build of about 0.1 us per triangle
    // a cluster holds at least a leaf's triangles: a collapsed (leaf) subtree
    // never spans the cluster frontier
    const int T = std::max(max_leaf, cluster_size > 0 ? cluster_size : std::max(kAutoClusterMin, n / kAutoClusterDiv));
    const float c_trav = 1.0f, c_isect = 1.0f;
    Tmp tmp(c);
    // boxes & centroids
    BoxArr tb;
    tb.lox = tmp.get<float>(n); tb.loy = tmp.get<float>(n); tb.loz = tmp.get<float>(n);
    tb.hix = tmp.get<float>(n); tb.hiy = tmp.get<float>(n); tb.hiz = tmp.get<float>(n);
    float* cen = tmp.get<float>(3 * (size_t)n);
    unsigned nb = nblk(n);
    uint32_t* partial = tmp.get<uint32_t>(6 * (size_t)nb);
    uint32_t* cb6 = tmp.get<uint32_t>(6);
    uint64_t* keys = tmp.get<uint64_t>(n);
    uint64_t* keys2 = tmp.get<uint64_t>(n);
    uint32_t* vals = tmp.get<uint32_t>(n);
    int32_t* kl = tmp.get<int32_t>(n);
    int32_t* kr = tmp.get<int32_t>(n);
    int32_t* kfirst = tmp.get<int32_t>(n);
    int32_t* par_int = tmp.get<int32_t>(n);
    int32_t* par_leaf = tmp.get<int32_t>(n);
    KArr K;
    K.box.lox = tmp.get<float>(n); K.box.loy = tmp.get<float>(n); K.box.loz = tmp.get<float>(n);
    K.box.hix = tmp.get<float>(n); K.box.hiy = tmp.get<float>(n); K.box.hiz = tmp.get<float>(n);
    K.count = tmp.get<int32_t>(n); K.cost = tmp.get<float>(n); K.collapse = tmp.get<int32_t>(n);
    int32_t* tri_mesh = (int32_t*)s->tri_mesh.p;
    if (tmp.failed || !cen || !K.collapse) return set_err(c, SRT_E_OOM, "BVH build scratch exceeds the HBM budget");

    // final outputs
    if (dev_realloc(c, s->order, (size_t)n * 4) || dev_realloc(c, s->tris, (size_t)n * 48)) return SRT_E_OOM;
    uint32_t* order = (uint32_t*)s->order.p;

    k_tri_boxes<<<nb, 256, 0, st>>>((const float*)s->pos.p, (const uint32_t*)s->idx.p, n, tb, cen, partial);
    k_reduce_bounds<<<1, 256, 0, st>>>(partial, (int)nb, cb6);
    k_morton<<<nb, 256, 0, st>>>(cen, cb6, n, keys, vals);
    {
        size_t tbytes = 0;
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, keys, keys2, vals, order, n, 0, 63, st);
        void* tstore = tmp.get<char>(tbytes);
        if (tmp.failed || !tstore) return set_err(c, SRT_E_OOM, "sort scratch");
        if (int r = hip_check(c, hipcub::DeviceRadixSort::SortPairs(tstore, tbytes, keys, keys2, vals, order, n, 0, 63, st), "radix sort"))
            return r;
    }
    k_fill_i32<<<nblk(n), 256, 0, st>>>(par_int, n, -1);
    if (n > 1) {
        k_karras<<<nblk(n - 1), 256, 0, st>>>(keys2, n, kl, kr, kfirst, par_int, par_leaf);
        // bottom-up by depth: sort the n - 1 radix-tree nodes by depth, then one launch per depth
        const int m = n - 1;
        uint32_t* dkey = tmp.get<uint32_t>(m);
        uint32_t* dkey2 = tmp.get<uint32_t>(m);
        uint32_t* didx = tmp.get<uint32_t>(m);
        uint32_t* dids = tmp.get<uint32_t>(m);
        int32_t* dfirst = tmp.get<int32_t>(kMaxRadixDepth + 1);
        if (tmp.failed || !dkey || !dkey2 || !didx || !dids || !dfirst) return set_err(c, SRT_E_OOM, "depth scratch");
        k_radix_depth<<<nblk(m), 256, 0, st>>>(m, par_int, dkey, didx);
        size_t tbytes = 0;
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, dkey, dkey2, didx, dids, m, 0, 16, st);
        void* tstore = tmp.get<char>(tbytes);
        if (tmp.failed || !tstore) return set_err(c, SRT_E_OOM, "sort scratch");
        if (int r = hip_check(c, hipcub::DeviceRadixSort::SortPairs(tstore, tbytes, dkey, dkey2, didx, dids, m, 0, 16, st),
                              "depth sort"))
            return r;
        uint32_t dmax = 0;
        if (int r = hip_check(c, hipMemcpyAsync(&dmax, dkey2 + (m - 1), 4, hipMemcpyDeviceToHost, st), "copy")) return r;
        if (int r = hip_check(c, hipStreamSynchronize(st), "depth sort")) return r;
        if (dmax >= (uint32_t)kMaxRadixDepth)
            return set_err(c, SRT_E_LIMIT, "radix tree depth %u exceeds %d", dmax, kMaxRadixDepth);
        k_fill_i32<<<nblk(dmax + 1), 256, 0, st>>>(dfirst, (int64_t)dmax + 1, -1);
        k_level_first<<<nblk(m), 256, 0, st>>>(m, dkey2, dfirst);
        std::vector<int32_t> df(dmax + 2);
        if (int r = hip_check(c, hipMemcpyAsync(df.data(), dfirst, 4 * (size_t)(dmax + 1), hipMemcpyDeviceToHost, st), "copy"))
            return r;
        if (int r = hip_check(c, hipStreamSynchronize(st), "depth order")) return r;
        df[dmax + 1] = m;
        for (int D = (int)dmax; D >= 0; D--) {
            if (df[D] < 0) return set_err(c, SRT_E_STATE, "radix tree depth %d is empty", D);
            const int count = df[D + 1] - df[D];
            k_bottom_up_level<<<nblk(count), 256, 0, st>>>(dids + df[D], count, order, tb, kl, kr, K, max_leaf, c_trav,
                                                           c_isect);
        }
    }
    // host needs a few scalars: root count/collapse
    int32_t root_count = 1, root_collapse = 0, root_first = 0;
    if (n > 1) {
        hip_check(c, hipMemcpyAsync(&root_count, K.count, 4, hipMemcpyDeviceToHost, st), "copy");
        hip_check(c, hipMemcpyAsync(&root_collapse, K.collapse, 4, hipMemcpyDeviceToHost, st), "copy");
        hip_check(c, hipMemcpyAsync(&root_first, kfirst, 4, hipMemcpyDeviceToHost, st), "copy");
    }
    if (int r = hip_check(c, hipStreamSynchronize(st), "bvh build (sort/tree)")) return r;

    // clusters
    int m = 1;
    int32_t* clusters = nullptr;
    if (n > 1 && n > T) {
        int32_t* mark = tmp.get<int32_t>(n);
        clusters = tmp.get<int32_t>(n);
        int32_t* nsel = tmp.get<int32_t>(1);
        if (tmp.failed || !nsel) return set_err(c, SRT_E_OOM, "cluster scratch");
        k_fill_i32<<<nblk(n), 256, 0, st>>>(mark, n, 0x7fffffff);
        k_frontier<<<nblk(2 * (int64_t)n - 1), 256, 0, st>>>(n, T, kfirst, K.count, par_int, par_leaf, mark);
        struct NotSentinel {
            __host__ __device__ bool operator()(const int32_t& x) const { return x != 0x7fffffff; }
        };
        size_t tbytes = 0;
        (void)hipcub::DeviceSelect::If(nullptr, tbytes, mark, clusters, nsel, n, NotSentinel(), st);
        void* tstore = tmp.get<char>(tbytes);
        if (tmp.failed || !tstore) return set_err(c, SRT_E_OOM, "select scratch");
        if (int r = hip_check(c, hipcub::DeviceSelect::If(tstore, tbytes, mark, clusters, nsel, n, NotSentinel(), st), "select"))
            return r;
        hip_check(c, hipMemcpyAsync(&m, nsel, 4, hipMemcpyDeviceToHost, st), "copy");
        if (int r = hip_check(c, hipStreamSynchronize(st), "cluster select")) return r;
    }
    const int top = m - 1;
    const int num_nodes = std::max(top + std::max(n - 1, 0), 1);
    if (dev_realloc(c, s->nodes, (size_t)num_nodes * sizeof(BvhNode)) ||
        dev_realloc(c, s->node_parent, (size_t)num_nodes * 4) || dev_realloc(c, s->level_nodes, (size_t)num_nodes * 4) ||
        dev_realloc(c, s->node_valid, (size_t)num_nodes * 4))
        return SRT_E_OOM;
    BvhNode* nodes = (BvhNode*)s->nodes.p;
    int32_t* valid = (int32_t*)s->node_valid.p;
    k_fill_i32<<<nblk((int64_t)num_nodes * 16), 256, 0, st>>>((int32_t*)nodes, (int64_t)num_nodes * 16, 0);
    k_fill_i32<<<nblk(num_nodes), 256, 0, st>>>(valid, num_nodes, 0);

    int32_t root;
    if (m > 1) {
        Items it, it2;
        auto mk_items = [&](Items& x) {
            x.box.lox = tmp.get<float>(m); x.box.loy = tmp.get<float>(m); x.box.loz = tmp.get<float>(m);
            x.box.hix = tmp.get<float>(m); x.box.hiy = tmp.get<float>(m); x.box.hiz = tmp.get<float>(m);
            x.cx = tmp.get<float>(m); x.cy = tmp.get<float>(m); x.cz = tmp.get<float>(m);
            x.count = tmp.get<int32_t>(m); x.ref = tmp.get<int32_t>(m);
        };
        mk_items(it);
        mk_items(it2);
        int32_t* lb[2] = {tmp.get<int32_t>(m), tmp.get<int32_t>(m)};
        int32_t* le[2] = {tmp.get<int32_t>(m), tmp.get<int32_t>(m)};
        int32_t* ln[2] = {tmp.get<int32_t>(m), tmp.get<int32_t>(m)};
        TopSplit* sp = tmp.get<TopSplit>(m);
        WideAcc* wacc = tmp.get<WideAcc>(1);
        int32_t* link_inner = tmp.get<int32_t>(m);
        int32_t* link_first = tmp.get<int32_t>(m);
        size_t link_bytes = 0;
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, link_bytes, link_inner, link_first, m, st);
        void* link_tmp = tmp.get<char>(link_bytes);
        int32_t* chunk_left = tmp.get<int32_t>((size_t)m / kWideChunk + kWideTasks + 1);
        int32_t* ctr = tmp.get<int32_t>(2);     // next free node id, next level size
        if (tmp.failed || !ctr || !link_tmp) return set_err(c, SRT_E_OOM, "top-level scratch");
        k_items<<<nblk(m), 256, 0, st>>>(m, clusters, order, tb, K, kfirst, top, it);
        // level 0: the root (node 0) over all items; ids 1.. follow breadth-first
        k_fill_i32<<<1, 64, 0, st>>>(lb[0], 1, 0);
        k_fill_i32<<<1, 64, 0, st>>>(le[0], 1, m);
        k_fill_i32<<<1, 64, 0, st>>>(ln[0], 1, 0);
        k_fill_i32<<<1, 64, 0, st>>>(ctr, 1, 1);
        int ntask = 1, cur = 0;
        while (ntask > 0) {
            if (ntask <= kWideTasks) {
                // chunks of the level: at most m / kWideChunk + one partial chunk per node
                const int grid = std::min(kWideMaxGrid, m / kWideChunk + ntask);
                k_wide_init<<<8, 256, 0, st>>>(ntask, wacc);
                k_wide_bounds<<<grid, 256, 0, st>>>(ntask, lb[cur], le[cur], it, wacc);
                k_wide_bins<<<grid, 256, 0, st>>>(ntask, lb[cur], le[cur], it, wacc);
                k_wide_sah<<<ntask, 64, 0, st>>>(wacc);
                k_wide_count<<<grid, 256, 0, st>>>(ntask, lb[cur], le[cur], it, wacc, chunk_left);
                k_wide_scatter<<<grid, 256, 0, st>>>(ntask, lb[cur], le[cur], it, it2, wacc, chunk_left);
                k_wide_boxes<<<grid, 256, 0, st>>>(ntask, lb[cur], le[cur], it, it2, wacc);
                k_wide_final<<<ntask, 64, 0, st>>>(ntask, lb[cur], le[cur], wacc, sp);
            } else if (ntask <= kWaveSplitMin) {
                k_top_split<<<ntask, kSplitThreads, 0, st>>>(it, it2, lb[cur], le[cur], sp);
            } else {
                k_top_split_wave<<<(ntask + kSplitWaves - 1) / kSplitWaves, 256, 0, st>>>(ntask, it, it2, lb[cur], le[cur],
                                                                                           sp);
            }
            if (ntask <= kTopThreads) {
                k_top_link<<<1, kTopThreads, 0, st>>>(ntask, lb[cur], le[cur], ln[cur], sp, it.ref, nodes, lb[cur ^ 1],
                                                      le[cur ^ 1], ln[cur ^ 1], ctr, ctr + 1);
            } else {
                k_link_count<<<nblk(ntask), 256, 0, st>>>(ntask, lb[cur], le[cur], sp, link_inner);
                if (int r = hip_check(c, hipcub::DeviceScan::ExclusiveSum(link_tmp, link_bytes, link_inner, link_first, ntask, st),
                                      "link scan"))
                    return r;
                k_link_write<<<nblk(ntask), 256, 0, st>>>(ntask, lb[cur], le[cur], ln[cur], sp, it.ref, link_first, nodes,
                                                          lb[cur ^ 1], le[cur ^ 1], ln[cur ^ 1], ctr);
                k_link_finish<<<1, 64, 0, st>>>(ntask, link_inner, link_first, ctr, ctr + 1);
            }
            int32_t next = 0;
            if (int r = hip_check(c, hipMemcpyAsync(&next, ctr + 1, 4, hipMemcpyDeviceToHost, st), "top level")) return r;
            if (int r = hip_check(c, hipStreamSynchronize(st), "top level")) return r;
            if (next > m) return set_err(c, SRT_E_LIMIT, "top-level build overflow");
            ntask = next;
            cur ^= 1;
        }
        k_fill_i32<<<nblk(top), 256, 0, st>>>(valid, top, 1);
        root = 0;
    } else if (n == 1) {
        root = leaf_ref(0u, 1u);
    } else {
        root = root_collapse ? leaf_ref((uint32_t)root_first, (uint32_t)root_count) : top + 0;
    }
    if (n > 1)
        k_emit<<<nblk(n - 1), 256, 0, st>>>(n, T, top, max_leaf, kl, kr, kfirst, par_int, order, tb, K, nodes, valid);
    if (dev_realloc(c, s->tri_rec, (size_t)n * 8 * sizeof(float4))) return SRT_E_OOM;
    launch_tris(s, st);
    int32_t* parent = (int32_t*)s->node_parent.p;
    k_fill_i32<<<nblk(num_nodes), 256, 0, st>>>(parent, num_nodes, -1);
    k_final_parents<<<nblk(num_nodes), 256, 0, st>>>(num_nodes, nodes, valid, parent);
    int32_t* dmax = tmp.get<int32_t>(1);
    k_fill_i32<<<1, 64, 0, st>>>(dmax, 1, 0);
    k_depth<<<nblk(num_nodes), 256, 0, st>>>(num_nodes, nodes, valid, parent, dmax);
    int32_t depth = 0;
    hip_check(c, hipMemcpyAsync(&depth, dmax, 4, hipMemcpyDeviceToHost, st), "copy");
    if (int r = hip_check(c, hipGetLastError(), "bvh launch")) return r;
    if (int r = hip_check(c, hipStreamSynchronize(st), "bvh build")) return r;
    if (root < 0) depth = 0;
    if (depth > kMaxDepth) return set_err(c, SRT_E_LIMIT, "BVH depth %d exceeds the traversal limit %d", depth, kMaxDepth);
    // nodes by level for the refit: sort (level, node) pairs, invalid nodes last
    s->level_first.assign(1, 0);
    if (root >= 0) {
        uint32_t* lkey = tmp.get<uint32_t>(num_nodes);
        uint32_t* lkey2 = tmp.get<uint32_t>(num_nodes);
        uint32_t* lidx = tmp.get<uint32_t>(num_nodes);
        int32_t* first = tmp.get<int32_t>(256);
        if (tmp.failed || !lkey || !lkey2 || !lidx || !first) return set_err(c, SRT_E_OOM, "level scratch");
        k_node_level<<<nblk(num_nodes), 256, 0, st>>>(num_nodes, valid, parent, lkey, lidx);
        size_t tbytes = 0;
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, lkey, lkey2, lidx, (uint32_t*)s->level_nodes.p, num_nodes,
                                                 0, 8, st);
        void* tstore = tmp.get<char>(tbytes);
        if (tmp.failed || !tstore) return set_err(c, SRT_E_OOM, "sort scratch");
        if (int r = hip_check(c, hipcub::DeviceRadixSort::SortPairs(tstore, tbytes, lkey, lkey2, lidx,
                                                                    (uint32_t*)s->level_nodes.p, num_nodes, 0, 8, st),
                              "level sort"))
            return r;
        k_fill_i32<<<1, 256, 0, st>>>(first, 256, -1);
        k_level_first<<<nblk(num_nodes), 256, 0, st>>>(num_nodes, lkey2, first);
        std::vector<int32_t> hf(256);
        if (int r = hip_check(c, hipMemcpyAsync(hf.data(), first, 256 * 4, hipMemcpyDeviceToHost, st), "copy")) return r;
        if (int r = hip_check(c, hipStreamSynchronize(st), "level order")) return r;
        // levels 0 .. depth-1 hold the inner nodes; 0xff marks unused node slots
        s->level_first.assign((size_t)depth + 1, 0);
        for (int L = 0; L < depth; L++) {
            if (hf[L] < 0) return set_err(c, SRT_E_STATE, "BVH level %d is empty", L);
            s->level_first[L] = hf[L];
        }
        for (int L = depth; L < 255; L++)
            if (hf[L] >= 0) return set_err(c, SRT_E_STATE, "BVH node below the deepest level");
        s->level_first[depth] = hf[255] >= 0 ? hf[255] : num_nodes;
    }
    s->root = root;
    s->num_nodes = num_nodes;
    s->max_depth = depth;
    s->num_clusters = m;
    s->num_top = top;
    s->cluster_size = T;
    s->max_leaf = max_leaf;
    if (int r = bvh4_build(c, s)) return r;
    s->has_bvh = true;
    return SRT_OK;
}

int bvh4_build(srt_ctx* c, SceneRec* s) {
    s->num_nodes4 = 0;
    s->depth4 = 0;
    if (s->root < 0) {            // a single leaf
        dev_free(c, s->nodes4);
        s->root4 = s->root;
        return SRT_OK;
    }
    hipStream_t st = c->stream;
    const int n2 = s->num_nodes;
    if (dev_realloc(c, s->nodes4, (size_t)n2 * sizeof(Bvh4Node))) return SRT_E_OOM;
    Tmp tmp(c);
    int32_t* lev[2] = {tmp.get<int32_t>(n2), tmp.get<int32_t>(n2)};
    uint32_t* ninner = tmp.get<uint32_t>(n2);
    uint32_t* off = tmp.get<uint32_t>(n2);
    uint32_t* tot = tmp.get<uint32_t>(2);
    if (!lev[0] || !lev[1] || !ninner || !off || !tot) return SRT_E_OOM;
    size_t scan_bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, ninner, off, n2, st);
    void* scan_tmp = tmp.get<uint8_t>(scan_bytes);
    C4Dp dp;
    dp.W = tmp.get<float>(n2);
    dp.G = tmp.get<float>(3 * (size_t)n2);
    dp.J = tmp.get<uint32_t>(n2);
    if (!scan_tmp || !dp.W || !dp.G || !dp.J) return SRT_E_OOM;
    // the dynamic program, deepest level first (levels 0 .. depth-1 hold the inner nodes)
    const int nlev = (int)s->level_first.size() - 1;
    if (nlev < 1) return set_err(c, SRT_E_STATE, "4-wide collapse without the binary tree's levels");
    for (int L = nlev - 1; L >= 0; L--) {
        const int first = s->level_first[L], count = s->level_first[L + 1] - first;
        k_c4_dp<<<nblk(count), 256, 0, st>>>((const uint32_t*)s->level_nodes.p + first, count, (const BvhNode*)s->nodes.p,
                                             dp);
    }
    k_fill_i32<<<1, 64, 0, st>>>(lev[0], 1, s->root);
    Bvh4Node* out = (Bvh4Node*)s->nodes4.p;
    int base = 0, nf = 1, cur = 0, levels = 0;
    while (nf > 0) {
        if (base + nf > n2) return set_err(c, SRT_E_LIMIT, "4-wide collapse exceeded the binary node count");
        if (!nodes4_fit((int64_t)base + nf, c->max_wide_nodes)) {
            // beyond 32-bit buffer offsets: keep the binary tree only
            dev_free(c, s->nodes4);
            s->root4 = s->root;
            s->num_nodes4 = 0;
            s->depth4 = 0;
            return SRT_OK;
        }
        k_c4_count<<<nblk(nf), 256, 0, st>>>(nf, lev[cur], (const BvhNode*)s->nodes.p, dp, ninner);
        if (int r = hip_check(c, hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, ninner, off, nf, st), "scan"))
            return r;
        k_c4_emit<<<nblk(nf), 256, 0, st>>>(nf, lev[cur], (const BvhNode*)s->nodes.p, dp, off, base, base + nf, out,
                                            lev[cur ^ 1]);
        uint32_t last[2] = {0, 0};
        if (int r = hip_check(c, hipMemcpyAsync(&last[0], off + nf - 1, 4, hipMemcpyDeviceToHost, st), "count")) return r;
        if (int r = hip_check(c, hipMemcpyAsync(&last[1], ninner + nf - 1, 4, hipMemcpyDeviceToHost, st), "count")) return r;
        if (int r = hip_check(c, hipStreamSynchronize(st), "collapse")) return r;
        base += nf;
        nf = (int)(last[0] + last[1]);
        cur ^= 1;
        levels++;
    }
    s->root4 = 0;
    s->num_nodes4 = base;
    s->depth4 = levels;
    return hip_check(c, hipGetLastError(), "collapse launch");
}

int bvh_refit(srt_ctx* c, SceneRec* s) {
    if (!s->has_bvh) return set_err(c, SRT_E_STATE, "refit before build");
    hipStream_t st = c->stream;
    const int n = s->nf;
    launch_tris(s, st);
    if (s->root >= 0) {
        const int levels = (int)s->level_first.size() - 1;
        for (int L = levels - 1; L >= 0; L--) {
            const int first = s->level_first[L], count = s->level_first[L + 1] - first;
            k_refit_level<<<nblk(count), 256, 0, st>>>((const uint32_t*)s->level_nodes.p + first, count,
                                                       (BvhNode*)s->nodes.p, (const float4*)s->tris.p);
        }
        if (s->num_nodes4 > 0)
            k_refit4<<<nblk(s->num_nodes4), 256, 0, st>>>(s->num_nodes4, (const BvhNode*)s->nodes.p,
                                                          (Bvh4Node*)s->nodes4.p);
    }
    if (int r = hip_check(c, hipGetLastError(), "refit launch")) return r;
    return hip_check(c, hipStreamSynchronize(st), "refit");
}

int skin_vertices(srt_ctx* c, SceneRec* s, const float* bones_3x4, int nb) {
    if (!s->bone_ids.p) return set_err(c, SRT_E_STATE, "no skin weights set on this scene");
    if (nb < 0 || nb > 4096) return set_err(c, SRT_E_LIMIT, "bone count %d outside [0, 4096]", nb);
    hipStream_t st = c->stream;
    if (s->nbones < nb || !s->bones.p) {
        dev_free(c, s->bones);
        if (dev_alloc(c, s->bones, (size_t)std::max(nb, 1) * 48)) return SRT_E_OOM;
        s->nbones = nb;
    }
    if (nb > 0)
        if (int r = hip_check(c, hipMemcpyAsync(s->bones.p, bones_3x4, (size_t)nb * 48, hipMemcpyHostToDevice, st), "bones upload"))
            return r;
    hipEvent_t ev;
    prof_begin(c, "skin", &ev);
    k_skin<<<nblk(s->nv), 256, 0, st>>>(s->nv, (const float*)s->orig_pos.p, (const float*)s->orig_nrm.p,
                                        (const float*)s->orig_tan.p, (const int32_t*)s->bone_ids.p, (const float*)s->bone_w.p,
                                        nb, (const float*)s->bones.p, (float*)s->pos.p, (float*)s->nrm.p, (float*)s->tan.p);
    prof_end(c, "skin", ev, s->nv);
    return hip_check(c, hipGetLastError(), "skin launch");
}

}  // namespace srt
