// Internal data layout of the SRT HIP library (device buffers in HBM and the
// host-side bookkeeping around them).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <map>
#include <mutex>
#include <atomic>
#include "dev_math.h"
#include "limits.h"
#include "../../include/srt_hip.h"

namespace srt {

// ---------------------------------------------------------------------------
// BVH node: 64 bytes, both children's boxes stored in the parent so one node
// fetch (four 16-byte loads) tests two boxes.
//   a = (L.lo.x, L.hi.x, L.lo.y, L.hi.y)
//   b = (R.lo.x, R.hi.x, R.lo.y, R.hi.y)
//   c = (L.lo.z, L.hi.z, R.lo.z, R.hi.z)
//   d = (left ref, right ref, 0, 0)
// A ref >= 0 is a node index; a ref < 0 is a leaf:
//   bit 31 set | first sorted triangle << 3 | (count - 1), count <= 8.
// ---------------------------------------------------------------------------
struct alignas(16) BvhNode {
    float4 a, b, c;
    int4 d;
};
static_assert(sizeof(BvhNode) == 64, "node is 64 bytes");

// 4-wide node (128 B) collapsed from the binary tree: the four child boxes as
// one float4 per bound (component k = child k), the child refs (node >= 0,
// leaf < 0, as in BvhNode) and, for refit, the binary-tree slot each child box
// is copied from (parent << 1 | side; -1 for an empty child, whose box is
// empty: lo = +inf, hi = -inf, so no ray hits it).
struct alignas(16) Bvh4Node {
    float4 lox, hix, loy, hiy, loz, hiz;
    int4 ref;
    int4 src;
};
static_assert(sizeof(Bvh4Node) == 128, "wide node is 128 bytes");

__host__ __device__ inline int32_t leaf_ref(uint32_t first, uint32_t count) {
    return (int32_t)(0x80000000u | (first << 3) | (count - 1u));
}
__host__ __device__ inline uint32_t leaf_first(int32_t r) { return ((uint32_t)r & 0x7fffffffu) >> 3; }
__host__ __device__ inline uint32_t leaf_count(int32_t r) { return ((uint32_t)r & 7u) + 1u; }

constexpr int kMaxLeaf = 8;
constexpr int kDefaultMaxLeaf = 2;     // triangles per leaf when the caller gives none
constexpr int kAutoClusterMin = 4, kAutoClusterDiv = 262144;   // automatic top-level cluster size
constexpr int kMaxDepth = 96;      // checked after every build
constexpr int kTraceBlock = 128;   // traversal workgroup size (2 waves)

// traversal launch settings (context options "trace_stack_lds", "trace_refill",
// "trace_blocks_per_cu")
struct TraceCfg {
    int32_t lds_depth;   // stack entries per lane kept in LDS, 4..64
    int32_t refill;      // finished lanes that make a wave fetch new rays, 1..64
    int32_t wide;        // 1: traverse the 4-wide tree, 0: the binary tree
    int32_t inner_exit;  // leave the inner-node loop once at most this many lanes still search a leaf
    int32_t packet;      // > 0: each wave traverses its 64 rays as one packet (coherent camera rays),
                         // testing postponed leaves once this many lanes hold one
};

// ---------------------------------------------------------------------------
// device view of one resident scene
// ---------------------------------------------------------------------------
// Shading record of sorted triangle k: 8 float4 at tri_rec + 8 k, one
// 128-byte line, so a hit reads its triangle with one line fetch.
//   [0] v0.xyz, material id   [1] v1.xyz, n0.x    [2] v2.xyz, n0.y
//   [3] n0.z, n1.xyz          [4] n2.xyz, t0.x    [5] t0.yz, t1.xy
//   [6] t1.z, t2.xyz          [7] unused (0)
// (v: vertex positions, n: vertex normals, t: vertex tangents; the parts
// whose flag is missing are 0)
enum { REC_NORMALS = 1, REC_TANGENTS = 2 };
struct DevScene {
    const float4* tris;        // 3 float4 per sorted triangle: (v0, prim) (v1, mesh) (v2, material)
    const float4* tri_rec;     // 8 float4 (one 128-byte line) per sorted triangle: its shading record
    int32_t rec_flags;         // REC_NORMALS | REC_TANGENTS: which parts of the records are filled
    const BvhNode* nodes;
    int32_t root;
    const Bvh4Node* nodes4;    // the same tree, 4 wide
    int32_t root4;
    int32_t num_nodes4;
    int32_t num_tris;
    const uint32_t* idx;       // 3 per triangle (original order)
    const float* pos;          // 3 per vertex
    const float* nrm;          // 3 per vertex or null
    const float* tan;          // 3 per vertex or null
    const int32_t* mesh_mat;   // per mesh
    const int32_t* mesh_first; // num_meshes + 1
    const float* area_cdf;     // per triangle, per-mesh segments
    const MatRec* mats;
    const float* light_cdf;    // per light
    const int32_t* light_mesh; // per light
    int32_t num_lights;
    float light_wsum;
    float env[3];
    // instanced scenes (num_inst > 0): meshes in object space, the 4-wide
    // nodes hold every mesh's tree and the top-level tree (root4) over the
    // instances; transforms as three float4 rows (3x4 row-major)
    int32_t num_inst;
    const float4* inst_xf;     // 3 per instance: object -> world
    const float4* inst_inv;    // 3 per instance: world -> object
    const int32_t* inst_root;  // per instance: root ref of its mesh's tree in nodes4
    const int32_t* light_inst; // per light: its instance
};

// Node refs of the top-level tree of an instanced scene: an instance leaf is
// kInstLeaf | instance, and the return marker pushed when a ray enters an
// instance is kInstReturn (triangle leaves of instanced scenes keep bit 30
// clear: fewer than 2^27 triangles).
constexpr int32_t kInstLeaf = (int32_t)0xC0000000u;
constexpr int32_t kInstReturn = -1;
constexpr int64_t kMaxInstTriangles = (1ll << 27) - 1;
__host__ __device__ inline bool is_inst_ref(int32_t r) { return r < 0 && (r & 0x40000000); }

// camera in the form the kernels use (computed on the host in float)
struct DevCamera {
    float pos[3], X[3], Y[3], Z[3];
    float nw, nh;
};

struct DevParams {
    int32_t spp, max_depth, rr_depth, use_nee, use_mis;
    uint32_t seed;
    int32_t width, height;
    int32_t mlp_mode;      // 0 plain, 1 query the radiance MLP at vertex 2, 2 record training samples
};

// A segment is one request's contiguous range of paths inside a chunk (path
// id = sample * W * H + pixel, in increasing order per request).  Segments of
// one chunk share the path queues; each keeps its own camera, resolution,
// sample count, depth settings and film.  A single render is one segment per
// chunk, so a request renders the same bits alone or batched with others.
struct Segment {
    int64_t path_begin;
    int32_t n;               // paths of this segment in the chunk
    int32_t slot_offset;     // first slot of the segment in the chunk
    int32_t wh, width, height, spp;
    uint32_t seed;
    int32_t max_depth, rr_depth, use_nee, use_mis, mlp_mode;
    int32_t last;            // the request's final paths: write film_out
    DevCamera cam;
    float* film_acc;         // 3 floats per pixel, zeroed before the first chunk
    float4* film_out;
};
constexpr int kMaxSegments = 255;

// Radiance-MLP hooks of the integrator.  Vertex 2 is the second surface hit
// (the hit of the trace at depth 1).  Mode 1: after NEE at vertex 2 the path
// stops and queues (features, throughput); the MLP's prediction times the
// throughput is added once the bounce's shadow rays are done.  Mode 2: the path
// continues as in the plain tracer and records (features, throughput) at vertex
// 2; the radiance gathered after vertex 2's NEE, divided by that throughput,
// is the training target.
struct MlpIO {
    float* q_in9;          // mode 1 queue: features
    float* q_beta;         // mode 1 queue: throughput (3)
    uint32_t* q_slot;      // mode 1 queue: path slot
    float* qs_in9;         // the same, as appended: partitioned like the path queues, then
    float* qs_beta;        // compacted into q_* (k_mlp_compact) for the MLP
    uint32_t* qs_slot;
    float* rec_in9;        // mode 2, per slot
    float* rec_beta;       // mode 2, per slot
    uint32_t* rec_valid;   // mode 2, per slot
    float lo[3], inv_ext[3];
};

// ---------------------------------------------------------------------------
// Structure-of-arrays path queue.  Entry i of a queue is one live path
// segment: its ray and the path state that travels with it.
// ---------------------------------------------------------------------------
struct PathQueue {
    float *ox, *oy, *oz, *dx, *dy, *dz;   // ray
    float *br, *bg, *bb;                  // throughput beta
    float* pdf;                           // scatter pdf of the previous bounce
    uint32_t* flags;                      // bit0: previous bounce specular
    uint64_t* rng;                        // PCG32 state (stream implied by the sample)
    uint32_t* slot;                       // path slot in the current chunk
};
struct HitBuf {                           // (shading rebuilds the hit point from u, v)
    float *u, *v;
    int32_t* tri;                         // sorted triangle position or -1
    int32_t* inst = nullptr;              // instanced scenes: the instance hit
};
struct ShadowQueue {
    float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
    float *cr, *cg, *cb;                  // contribution if unoccluded
    uint32_t* slot;
};

// Counters that live in device memory between the kernels of one chunk.  Each
// sits on its own 256-byte line: atomics on one line from the whole GPU are
// serialised (about 88 M per second on MI355X, whatever the number of
// counters on the line), so counters used together must not share a line.
//
// The path queues, their hit records and the shadow queue are cut into kParts
// partitions with a size counter each: partition p holds entries
// [p cap, p cap + size_p), cap = part_cap(chunk paths).  A stage reads its
// input queue as the concatenation of the partitions, cut into kParts
// contiguous pieces of equal length (QueueView): workgroup b works on piece
// b mod kParts and appends what it produces to partition b mod kParts of the
// next queue.  The appends of a stage thus spread over kParts counters, the
// pieces keep the work balanced and spatially coherent (camera paths fill the
// partitions in slot order), and no partition can outgrow cap: a piece holds
// at most ceil(paths / kParts) <= cap entries and each entry appends at most
// one entry to each queue.  Fetch counters have one counter per piece.
constexpr int kCntLine = 64;           // 32-bit words per 256-byte line
constexpr int kParts = 32;             // queue partitions = fetch pieces
enum {
    CNT_Q0 = 0,                                        // kParts lines: sizes of path queue 0 (cnt_q)
    CNT_Q1 = CNT_Q0 + kParts * kCntLine,               // kParts lines: path queue 1
    CNT_SHADOW = CNT_Q1 + kParts * kCntLine,           // kParts lines: shadow queue
    CNT_FETCH_T = CNT_SHADOW + kParts * kCntLine,      // kParts lines: next queue entry to trace (cleared by raygen and shade)
    CNT_FETCH_S = CNT_FETCH_T + kParts * kCntLine,     // kParts lines: next shadow entry to trace (cleared by trace)
    CNT_FETCH_Q = CNT_FETCH_S + kParts * kCntLine,     // kParts lines: next standalone ray query
    CNT_MLPQ_PARTS = CNT_FETCH_Q + kParts * kCntLine,  // kParts lines: MLP query appends per partition
    CNT_MLPQ = CNT_MLPQ_PARTS + kParts * kCntLine,     // one line: MLP queries after compaction (dense)
    CNT_WORDS = CNT_MLPQ + kCntLine
};
// statistics (unsigned 64-bit) after the counters, then the end of the buffer
constexpr size_t kTotalsOffset = (size_t)CNT_WORDS * 4;
constexpr size_t kCountersBytes = kTotalsOffset + 256;
// first size counter of path queue qi (0 or 1)
__host__ __device__ inline int cnt_q(int qi) { return qi ? CNT_Q1 : CNT_Q0; }
// entries per partition for a chunk of n paths (a multiple of 64)
__host__ __device__ inline uint32_t part_cap(uint32_t n) {
    const uint32_t turns = (n + 63u) / 64u;
    return (turns + (uint32_t)kParts - 1u) / (uint32_t)kParts * 64u;
}
// camera paths in partition p: slots [p cap, min((p + 1) cap, n)), at their own index
__host__ __device__ inline uint32_t part_size(uint32_t n, uint32_t cap, uint32_t p) {
    const uint64_t first = (uint64_t)p * cap;
    if (first >= n) return 0u;
    const uint64_t left = n - first;
    return left < cap ? (uint32_t)left : cap;
}

// ---------------------------------------------------------------------------
// host-side records
// ---------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct SceneRec {
    uint64_t handle = 0;
    int32_t nv = 0, nf = 0, nmesh = 0, nmat = 0, nlights = 0;
    bool has_nrm = false, has_tan = false;
    float env[3] = {0, 0, 0};
    float light_wsum = 0.0f;
    // device arrays
    DevBuf pos, nrm, tan, idx, mesh_mat, mesh_first, area_cdf, mats, light_cdf, light_mesh, tri_mesh;
    DevBuf orig_pos, orig_nrm, orig_tan, bone_ids, bone_w, bones;    // skinning
    int32_t nbones = 0;
    // BVH
    float bounds_lo[3] = {0, 0, 0}, bounds_hi[3] = {0, 0, 0};   // scene box (MLP input normalisation)
    bool has_bvh = false;
    DevBuf nodes, tris, order, node_parent, level_nodes, node_valid;
    // refit order: the binary-tree nodes of level L (root = 0) are
    // level_nodes[level_first[L] .. level_first[L + 1])
    std::vector<int32_t> level_first;
    DevBuf tri_rec;              // per sorted triangle: 128-byte shading record (positions, material, normals, tangents)
    int32_t root = 0, num_nodes = 0, max_depth = 0, num_clusters = 0, num_top = 0, cluster_size = 0;
    DevBuf nodes4;
    int32_t root4 = 0, num_nodes4 = 0, depth4 = 0;     // 4-wide tree: nodes, levels
    int32_t max_leaf = 4;
    // host copies needed for light CDF construction
    std::vector<float> h_mesh_area;
    std::vector<float> h_mesh_box;   // 6 per mesh: object-space lo xyz, hi xyz (at upload)
    // instanced scenes: each mesh's distinct vertex positions (object space),
    // mesh m's at [h_mesh_pts_first[m], h_mesh_pts_first[m + 1]) (xyz triples)
    std::vector<float> h_mesh_pts;
    std::vector<int64_t> h_mesh_pts_first;
    // instances (srt_scene_set_instances): meshes in object space
    int32_t ninst = 0;
    std::vector<int32_t> h_inst_mesh;
    std::vector<float> h_inst_xf, h_inst_inv;   // 12 per instance
    DevBuf inst_xf, inst_inv, inst_root, light_inst;
    std::vector<int32_t> h_mesh_root4;          // per mesh: root ref of its tree in nodes4
    // top-level tree (host-built, refit on the host after transform updates):
    // nodes4[tlas_first ..] ; child c of top node k is top node or instance ref
    int32_t tlas_first = 0;
    std::vector<Bvh4Node> h_tlas;
    std::vector<int32_t> h_tlas_order;          // top nodes in post order (children first)
};

struct KernelStat {
    double ms = 0.0;
    int64_t launches = 0;
    int64_t items = 0;
};

}  // namespace srt

struct srt_ctx {
    int device = 0;
    size_t budget = 0;
    std::atomic<size_t> used{0};          // the serving thread allocates films too
    hipStream_t stream = nullptr;
    std::string err;
    std::mutex mu;                        // guards err
    uint64_t next_scene = 1;
    std::map<uint64_t, srt::SceneRec*> scenes;
    // render pool
    int64_t pool_cap = 0;                 // path slots allocated
    int64_t pool_request = 1ll << 27;      // ~180 B per slot: 24.2 GB covers a 1080p x 64 spp frame in one chunk
    srt::DevBuf q[2][14], hits[5], shq[11], Lbuf[3], counters, film_acc, film_out, gstack;
    // device chunks of BVH-build scratch kept between builds (a rebuild then
    // allocates and frees nothing); released when an allocation would exceed
    // the budget or fail, and at srt_destroy
    std::vector<srt::DevBuf> build_scratch;
    std::mutex scratch_mu;
    bool keep_build_scratch = true;
    int64_t film_pixels = 0;
    size_t gstack_entries = 0;            // int32 entries allocated in gstack
    int32_t trace_lds_depth = 16, trace_refill = 40, shadow_refill = 32, trace_blocks_per_cu = 0;   // 0: occupancy limit
    int32_t trace_bvh_width = 4;          // traversal over the 4-wide (4) or binary (2) tree
    int32_t trace_inner_exit = 18;        // stragglers resume after the leaf step
    int32_t trace_packet = 0;             // > 0: camera rays traversed as 64-ray packets (0: per ray)
    int64_t max_wide_nodes = srt::kMaxNodes4;   // 4-wide trees above this many nodes are traversed in binary form
    // serving order: remaining paths at 2 Gpaths/s minus serve_age/1000 of the
    // time waited (0: shortest remaining first; 1000: earliest due time)
    int32_t serve_age_pm = 250;
    // profiling
    bool profile = false;
    std::map<std::string, srt::KernelStat> stats;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    // last render counters
    int64_t last_segments = 0, last_shadow_rays = 0;
    int64_t last_visits[4] = {0, 0, 0, 0};   // node visits / triangle tests: closest, shadow
    bool count_visits = false;
    std::map<void*, size_t> user_bufs;    // srt_device_alloc allocations
    srt::DevBuf mlpbuf[13], scan_tmp, train_in9, train_t;   // MLP hooks of the integrator
    srt::DevBuf segbuf;                   // device copy of the chunk's segment table
    int64_t mlp_pool_cap = 0, train_cap = 0;
    void* serve = nullptr;                // serving loop state; non-null while a dispatcher runs
    void* mlp = nullptr;                  // the radiance MLP in use (srt_mlp_* calls; the integrator binds a scene's)
    std::map<uint64_t, void*> mlps;       // radiance MLPs by scene handle; 0: the context's own
    uint64_t mlp_sel = 0;                 // the MLP srt_mlp_* calls act on (srt_mlp_select)
    void (*serve_free)(srt_ctx*) = nullptr;
    void (*mlp_free)(srt_ctx*) = nullptr;
};

// Entry points other than the serving ones refuse to run while the serving
// thread owns the context's stream and path pool.
#define SRT_IDLE(c)                                                                                   \
    do {                                                                                              \
        if ((c) && (c)->serve) return srt::set_err((c), SRT_E_STATE, "the context is serving; call srt_serve_stop first"); \
    } while (0)

namespace srt {
// allocation helpers (budget-checked)
int dev_alloc(srt_ctx* c, DevBuf& b, size_t bytes);
void dev_free(srt_ctx* c, DevBuf& b);
// keeps b when it already holds exactly `bytes`, else frees and allocates
int dev_realloc(srt_ctx* c, DevBuf& b, size_t bytes);
// frees the kept build-scratch chunks; returns whether there were any
bool release_build_scratch(srt_ctx* c);
int set_err(srt_ctx* c, int code, const char* fmt, ...);
int hip_check(srt_ctx* c, hipError_t e, const char* what);
// kernel timing
void prof_begin(srt_ctx* c, const char* name, hipEvent_t* ev);
void prof_end(srt_ctx* c, const char* name, hipEvent_t ev0, int64_t items);
void prof_collect(srt_ctx* c);

// BVH (bvh.hip)
int bvh_build(srt_ctx* c, SceneRec* s, int max_leaf, int cluster_size);
int bvh_refit(srt_ctx* c, SceneRec* s);
int bvh4_build(srt_ctx* c, SceneRec* s);
// instanced scenes (instances.hip)
int inst_build(srt_ctx* c, SceneRec* s, int max_leaf, int cluster_size);
int inst_update(srt_ctx* c, SceneRec* s, const float* xf);
void inst_scene_box(SceneRec* s);
bool affine_inverse(const float m[12], float out[12]);
int skin_vertices(srt_ctx* c, SceneRec* s, const float* bones_3x4, int nb);
// wavefront integrator (wavefront.hip)
// training samples produced by a mode-2 render
struct TrainOut {
    float* in9;        // device [cap][9]
    float* target;     // device [cap][3]
    int64_t cap;
    int64_t count;     // samples written
};
int render_wavefront(srt_ctx* c, SceneRec* s, const DevCamera& cam, const DevParams& p, float* film_dev,
                     TrainOut* train = nullptr);
int run_chunk(srt_ctx* c, SceneRec* s, const Segment* segs_host, int nseg, TrainOut* train = nullptr);
int render_simple(srt_ctx* c, SceneRec* s, const DevCamera& cam, int type, int spp, uint32_t seed, int width, int height,
                  float* out_dev);
int render_light(srt_ctx* c, SceneRec* s, const DevCamera& cam, int spp, int max_depth, int rr_depth, uint32_t seed,
                 int width, int height, float* film_dev);
// MLP (mlp.hip)
int mlp_infer_dev(srt_ctx* c, const float* in9, int n, float* out3);
int mlp_query_dev(srt_ctx* c, const float* in9, const uint32_t* count_dev, int max_n, float* out3);
int mlp_train_dev(srt_ctx* c, const float* in9, const float* target, int n, float* loss_host);
bool mlp_ready(srt_ctx* c);
// the MLP that renders of scene h query: h's own, else the context's own (or null)
void* mlp_for_scene(srt_ctx* c, uint64_t h);
void mlp_release_scene(srt_ctx* c, uint64_t h);
// makes scene h's MLP the one in use for a render or a training frame
struct MlpBind {
    srt_ctx* c;
    void* prev;
    MlpBind(srt_ctx* ctx, uint64_t h) : c(ctx), prev(ctx->mlp) { c->mlp = mlp_for_scene(c, h); }
    ~MlpBind() { c->mlp = prev; }
};
int trace_rays(srt_ctx* c, SceneRec* s, int n, const float* o_dev, const float* d_dev, float tmin, const float* tmax_dev,
               int any, float* t, float* u, float* v, int32_t* prim, int32_t* inst = nullptr);
int noalloc_init(srt_ctx* c);
int trace_rays_noalloc(srt_ctx* c, SceneRec* s, int n, const float* o_dev, const float* d_dev, float tmin,
                       const float* tmax_dev, int any, float* t, float* u, float* v, int32_t* prim);
int tonemap_dev(srt_ctx* c, float* rgba, int n, int type, float exposure, int use_gamma);
int accumulate_dev(srt_ctx* c, float* acc, float* rgba, int n, int frame_count);
int denoise_dev(srt_ctx* c, float* rgba, float* tmp, int W, int H);
DevScene dev_scene(const SceneRec* s);
DevCamera make_camera(const float pos[3], const float target[3], const float up[3], float fov, int W, int H);
}  // namespace srt
