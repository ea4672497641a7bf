// Wavefront path tracer for gfx950: raygen -> traverse -> shade (Disney BSDF
// sample/eval + next-event estimation) -> shadow traverse -> finalize, plus the
// tone-mapping and accumulation post-passes.
//
// Per-path semantics follow the reference megakernel
// (src/shader/integrator/path_tracer.cu:50-169) rule for rule; the stage split
// follows the reference wavefront pass (src/passes/wavefront/wavefront.cu:40-300)
// but every stage runs on compacted structure-of-arrays queues:
//   - one path slot per (pixel, sample); its PCG32 stream is keyed by
//     (pixel, seed, sample), so results do not depend on queue order;
//   - surviving paths and shadow rays are appended with one atomic per wave
//     (64-lane ballot + popcount prefix);
//   - radiance is accumulated per slot (no float atomics) and reduced per pixel
//     in sample order, reproducing the megakernel's summation order.
#include "srt_internal.h"
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <cstdio>
#include <cmath>
#include <mutex>
#include <algorithm>

namespace srt {

// ---------------------------------------------------------------------------
// traversal
// ---------------------------------------------------------------------------
struct RayPre {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
SRT_D float sel(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

SRT_D RayPre ray_pre(f3 d) {
    RayPre r;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    float dkz = sel(d, kz);
    if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = sel(d, kx) / dkz;
    r.Sy = sel(d, ky) / dkz;
    r.Sz = 1.0f / dkz;
    return r;
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013).  Returns
// t with barycentrics u (weight of v1) and v (weight of v2).
SRT_D bool intersect_tri(const RayPre& r, f3 o, f3 v0, f3 v1, f3 v2, float tmin, float tmax, float* t_out, float* u_out,
                         float* v_out) {
    f3 A = v0 - o, B = v1 - o, C = v2 - o;
    float Akz = sel(A, r.kz), Bkz = sel(B, r.kz), Ckz = sel(C, r.kz);
    float Ax = sel(A, r.kx) - r.Sx * Akz;
    float Ay = sel(A, r.ky) - r.Sy * Akz;
    float Bx = sel(B, r.kx) - r.Sx * Bkz;
    float By = sel(B, r.ky) - r.Sy * Bkz;
    float Cx = sel(C, r.kx) - r.Sx * Ckz;
    float Cy = sel(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Akz;
    float Bz = r.Sz * Bkz;
    float Cz = r.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    *t_out = t;
    *u_out = V * rcp;
    *v_out = W * rcp;
    return true;
}

constexpr float kRobust = 1.0000004f;   // 1 + 2 gamma_3 (Ize 2013): conservative slab test
// Closest-hit traversals cull boxes at the best hit distance so far times
// 1 + 2^-12 (any-hit traversals at tmax, as the slab test's far plane): a
// triangle's computed hit distance can lie below its own box's computed entry
// distance by more than the slab test's rounding (the watertight test's t
// carries the error of the shear and of T / det), so without the margin a
// triangle hit at nearly the same distance as the best one (coplanar faces)
// could be culled or kept depending on the visiting order.  With it, every
// triangle whose computed t is at most the best distance is tested, and the
// (t, primitive) minimum does not depend on the order.
constexpr float kCullTol = 1.000244140625f;

// ---------------------------------------------------------------------------
// wave-level queue append: one atomic per wave
// ---------------------------------------------------------------------------
SRT_D int wave_append(bool pred, uint32_t* counter) {
    // works on the lanes active at the call: the lowest active lane does the
    // atomic for all of them and readfirstlane broadcasts the base
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return -1;
    const int lane = __lane_id();
    const int first = __ffsll((unsigned long long)__builtin_amdgcn_read_exec()) - 1;
    uint32_t base = 0;
    if (lane == first) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __builtin_amdgcn_readfirstlane(base);
    const uint32_t prefix = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    return pred ? (int)(base + prefix) : -1;
}

// Zeroes the kParts counters (one line apart) that start at `first`; run by
// the threads 0..kParts-1 of one workgroup.
SRT_D void clear_parts(uint32_t* first) {
    if (threadIdx.x < (unsigned)kParts) first[threadIdx.x * kCntLine] = 0u;
}
// Partition of this workgroup in a partitioned queue (grid: a multiple of kParts)
SRT_D uint32_t my_part() { return blockIdx.x % (uint32_t)kParts; }

// One ray's traversal state, kept in registers while its lane walks the BVH.
struct TravState {
    f3 o, inv;
    RayPre rp;
    float tmin, best_t, bu, bv;
    float cull_t;      // boxes entered beyond this are skipped: best_t * kCullTol (any hit: tmax * kRobust)
    int32_t best_prim, best_pos;
    int32_t inst, best_inst;   // instanced scenes: the instance the ray is in (-1: world), the one hit
    int32_t cur;       // next node (>= 0) or leaf (< 0) to visit; valid while cur_ok
    int32_t pend;      // postponed leaf; valid while has_pend
    int32_t sp;        // stack depth
    int32_t top;       // the stack's top entry (valid while sp > 0), kept in a register
    bool sx, sy, sz;   // direction signs select the near/far slab planes
    bool cur_ok, has_pend;
    uint32_t onx, ony, onz;   // byte offsets of the near planes in a Bvh4Node (far plane: offset ^ 16)
};

// The ray a lane traverses with: origin, reciprocal direction, the
// watertight test's shear, the direction signs that pick near/far planes.
SRT_D void set_planes(TravState& st) {
    // a ray lying in a slab plane with a zero direction component yields NaN
    // slab distances, which fmaxf/fminf ignore
    st.sx = st.inv.x < 0.0f; st.sy = st.inv.y < 0.0f; st.sz = st.inv.z < 0.0f;
    st.onx = st.sx ? 16u : 0u;    // lox 0, hix 16, loy 32, hiy 48, loz 64, hiz 80
    st.ony = st.sy ? 48u : 32u;
    st.onz = st.sz ? 80u : 64u;
}
SRT_D void set_ray(TravState& st, f3 o, f3 d) {
    st.o = o;
    st.rp = ray_pre(d);
    st.inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    set_planes(st);
}

// Traversal kernels of instanced scenes keep each lane's world ray in LDS
// (ahead of the stack, lane-interleaved): origin and direction, from which
// an instance's object-space ray is computed on entry, and the reciprocal
// direction and triangle-test shear, which a ray leaving an instance gets
// back as they were (no divisions on the way out).
extern __shared__ int32_t trav_lds[];    // dynamic LDS of the traversal kernels
constexpr int kStashWords = 13;
constexpr int kRayStash = kStashWords * kTraceBlock;   // int32 entries of the stash
SRT_D void stash_ray(const TravState& st, f3 d) {
    float* r = reinterpret_cast<float*>(trav_lds) + threadIdx.x;
    const float w[kStashWords] = {st.o.x, st.o.y, st.o.z, d.x, d.y, d.z, st.inv.x, st.inv.y, st.inv.z,
                                  st.rp.Sx, st.rp.Sy, st.rp.Sz,
                                  __int_as_float(st.rp.kx | (st.rp.ky << 2) | (st.rp.kz << 4))};
    for (int j = 0; j < kStashWords; j++) r[j * kTraceBlock] = w[j];
}
SRT_D void unstash_ray(f3& o, f3& d) {
    const float* r = reinterpret_cast<const float*>(trav_lds) + threadIdx.x;
    o = mk(r[0], r[kTraceBlock], r[2 * kTraceBlock]);
    d = mk(r[3 * kTraceBlock], r[4 * kTraceBlock], r[5 * kTraceBlock]);
}
SRT_D void restore_world_ray(TravState& st) {
    const float* r = reinterpret_cast<const float*>(trav_lds) + threadIdx.x;
    st.o = mk(r[0], r[kTraceBlock], r[2 * kTraceBlock]);
    st.inv = mk(r[6 * kTraceBlock], r[7 * kTraceBlock], r[8 * kTraceBlock]);
    st.rp.Sx = r[9 * kTraceBlock]; st.rp.Sy = r[10 * kTraceBlock]; st.rp.Sz = r[11 * kTraceBlock];
    const int kp = __float_as_int(r[12 * kTraceBlock]);
    st.rp.kx = kp & 3; st.rp.ky = (kp >> 2) & 3; st.rp.kz = kp >> 4;
    set_planes(st);
}

template <bool INST = false>
SRT_D void trav_begin(TravState& st, f3 o, f3 d, float tmin, float tmax, int32_t root, bool closest = true) {
    set_ray(st, o, d);
    if (INST) stash_ray(st, d);
    st.tmin = tmin;
    st.best_t = tmax;
    // any-hit rays cull at tmax (their answer does not depend on the order)
    st.cull_t = tmax * (closest ? kCullTol : kRobust);
    st.best_prim = 0x7fffffff;
    st.best_pos = -1;
    st.inst = -1;
    st.best_inst = -1;
    st.bu = 0.0f; st.bv = 0.0f;
    st.cur = root;
    st.pend = 0;
    st.sp = 0;
    st.top = 0;
    st.cur_ok = true;
    st.has_pend = false;
}

// Traversal stack: the first `lds_depth` entries of each lane live in LDS
// (lane-interleaved: entry e of thread t at lds[e * blockDim + t]); deeper
// entries go to a per-thread global area (entry e at gst[(e - lds_depth) *
// gstride]).  The build guarantees a tree depth <= kMaxDepth.
// The traversal kernels run kTraceBlock threads per workgroup, so entry sp of a
// lane sits at trav_lds[sp * kTraceBlock + threadIdx.x] (a shift, no multiply).
struct TravStack {
    int32_t* gst;
    size_t gstride;
    int lds_depth;
    int lane0;           // this thread's first LDS entry (threadIdx.x)
    int inner_exit;
};
SRT_D void stk_push(const TravStack& k, int32_t sp, int32_t v) {
    if (sp < k.lds_depth) trav_lds[k.lane0 + sp * kTraceBlock] = v;
    else k.gst[(size_t)(sp - k.lds_depth) * k.gstride] = v;
}
SRT_D int32_t stk_get(const TravStack& k, int32_t sp) {
    if (sp < k.lds_depth) return trav_lds[k.lane0 + sp * kTraceBlock];
    int32_t v = k.gst[(size_t)(sp - k.lds_depth) * k.gstride];
    // keeps the compiler from merging the two loads into one generic (flat) load
    __asm__ volatile("" : "+v"(v));
    return v;
}
// A pop takes the top from its register, so the next node's fetch does not
// wait for a stack read; the read of the entry below (for the next pop) then
// overlaps that fetch.  Every push also sets the register (push_top).
SRT_D void stk_pop(const TravStack& k, TravState& st) {
    if (st.sp == 0) {
        st.cur_ok = false;
    } else {
        st.sp--;
        st.cur = st.top;
        if (st.sp > 0) st.top = stk_get(k, st.sp - 1);
    }
}
SRT_D void push_top(const TravStack& k, TravState& st, int32_t v) {
    stk_push(k, st.sp, v);
    st.sp++;
    st.top = v;
}

// Inner nodes, with speculative leaf postponement (Aila & Laine 2009): a lane
// that reaches a leaf keeps it in `pend` and continues with the next stack
// entry; it stops at its second leaf or when the stack is empty.  The wave
// leaves the loop once every lane still in it holds a postponed leaf, so the
// triangle tests that follow run with most lanes active.  Both child boxes are
// tested with the robust slab test (Ize 2013); when both are hit the nearer
// child is visited first and the other is pushed.
template <bool COUNT>
SRT_D void trav_inner(const DevScene& S, TravState& st, const TravStack& k, uint32_t& nv) {
    while (st.cur_ok) {
        if (st.cur < 0) {
            if (st.has_pend) break;               // second leaf: wait here
            st.pend = st.cur;
            st.has_pend = true;
            stk_pop(k, st);
        } else {
            if (COUNT) nv++;
            const BvhNode* np = S.nodes + st.cur;
            const float4 na = np->a, nb = np->b, nc = np->c;
            const int4 nd = np->d;
            const f3 o = st.o, inv = st.inv;
            const bool sx = st.sx, sy = st.sy, sz = st.sz;
            float lnx = ((sx ? na.y : na.x) - o.x) * inv.x, lfx = ((sx ? na.x : na.y) - o.x) * inv.x;
            float lny = ((sy ? na.w : na.z) - o.y) * inv.y, lfy = ((sy ? na.z : na.w) - o.y) * inv.y;
            float lnz = ((sz ? nc.y : nc.x) - o.z) * inv.z, lfz = ((sz ? nc.x : nc.y) - o.z) * inv.z;
            float rnx = ((sx ? nb.y : nb.x) - o.x) * inv.x, rfx = ((sx ? nb.x : nb.y) - o.x) * inv.x;
            float rny = ((sy ? nb.w : nb.z) - o.y) * inv.y, rfy = ((sy ? nb.z : nb.w) - o.y) * inv.y;
            float rnz = ((sz ? nc.w : nc.z) - o.z) * inv.z, rfz = ((sz ? nc.z : nc.w) - o.z) * inv.z;
            float tln = fmaxf(fmaxf(fmaxf(st.tmin, lnx), lny), lnz);
            float tlf = fminf(fminf(fminf(lfx, lfy), lfz) * kRobust, st.cull_t);
            float trn = fmaxf(fmaxf(fmaxf(st.tmin, rnx), rny), rnz);
            float trf = fminf(fminf(fminf(rfx, rfy), rfz) * kRobust, st.cull_t);
            const bool hl = tln <= tlf, hr = trn <= trf;
            const bool swap = hr && (!hl || trn < tln);        // right child first
            const int32_t first = swap ? nd.y : nd.x, second = swap ? nd.x : nd.y;
            if (hl && hr) {
                push_top(k, st, second);
            }
            if (hl || hr) st.cur = first;
            else stk_pop(k, st);
        }
        if (__all(st.has_pend)) break;
    }
}

// Slab test of one box given its near and far planes (same arithmetic as the
// binary step).
SRT_D bool slab(const TravState& st, float bnx, float bfx, float bny, float bfy, float bnz, float bfz, float* tn) {
    const f3 o = st.o, inv = st.inv;
    const float nx = (bnx - o.x) * inv.x, fx = (bfx - o.x) * inv.x;
    const float ny = (bny - o.y) * inv.y, fy = (bfy - o.y) * inv.y;
    const float nz = (bnz - o.z) * inv.z, fz = (bfz - o.z) * inv.z;
    const float tnear = fmaxf(fmaxf(fmaxf(st.tmin, nx), ny), nz);
    const float tfar = fminf(fminf(fminf(fx, fy), fz) * kRobust, st.cull_t);
    *tn = tnear;
    return tnear <= tfar;
}
SRT_D void ucswap(uint32_t& ka, int32_t& ra, uint32_t& kb, int32_t& rb) {
    const bool s = kb < ka;
    const uint32_t k = s ? kb : ka;
    const int32_t r = s ? rb : ra;
    kb = s ? ka : kb;
    rb = s ? ra : rb;
    ka = k;
    ra = r;
}
SRT_D void cswap(float& ka, int32_t& ra, float& kb, int32_t& rb) {
    const bool s = kb < ka;
    const float k = s ? kb : ka;
    const int32_t r = s ? rb : ra;
    kb = s ? ka : kb;
    rb = s ? ra : rb;
    ka = k;
    ra = r;
}

// 16-byte loads from the 4-wide node array through a buffer resource: one
// SGPR base and a 32-bit per-lane byte offset (instead of a 64-bit address per
// load); offsets outside the array read zeros.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
SRT_D __amdgpu_buffer_rsrc_t node4_rsrc(const DevScene& S) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)S.nodes4, (short)0, S.num_nodes4 * (int)sizeof(Bvh4Node), 0x00020000);
}
SRT_D float4 ld4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}
SRT_D int4 ld4i(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return make_int4((int)v.x, (int)v.y, (int)v.z, (int)v.w);
}

// Inner nodes of the 4-wide tree, with the same leaf postponement as the
// binary step: the hit children are ordered by entry distance, the nearest is
// visited next and the others are pushed farthest first.  The wave leaves the
// loop once at most `inner_exit` of its lanes still search for a first leaf:
// those lanes keep their state and continue after the leaf step, instead of
// every other lane waiting for them.
// Instanced scenes (INST): a top-level leaf kInstLeaf | i is postponed like a
// triangle leaf; the step after the loop (trav_pend) moves the lane's ray
// into instance i's object space (inverse transform, applied as
// apply_point / apply_vector in float, with the divisions of the new
// reciprocal direction and shear), so the lanes of a wave enter their
// instances together.  Popping kInstReturn restores the world ray from the
// LDS stash, after the lane's postponed leaf (which lies in that instance)
// has been tested.
SRT_D f3 xf_point(float4 r0, float4 r1, float4 r2, f3 p) {
    return mk(r0.x * p.x + r0.y * p.y + r0.z * p.z + r0.w, r1.x * p.x + r1.y * p.y + r1.z * p.z + r1.w,
              r2.x * p.x + r2.y * p.y + r2.z * p.z + r2.w);
}
SRT_D f3 xf_vector(float4 r0, float4 r1, float4 r2, f3 v) {
    return mk(r0.x * v.x + r0.y * v.y + r0.z * v.z, r1.x * v.x + r1.y * v.y + r1.z * v.z,
              r2.x * v.x + r2.y * v.y + r2.z * v.z);
}
template <bool COUNT, bool INST = false, bool ANY = false>
SRT_D void trav_inner4(const DevScene& S, TravState& st, const TravStack& k, uint32_t& nv, int inner_exit) {
    const __amdgpu_buffer_rsrc_t rsrc = node4_rsrc(S);
    while (st.cur_ok) {
        if (INST && is_inst_ref(st.cur)) {
            if (st.cur == kInstReturn) {
                if (st.has_pend) break;           // test the postponed leaf in this instance's space first
                restore_world_ray(st);
                st.inst = -1;
                stk_pop(k, st);
            } else {
                if (st.has_pend) break;           // second postponed item: wait here
                st.pend = st.cur;
                st.has_pend = true;
                stk_pop(k, st);
            }
        } else if (st.cur < 0) {
            if (st.has_pend) break;
            st.pend = st.cur;
            st.has_pend = true;
            stk_pop(k, st);
        } else {
            if (COUNT) nv++;
            // the ray's near and far planes are picked by address (per-ray offsets)
            const uint32_t nb = (uint32_t)st.cur * (uint32_t)sizeof(Bvh4Node);
            const float4 nx = ld4(rsrc, nb + st.onx), fx = ld4(rsrc, nb + (st.onx ^ 16u));
            const float4 ny = ld4(rsrc, nb + st.ony), fy = ld4(rsrc, nb + (st.ony ^ 16u));
            const float4 nz = ld4(rsrc, nb + st.onz), fz = ld4(rsrc, nb + (st.onz ^ 16u));
            int4 rf = ld4i(rsrc, nb + 96u);
            // the child refs lie in the same 128-byte line as the planes: have them
            // arrive with the planes, instead of a second dependent load after the
            // box tests (where the compiler would otherwise move it)
            __asm__ volatile("" : "+v"(rf.x), "+v"(rf.y), "+v"(rf.z), "+v"(rf.w));
            float k0, k1, k2, k3;
            const bool h0 = slab(st, nx.x, fx.x, ny.x, fy.x, nz.x, fz.x, &k0);
            const bool h1 = slab(st, nx.y, fx.y, ny.y, fy.y, nz.y, fz.y, &k1);
            const bool h2 = slab(st, nx.z, fx.z, ny.z, fy.z, nz.z, fz.z, &k2);
            const bool h3 = slab(st, nx.w, fx.w, ny.w, fy.w, nz.w, fz.w, &k3);
            if (ANY && !INST) {
                // any hit: the hit children in slot order (largest box first), no sort;
                // instanced scenes keep the distance order (their top-level slots are
                // not ordered, and entering an instance costs a ray transform)
                if (!(h0 || h1 || h2 || h3)) {
                    stk_pop(k, st);
                } else {
                    if (h3 && (h0 || h1 || h2)) { stk_push(k, st.sp, rf.w); st.sp++; st.top = rf.w; }
                    if (h2 && (h0 || h1)) { stk_push(k, st.sp, rf.z); st.sp++; st.top = rf.z; }
                    if (h1 && h0) push_top(k, st, rf.y);
                    st.cur = h0 ? rf.x : (h1 ? rf.y : (h2 ? rf.z : rf.w));
                }
                if (__popcll(__ballot(!st.has_pend)) <= (unsigned)inner_exit) break;
                continue;
            }
            const float inf = __builtin_huge_valf();
            k0 = h0 ? k0 : inf; k1 = h1 ? k1 : inf; k2 = h2 ? k2 : inf; k3 = h3 ? k3 : inf;
            int32_t r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
            const int nh = (int)h0 + (int)h1 + (int)h2 + (int)h3;
            cswap(k0, r0, k1, r1);
            cswap(k2, r2, k3, r3);
            cswap(k0, r0, k2, r2);
            cswap(k1, r1, k3, r3);
            cswap(k1, r1, k2, r2);
            if (nh == 0) {
                stk_pop(k, st);
            } else {
                if (nh > 3) { stk_push(k, st.sp, r3); st.sp++; }
                if (nh > 2) { stk_push(k, st.sp, r2); st.sp++; }
                if (nh > 1) push_top(k, st, r1);
                st.cur = r0;
            }
        }
        if (__popcll(__ballot(!st.has_pend)) <= (unsigned)inner_exit) break;
    }
}

// The triangles of the postponed leaf.  Closest hit: (t, primitive id)
// lexicographic minimum, so the result does not depend on the visiting order.
// Any hit: the first hit ends the traversal; returns true.
template <bool ANY, bool COUNT, bool INST = false>
SRT_D bool trav_leaf(const DevScene& S, TravState& st, uint32_t& nt) {
    st.has_pend = false;
    const uint32_t f = leaf_first(st.pend), c = leaf_count(st.pend);
    for (uint32_t j = f; j < f + c; j++) {
        const float4* tp = S.tris + 3 * (size_t)j;
        const float4 A = tp[0], B = tp[1], C = tp[2];
        if (COUNT) nt++;
        float t, u, v;
        if (!intersect_tri(st.rp, st.o, mk(A.x, A.y, A.z), mk(B.x, B.y, B.z), mk(C.x, C.y, C.z), st.tmin, st.best_t, &t,
                           &u, &v))
            continue;
        if (ANY) {
            st.cur_ok = false;
            return true;
        }
        const int32_t prim = __float_as_int(A.w);
        // (t, primitive, instance) lexicographic: instances of one mesh share primitive ids
        if (t < st.best_t ||
            (t == st.best_t && (prim < st.best_prim || (INST && prim == st.best_prim && st.inst < st.best_inst)))) {
            st.best_t = t; st.best_prim = prim; st.best_pos = (int32_t)j; st.bu = u; st.bv = v;
            st.cull_t = t * kCullTol;
            if (INST) st.best_inst = st.inst;
        }
    }
    return false;
}

// The postponed item of a lane: a triangle leaf (trav_leaf), or, in an
// instanced scene, an instance to enter: the lane's current world-space
// entry (if any) and kInstReturn are pushed, and the ray continues at the
// instance's mesh tree in object space.
template <bool ANY, bool COUNT, bool INST>
SRT_D bool trav_pend(const DevScene& S, TravState& st, const TravStack& k, uint32_t& nt) {
    if (INST && is_inst_ref(st.pend)) {
        st.has_pend = false;
        const int32_t i = st.pend & 0x3fffffff;
        f3 o, d;
        unstash_ray(o, d);
        const float4* M = S.inst_inv + 3 * (size_t)i;
        const float4 r0 = M[0], r1 = M[1], r2 = M[2];
        set_ray(st, xf_point(r0, r1, r2, o), xf_vector(r0, r1, r2, d));
        st.inst = i;
        if (st.cur_ok) { stk_push(k, st.sp, st.cur); st.sp++; }
        push_top(k, st, kInstReturn);
        st.cur = S.inst_root[i];
        st.cur_ok = true;
        return false;
    }
    return trav_leaf<ANY, COUNT, INST>(S, st, nt);
}

// Persistent traversal with dynamic ray fetch.  Each lane holds one ray; when
// at least `refill` lanes of the wave (or all of them) are idle, the wave
// stores the finished lanes' results and takes that many new ray indices from
// `fetch` with one atomic, so lanes do not idle while the wave's slowest ray
// finishes, and the stores' appends and memory traffic happen once per refill.
// The wave works on indices [lo, hi), taken from `fetch` (lo + fetch), of
// entries view.entry(index); the workgroups of a grid work on separate ranges
// with separate counters (queue_piece, dense_range), so the fetch atomics
// spread over lines.
// IO supplies
//   void load(uint32_t i, TravState&, int32_t root)
//   void store(uint32_t i, const TravState&, bool occluded)   (on the storing lanes)
// Every wave leaves the loop once `fetch` has passed hi - lo and its lanes
// have finished and stored.
template <bool ANY, bool COUNT, bool WIDE, class IO, class View, bool INST = false>
SRT_D void trace_persistent(const DevScene& S, const View& view, uint32_t lo, uint32_t hi, uint32_t* fetch, int refill,
                            const TravStack& k, IO& io, uint32_t& nv, uint32_t& nt) {
    static_assert(!INST || WIDE, "instanced scenes are traversed in the 4-wide form");
    const int32_t root = WIDE ? S.root4 : S.root;
    const int lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    TravState st;
    st.cur_ok = false;
    st.has_pend = false;
    uint32_t ri = 0;
    bool has = false, done = false, exhausted = false, occl = false;
    for (;;) {
        const bool idle = !has && !exhausted;
        const unsigned long long m = __ballot(idle);
        unsigned long long act = __ballot(has);
        if (m != 0ull && (act == 0ull || __popcll(m) >= refill)) {
            if (done) {
                io.store(ri, st, occl);
                done = false;
            }
            const int leader = __ffsll(m) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(fetch, (uint32_t)__popcll(m));
            base = __shfl(base, leader);
            if (idle) {
                const uint32_t idx = lo + base + (uint32_t)__popcll(m & below);
                if (idx < hi) {
                    ri = view.entry(idx);
                    has = true;
                    occl = false;
                    io.template load<INST>(ri, st, root);
                } else {
                    exhausted = true;
                }
            }
            act = __ballot(has);
        }
        if (act == 0ull) break;
        if (has) {
            if (WIDE) trav_inner4<COUNT, INST, ANY>(S, st, k, nv, k.inner_exit);
            else trav_inner<COUNT>(S, st, k, nv);
            if (st.has_pend) occl = trav_pend<ANY, COUNT, INST>(S, st, k, nt);
            if (occl || (!st.cur_ok && !st.has_pend)) {
                has = false;
                done = true;
            }
        }
    }
}

// Packet traversal of coherent rays (camera rays: the 64 paths of a wave are
// one 8x8 pixel tile): the wave walks one node sequence for its 64 rays.  At
// an inner node every lane tests its ray against the four child boxes (the
// same slab test as trav_inner4).  An inner child is visited when any lane's
// ray hits it, nearest first in the order of the first lane that hits one;
// the wave's stack, that order and the loop control are wave-uniform (SGPRs,
// a scalar sort of the entry distances' bit patterns, which order as unsigned
// integers since tnear >= tmin > 0), so a node step costs the slab tests and a
// few cross-lane reads instead of a per-lane sort and stack.  A leaf child is
// not visited by the wave: each lane whose ray hits it appends it to its own
// list of postponed leaves, and the lanes test their lists together once
// `flush` lanes hold one (or a list could overflow, or the traversal ends), so
// the triangle tests run with most lanes active and each lane tests only the
// leaves its own box tests admit.  The closest hit is the same as with
// per-ray traversal: each lane tests every leaf whose box its ray enters
// before its current closest hit, and the (t, primitive) minimum does not
// depend on the order.  The wave's stack holds at most 3 entries per tree
// level; it lives in the first rows of the wave's LDS stack area (entry e at
// row e / 64, lane e mod 64), the leaf lists in its last kPacketLeafRows rows
// (the host enables packets only when both fit).
// The wave takes 64 consecutive entries of [lo, hi) at a time from `fetch`,
// asking for the next 64 when it starts on the current ones.
constexpr int kPacketLeafRows = 8;       // postponed leaves per lane
template <class IO, class View>
SRT_D void trace_packet(const DevScene& S, const View& view, uint32_t lo, uint32_t hi, uint32_t* fetch, int lds_depth,
                        int flush, IO& io) {
    const __amdgpu_buffer_rsrc_t rsrc = node4_rsrc(S);
    const int lane = __lane_id();
    int32_t* wst = trav_lds + (threadIdx.x & ~63u);
    int32_t* lf = trav_lds + (lds_depth - kPacketLeafRows) * kTraceBlock + threadIdx.x;
    const uint32_t n = hi - lo;
    const float inf = __builtin_huge_valf();
    uint32_t nt = 0;
    uint32_t bn = 0;                      // the next batch: fetched while the current one is traversed
    if (lane == 0) bn = atomicAdd(fetch, 64u);
    for (;;) {
        const uint32_t b = __builtin_amdgcn_readfirstlane(bn);
        if (b >= n) break;
        if (lane == 0) bn = atomicAdd(fetch, 64u);
        const uint32_t j = b + (uint32_t)lane;
        const bool valid = j < n;
        const uint32_t ri = view.entry(lo + (valid ? j : n - 1u));   // lanes past the end repeat the last ray
        TravState st;
        io.template load<false>(ri, st, S.root4);
        int cnt = 0;                      // this lane's postponed leaves, at lf[i * kTraceBlock]
        auto test_leaves = [&]() {
            for (int i = 0; __ballot(i < cnt) != 0ull; i++) {
                if (i < cnt) {
                    st.pend = lf[i * kTraceBlock];
                    (void)trav_leaf<false, false, false>(S, st, nt);
                }
            }
            cnt = 0;
        };
        int32_t cur = S.root4;
        int32_t sp = 0;
        if (cur < 0) {                    // the whole tree is one leaf
            st.pend = cur;
            (void)trav_leaf<false, false, false>(S, st, nt);
        }
        while (cur >= 0) {
            const uint32_t nb = (uint32_t)cur * (uint32_t)sizeof(Bvh4Node);
            const float4 nx = ld4(rsrc, nb + st.onx), fx = ld4(rsrc, nb + (st.onx ^ 16u));
            const float4 ny = ld4(rsrc, nb + st.ony), fy = ld4(rsrc, nb + (st.ony ^ 16u));
            const float4 nz = ld4(rsrc, nb + st.onz), fz = ld4(rsrc, nb + (st.onz ^ 16u));
            const int4 rf = ld4i(rsrc, nb + 96u);
            float k0, k1, k2, k3;
            const bool h0 = slab(st, nx.x, fx.x, ny.x, fy.x, nz.x, fz.x, &k0);
            const bool h1 = slab(st, nx.y, fx.y, ny.y, fy.y, nz.y, fz.y, &k1);
            const bool h2 = slab(st, nx.z, fx.z, ny.z, fy.z, nz.z, fz.z, &k2);
            const bool h3 = slab(st, nx.w, fx.w, ny.w, fy.w, nz.w, fz.w, &k3);
            int32_t r0 = __builtin_amdgcn_readfirstlane(rf.x), r1 = __builtin_amdgcn_readfirstlane(rf.y);
            int32_t r2 = __builtin_amdgcn_readfirstlane(rf.z), r3 = __builtin_amdgcn_readfirstlane(rf.w);
            // leaf children go to the lists of the lanes that hit them
            if (r0 < 0 && h0) { lf[cnt * kTraceBlock] = r0; cnt++; }
            if (r1 < 0 && h1) { lf[cnt * kTraceBlock] = r1; cnt++; }
            if (r2 < 0 && h2) { lf[cnt * kTraceBlock] = r2; cnt++; }
            if (r3 < 0 && h3) { lf[cnt * kTraceBlock] = r3; cnt++; }
            const unsigned long long m0 = r0 >= 0 ? __ballot(h0) : 0ull, m1 = r1 >= 0 ? __ballot(h1) : 0ull;
            const unsigned long long m2 = r2 >= 0 ? __ballot(h2) : 0ull, m3 = r3 >= 0 ? __ballot(h3) : 0ull;
            const unsigned long long any = m0 | m1 | m2 | m3;
            if (any != 0ull) {
                // keys: the first such lane's entry distances (+inf for a child only
                // other lanes hit), absent children last
                const int ld = __builtin_ctzll(any);
                uint32_t q0 = __builtin_amdgcn_readlane(__float_as_uint(h0 ? k0 : inf), ld);
                uint32_t q1 = __builtin_amdgcn_readlane(__float_as_uint(h1 ? k1 : inf), ld);
                uint32_t q2 = __builtin_amdgcn_readlane(__float_as_uint(h2 ? k2 : inf), ld);
                uint32_t q3 = __builtin_amdgcn_readlane(__float_as_uint(h3 ? k3 : inf), ld);
                q0 = m0 ? q0 : 0xffffffffu; q1 = m1 ? q1 : 0xffffffffu;
                q2 = m2 ? q2 : 0xffffffffu; q3 = m3 ? q3 : 0xffffffffu;
                const int nin = (m0 ? 1 : 0) + (m1 ? 1 : 0) + (m2 ? 1 : 0) + (m3 ? 1 : 0);
                ucswap(q0, r0, q1, r1);
                ucswap(q2, r2, q3, r3);
                ucswap(q0, r0, q2, r2);
                ucswap(q1, r1, q3, r3);
                ucswap(q1, r1, q2, r2);
                // push the others farthest first: entry sp + i holds sorted child nin - 1 - i
                const int np = nin - 1;
                if (lane < np) {
                    const int c = np - lane;        // 1 .. 3
                    const int32_t v = c == 1 ? r1 : (c == 2 ? r2 : r3);
                    const int e = sp + lane;
                    wst[(e >> 6) * kTraceBlock + (e & 63)] = v;
                }
                __builtin_amdgcn_wave_barrier();
                sp += np;
                cur = r0;
            } else if (sp > 0) {
                sp--;
                cur = __builtin_amdgcn_readfirstlane(wst[(sp >> 6) * kTraceBlock + (sp & 63)]);
            } else {
                cur = -1;                     // traversal done
            }
            // test the postponed leaves when enough lanes hold one, when a list
            // could not take another node's four, or at the end
            const unsigned long long pend = __ballot(cnt > 0);
            if (pend != 0ull &&
                (cur < 0 || __popcll(pend) >= (unsigned)flush || __ballot(cnt > kPacketLeafRows - 4) != 0ull))
                test_leaves();
        }
        if (valid) io.store(ri, st, false);
        __builtin_amdgcn_wave_barrier();
    }
}

// A partitioned queue read as the concatenation of its partitions: prefix[p]
// entries come before partition p, prefix[kParts] is the total (in LDS).
struct QueueView {
    const uint32_t* prefix;
    uint32_t cap;
    uint32_t p0;        // partition of the first entry of this workgroup's piece
    // queue entry of concatenated index j (prefix[p0] <= j < total)
    SRT_D uint32_t entry(uint32_t j) const {
        uint32_t p = p0;
        while (j >= prefix[p + 1]) ++p;
        return p * cap + (j - prefix[p]);
    }
};
// Workgroup prologue, run by all its threads: the prefix sums of the kParts
// size counters at counts[cnt] into lds[0..kParts], and this workgroup's piece
// [lo, hi) of the concatenation (piece blockIdx mod kParts of kParts equal ones).
SRT_D QueueView queue_piece(const uint32_t* counts, int cnt, uint32_t cap, uint32_t* lds, uint32_t& lo, uint32_t& hi) {
    if (threadIdx.x < 64u) {
        uint32_t v = threadIdx.x < (unsigned)kParts ? counts[cnt + threadIdx.x * kCntLine] : 0u;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(v, d);
            if ((int)threadIdx.x >= d) v += t;
        }
        if (threadIdx.x < (unsigned)kParts) lds[threadIdx.x + 1] = v;
        if (threadIdx.x == 0) lds[0] = 0u;
    }
    __syncthreads();
    const uint32_t total = lds[kParts];
    const uint32_t piece = my_part();
    lo = (uint32_t)((uint64_t)total * piece / kParts);
    hi = (uint32_t)((uint64_t)total * (piece + 1u) / kParts);
    QueueView v{lds, cap, 0u};
    while (v.p0 + 1u < (unsigned)kParts && lo >= lds[v.p0 + 1]) ++v.p0;
    return v;
}
struct IdentityEntry {
    SRT_D uint32_t entry(uint32_t j) const { return j; }
};
// A dense array [0, n): workgroup b works on piece b mod P of P = min(kParts,
// grid size) contiguous pieces, with the piece's counter at fetch + piece lines.
SRT_D uint32_t* dense_range(uint32_t n, uint32_t* fetch, uint32_t& lo, uint32_t& hi) {
    const uint32_t np = gridDim.x < (unsigned)kParts ? gridDim.x : (uint32_t)kParts;
    const uint32_t piece = blockIdx.x % np;
    lo = (uint32_t)((uint64_t)n * piece / np);
    hi = (uint32_t)((uint64_t)n * (piece + 1u) / np);
    return fetch + piece * kCntLine;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Segment lookup: the segment whose slot range holds slot i (segments are in
// increasing slot_offset order)
SRT_D int find_segment(const Segment* segs, int nseg, uint32_t i) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if ((uint32_t)segs[mid].slot_offset <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Pixel order of the path ids: inside each sample plane, pixels follow 8x8
// tiles (bands of 8 rows, each cut into blocks of 8 columns, row-major inside
// a block; the last band and block may be narrower), so the 64 paths of a wave
// cover a compact tile.  tile_pixel is a bijection of [0, W H).  Results do
// not depend on the order: each path's random stream is
// keyed by (pixel, sample) and each pixel sums its samples in sample order.
// The same bijection as column x and row y; the full 8x8 tiles of the image
// (all but the last band and block) divide by shifts.
SRT_D void tile_xy(uint32_t rank, uint32_t W, uint32_t H, uint32_t& x, uint32_t& y) {
    const uint32_t band = rank / (8u * W);
    const uint32_t r = rank - band * 8u * W;
    const bool full_band = H - 8u * band >= 8u;
    const uint32_t b = full_band ? r >> 6 : r / (8u * (H - 8u * band));
    const uint32_t bh = full_band ? 8u : H - 8u * band;
    const uint32_t r2 = r - b * 8u * bh;
    const bool full_block = W - 8u * b >= 8u;
    const uint32_t ly = full_block ? r2 >> 3 : r2 / (W - 8u * b);
    const uint32_t lx = r2 - ly * (full_block ? 8u : W - 8u * b);
    x = b * 8u + lx;
    y = band * 8u + ly;
}
SRT_D uint32_t tile_pixel(uint32_t rank, uint32_t W, uint32_t H) {
    const uint32_t band = rank / (8u * W);
    const uint32_t bh = (H - 8u * band < 8u) ? H - 8u * band : 8u;
    const uint32_t r = rank - band * 8u * W;
    const uint32_t b = r / (8u * bh);
    const uint32_t bw = (W - 8u * b < 8u) ? W - 8u * b : 8u;
    const uint32_t r2 = r - b * 8u * bh;
    const uint32_t ly = r2 / bw, lx = r2 - ly * bw;
    return (band * 8u + ly) * W + b * 8u + lx;
}

// Camera ray of one sample of one pixel (path_tracer.cu:66-68,
// camera.h:107-117, :27-32): two jitter draws from the (pixel, sample)
// stream; rng is left after them.
SRT_D void camera_ray_at(const DevCamera& cam, int width, int height, int ix, int iy, uint32_t pixel, uint32_t seed,
                         uint32_t sample, f3& ro, f3& rd, Pcg32& rng) {
    rng = path_rng(pixel, seed, sample);
    float r0 = rng.next_float();
    float xx = ((float)ix + r0) / (float)width;
    float r1 = rng.next_float();
    float yy = ((float)iy + r1) / (float)height;
    float px = (xx - 0.5f) * cam.nw, py = (yy - 0.5f) * cam.nh;
    f3 dir = normalize(mk(px, py, 1.0f));
    f3 X = ld3(cam.X), Y = ld3(cam.Y), Z = ld3(cam.Z), CP = ld3(cam.pos);
    f3 zero = mk(0.0f);
    ro = zero.x * X + zero.y * Y + zero.z * Z + CP;
    rd = dir.x * X + dir.y * Y + dir.z * Z;
}
SRT_D void camera_ray(const DevCamera& cam, int width, int height, uint32_t pixel, uint32_t seed, uint32_t sample, f3& ro,
                      f3& rd, Pcg32& rng) {
    camera_ray_at(cam, width, height, (int)(pixel % (uint32_t)width), (int)(pixel / (uint32_t)width), pixel, seed, sample,
                  ro, rd, rng);
}

// Sample index of path id pid (>= 0) in planes of wh > 0 pixels: a 32-bit
// division when pid fits in 32 bits (every frame below 2^32 paths), else 64-bit.
SRT_D uint32_t path_sample(int64_t pid, int32_t wh) {
    if (((uint64_t)pid >> 32) == 0u) return (uint32_t)pid / (uint32_t)wh;
    return (uint32_t)(pid / wh);
}

// The state of path `slot` at its first bounce (the camera ray of its pixel
// and sample, path_tracer.cu:66-68).  Slot i belongs to segment k: path id
// pid = path_begin + (i - slot_offset), sample = pid / (W H), pixel =
// tile_pixel(pid mod (W H)).  Throughput 1, scatter pdf 1, flags: specular
// (bit 0) and the segment (bits 8..); rng: the stream after the two jitter
// draws.  The first bounce's queue is the slots themselves (entry = slot, in
// slot order over the partitions): the depth-0 trace and shade stages compute
// this state where they need it instead of reading it from a queue.
struct PathStart {
    f3 o, d;
    uint32_t flags;
    uint64_t rng;
};
SRT_D PathStart path_start(const Segment* segs, int nseg, uint32_t slot) {
    const int k = find_segment(segs, nseg, slot);
    const Segment& sg = segs[k];
    const int64_t pid = sg.path_begin + (slot - sg.slot_offset);
    const uint32_t sample = path_sample(pid, sg.wh);
    uint32_t x, y;
    tile_xy((uint32_t)(pid - (int64_t)sample * sg.wh), (uint32_t)sg.width, (uint32_t)sg.height, x, y);
    PathStart p;
    Pcg32 rng;
    camera_ray_at(sg.cam, sg.width, sg.height, (int)x, (int)y, y * (uint32_t)sg.width + x, sg.seed, sample, p.o, p.d, rng);
    p.flags = 1u | ((uint32_t)k << 8);
    p.rng = rng.state;
    return p;
}
// Sizes of the first bounce's queue partitions (n slots, partition p holds
// slots [p cap, p cap + size)); clears the traversal fetch counters.
__global__ void k_path_counts(uint32_t n, uint32_t cap, uint32_t* counts) {
    clear_parts(counts + CNT_FETCH_T);
    if (threadIdx.x < (unsigned)kParts) counts[CNT_Q0 + threadIdx.x * kCntLine] = part_size(n, cap, threadIdx.x);
}

// Ray sources/sinks of the persistent traversal loop.
struct QueueIO {       // closest hit of queue entries -> hit buffer (same index)
    PathQueue q;
    HitBuf h;
    template <bool INST>
    SRT_D void load(uint32_t i, TravState& st, int32_t root) const {
        trav_begin<INST>(st, mk(q.ox[i], q.oy[i], q.oz[i]), mk(q.dx[i], q.dy[i], q.dz[i]), 1e-3f, 1e16f, root);
    }
    SRT_D void store(uint32_t i, const TravState& st, bool) const {
        const bool hit = st.best_pos >= 0;
        h.u[i] = hit ? st.bu : 0.0f;
        h.v[i] = hit ? st.bv : 0.0f;
        h.tri[i] = hit ? st.best_pos : -1;
        if (h.inst) h.inst[i] = hit ? st.best_inst : -1;
    }
};
struct CamIO {         // closest hit of the first bounce's camera rays (entry = slot) -> hit buffer
    const Segment* segs;
    int nseg;
    HitBuf h;
    template <bool INST>
    SRT_D void load(uint32_t i, TravState& st, int32_t root) const {
        const PathStart p = path_start(segs, nseg, i);
        trav_begin<INST>(st, p.o, p.d, 1e-3f, 1e16f, root);
    }
    SRT_D void store(uint32_t i, const TravState& st, bool occluded) const {
        const QueueIO q{PathQueue{}, h};
        q.store(i, st, occluded);
    }
};
// L[slot] += c for the three channels; the three loads go out before the stores
SRT_D void add_rgb(float* Lr, float* Lg, float* Lb, uint32_t slot, float r, float g, float b) {
    const float r0 = Lr[slot], g0 = Lg[slot], b0 = Lb[slot];
    Lr[slot] = r0 + r;
    Lg[slot] = g0 + g;
    Lb[slot] = b0 + b;
}

struct ShadowIO {      // any hit in [1e-3, tmax]; unoccluded rays add their contribution
    ShadowQueue sq;
    float *Lr, *Lg, *Lb;
    template <bool INST>
    SRT_D void load(uint32_t i, TravState& st, int32_t root) const {
        trav_begin<INST>(st, mk(sq.ox[i], sq.oy[i], sq.oz[i]), mk(sq.dx[i], sq.dy[i], sq.dz[i]), 1e-3f, sq.tmax[i], root,
                         false);
    }
    SRT_D void store(uint32_t i, const TravState&, bool occluded) const {
        if (occluded) return;
        const uint32_t slot = sq.slot[i];     // one shadow ray per slot per bounce: no write conflicts
        add_rgb(Lr, Lg, Lb, slot, sq.cr[i], sq.cg[i], sq.cb[i]);
    }
};

// INST: the LDS stack sits after the world-ray stash
template <bool INST = false>
SRT_D TravStack make_stack(int32_t* gstack, TraceCfg cfg) {
    const size_t gtid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    TravStack k;
    k.lane0 = (int)threadIdx.x + (INST ? kRayStash : 0);
    k.gst = gstack + gtid;
    k.gstride = (size_t)gridDim.x * blockDim.x;
    k.lds_depth = cfg.lds_depth;
    k.inner_exit = cfg.inner_exit;
    return k;
}

// Closest hit for every entry of queue `qi`; also clears the counters that the
// following shade and shadow stages append to or fetch from.  COUNT: also
// total node visits and triangle tests (an untimed pass used for the
// algorithmic byte count).
template <bool COUNT, bool WIDE, bool INST, class IO>
SRT_D void trace_stage(const DevScene& S, IO& io, uint32_t* counts, int qi, uint32_t cap, int32_t* gstack,
                       unsigned long long* totals, TraceCfg cfg) {
    __shared__ uint32_t prefix[kParts + 1];
    uint32_t lo, hi;
    const QueueView view = queue_piece(counts, cnt_q(qi), cap, prefix, lo, hi);
    uint32_t* fetch = counts + CNT_FETCH_T + my_part() * kCntLine;
    if (blockIdx.x == 0) {
        clear_parts(counts + CNT_FETCH_S);
        clear_parts(counts + cnt_q(qi ^ 1));
        clear_parts(counts + CNT_SHADOW);
        clear_parts(counts + CNT_MLPQ_PARTS);
        if (threadIdx.x == 0) counts[CNT_MLPQ] = 0u;
    }
    if (blockIdx.x < (unsigned)kParts && threadIdx.x == 0) atomicAdd(&totals[0], (unsigned long long)(hi - lo));
    const TravStack k = make_stack<INST>(gstack, cfg);
    uint32_t nv = 0, nt = 0;
    trace_persistent<false, COUNT, WIDE, IO, QueueView, INST>(S, view, lo, hi, fetch, cfg.refill, k, io, nv, nt);
    if (COUNT) {
        atomicAdd(&totals[2], (unsigned long long)nv);
        atomicAdd(&totals[3], (unsigned long long)nt);
    }
}
template <bool COUNT, bool WIDE, bool INST = false>
__global__ void __launch_bounds__(kTraceBlock) k_trace(DevScene S, PathQueue q, HitBuf h, uint32_t* counts, int qi,
                                                      uint32_t cap, int32_t* gstack, unsigned long long* totals,
                                                      TraceCfg cfg) {
    QueueIO io{q, h};
    trace_stage<COUNT, WIDE, INST, QueueIO>(S, io, counts, qi, cap, gstack, totals, cfg);
}
// The first bounce (queue 0): the same, with each camera ray computed from
// its slot (path_start) instead of read from the queue.
template <bool COUNT, bool WIDE, bool INST = false>
__global__ void __launch_bounds__(kTraceBlock) k_trace_cam(DevScene S, const Segment* segs, int nseg, HitBuf h,
                                                          uint32_t* counts, uint32_t cap, int32_t* gstack,
                                                          unsigned long long* totals, TraceCfg cfg) {
    CamIO io{segs, nseg, h};
    trace_stage<COUNT, WIDE, INST, CamIO>(S, io, counts, 0, cap, gstack, totals, cfg);
}

// Closest hit of the first bounce's camera rays as 64-ray packets
// (trace_packet); the same contract as k_trace_cam<false, true>.
__global__ void __launch_bounds__(kTraceBlock) k_trace_packet(DevScene S, const Segment* segs, int nseg, HitBuf h,
                                                             uint32_t* counts, uint32_t cap, unsigned long long* totals,
                                                             TraceCfg cfg) {
    constexpr int qi = 0;
    __shared__ uint32_t prefix[kParts + 1];
    uint32_t lo, hi;
    const QueueView view = queue_piece(counts, cnt_q(qi), cap, prefix, lo, hi);
    uint32_t* fetch = counts + CNT_FETCH_T + my_part() * kCntLine;
    if (blockIdx.x == 0) {
        clear_parts(counts + CNT_FETCH_S);
        clear_parts(counts + cnt_q(qi ^ 1));
        clear_parts(counts + CNT_SHADOW);
        clear_parts(counts + CNT_MLPQ_PARTS);
        if (threadIdx.x == 0) counts[CNT_MLPQ] = 0u;
    }
    if (blockIdx.x < (unsigned)kParts && threadIdx.x == 0) atomicAdd(&totals[0], (unsigned long long)(hi - lo));
    CamIO io{segs, nseg, h};
    trace_packet<CamIO, QueueView>(S, view, lo, hi, fetch, cfg.lds_depth, cfg.packet, io);
}

struct LightSample { f3 pos, normal, emission; float pdf; };

// light.h:61-76 and gmesh.h:45-80, with the float lower CDF bound (D2)
SRT_D LightSample light_sample(const DevScene& S, f2 rnd) {
    int idx = binary_search(S.light_cdf, S.num_lights, rnd.x);
    float lower = (idx == 0) ? 0.0f : S.light_cdf[idx - 1];
    rnd.x = (rnd.x - lower) / (S.light_cdf[idx] - lower);
    int mesh = S.light_mesh[idx];
    int f0 = S.mesh_first[mesh], fn = S.mesh_first[mesh + 1] - f0;
    const float* ac = S.area_cdf + f0;
    int f = binary_search(ac, fn, rnd.x);
    float lo2 = (f == 0) ? 0.0f : ac[f - 1];
    rnd.x = (rnd.x - lo2) / (ac[f] - lo2);
    f2 uv = uniform_sample_triangle(rnd);
    int prim = f0 + f;
    const uint32_t* ix = S.idx + 3 * (size_t)prim;
    f3 v0 = ld3(S.pos + 3 * (size_t)ix[0]), v1 = ld3(S.pos + 3 * (size_t)ix[1]), v2 = ld3(S.pos + 3 * (size_t)ix[2]);
    float w = 1.0f - uv.x - uv.y;
    LightSample ls;
    ls.pos = v0 * w + v1 * uv.x + v2 * uv.y;
    f3 normal = normalize(cross(v1 - v0, v2 - v0));
    if (S.nrm) {
        f3 n0 = ld3(S.nrm + 3 * (size_t)ix[0]), n1 = ld3(S.nrm + 3 * (size_t)ix[1]), n2 = ld3(S.nrm + 3 * (size_t)ix[2]);
        f3 sn = n0 * w + n1 * uv.x + n2 * uv.y;
        if (dot(normal, sn) < 0.0f) normal = -normal;
    }
    ls.normal = normal;
    if (S.num_inst > 0) {
        // GInstance::sample (gmesh.h:77-78): apply_point / apply_vector, no renormalising
        const float4* T = S.inst_xf + 3 * (size_t)S.light_inst[idx];
        const float4 r0 = T[0], r1 = T[1], r2 = T[2];
        ls.pos = xf_point(r0, r1, r2, ls.pos);
        ls.normal = xf_vector(r0, r1, r2, ls.normal);
    }
    const MatRec& m = S.mats[S.mesh_mat[mesh]];
    ls.emission = mk(m.emission_color[0], m.emission_color[1], m.emission_color[2]) * m.intensity;
    ls.pdf = m.intensity / S.light_wsum;
    return ls;
}

SRT_D bool finite3(f3 v) {
    return !(__builtin_isnan(v.x) || __builtin_isinf(v.x) || __builtin_isnan(v.y) || __builtin_isinf(v.y) ||
             __builtin_isnan(v.z) || __builtin_isinf(v.z));
}

// One bounce for every entry of queue `qi` (path_tracer.cu:80-161).
// Shading of one hit (queue entry i, hit record i): hit reconstruction,
// emitter term with MIS, next-event estimation, the radiance-MLP hooks, BSDF
// sampling and Russian roulette.  Appends go to the lanes active at the call.
// First-hit geometry of sorted triangle tri at barycentrics (u, v)
// (hit_info.h:27-65, identity instance transform): position, the geometric
// normal turned to the side of the interpolated shading normal, the shading
// normal, and the material id.
struct HitGeom {
    f3 pos, normal, sn;
    float w;
    int32_t matid;
};
SRT_D HitGeom hit_geometry(const DevScene& S, int32_t tri, float u, float v) {
    HitGeom g;
    // the first five parts of the shading record, read together (one line)
    const float4* r = S.tri_rec + 8 * (size_t)tri;
    const float4 R0 = r[0], R1 = r[1], R2 = r[2], R3 = r[3], R4 = r[4];
    g.matid = __float_as_int(R0.w);
    const f3 v0 = mk(R0.x, R0.y, R0.z), v1 = mk(R1.x, R1.y, R1.z), v2 = mk(R2.x, R2.y, R2.z);
    g.w = 1.0f - u - v;
    g.pos = v0 * g.w + v1 * u + v2 * v;
    g.normal = normalize(cross(v1 - v0, v2 - v0));
    g.sn = g.normal;
    if (S.rec_flags & REC_NORMALS) {
        const f3 n0 = mk(R1.w, R2.w, R3.x), n1 = mk(R3.y, R3.z, R3.w), n2 = mk(R4.x, R4.y, R4.z);
        g.sn = normalize(n0 * g.w + n1 * u + n2 * v);
    }
    if (dot(g.normal, g.sn) < 0.0f) g.normal = -g.normal;
    return g;
}

// Instanced scenes (hit_info.h:58-62): the hit position by apply_point and the
// geometric normal by apply_vector of the instance transform (no
// renormalising); returns the transform for the shading frame.
SRT_D void inst_hit(const DevScene& S, int32_t inst, HitGeom& hg, float4* X = nullptr) {
    const float4* T = S.inst_xf + 3 * (size_t)inst;
    const float4 X0 = T[0], X1 = T[1], X2 = T[2];
    hg.pos = xf_point(X0, X1, X2, hg.pos);
    hg.normal = xf_vector(X0, X1, X2, hg.normal);
    if (X) { X[0] = X0; X[1] = X1; X[2] = X2; }
}
// Transform::apply_onb (transform.h:72-75): the frame axis by axis
SRT_D void inst_onb(const float4* X, Onb& onb) {
    onb.z = xf_vector(X[0], X[1], X[2], onb.z);
    onb.x = xf_vector(X[0], X[1], X[2], onb.x);
    onb.y = xf_vector(X[0], X[1], X[2], onb.y);
}

// The leading 64 bytes of a segment (path range and depth settings), read
// together as eight 8-byte parts.
struct SegHead {
    int64_t path_begin;
    int32_t n, slot_offset, wh, width, height, spp;
    uint32_t seed;
    int32_t max_depth, rr_depth, use_nee, use_mis, mlp_mode, last;
    float cam0;
};
static_assert(sizeof(SegHead) == 64 && offsetof(Segment, last) == offsetof(SegHead, last) &&
                  offsetof(Segment, mlp_mode) == offsetof(SegHead, mlp_mode) && sizeof(Segment) >= 64 &&
                  alignof(Segment) >= 8,
              "SegHead mirrors the head of Segment");
SRT_D SegHead load_seg_head(const Segment* segs, uint32_t k) {
    const uint2* p = reinterpret_cast<const uint2*>(segs + k);
    const uint2 q[8] = {p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]};
    SegHead h;
    __builtin_memcpy(&h, q, sizeof(SegHead));
    return h;
}

// Tangent for the shading frame (hit_info.h:51-55): the interpolated vertex
// tangent when the scene has tangents, else an axis picked from the shading
// normal.
SRT_D f3 hit_tangent(const DevScene& S, int32_t tri, const HitGeom& g, float u, float v) {
    if (S.rec_flags & REC_TANGENTS) {
        const float4* r = S.tri_rec + 8 * (size_t)tri;
        const float4 R4 = r[4], R5 = r[5], R6 = r[6];
        const f3 t0 = mk(R4.w, R5.x, R5.y), t1 = mk(R5.z, R5.w, R6.x), t2 = mk(R6.y, R6.z, R6.w);
        return t0 * g.w + t1 * u + t2 * v;
    }
    return (fabsf(g.sn.x) > fabsf(g.sn.y)) ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
}

// A material record by value: its five 16-byte parts are read together.
SRT_D MatRec load_mat(const MatRec* mats, int32_t id) {
    const float4* p = reinterpret_cast<const float4*>(mats + id);
    const float4 q[5] = {p[0], p[1], p[2], p[3], p[4]};
    MatRec m;
    __builtin_memcpy(&m, q, sizeof(MatRec));
    return m;
}

// CAM: the first bounce, whose entries are the slots and whose path state is
// path_start's; it also sets the path's radiance (no earlier stage does).
template <bool CAM>
SRT_D void shade_hit(const DevScene& S, const Segment* segs, int nseg, int depth, const PathQueue& qin, const HitBuf& h,
                     const PathQueue& qout, const ShadowQueue& sq, float* Lr, float* Lg, float* Lb, uint32_t* counts,
                     int qi, uint32_t part, uint32_t cap, const MlpIO& mio, uint32_t i) {
    PathStart ps{};
    if constexpr (CAM) ps = path_start(segs, nseg, i);
    const uint32_t slot = CAM ? i : qin.slot[i];
    const f3 rd = CAM ? ps.d : mk(qin.dx[i], qin.dy[i], qin.dz[i]);
    f3 beta = CAM ? mk(1.0f) : mk(qin.br[i], qin.bg[i], qin.bb[i]);
    if constexpr (CAM) {
        Lr[slot] = 0.0f; Lg[slot] = 0.0f; Lb[slot] = 0.0f;
    }
    const int32_t tri = h.tri[i];
    const uint32_t flags = CAM ? ps.flags : qin.flags[i];
    const bool specular = (flags & 1u) != 0u;
    const uint32_t segk = flags >> 8;
    const SegHead P = load_seg_head(segs, segk);
    const float u = h.u[i], v = h.v[i];
    HitGeom hg = hit_geometry(S, tri, u, v);
    float4 X[3];               // instance transform (instanced scenes)
    if (S.num_inst > 0) inst_hit(S, h.inst[i], hg, X);
    const float w = hg.w;
    const f3 pos = hg.pos, sn = hg.sn;
    f3 normal = hg.normal;
    const MatRec m = load_mat(S.mats, hg.matid);
    const f3 color = mk(m.base_color[0], m.base_color[1], m.base_color[2]);
    if (mat_is_emissive(m)) {
        // emitter hit with MIS (path_tracer.cu:88-106)
        const float cos_i = dot(normal, -rd);
        if (cos_i > 1e-4f) {
            float mis = 1.0f;
            if (P.use_nee && !specular) {
                mis = 0.0f;
                if (P.use_mis) {
                    const f3 ro = CAM ? ps.o : mk(qin.ox[i], qin.oy[i], qin.oz[i]);
                    const float t2 = dot(pos - ro, pos - ro);
                    const float lpdf = (m.intensity / S.light_wsum) * t2 / cos_i;
                    const float scatter_pdf = CAM ? 1.0f : qin.pdf[i];
                    mis = scatter_pdf / (lpdf + scatter_pdf);
                }
            }
            const f3 Le = mk(m.emission_color[0], m.emission_color[1], m.emission_color[2]) * m.intensity;
            const f3 add = beta * Le * mis;
            add_rgb(Lr, Lg, Lb, slot, add.x, add.y, add.z);
        }
        return;
    }
    const f3 tg = hit_tangent(S, tri, hg, u, v);
    Onb onb = make_onb(sn, tg);
    if (S.num_inst > 0) inst_onb(X, onb);
    const ShadePoint spt = shade_point(m, -rd, onb, color);
    if (dot(normal, -rd) < 0.0f) normal = -normal;
    const bool last = (depth + 1 >= P.max_depth);
    const int64_t pid = P.path_begin + (int64_t)(slot - (uint32_t)P.slot_offset);
    const uint32_t sample = path_sample(pid, P.wh);
    Pcg32 rng = rng_resume(CAM ? ps.rng : qin.rng[i], sample);
    // next-event estimation (path_tracer.cu:113-142); the shadow ray and its
    // would-be contribution go to the shadow queue
    if (P.use_nee && !mat_is_specular(m)) {
        f2 lr;
        lr.x = rng.next_float();
        lr.y = rng.next_float();
        const LightSample ls = light_sample(S, lr);
        const f3 sdir = normalize(ls.pos - pos);
        const float cos_o = dot(normal, sdir);
        const float cos_light = dot(ls.normal, -sdir);
        const float t = length(ls.pos - pos);
        const bool emit = (cos_light > 1e-4f) && (ls.pdf > 1e-4f) && (t > 1e-3f) && (cos_o > 0.0f || mat_is_transmissive(m));
        f3 s_c = mk(0.0f);
        if (emit) {
            float mis = 1.0f;
            const float lpdf = ls.pdf * t * t / cos_light;
            if (P.use_mis) {
                const float mpdf = sp_pdf(m, spt, sdir, color);
                mis = lpdf / (lpdf + mpdf);
            }
            const f3 f = sp_eval(m, spt, sdir, color);
            s_c = beta * f * fabsf(cos_o) * ls.emission * mis / lpdf;
        }
        const int ja = wave_append(emit, counts + CNT_SHADOW + part * kCntLine);
        if (ja >= 0) {
            const uint32_t js = part * cap + (uint32_t)ja;
            sq.ox[js] = pos.x; sq.oy[js] = pos.y; sq.oz[js] = pos.z;
            sq.dx[js] = sdir.x; sq.dy[js] = sdir.y; sq.dz[js] = sdir.z;
            sq.tmax[js] = t - 1e-3f;
            sq.cr[js] = s_c.x; sq.cg[js] = s_c.y; sq.cb[js] = s_c.z;
            sq.slot[js] = slot;
        }
    }
    if (P.mlp_mode != 0 && depth == 1 && !last) {
        // radiance-MLP hooks at vertex 2; features: position in the scene
        // box, (n+1)/2, (w_o+1)/2
        float* q9;
        if (P.mlp_mode == 1) {
            // the path stops here; the prediction is added after the bounce's shadow rays
            const int ja = wave_append(true, counts + CNT_MLPQ_PARTS + part * kCntLine);
            q9 = mio.qs_in9 + (size_t)(part * cap + (uint32_t)ja) * 9;
        } else {
            q9 = mio.rec_in9 + (size_t)slot * 9;
        }
        q9[0] = (pos.x - mio.lo[0]) * mio.inv_ext[0];
        q9[1] = (pos.y - mio.lo[1]) * mio.inv_ext[1];
        q9[2] = (pos.z - mio.lo[2]) * mio.inv_ext[2];
        q9[3] = (normal.x + 1.0f) * 0.5f; q9[4] = (normal.y + 1.0f) * 0.5f; q9[5] = (normal.z + 1.0f) * 0.5f;
        q9[6] = (-rd.x + 1.0f) * 0.5f; q9[7] = (-rd.y + 1.0f) * 0.5f; q9[8] = (-rd.z + 1.0f) * 0.5f;
        if (P.mlp_mode == 1) {
            const size_t jq = (size_t)(q9 - mio.qs_in9) / 9;
            mio.qs_beta[jq * 3] = beta.x;
            mio.qs_beta[jq * 3 + 1] = beta.y;
            mio.qs_beta[jq * 3 + 2] = beta.z;
            mio.qs_slot[jq] = slot;
            return;
        }
        mio.rec_beta[(size_t)slot * 3] = beta.x;
        mio.rec_beta[(size_t)slot * 3 + 1] = beta.y;
        mio.rec_beta[(size_t)slot * 3 + 2] = beta.z;
        mio.rec_valid[slot] = 1u;
    }
    if (last) return;
    // BSDF sampling, throughput update, Russian roulette (path_tracer.cu:145-161)
    f2 br;
    br.x = rng.next_float();
    br.y = rng.next_float();
    const BxSample ms = sp_sample(m, spt, br, color);
    bool alive = !(ms.pdf <= 1e-4f);
    if (alive && dot(ms.wi, normal) < 0.0f && !mat_is_transmissive(m)) alive = false;
    if (alive) {
        beta = beta * (ms.f / ms.pdf);
        if (!ms.delta) beta = beta * fabsf(dot(ms.wi, normal));
        if (depth >= P.rr_depth) {
            const float q = fmaxf(fmaxf(beta.x, fmaxf(beta.y, beta.z)), 0.05f);
            if (rng.next_float() > q) alive = false;
            else beta = beta / q;
        }
    }
    const int ja = wave_append(alive, counts + cnt_q(qi ^ 1) + part * kCntLine);
    if (ja >= 0) {
        const uint32_t jn = part * cap + (uint32_t)ja;
        qout.ox[jn] = pos.x; qout.oy[jn] = pos.y; qout.oz[jn] = pos.z;
        qout.dx[jn] = ms.wi.x; qout.dy[jn] = ms.wi.y; qout.dz[jn] = ms.wi.z;
        qout.br[jn] = beta.x; qout.bg[jn] = beta.y; qout.bb[jn] = beta.z;
        qout.pdf[jn] = ms.pdf;
        qout.flags[jn] = (mat_is_specular(m) ? 1u : 0u) | (segk << 8);
        qout.rng[jn] = rng.state;
        qout.slot[jn] = slot;
    }
}

// Per entry of queue `qi`: a miss adds beta * env (path_tracer.cu:81-85) and
// ends the path; a hit is shaded.  About two thirds of the entries of an open
// scene are misses, so each wave first gathers the indices of 64 hits in its
// own LDS area (ballot + prefix count, no global atomics) and then shades them
// with every lane busy.
constexpr int kShadeBlock = 256;
// orders one wave's LDS accesses across lanes (LDS executes a wave's operations
// in order; this keeps the compiler from moving them across the point)
SRT_D void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
template <bool CAM>
__global__ void __launch_bounds__(kShadeBlock, 4) k_shade(DevScene S, const Segment* segs, int nseg, int depth,
                                                         PathQueue qin, HitBuf h, PathQueue qout, ShadowQueue sq,
                                                         float* Lr, float* Lg, float* Lb, uint32_t* counts, int qi,
                                                         uint32_t cap, MlpIO mio) {
    __shared__ uint32_t gather[kShadeBlock / 64][128];
    const int lane = threadIdx.x & 63;
    uint32_t* buf = gather[threadIdx.x >> 6];
    const unsigned long long below = (1ull << lane) - 1ull;
    // the waves of the workgroups of one piece share its entries; their
    // appends go to partition `part`
    __shared__ uint32_t prefix[kParts + 1];
    uint32_t lo, hi;
    const QueueView view = queue_piece(counts, cnt_q(qi), cap, prefix, lo, hi);
    const uint32_t part = my_part();
    const uint32_t count = hi - lo;
    const int wave = (int)((blockIdx.x / (unsigned)kParts) * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int nwaves = (int)((gridDim.x / (unsigned)kParts) * (blockDim.x >> 6));
    if (blockIdx.x == 0) clear_parts(counts + CNT_FETCH_T);   // the next trace stage fetches from 0
    const f3 env = mk(S.env[0], S.env[1], S.env[2]);
    uint32_t nbuf = 0;                     // gathered hits (wave-uniform)
    uint32_t base = (uint32_t)wave * 64u;
    for (;;) {
        // gather until 64 hits are waiting or the partition is exhausted
        while (nbuf < 64u && base < count) {
            const uint32_t r = base + (uint32_t)lane;
            const bool valid = r < count;
            const uint32_t i = valid ? view.entry(lo + r) : 0u;
            const int32_t tri = valid ? h.tri[i] : -1;
            if (valid && tri < 0) {
                if constexpr (CAM) {
                    // a camera ray's miss: its radiance is beta (1) times the environment
                    Lr[i] = 0.0f + env.x; Lg[i] = 0.0f + env.y; Lb[i] = 0.0f + env.z;
                } else {
                    const uint32_t slot = qin.slot[i];
                    const f3 add = mk(qin.br[i], qin.bg[i], qin.bb[i]) * env;
                    add_rgb(Lr, Lg, Lb, slot, add.x, add.y, add.z);
                }
            }
            const bool hit = valid && tri >= 0;
            const unsigned long long m = __ballot(hit);
            if (hit) buf[nbuf + (uint32_t)__popcll(m & below)] = i;
            nbuf += (uint32_t)__popcll(m);
            base += (uint32_t)nwaves * 64u;
        }
        if (nbuf == 0u) break;
        wave_sync_lds();
        const uint32_t take = nbuf < 64u ? nbuf : 64u;
        const uint32_t e = buf[lane];
        const uint32_t rest = (lane + 64u < nbuf) ? buf[lane + 64] : 0u;
        wave_sync_lds();
        if (lane + 64u < nbuf) buf[lane] = rest;
        nbuf -= take;
        if ((uint32_t)lane < take)
            shade_hit<CAM>(S, segs, nseg, depth, qin, h, qout, sq, Lr, Lg, Lb, counts, qi, part, cap, mio, e);
        wave_sync_lds();
    }
}

// Shadow rays of the bounce (fetch counter cleared by the trace stage).
template <bool COUNT, bool WIDE, bool INST = false>
__global__ void __launch_bounds__(kTraceBlock) k_shadow(DevScene S, ShadowQueue sq, float* Lr, float* Lg, float* Lb,
                                                       uint32_t* counts, uint32_t cap, int32_t* gstack,
                                                       unsigned long long* totals, TraceCfg cfg) {
    __shared__ uint32_t prefix[kParts + 1];
    uint32_t lo, hi;
    const QueueView view = queue_piece(counts, CNT_SHADOW, cap, prefix, lo, hi);
    uint32_t* fetch = counts + CNT_FETCH_S + my_part() * kCntLine;
    if (blockIdx.x < (unsigned)kParts && threadIdx.x == 0) atomicAdd(&totals[1], (unsigned long long)(hi - lo));
    const TravStack k = make_stack<INST>(gstack, cfg);
    ShadowIO io{sq, Lr, Lg, Lb};
    uint32_t nv = 0, nt = 0;
    trace_persistent<true, COUNT, WIDE, ShadowIO, QueueView, INST>(S, view, lo, hi, fetch, cfg.refill, k, io, nv, nt);
    if (COUNT) {
        atomicAdd(&totals[4], (unsigned long long)nv);
        atomicAdd(&totals[5], (unsigned long long)nt);
    }
}

// MLP mode 1: the queries appended to the partitions, moved to [0, total) in
// the order of the partitions' concatenation; counts[CNT_MLPQ] = total
__global__ void __launch_bounds__(256) k_mlp_compact(uint32_t* counts, uint32_t cap, MlpIO mio) {
    __shared__ uint32_t prefix[kParts + 1];
    uint32_t lo, hi;
    const QueueView view = queue_piece(counts, CNT_MLPQ_PARTS, cap, prefix, lo, hi);
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[CNT_MLPQ] = prefix[kParts];
    const uint32_t step = (gridDim.x / (unsigned)kParts) * blockDim.x;
    for (uint32_t r = (blockIdx.x / (unsigned)kParts) * blockDim.x + threadIdx.x; r < hi - lo; r += step) {
        const size_t j = lo + r, e = view.entry(lo + r);
        for (int k = 0; k < 9; k++) mio.q_in9[j * 9 + k] = mio.qs_in9[e * 9 + k];
        for (int k = 0; k < 3; k++) mio.q_beta[j * 3 + k] = mio.qs_beta[e * 3 + k];
        mio.q_slot[j] = mio.qs_slot[e];
    }
}

// MLP mode 1: L += beta * max(prediction, 0) for the queued vertex-2 paths
__global__ void __launch_bounds__(256) k_mlp_apply(const uint32_t* counts, const float* pred, const float* beta,
                                                   const uint32_t* slot, float* Lr, float* Lg, float* Lb) {
    const uint32_t n = counts[CNT_MLPQ];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t sl = slot[i];
        Lr[sl] = Lr[sl] + beta[3 * (size_t)i] * fmaxf(pred[3 * (size_t)i], 0.0f);
        Lg[sl] = Lg[sl] + beta[3 * (size_t)i + 1] * fmaxf(pred[3 * (size_t)i + 1], 0.0f);
        Lb[sl] = Lb[sl] + beta[3 * (size_t)i + 2] * fmaxf(pred[3 * (size_t)i + 2], 0.0f);
    }
}

// MLP mode 2: radiance gathered so far (after vertex 2's NEE) per slot
__global__ void __launch_bounds__(256) k_snapshot(int n, const float* Lr, const float* Lg, const float* Lb, float* snap) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    snap[3 * (size_t)i] = Lr[i]; snap[3 * (size_t)i + 1] = Lg[i]; snap[3 * (size_t)i + 2] = Lb[i];
}

// MLP mode 2: per-slot target (L - L_snapshot) / beta, flag valid finite samples
__global__ void __launch_bounds__(256) k_targets(int n, const float* Lr, const float* Lg, const float* Lb, const float* snap,
                                                 const float* rec_beta, const uint32_t* rec_valid, float* target,
                                                 uint32_t* flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float L[3] = {Lr[i], Lg[i], Lb[i]};
    bool ok = rec_valid[i] != 0u;
    for (int k = 0; k < 3; k++) {
        float b = rec_beta[3 * (size_t)i + k];
        float t = (b > 0.0f) ? (L[k] - snap[3 * (size_t)i + k]) / b : 0.0f;
        if (__builtin_isnan(t) || __builtin_isinf(t)) ok = false;
        target[3 * (size_t)i + k] = t;
    }
    flag[i] = ok ? 1u : 0u;
}

// Per pixel: add this chunk's samples in sample order (path_tracer.cu:164-165),
// dropping non-finite samples but keeping the 1/spp weight.  One grid row per
// segment; a request's film is written when its last segment completes.
__global__ void __launch_bounds__(256) k_finalize(const float* Lr, const float* Lg, const float* Lb, const Segment* segs) {
    const Segment& sg = segs[blockIdx.y];
    int q = blockIdx.x * blockDim.x + threadIdx.x;     // tile rank of the pixel
    if (q >= sg.wh) return;
    const uint32_t px = tile_pixel((uint32_t)q, (uint32_t)sg.width, (uint32_t)sg.height);
    const int64_t end = sg.path_begin + sg.n;
    int64_t s0 = (sg.path_begin - q + sg.wh - 1) / sg.wh;
    if (s0 < 0) s0 = 0;
    float* acc = sg.film_acc;
    float ax = acc[3 * (size_t)px], ay = acc[3 * (size_t)px + 1], az = acc[3 * (size_t)px + 2];
    const float fspp = (float)sg.spp;
    for (int64_t s = s0;; s++) {
        int64_t pid = s * sg.wh + q;
        if (pid >= end) break;
        size_t slot = (size_t)(sg.slot_offset + (pid - sg.path_begin));
        f3 L = mk(Lr[slot], Lg[slot], Lb[slot]);
        if (finite3(L)) {
            ax = ax + L.x / fspp;
            ay = ay + L.y / fspp;
            az = az + L.z / fspp;
        }
    }
    acc[3 * (size_t)px] = ax; acc[3 * (size_t)px + 1] = ay; acc[3 * (size_t)px + 2] = az;
    if (sg.last && sg.film_out) sg.film_out[px] = make_float4(ax, ay, az, 1.0f);
}

__global__ void k_zero(float* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// ---------------------------------------------------------------------------
// standalone ray queries (first-hit parity, shadow parity)
// ---------------------------------------------------------------------------
struct RaysIO {        // standalone queries on flat arrays
    const float *o3, *d3, *tmax;
    float tmin;
    float *t, *u, *v;
    int32_t* prim;
    int32_t* inst = nullptr;   // instanced scenes: the instance hit (may be null)
    const float4* tris;
    int any;
    template <bool INST>
    SRT_D void load(uint32_t i, TravState& st, int32_t root) const {
        trav_begin<INST>(st, ld3(o3 + 3 * (size_t)i), ld3(d3 + 3 * (size_t)i), tmin, tmax[i], root, !any);
    }
    SRT_D void store(uint32_t i, const TravState& st, bool occluded) const {
        if (any) {
            prim[i] = occluded ? 1 : 0;
            return;
        }
        const bool hit = st.best_pos >= 0;
        t[i] = hit ? st.best_t : 0.0f;
        u[i] = hit ? st.bu : 0.0f;
        v[i] = hit ? st.bv : 0.0f;
        prim[i] = hit ? __float_as_int(tris[3 * (size_t)st.best_pos].w) : -1;
        if (inst) inst[i] = hit ? st.best_inst : -1;
    }
};

template <bool INST = false>
__global__ void __launch_bounds__(kTraceBlock) k_trace_rays(DevScene S, uint32_t n, RaysIO io, uint32_t* fetch,
                                                           int32_t* gstack, TraceCfg cfg) {
    const TravStack k = make_stack<INST>(gstack, cfg);
    uint32_t nv = 0, nt = 0, lo, hi;
    fetch = dense_range(n, fetch, lo, hi);
    const IdentityEntry view;
    if (INST) {
        if (io.any) trace_persistent<true, false, true, RaysIO, IdentityEntry, true>(S, view, lo, hi, fetch, cfg.refill, k, io, nv, nt);
        else trace_persistent<false, false, true, RaysIO, IdentityEntry, true>(S, view, lo, hi, fetch, cfg.refill, k, io, nv, nt);
    } else if (cfg.wide) {
        if (io.any) trace_persistent<true, false, true>(S, view, lo, hi, fetch, cfg.refill, k, io, nv, nt);
        else trace_persistent<false, false, true>(S, view, lo, hi, fetch, cfg.refill, k, io, nv, nt);
    } else {
        if (io.any) trace_persistent<true, false, false>(S, view, lo, hi, fetch, cfg.refill, k, io, nv, nt);
        else trace_persistent<false, false, false>(S, view, lo, hi, fetch, cfg.refill, k, io, nv, nt);
    }
}

// camera rays of one sample for every pixel (index = pixel) -> hit buffer
struct SimpleIO {
    DevCamera cam;
    int width, height;
    uint32_t seed, sample;
    HitBuf h;
    template <bool INST>
    SRT_D void load(uint32_t i, TravState& st, int32_t root) const {
        f3 ro, rd;
        Pcg32 rng;
        camera_ray(cam, width, height, i, seed, sample, ro, rd, rng);
        trav_begin<INST>(st, ro, rd, 1e-3f, 1e16f, root);
    }
    SRT_D void store(uint32_t i, const TravState& st, bool) const {
        const bool hit = st.best_pos >= 0;
        h.u[i] = hit ? st.bu : 0.0f;
        h.v[i] = hit ? st.bv : 0.0f;
        h.tri[i] = hit ? st.best_pos : -1;
        if (h.inst) h.inst[i] = hit ? st.best_inst : -1;
    }
};

template <bool WIDE, bool INST>
__global__ void __launch_bounds__(kTraceBlock) k_simple_trace(DevScene S, uint32_t n, SimpleIO io, uint32_t* fetch,
                                                             int32_t* gstack, TraceCfg cfg) {
    const TravStack k = make_stack<INST>(gstack, cfg);
    uint32_t nv = 0, nt = 0, lo, hi;
    fetch = dense_range(n, fetch, lo, hi);
    trace_persistent<false, false, WIDE, SimpleIO, IdentityEntry, INST>(S, IdentityEntry{}, lo, hi, fetch, cfg.refill, k,
                                                                        io, nv, nt);
}

// the Simple pass types (SimpleParams::Type, params/npr_params.h)
enum SimpleType { kSimpleDepth = 0, kSimpleNormal = 1, kSimpleBaseColor = 2, kSimpleAmbient = 3, kSimpleFace = 4 };

// first-hit colouring of pixel i's ray of sample io.sample, added as L / spp
// to the pixel (simple.cu:62-95); misses are white
__global__ void __launch_bounds__(256) k_simple_accum(DevScene S, SimpleIO io, int spp, int type, float min_depth,
                                                     float max_depth, float4* acc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)io.width * (uint32_t)io.height) return;
    f3 ro, rd;
    Pcg32 rng;
    camera_ray(io.cam, io.width, io.height, i, io.seed, io.sample, ro, rd, rng);
    f3 L = mk(1.0f);
    const int32_t tri = io.h.tri[i];
    if (tri >= 0) {
        HitGeom hg = hit_geometry(S, tri, io.h.u[i], io.h.v[i]);
        if (S.num_inst > 0) inst_hit(S, io.h.inst[i], hg);
        const MatRec& m = S.mats[hg.matid];
        const f3 color = mk(m.base_color[0], m.base_color[1], m.base_color[2]);
        if (type == kSimpleDepth) {
            const float d = length(hg.pos - ld3(io.cam.pos));
            const float ratio = (d - min_depth) / (max_depth - min_depth);
            L = mk(ratio * 0.8f + 0.2f);
        } else if (type == kSimpleNormal) {
            L = (hg.normal + mk(1.0f)) * 0.5f;
        } else if (type == kSimpleBaseColor) {
            L = color;
        } else if (type == kSimpleAmbient) {
            L = color * (fabsf(dot(hg.normal, rd)) * 0.5f + 0.5f);
        } else {
            L = dot(hg.normal, rd) > 0.0f ? mk(1.0f, 0.85f, 0.0f) : mk(0.34f, 0.73f, 0.76f);
        }
    }
    const float fs = (float)spp;
    float4 a = io.sample == 0 ? make_float4(0.0f, 0.0f, 0.0f, 1.0f) : acc[i];
    a.x = a.x + L.x / fs;
    a.y = a.y + L.y / fs;
    a.z = a.z + L.z / fs;
    a.w = 1.0f;
    acc[i] = a;
}

// ---------------------------------------------------------------------------
// LightTracer pass (shader/integrator/light_tracer.cu:57-176,
// passes/integrator/light_tracer.cu:13-37): particles leave the area lights
// (cosine-distributed about the light normal), bounce through the scene, and
// splat into the film through the pinhole camera: at every non-specular
// vertex by a camera connection (a shadow ray toward the camera), and a
// specular particle that passes through the camera position directly.  The
// splat weights are the reference's own (aspect ratio, the previous
// direction's pdf pw), reproduced as written.  Splats are float atomic adds,
// so a pixel's sum order varies between runs (the reference's as well).
// ---------------------------------------------------------------------------
struct LtParams {
    DevCamera cam;
    int32_t width, height, spp, rr_depth;
    uint32_t seed;
    int64_t path_begin;      // path id of slot 0 of the chunk: pid = sample * W * H + launch index
    float aspect;            // width / height
};

// camera.h:143-148 get_xy; the pixel the direction `dir` (pointing away from
// the camera's image) lands in.  x < 1 can still round x * W up to W: such
// splats go to the last column / row.
SRT_D bool camera_pixel(const DevCamera& cam, f3 dir, int W, int H, int* idx) {
    const f3 d = -dir;
    const f3 l = mk(dot(d, ld3(cam.X)), dot(d, ld3(cam.Y)), dot(d, ld3(cam.Z)));
    const float px = l.x / l.z, py = l.y / l.z;
    const float x = px / cam.nw + 0.5f, y = py / cam.nh + 0.5f;
    if (!(x >= 0.0f && x < 1.0f && y >= 0.0f && y < 1.0f)) return false;
    const int ix = min((int)(x * (float)W), W - 1), iy = min((int)(y * (float)H), H - 1);
    *idx = ix + iy * W;
    return true;
}

SRT_D void splat(float4* film, int idx, f3 L) {
    atomicAdd(&film[idx].x, L.x);
    atomicAdd(&film[idx].y, L.y);
    atomicAdd(&film[idx].z, L.z);
}

__global__ void __launch_bounds__(256) k_lt_raygen(DevScene S, LtParams P, PathQueue q, int n, uint32_t cap,
                                                   uint32_t* counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0) {
        clear_parts(counts + CNT_FETCH_T);
        if (threadIdx.x < (unsigned)kParts)
            counts[CNT_Q0 + threadIdx.x * kCntLine] = part_size((uint32_t)n, cap, threadIdx.x);
    }
    if (i >= n) return;
    const int64_t wh = (int64_t)P.width * P.height;
    const int64_t pid = P.path_begin + i;
    const uint32_t sample = (uint32_t)(pid / wh);
    const uint32_t idx = (uint32_t)(pid - (int64_t)sample * wh);
    Pcg32 rng = path_rng(idx, P.seed, sample);
    f2 lr;
    lr.x = rng.next_float();
    lr.y = rng.next_float();
    const LightSample ls = light_sample(S, lr);
    f2 sr;
    sr.x = rng.next_float();
    sr.y = rng.next_float();
    const f3 wi = normalize(ls.normal + uniform_sample_sphere(sr));
    const float cos_theta = dot(wi, ls.normal);
    const float pw = cosine_hemisphere_pdf(cos_theta);
    const f3 beta = ls.emission * cos_theta / ls.pdf;
    const uint32_t e = (uint32_t)i;           // partition i / cap, in path order
    q.ox[e] = ls.pos.x; q.oy[e] = ls.pos.y; q.oz[e] = ls.pos.z;
    q.dx[e] = wi.x; q.dy[e] = wi.y; q.dz[e] = wi.z;
    q.br[e] = beta.x; q.bg[e] = beta.y; q.bb[e] = beta.z;
    q.pdf[e] = pw;
    q.flags[e] = 1u;                 // bit 0: the last scattering was specular (the light counts as one)
    q.rng[e] = rng.state;
    q.slot[e] = (uint32_t)i;
}

__global__ void __launch_bounds__(256) k_lt_shade(DevScene S, LtParams P, int depth, PathQueue qin, HitBuf h,
                                                 PathQueue qout, ShadowQueue sq, float4* film, uint32_t* counts, int qi,
                                                 uint32_t cap) {
    __shared__ uint32_t prefix[kParts + 1];
    uint32_t lo, hi;
    const QueueView view = queue_piece(counts, cnt_q(qi), cap, prefix, lo, hi);
    const uint32_t part = my_part();
    const uint32_t first = part * cap;     // of the output partition
    if (blockIdx.x == 0) clear_parts(counts + CNT_FETCH_T);   // the next trace stage fetches from 0
    const f3 cam_pos = ld3(P.cam.pos);
    const int64_t wh = (int64_t)P.width * P.height;
    const uint32_t step = (gridDim.x / (unsigned)kParts) * blockDim.x;
    for (uint32_t r = (blockIdx.x / (unsigned)kParts) * blockDim.x + threadIdx.x; r < hi - lo; r += step) {
        const uint32_t i = view.entry(lo + r);
        const f3 ro = mk(qin.ox[i], qin.oy[i], qin.oz[i]), rd = mk(qin.dx[i], qin.dy[i], qin.dz[i]);
        f3 beta = mk(qin.br[i], qin.bg[i], qin.bb[i]);
        float pw = qin.pdf[i];
        const uint32_t slot = qin.slot[i];
        const uint32_t sample = (uint32_t)((P.path_begin + slot) / wh);
        const int32_t tri = h.tri[i];
        const bool hit = tri >= 0;
        HitGeom hg;
        float4 X[3];
        if (hit) {
            hg = hit_geometry(S, tri, h.u[i], h.v[i]);
            if (S.num_inst > 0) inst_hit(S, h.inst[i], hg, X);
        }
        // a specular particle through the camera position (light_tracer.cu:89-110)
        if (qin.flags[i] & 1u) {
            f3 to_camera = cam_pos - ro;
            const float t = length(to_camera);
            to_camera = to_camera / t;
            if ((!hit || t < length(hg.pos - ro)) && dot(to_camera, rd) > 1.0f - 1e-6f) {
                int px;
                if (camera_pixel(P.cam, rd, P.width, P.height, &px)) splat(film, px, beta / pw / (float)P.spp);
            }
        }
        if (!hit) continue;
        const MatRec& m = S.mats[hg.matid];
        if (mat_is_emissive(m)) continue;
        f3 normal = hg.normal;
        if (dot(-rd, normal) < 0.0f) normal = -normal;
        const f3 color = mk(m.base_color[0], m.base_color[1], m.base_color[2]);
        const f3 tg = hit_tangent(S, tri, hg, h.u[i], h.v[i]);
        Onb onb = make_onb(hg.sn, tg);
        if (S.num_inst > 0) inst_onb(X, onb);
        const ShadePoint spt = shade_point(m, -rd, onb, color);
        Pcg32 rng = rng_resume(qin.rng[i], sample);
        // camera connection (light_tracer.cu:124-142): the contribution waits in the shadow queue
        bool emit = false;
        f3 sdir = mk(0.0f), L = mk(0.0f);
        float t = 0.0f;
        int px = 0;
        if (!mat_is_specular(m)) {
            sdir = normalize(cam_pos - hg.pos);
            if (camera_pixel(P.cam, sdir, P.width, P.height, &px)) {
                t = length(cam_pos - hg.pos);
                const f3 f = sp_eval(m, spt, sdir, color);
                L = P.aspect * beta * f * dot(sdir, normal) / t / t / pw;
                L = L / (float)P.spp;
                emit = true;
            }
        }
        const int ja = wave_append(emit, counts + CNT_SHADOW + part * kCntLine);
        if (ja >= 0) {
            const uint32_t js = first + (uint32_t)ja;
            sq.ox[js] = hg.pos.x; sq.oy[js] = hg.pos.y; sq.oz[js] = hg.pos.z;
            sq.dx[js] = sdir.x; sq.dy[js] = sdir.y; sq.dz[js] = sdir.z;
            sq.tmax[js] = t - 1e-3f;
            sq.cr[js] = L.x; sq.cg[js] = L.y; sq.cb[js] = L.z;
            sq.slot[js] = (uint32_t)px;
        }
        // next direction (light_tracer.cu:145-165)
        f2 br;
        br.x = rng.next_float();
        br.y = rng.next_float();
        const BxSample ms = sp_sample(m, spt, br, color);
        bool alive = !(ms.pdf <= 1e-5f);
        bool spec = false;
        if (alive) {
            const f3 f = sp_eval(m, spt, ms.wi, color);
            beta = beta * (f * fabsf(dot(ms.wi, normal)) / pw);
            pw = ms.pdf;
            spec = mat_is_specular(m);
            if (depth >= P.rr_depth) {
                const float p = fminf(fmaxf(beta.x, fmaxf(beta.y, beta.z)), 0.95f);
                if (rng.next_float() > p) alive = false;
                else beta = beta / p;
            }
        }
        const int ja2 = wave_append(alive, counts + cnt_q(qi ^ 1) + part * kCntLine);
        if (ja2 >= 0) {
            const uint32_t jn = first + (uint32_t)ja2;
            qout.ox[jn] = hg.pos.x; qout.oy[jn] = hg.pos.y; qout.oz[jn] = hg.pos.z;
            qout.dx[jn] = ms.wi.x; qout.dy[jn] = ms.wi.y; qout.dz[jn] = ms.wi.z;
            qout.br[jn] = beta.x; qout.bg[jn] = beta.y; qout.bb[jn] = beta.z;
            qout.pdf[jn] = pw;
            qout.flags[jn] = spec ? 1u : 0u;
            qout.rng[jn] = rng.state;
            qout.slot[jn] = slot;
        }
    }
}

struct LtShadowIO {    // camera connections: unoccluded ones splat their contribution
    ShadowQueue sq;
    float4* film;
    template <bool INST>
    SRT_D void load(uint32_t i, TravState& st, int32_t root) const {
        trav_begin<INST>(st, mk(sq.ox[i], sq.oy[i], sq.oz[i]), mk(sq.dx[i], sq.dy[i], sq.dz[i]), 1e-3f, sq.tmax[i], root,
                         false);
    }
    SRT_D void store(uint32_t i, const TravState&, bool occluded) const {
        if (!occluded) splat(film, (int)sq.slot[i], mk(sq.cr[i], sq.cg[i], sq.cb[i]));
    }
};

template <bool WIDE, bool INST>
__global__ void __launch_bounds__(kTraceBlock) k_lt_shadow(DevScene S, ShadowQueue sq, float4* film, uint32_t* counts,
                                                          uint32_t cap, int32_t* gstack, TraceCfg cfg) {
    __shared__ uint32_t prefix[kParts + 1];
    uint32_t lo, hi;
    const QueueView view = queue_piece(counts, CNT_SHADOW, cap, prefix, lo, hi);
    uint32_t* fetch = counts + CNT_FETCH_S + my_part() * kCntLine;
    const TravStack k = make_stack<INST>(gstack, cfg);
    LtShadowIO io{sq, film};
    uint32_t nv = 0, nt = 0;
    trace_persistent<true, false, WIDE, LtShadowIO, QueueView, INST>(S, view, lo, hi, fetch, cfg.refill, k, io, nv, nt);
}

__global__ void __launch_bounds__(256) k_film_clear(float4* film, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) film[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
}

__global__ void k_clear_parts(uint32_t* first) { clear_parts(first); }

// ---------------------------------------------------------------------------
// post-processing (tonemapping.cu:14-68, accumulate.cu:22-26)
// ---------------------------------------------------------------------------
__global__ void k_tonemap(float4* px, int64_t n, int type, float ex, int use_gamma) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = px[i];
    f3 c = mk(p.x, p.y, p.z) * ex;
    switch (type) {
    case 1: c = mk(clampf(c.x, 0.0f, 1.0f), clampf(c.y, 0.0f, 1.0f), clampf(c.z, 0.0f, 1.0f)); break;
    case 2: { float L = luminance(c); c = c / (1.0f + L); break; }
    case 3: {
        const float A = 0.15f, B = 0.50f, Cc = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
        float cx = ((c.x * (A * c.x + Cc * B) + D * E) / (c.x * (A * c.x + B) + D * F)) - E / F;
        float cy = ((c.y * (A * c.y + Cc * B) + D * E) / (c.y * (A * c.y + B) + D * F)) - E / F;
        float cz = ((c.z * (A * c.z + Cc * B) + D * E) / (c.z * (A * c.z + B) + D * F)) - E / F;
        const float Wt = 11.2f;
        float white = ((Wt * (A * Wt + Cc * B) + D * E) / (Wt * (A * Wt + B) + D * F)) - E / F;
        c = mk(cx, cy, cz) / white;
        break;
    }
    case 4: {
        c = c * 0.6f;
        const float a = 2.51f, b = 0.03f, cc = 2.43f, d = 0.59f, e = 0.14f;
        float cx = (c.x * (a * c.x + b)) / (c.x * (cc * c.x + d) + e);
        float cy = (c.y * (a * c.y + b)) / (c.y * (cc * c.y + d) + e);
        float cz = (c.z * (a * c.z + b)) / (c.z * (cc * c.z + d) + e);
        c = mk(clampf(cx, 0.0f, 1.0f), clampf(cy, 0.0f, 1.0f), clampf(cz, 0.0f, 1.0f));
        break;
    }
    default: break;
    }
    if (use_gamma) {
        const float g = 1.0f / 2.2f;
        c = mk(pt_powf(c.x, g), pt_powf(c.y, g), pt_powf(c.z, g));
    }
    px[i] = make_float4(c.x, c.y, c.z, 1.0f);
}

__global__ void k_accumulate(float4* acc, float4* px, int64_t n, int frame_count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float w = 1.0f / (float)(frame_count + 1);
    float4 a = acc[i], p = px[i];
    float4 r = make_float4(a.x * (1.0f - w) + p.x * w, a.y * (1.0f - w) + p.y * w, a.z * (1.0f - w) + p.z * w,
                           a.w * (1.0f - w) + p.w * w);
    acc[i] = r;
    px[i] = r;
}

// Denoise pass.  The reference runs the OptiX HDR denoiser (a closed network)
// on the colour alone, no albedo/normal guide layers
// (passes/postprocess/denoise.cu:54-114); this pass is the edge-avoiding
// a-trous wavelet filter of Dammertz et al. (HPG 2010) with the same single
// input: kDenoisePasses passes of the B3-spline kernel (1, 4, 6, 4, 1) / 16 at
// tap spacing 2^i, each tap weighted by exp(-|g_p - g_q|^2 / sigma_i^2) on the
// log colour g = log(1 + c) of the pass's input (HDR-robust: a difference is
// a ratio), sigma_i = kDenoiseSigma 2^-i; taps outside the image are skipped
// and the weights renormalised.  Alpha is kept.
constexpr int kDenoisePasses = 5;
constexpr float kDenoiseSigma = 0.2f;

// the edge-stopping guide of one pass: log(1 + c) per channel
__global__ void __launch_bounds__(256) k_denoise_guide(const float4* in, float4* g, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = in[i];
    g[i] = make_float4(pt_logf(1.0f + c.x), pt_logf(1.0f + c.y), pt_logf(1.0f + c.z), 0.0f);
}

__global__ void __launch_bounds__(256) k_atrous(const float4* in, const float4* g, float4* out, int W, int H, int step,
                                               float inv_s2) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const float h[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    const float4 cp = in[(size_t)y * W + x];
    const float4 g4 = g[(size_t)y * W + x];
    const f3 gp = mk(g4.x, g4.y, g4.z);
    f3 sum = mk(0.0f);
    float wsum = 0.0f;
    for (int j = -2; j <= 2; j++) {
        const int yy = y + j * step;
        if (yy < 0 || yy >= H) continue;
        for (int i = -2; i <= 2; i++) {
            const int xx = x + i * step;
            if (xx < 0 || xx >= W) continue;
            const float4 cq = in[(size_t)yy * W + xx];
            const float4 gq = g[(size_t)yy * W + xx];
            const f3 d = gp - mk(gq.x, gq.y, gq.z);
            const float w = (h[j + 2] * h[i + 2]) * pt_expf(-(dot(d, d) * inv_s2));
            sum = sum + mk(cq.x, cq.y, cq.z) * w;
            wsum = wsum + w;
        }
    }
    out[(size_t)y * W + x] = make_float4(sum.x / wsum, sum.y / wsum, sum.z / wsum, cp.w);
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
static int num_cus(srt_ctx* c) {
    static int cached = -1;
    if (cached < 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) return 256;
        cached = prop.multiProcessorCount;
    }
    return cached;
}

// The 4-wide tree is traversed when the context asks for it and the scene
// has one (a tree beyond 32-bit node offsets keeps only its binary form).
static bool use_wide(const srt_ctx* c, const SceneRec* s) {
    if (s && s->ninst > 0) return true;      // instanced scenes hold only the 4-wide form
    return c->trace_bvh_width == 4 && (s == nullptr || s->num_nodes4 > 0 || s->root < 0);
}

static TraceCfg trace_cfg(srt_ctx* c, const SceneRec* s) {
    TraceCfg cfg;
    cfg.lds_depth = c->trace_lds_depth;
    cfg.refill = c->trace_refill;
    cfg.wide = use_wide(c, s) ? 1 : 0;
    cfg.inner_exit = c->trace_inner_exit;
    cfg.packet = 0;
    return cfg;
}
// Packet traversal of a launch's camera rays (TraceCfg.packet: the number of
// lanes holding postponed leaves that makes a wave test them; 0: per-ray
// traversal): a 4-wide flat scene whose wave stack (3 entries per level) fits
// the rows of a wave's LDS stack area that the leaf lists leave.
static int trace_packet_ok(const srt_ctx* c, const SceneRec* s, const TraceCfg& cfg) {
    if (c->trace_packet <= 0 || !cfg.wide || s->ninst > 0 || s->num_nodes4 <= 0) return 0;
    const int64_t rows = cfg.lds_depth - kPacketLeafRows;
    return 3 * ((int64_t)s->depth4 + 1) <= 64 * rows ? c->trace_packet : 0;
}

static size_t trace_lds_bytes(const TraceCfg& cfg, bool inst = false) {
    return ((size_t)cfg.lds_depth * kTraceBlock + (inst ? kRayStash : 0)) * sizeof(int32_t);
}

// Persistent traversal grid: as many workgroups as are resident at once
// (occupancy of the traversal kernel with its LDS stack), capped by the work.
static int trace_grid(srt_ctx* c, int64_t n, const void* kernel = nullptr) {
    const TraceCfg cfg = trace_cfg(c, nullptr);
    int per_cu = c->trace_blocks_per_cu;
    if (per_cu <= 0) {
        int nb = 0;
        if (!kernel) kernel = reinterpret_cast<const void*>(k_trace<false, true>);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kTraceBlock, trace_lds_bytes(cfg)) != hipSuccess ||
            nb <= 0)
            nb = 8;
        per_cu = nb;
    }
    int64_t want = (n + kTraceBlock - 1) / kTraceBlock;
    int64_t cap = (int64_t)num_cus(c) * per_cu;
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (int)want;
}

// global part of the traversal stacks, per thread: the deepest stack minus the
// LDS entries.  A binary tree pushes at most one entry per level (depth + 1
// entries); the 4-wide tree at most three per level.
static int ensure_gstack(srt_ctx* c, const SceneRec* s, int64_t threads) {
    const int64_t deepest = use_wide(c, s) ? 3 * (int64_t)s->depth4 + 1 : (int64_t)s->max_depth + 1;
    const int64_t per = std::max<int64_t>(1, deepest - c->trace_lds_depth);
    const size_t need = (size_t)(threads * per);
    if (c->gstack_entries >= need) return SRT_OK;
    dev_free(c, c->gstack);
    c->gstack_entries = 0;
    if (int r = dev_alloc(c, c->gstack, need * sizeof(int32_t))) return r;
    c->gstack_entries = need;
    return SRT_OK;
}

// Grids of the kernels on partitioned queues: a multiple of kParts, so every
// partition has the same number of workgroups.
static int part_grid(int64_t g) {
    if (g < kParts) g = kParts;
    return (int)((g + kParts - 1) / kParts * kParts);
}

static int ensure_pool(srt_ctx* c, int64_t slots) {
    if (c->pool_cap >= slots) return SRT_OK;
    for (int k = 0; k < 2; k++)
        for (auto& b : c->q[k]) dev_free(c, b);
    for (auto& b : c->hits) dev_free(c, b);
    for (auto& b : c->shq) dev_free(c, b);
    for (auto& b : c->Lbuf) dev_free(c, b);
    c->pool_cap = 0;
    const size_t n = (size_t)slots;
    const size_t qn = (size_t)kParts * part_cap((uint32_t)slots);   // partitioned queue entries (>= n)
    for (int k = 0; k < 2; k++) {
        for (int j = 0; j < 14; j++) {
            size_t bytes = (j == 11) ? 8 * qn : 4 * qn;   // 11: rng (u64)
            if (j == 13) bytes = 0;                         // unused spare
            if (bytes && dev_alloc(c, c->q[k][j], bytes)) return SRT_E_OOM;
        }
    }
    for (int j = 1; j < 5; j++) if (dev_alloc(c, c->hits[j], 4 * qn)) return SRT_E_OOM;   // 1, 2: u, v; 3: triangle; 4: instance
    for (int j = 0; j < 11; j++) if (dev_alloc(c, c->shq[j], 4 * qn)) return SRT_E_OOM;
    for (int j = 0; j < 3; j++) if (dev_alloc(c, c->Lbuf[j], 4 * n)) return SRT_E_OOM;
    c->pool_cap = slots;
    return SRT_OK;
}

static PathQueue make_queue(srt_ctx* c, int k) {
    PathQueue q;
    auto f = [&](int j) { return (float*)c->q[k][j].p; };
    q.ox = f(0); q.oy = f(1); q.oz = f(2); q.dx = f(3); q.dy = f(4); q.dz = f(5);
    q.br = f(6); q.bg = f(7); q.bb = f(8); q.pdf = f(9);
    q.flags = (uint32_t*)c->q[k][10].p;
    q.rng = (uint64_t*)c->q[k][11].p;
    q.slot = (uint32_t*)c->q[k][12].p;
    return q;
}

DevCamera make_camera(const float pos[3], const float target[3], const float up[3], float fov, int W, int H) {
    // camera.h:41-53, 165-170; transform.h:136-145 (host float arithmetic)
    auto sub = [](const float* a, const float* b, float* o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; };
    auto norm = [](float* v) {
        float inv = 1.0f / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] = v[0] * inv; v[1] = v[1] * inv; v[2] = v[2] * inv;
    };
    auto crs = [](const float* a, const float* b, float* o) {
        o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
    };
    float z[3], x[3], y[3];
    sub(target, pos, z); norm(z);
    crs(up, z, x); norm(x);
    crs(z, x, y);
    norm(x); norm(y); norm(z);
    DevCamera cam;
    for (int i = 0; i < 3; i++) { cam.pos[i] = pos[i]; cam.X[i] = x[i]; cam.Y[i] = y[i]; cam.Z[i] = z[i]; }
    float aspect = (float)W / (float)H;
    float rad = (fov * 0.5f) * 3.14159265358979323846f / 180.0f;
    cam.nh = 2.0f * std::tan(rad);
    cam.nw = cam.nh * aspect;
    return cam;
}

__global__ void __launch_bounds__(256) k_scatter_train(int n, const uint32_t* flag, const uint32_t* pos, int64_t base,
                                                       int64_t cap, const float* rec_in9, const float* tgt, float* out_in9,
                                                       float* out_t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    int64_t j = base + pos[i];
    if (j >= cap) return;
    for (int k = 0; k < 9; k++) out_in9[j * 9 + k] = rec_in9[(size_t)i * 9 + k];
    for (int k = 0; k < 3; k++) out_t[j * 3 + k] = tgt[(size_t)i * 3 + k];
}

// per-slot buffers of the MLP hooks (allocated on first use, sized to the pool)
static int ensure_mlp_bufs(srt_ctx* c, int64_t slots) {
    if (c->mlp_pool_cap >= slots) return SRT_OK;
    for (auto& b : c->mlpbuf) dev_free(c, b);
    const size_t n = (size_t)slots;
    const size_t qn = (size_t)kParts * part_cap((uint32_t)slots);   // partitioned appends (>= n)
    const size_t sizes[13] = {36 * n, 12 * n, 4 * n, 12 * n, 36 * n, 12 * n, 4 * n, 12 * n, 4 * n, 4 * n,
                              36 * qn, 12 * qn, 4 * qn};
    // 0 q_in9, 1 q_beta, 2 q_slot, 3 pred, 4 rec_in9, 5 rec_beta, 6 rec_valid, 7 snap/target, 8 flag, 9 scan,
    // 10..12 the mode-1 queue as appended (q_in9, q_beta, q_slot per partition)
    for (int j = 0; j < 13; j++)
        if (int r = dev_alloc(c, c->mlpbuf[j], sizes[j])) return r;
    c->mlp_pool_cap = slots;
    return SRT_OK;
}

// Runs one chunk: segments (in slot order) of one scene share the path queues.
int run_chunk(srt_ctx* c, SceneRec* s, const Segment* segs_host, int nseg, TrainOut* train) {
    if (nseg <= 0 || nseg > kMaxSegments) return set_err(c, SRT_E_LIMIT, "bad segment count %d", nseg);
    int64_t n = 0;
    int max_depth = 0, mlp_mode = 0;
    bool any_nee = false;
    for (int k = 0; k < nseg; k++) {
        if (segs_host[k].slot_offset != n) return set_err(c, SRT_E_INVALID, "segments must be contiguous");
        n += segs_host[k].n;
        max_depth = std::max(max_depth, (int)segs_host[k].max_depth);
        any_nee |= segs_host[k].use_nee != 0;
        mlp_mode = std::max(mlp_mode, (int)segs_host[k].mlp_mode);
    }
    if (n <= 0 || n > 0x7fffffff) return set_err(c, SRT_E_LIMIT, "chunk of %lld paths", (long long)n);
    if (mlp_mode == 1 && !mlp_for_scene(c, s->handle)) return set_err(c, SRT_E_STATE, "mlp_mode 1 needs an MLP");
    const MlpBind bind(c, s->handle);       // mode 1 queries the scene's MLP (else the context's own)
    if (int r = ensure_pool(c, n)) return r;
    const int tgrid = part_grid(trace_grid(c, n));
    const int pgrid = part_grid(trace_grid(c, n, reinterpret_cast<const void*>(k_trace_packet)));
    const uint32_t cap = part_cap((uint32_t)n);
    if (int r = ensure_gstack(c, s, (int64_t)tgrid * kTraceBlock)) return r;
    const TraceCfg tcfg = trace_cfg(c, s);
    TraceCfg scfg = tcfg;
    scfg.refill = c->shadow_refill;
    const size_t tlds = trace_lds_bytes(tcfg);
    if (!c->segbuf.p) if (int r = dev_alloc(c, c->segbuf, sizeof(Segment) * kMaxSegments)) return r;
    hipStream_t st = c->stream;
    // the segment table is small; a synchronous copy keeps the host array free
    if (int r = hip_check(c, hipMemcpyAsync(c->segbuf.p, segs_host, sizeof(Segment) * nseg, hipMemcpyHostToDevice, st), "segments"))
        return r;
    const Segment* segs = (const Segment*)c->segbuf.p;
    MlpIO mio{};
    if (mlp_mode != 0) {
        if (int r = ensure_mlp_bufs(c, n)) return r;
        mio.q_in9 = (float*)c->mlpbuf[0].p;
        mio.q_beta = (float*)c->mlpbuf[1].p;
        mio.q_slot = (uint32_t*)c->mlpbuf[2].p;
        mio.qs_in9 = (float*)c->mlpbuf[10].p;
        mio.qs_beta = (float*)c->mlpbuf[11].p;
        mio.qs_slot = (uint32_t*)c->mlpbuf[12].p;
        mio.rec_in9 = (float*)c->mlpbuf[4].p;
        mio.rec_beta = (float*)c->mlpbuf[5].p;
        mio.rec_valid = (uint32_t*)c->mlpbuf[6].p;
        for (int i = 0; i < 3; i++) {
            float e = s->bounds_hi[i] - s->bounds_lo[i];
            mio.lo[i] = s->bounds_lo[i];
            mio.inv_ext[i] = e > 0.0f ? 1.0f / e : 1.0f;
        }
    }
    DevScene S = dev_scene(s);
    uint32_t* counts = (uint32_t*)c->counters.p;
    unsigned long long* totals = (unsigned long long*)((char*)c->counters.p + kTotalsOffset);
    PathQueue Q[2] = {make_queue(c, 0), make_queue(c, 1)};
    const bool inst = s->ninst > 0;
    HitBuf H;
    H.u = (float*)c->hits[1].p; H.v = (float*)c->hits[2].p; H.tri = (int32_t*)c->hits[3].p;
    H.inst = inst ? (int32_t*)c->hits[4].p : nullptr;
    const size_t ilds = trace_lds_bytes(tcfg, true);
    ShadowQueue SQ;
    {
        auto f = [&](int j) { return (float*)c->shq[j].p; };
        SQ.ox = f(0); SQ.oy = f(1); SQ.oz = f(2); SQ.dx = f(3); SQ.dy = f(4); SQ.dz = f(5); SQ.tmax = f(6);
        SQ.cr = f(7); SQ.cg = f(8); SQ.cb = f(9); SQ.slot = (uint32_t*)c->shq[10].p;
    }
    float* Lr = (float*)c->Lbuf[0].p; float* Lg = (float*)c->Lbuf[1].p; float* Lb = (float*)c->Lbuf[2].p;
    int32_t* gst = (int32_t*)c->gstack.p;
    const int shade_grid = part_grid(std::min<int64_t>((n + 255) / 256, (int64_t)num_cus(c) * 8));
    const int nblk = (int)((n + 255) / 256);
    hipEvent_t ev;
    // the first bounce's queue is the slots (its trace and shade stages compute
    // each path's camera ray and state); only its partition sizes are set here
    k_path_counts<<<1, 64, 0, st>>>((uint32_t)n, cap, counts);
    if (max_depth == 0)                   // no shade stage sets the radiance
        for (float* L : {Lr, Lg, Lb}) k_zero<<<nblk, 256, 0, st>>>(L, (size_t)n);
    if (mlp_mode == 2) k_zero<<<nblk, 256, 0, st>>>((float*)mio.rec_valid, (size_t)n);
    for (int depth = 0; depth < max_depth; depth++) {
        int qi = depth & 1;
        const bool cam = depth == 0;
        const bool cnt = c->count_visits;
        prof_begin(c, "trace", &ev);
        if (inst) {
            if (cam && cnt) k_trace_cam<true, true, true><<<tgrid, kTraceBlock, ilds, st>>>(S, segs, nseg, H, counts, cap, gst, totals, tcfg);
            else if (cam) k_trace_cam<false, true, true><<<tgrid, kTraceBlock, ilds, st>>>(S, segs, nseg, H, counts, cap, gst, totals, tcfg);
            else if (cnt) k_trace<true, true, true><<<tgrid, kTraceBlock, ilds, st>>>(S, Q[qi], H, counts, qi, cap, gst, totals, tcfg);
            else k_trace<false, true, true><<<tgrid, kTraceBlock, ilds, st>>>(S, Q[qi], H, counts, qi, cap, gst, totals, tcfg);
        } else if (tcfg.wide) {
            TraceCfg pcfg = tcfg;
            pcfg.packet = cam && !cnt ? trace_packet_ok(c, s, tcfg) : 0;
            if (cam && cnt) k_trace_cam<true, true><<<tgrid, kTraceBlock, tlds, st>>>(S, segs, nseg, H, counts, cap, gst, totals, tcfg);
            else if (pcfg.packet) k_trace_packet<<<pgrid, kTraceBlock, tlds, st>>>(S, segs, nseg, H, counts, cap, totals, pcfg);
            else if (cam) k_trace_cam<false, true><<<tgrid, kTraceBlock, tlds, st>>>(S, segs, nseg, H, counts, cap, gst, totals, tcfg);
            else if (cnt) k_trace<true, true><<<tgrid, kTraceBlock, tlds, st>>>(S, Q[qi], H, counts, qi, cap, gst, totals, tcfg);
            else k_trace<false, true><<<tgrid, kTraceBlock, tlds, st>>>(S, Q[qi], H, counts, qi, cap, gst, totals, tcfg);
        } else {
            if (cam && cnt) k_trace_cam<true, false><<<tgrid, kTraceBlock, tlds, st>>>(S, segs, nseg, H, counts, cap, gst, totals, tcfg);
            else if (cam) k_trace_cam<false, false><<<tgrid, kTraceBlock, tlds, st>>>(S, segs, nseg, H, counts, cap, gst, totals, tcfg);
            else if (cnt) k_trace<true, false><<<tgrid, kTraceBlock, tlds, st>>>(S, Q[qi], H, counts, qi, cap, gst, totals, tcfg);
            else k_trace<false, false><<<tgrid, kTraceBlock, tlds, st>>>(S, Q[qi], H, counts, qi, cap, gst, totals, tcfg);
        }
        prof_end(c, "trace", ev, 0);
        prof_begin(c, "shade", &ev);
        if (cam)
            k_shade<true><<<shade_grid, 256, 0, st>>>(S, segs, nseg, depth, Q[qi], H, Q[qi ^ 1], SQ, Lr, Lg, Lb, counts, qi, cap, mio);
        else
            k_shade<false><<<shade_grid, 256, 0, st>>>(S, segs, nseg, depth, Q[qi], H, Q[qi ^ 1], SQ, Lr, Lg, Lb, counts, qi, cap, mio);
        prof_end(c, "shade", ev, 0);
        if (any_nee) {
            prof_begin(c, "shadow", &ev);
            if (inst) {
                if (c->count_visits)
                    k_shadow<true, true, true><<<tgrid, kTraceBlock, ilds, st>>>(S, SQ, Lr, Lg, Lb, counts, cap, gst, totals, scfg);
                else k_shadow<false, true, true><<<tgrid, kTraceBlock, ilds, st>>>(S, SQ, Lr, Lg, Lb, counts, cap, gst, totals, scfg);
            } else if (tcfg.wide) {
                if (c->count_visits)
                    k_shadow<true, true><<<tgrid, kTraceBlock, tlds, st>>>(S, SQ, Lr, Lg, Lb, counts, cap, gst, totals, scfg);
                else k_shadow<false, true><<<tgrid, kTraceBlock, tlds, st>>>(S, SQ, Lr, Lg, Lb, counts, cap, gst, totals, scfg);
            } else {
                if (c->count_visits)
                    k_shadow<true, false><<<tgrid, kTraceBlock, tlds, st>>>(S, SQ, Lr, Lg, Lb, counts, cap, gst, totals, scfg);
                else k_shadow<false, false><<<tgrid, kTraceBlock, tlds, st>>>(S, SQ, Lr, Lg, Lb, counts, cap, gst, totals, scfg);
            }
            prof_end(c, "shadow", ev, 0);
        }
        if (depth == 1 && mlp_mode == 1) {
            // queued vertex-2 paths: prediction after the bounce's shadow rays (NEE first, as in the plain order)
            float* pred = (float*)c->mlpbuf[3].p;
            k_mlp_compact<<<shade_grid, 256, 0, st>>>(counts, cap, mio);
            if (int r = mlp_query_dev(c, mio.q_in9, counts + CNT_MLPQ, (int)n, pred)) return r;
            k_mlp_apply<<<shade_grid, 256, 0, st>>>(counts, pred, mio.q_beta, mio.q_slot, Lr, Lg, Lb);
        }
        if (depth == 1 && mlp_mode == 2) k_snapshot<<<nblk, 256, 0, st>>>((int)n, Lr, Lg, Lb, (float*)c->mlpbuf[7].p);
    }
    if (mlp_mode == 2 && train && train->count < train->cap) {
        float* snap_t = (float*)c->mlpbuf[7].p;     // snapshot, then overwritten by the targets
        uint32_t* flag = (uint32_t*)c->mlpbuf[8].p;
        uint32_t* pos = (uint32_t*)c->mlpbuf[9].p;
        k_targets<<<nblk, 256, 0, st>>>((int)n, Lr, Lg, Lb, snap_t, mio.rec_beta, mio.rec_valid, snap_t, flag);
        size_t tbytes = 0;
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tbytes, flag, pos, (int)n, st);
        if (c->scan_tmp.bytes < tbytes) {
            dev_free(c, c->scan_tmp);
            if (int r = dev_alloc(c, c->scan_tmp, tbytes)) return r;
        }
        if (int r = hip_check(c, hipcub::DeviceScan::ExclusiveSum(c->scan_tmp.p, tbytes, flag, pos, (int)n, st), "scan"))
            return r;
        k_scatter_train<<<nblk, 256, 0, st>>>((int)n, flag, pos, train->count, train->cap, mio.rec_in9, snap_t, train->in9,
                                              train->target);
        uint32_t lastpos = 0, lastflag = 0;
        if (int r = hip_check(c, hipMemcpyAsync(&lastpos, pos + n - 1, 4, hipMemcpyDeviceToHost, st), "count")) return r;
        if (int r = hip_check(c, hipMemcpyAsync(&lastflag, flag + n - 1, 4, hipMemcpyDeviceToHost, st), "count")) return r;
        if (int r = hip_check(c, hipStreamSynchronize(st), "train samples")) return r;
        train->count = std::min<int64_t>(train->cap, train->count + lastpos + lastflag);
    }
    int maxwh = 0;
    for (int k = 0; k < nseg; k++) maxwh = std::max(maxwh, (int)segs_host[k].wh);
    prof_begin(c, "finalize", &ev);
    k_finalize<<<dim3((unsigned)((maxwh + 255) / 256), (unsigned)nseg), 256, 0, st>>>(Lr, Lg, Lb, segs);
    prof_end(c, "finalize", ev, maxwh);
    // the host segment array may be reused by the caller after this returns
    return hip_check(c, hipStreamSynchronize(st), "chunk");
}

int render_wavefront(srt_ctx* c, SceneRec* s, const DevCamera& cam, const DevParams& p, float* film_dev, TrainOut* train) {
    const int64_t wh = (int64_t)p.width * p.height;
    const int64_t total = wh * p.spp;
    int64_t pool = c->pool_request;
    if (pool > total) pool = total;
    if (pool > 0x7fffffff) pool = 0x7fffffff;
    if (c->film_pixels < wh) {
        dev_free(c, c->film_acc);
        if (dev_alloc(c, c->film_acc, (size_t)wh * 3 * sizeof(float))) return SRT_E_OOM;
        c->film_pixels = wh;
    }
    hipStream_t st = c->stream;
    unsigned long long* totals = (unsigned long long*)((char*)c->counters.p + kTotalsOffset);
    {
        size_t n3 = (size_t)wh * 3;
        k_zero<<<(unsigned)((n3 + 255) / 256), 256, 0, st>>>((float*)c->film_acc.p, n3);
        k_zero<<<1, 64, 0, st>>>((float*)totals, 16);
    }
    if (train) train->count = 0;
    Segment sg;
    sg.slot_offset = 0;
    sg.wh = (int32_t)wh; sg.width = p.width; sg.height = p.height; sg.spp = p.spp;
    sg.seed = p.seed; sg.max_depth = p.max_depth; sg.rr_depth = p.rr_depth; sg.use_nee = p.use_nee;
    sg.use_mis = p.use_mis; sg.mlp_mode = p.mlp_mode;
    sg.cam = cam;
    sg.film_acc = (float*)c->film_acc.p;
    sg.film_out = (float4*)film_dev;
    for (int64_t begin = 0; begin < total; begin += pool) {
        sg.path_begin = begin;
        sg.n = (int32_t)((total - begin < pool) ? (total - begin) : pool);
        sg.last = (begin + pool >= total) ? 1 : 0;
        if (int r = run_chunk(c, s, &sg, 1, train)) return r;
    }
    if (int r = hip_check(c, hipGetLastError(), "render launch")) return r;
    unsigned long long tot[6] = {0, 0, 0, 0, 0, 0};
    if (int r = hip_check(c, hipMemcpyAsync(tot, totals, sizeof(tot), hipMemcpyDeviceToHost, st), "counters")) return r;
    if (int r = hip_check(c, hipStreamSynchronize(st), "render")) return r;
    c->last_segments = (int64_t)tot[0];
    c->last_shadow_rays = (int64_t)tot[1];
    for (int k = 0; k < 4; k++) c->last_visits[k] = (int64_t)tot[2 + k];
    prof_collect(c);
    return SRT_OK;
}

// Simple NPR pass (shader/npr/simple.cu:42-97, passes/npr/simple.cu:14-33):
// per sample, the camera rays of every pixel go through the persistent
// traversal, then each pixel adds L / spp of its first-hit colouring, in
// sample order.  min_depth / max_depth: the scene box's distances from the
// camera (aabb.h:24-41, computed on the host as passes/npr/simple.cu:28-30).
int render_simple(srt_ctx* c, SceneRec* s, const DevCamera& cam, int type, int spp, uint32_t seed, int width, int height,
                  float* out_dev) {
    const int64_t n = (int64_t)width * height;
    if (int r = ensure_pool(c, n)) return r;
    DevScene S = dev_scene(s);
    const int grid = trace_grid(c, n);
    if (int r = ensure_gstack(c, s, (int64_t)grid * kTraceBlock)) return r;
    const TraceCfg cfg = trace_cfg(c, s);
    float box_far[3], box_near[3];
    bool inside = true;
    for (int a = 0; a < 3; a++) {
        const float dlo = std::fabs(cam.pos[a] - s->bounds_lo[a]), dhi = std::fabs(cam.pos[a] - s->bounds_hi[a]);
        box_far[a] = std::fmax(dlo, dhi);
        box_near[a] = std::fmin(dlo, dhi);
        inside = inside && cam.pos[a] >= s->bounds_lo[a] && cam.pos[a] <= s->bounds_hi[a];
    }
    const float max_depth = std::sqrt(box_far[0] * box_far[0] + box_far[1] * box_far[1] + box_far[2] * box_far[2]);
    const float min_depth =
        inside ? 0.0f : std::sqrt(box_near[0] * box_near[0] + box_near[1] * box_near[1] + box_near[2] * box_near[2]);
    SimpleIO io;
    io.cam = cam;
    io.width = width;
    io.height = height;
    io.seed = seed;
    io.h.u = (float*)c->hits[1].p; io.h.v = (float*)c->hits[2].p;
    io.h.tri = (int32_t*)c->hits[3].p;
    const bool inst = s->ninst > 0;
    io.h.inst = inst ? (int32_t*)c->hits[4].p : nullptr;
    uint32_t* fetch = (uint32_t*)c->counters.p + CNT_FETCH_Q;
    hipEvent_t ev;
    prof_begin(c, "simple", &ev);
    for (int smp = 0; smp < spp; smp++) {
        io.sample = (uint32_t)smp;
        k_clear_parts<<<1, kParts, 0, c->stream>>>(fetch);
        if (inst)
            k_simple_trace<true, true><<<grid, kTraceBlock, trace_lds_bytes(cfg, true), c->stream>>>(
                S, (uint32_t)n, io, fetch, (int32_t*)c->gstack.p, cfg);
        else if (cfg.wide)
            k_simple_trace<true, false><<<grid, kTraceBlock, trace_lds_bytes(cfg), c->stream>>>(
                S, (uint32_t)n, io, fetch, (int32_t*)c->gstack.p, cfg);
        else
            k_simple_trace<false, false><<<grid, kTraceBlock, trace_lds_bytes(cfg), c->stream>>>(
                S, (uint32_t)n, io, fetch, (int32_t*)c->gstack.p, cfg);
        k_simple_accum<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(S, io, spp, type, min_depth, max_depth,
                                                                          (float4*)out_dev);
    }
    prof_end(c, "simple", ev, n * spp);
    if (int r = hip_check(c, hipGetLastError(), "simple launch")) return r;
    return hip_check(c, hipStreamSynchronize(c->stream), "simple");
}

// LightTracer pass: W x H x spp particles (launch index x sample), in chunks
// of the path pool; film_dev is cleared to (0, 0, 0, 1) first.
int render_light(srt_ctx* c, SceneRec* s, const DevCamera& cam, int spp, int max_depth, int rr_depth, uint32_t seed,
                 int width, int height, float* film_dev) {
    const int64_t wh = (int64_t)width * height;
    const int64_t total = wh * spp;
    int64_t pool = std::min<int64_t>(std::min<int64_t>(total, c->pool_request), 0x7fffffff);
    if (int r = ensure_pool(c, pool)) return r;
    const int tgrid = part_grid(trace_grid(c, pool));
    DevScene S = dev_scene(s);
    if (int r = ensure_gstack(c, s, (int64_t)tgrid * kTraceBlock)) return r;
    const TraceCfg tcfg = trace_cfg(c, s);
    TraceCfg scfg = tcfg;
    scfg.refill = c->shadow_refill;
    const bool inst = s->ninst > 0;
    const size_t tlds = trace_lds_bytes(tcfg, inst);
    hipStream_t st = c->stream;
    uint32_t* counts = (uint32_t*)c->counters.p;
    unsigned long long* totals = (unsigned long long*)((char*)c->counters.p + kTotalsOffset);
    PathQueue Q[2] = {make_queue(c, 0), make_queue(c, 1)};
    HitBuf H;
    H.u = (float*)c->hits[1].p; H.v = (float*)c->hits[2].p; H.tri = (int32_t*)c->hits[3].p;
    H.inst = inst ? (int32_t*)c->hits[4].p : nullptr;
    ShadowQueue SQ;
    {
        auto f = [&](int j) { return (float*)c->shq[j].p; };
        SQ.ox = f(0); SQ.oy = f(1); SQ.oz = f(2); SQ.dx = f(3); SQ.dy = f(4); SQ.dz = f(5); SQ.tmax = f(6);
        SQ.cr = f(7); SQ.cg = f(8); SQ.cb = f(9); SQ.slot = (uint32_t*)c->shq[10].p;
    }
    int32_t* gst = (int32_t*)c->gstack.p;
    float4* film = (float4*)film_dev;
    LtParams P;
    P.cam = cam;
    P.width = width; P.height = height; P.spp = spp; P.rr_depth = rr_depth; P.seed = seed;
    P.aspect = (float)width / (float)height;
    hipEvent_t ev;
    prof_begin(c, "light_tracer", &ev);
    k_film_clear<<<(unsigned)((wh + 255) / 256), 256, 0, st>>>(film, wh);
    for (int64_t p0 = 0; p0 < total; p0 += pool) {
        const int n = (int)std::min<int64_t>(pool, total - p0);
        P.path_begin = p0;
        const int shade_grid = part_grid(std::min<int64_t>((n + 255) / 256, (int64_t)num_cus(c) * 8));
        const uint32_t cap = part_cap((uint32_t)n);
        k_lt_raygen<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(S, P, Q[0], n, cap, counts);
        for (int depth = 0; depth < max_depth; depth++) {
            const int qi = depth & 1;
            if (inst)
                k_trace<false, true, true><<<tgrid, kTraceBlock, tlds, st>>>(S, Q[qi], H, counts, qi, cap, gst, totals, tcfg);
            else if (tcfg.wide)
                k_trace<false, true><<<tgrid, kTraceBlock, tlds, st>>>(S, Q[qi], H, counts, qi, cap, gst, totals, tcfg);
            else k_trace<false, false><<<tgrid, kTraceBlock, tlds, st>>>(S, Q[qi], H, counts, qi, cap, gst, totals, tcfg);
            k_lt_shade<<<shade_grid, 256, 0, st>>>(S, P, depth, Q[qi], H, Q[qi ^ 1], SQ, film, counts, qi, cap);
            if (inst) k_lt_shadow<true, true><<<tgrid, kTraceBlock, tlds, st>>>(S, SQ, film, counts, cap, gst, scfg);
            else if (tcfg.wide) k_lt_shadow<true, false><<<tgrid, kTraceBlock, tlds, st>>>(S, SQ, film, counts, cap, gst, scfg);
            else k_lt_shadow<false, false><<<tgrid, kTraceBlock, tlds, st>>>(S, SQ, film, counts, cap, gst, scfg);
        }
    }
    prof_end(c, "light_tracer", ev, total);
    if (int r = hip_check(c, hipGetLastError(), "light tracer launch")) return r;
    return hip_check(c, hipStreamSynchronize(st), "light tracer");
}

int trace_rays(srt_ctx* c, SceneRec* s, int n, const float* o_dev, const float* d_dev, float tmin, const float* tmax_dev,
               int any, float* t, float* u, float* v, int32_t* prim, int32_t* inst) {
    DevScene S = dev_scene(s);
    int grid = trace_grid(c, n);
    if (int r = ensure_gstack(c, s, (int64_t)grid * kTraceBlock)) return r;
    const TraceCfg cfg = trace_cfg(c, s);
    RaysIO io;
    io.o3 = o_dev; io.d3 = d_dev; io.tmax = tmax_dev; io.tmin = tmin;
    io.t = t; io.u = u; io.v = v; io.prim = prim; io.tris = S.tris; io.any = any;
    io.inst = inst;
    uint32_t* fetch = (uint32_t*)c->counters.p + CNT_FETCH_Q;
    k_clear_parts<<<1, kParts, 0, c->stream>>>(fetch);
    if (s->ninst > 0)
        k_trace_rays<true><<<grid, kTraceBlock, trace_lds_bytes(cfg, true), c->stream>>>(S, (uint32_t)n, io, fetch,
                                                                                       (int32_t*)c->gstack.p, cfg);
    else
        k_trace_rays<false><<<grid, kTraceBlock, trace_lds_bytes(cfg), c->stream>>>(S, (uint32_t)n, io, fetch,
                                                                                  (int32_t*)c->gstack.p, cfg);
    if (int r = hip_check(c, hipGetLastError(), "trace launch")) return r;
    return hip_check(c, hipStreamSynchronize(c->stream), "trace");
}

// ---------------------------------------------------------------------------
// Ray queries that allocate nothing (srt_trace_rays_noalloc): the traversal's
// fetch counters and stack spill area live in `scratch`, a file-scope device
// array of kNoallocScratchBytes on every device the library runs on, instead
// of context allocations.  Host code never uses the array's name as an
// address: noalloc_init looks its device address up once per device
// (hipGetSymbolAddress, with that device current) and every launch passes
// that address.  The array keeps the last call's contents; the counters are
// cleared by a fill kernel before each traversal, and calls that overlap on
// one device race on it.
// ---------------------------------------------------------------------------
constexpr size_t kNoallocScratchBytes = 32ull << 20;               // 32 MiB per device
constexpr size_t kNoallocCounterBytes = (size_t)kParts * kCntLine * 4;   // one fetch counter per piece, 8 KiB
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];
constexpr int kNoallocMaxDevices = 64;
static void* g_scratch_ptr[kNoallocMaxDevices];
static std::mutex g_scratch_mu;

int noalloc_init(srt_ctx* c) {
    if (c->device < 0 || c->device >= kNoallocMaxDevices)
        return set_err(c, SRT_E_LIMIT, "device %d beyond the %d devices of the scratch table", c->device,
                       kNoallocMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[c->device]) return SRT_OK;
    void* ptr = nullptr;
    if (int r = hip_check(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "scratch symbol address")) return r;
    if (!ptr) return set_err(c, SRT_E_HIP, "scratch symbol has no device address");
    g_scratch_ptr[c->device] = ptr;
    return SRT_OK;
}

// Closest or any hit of n rays given in device memory, into device memory;
// the same kernels and results as trace_rays.  The persistent grid is sized
// so that its stack spill fits the scratch array after the counters.
int trace_rays_noalloc(srt_ctx* c, SceneRec* s, int n, const float* o_dev, const float* d_dev, float tmin,
                       const float* tmax_dev, int any, float* t, float* u, float* v, int32_t* prim) {
    void* base;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = (c->device >= 0 && c->device < kNoallocMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
    }
    if (!base) return set_err(c, SRT_E_STATE, "scratch not set up on device %d", c->device);
    DevScene S = dev_scene(s);
    const TraceCfg cfg = trace_cfg(c, s);
    const int64_t deepest = use_wide(c, s) ? 3 * (int64_t)s->depth4 + 1 : (int64_t)s->max_depth + 1;
    const int64_t per_thread = std::max<int64_t>(1, deepest - cfg.lds_depth);     // int32 spill entries
    const int64_t spill_entries = (int64_t)((kNoallocScratchBytes - kNoallocCounterBytes) / sizeof(int32_t));
    const int64_t max_blocks = spill_entries / (per_thread * kTraceBlock);
    if (max_blocks < 1)
        return set_err(c, SRT_E_LIMIT, "a traversal stack of %lld entries per lane exceeds the %zu-byte scratch",
                       (long long)deepest, kNoallocScratchBytes);
    int grid = trace_grid(c, n);
    if (grid > max_blocks) grid = (int)max_blocks;
    RaysIO io;
    io.o3 = o_dev; io.d3 = d_dev; io.tmax = tmax_dev; io.tmin = tmin;
    io.t = t; io.u = u; io.v = v; io.prim = prim; io.tris = S.tris; io.any = any;
    uint32_t* fetch = (uint32_t*)base;
    int32_t* gstack = (int32_t*)((unsigned char*)base + kNoallocCounterBytes);
    k_clear_parts<<<1, kParts, 0, c->stream>>>(fetch);
    if (s->ninst > 0)
        k_trace_rays<true><<<grid, kTraceBlock, trace_lds_bytes(cfg, true), c->stream>>>(S, (uint32_t)n, io, fetch, gstack, cfg);
    else
        k_trace_rays<false><<<grid, kTraceBlock, trace_lds_bytes(cfg), c->stream>>>(S, (uint32_t)n, io, fetch, gstack, cfg);
    if (int r = hip_check(c, hipGetLastError(), "trace launch")) return r;
    return hip_check(c, hipStreamSynchronize(c->stream), "trace");
}

int tonemap_dev(srt_ctx* c, float* rgba, int n, int type, float exposure, int use_gamma) {
    float ex = std::pow(2.0f, exposure);   // host, as tonemapping.cu:11
    hipEvent_t ev;
    prof_begin(c, "tonemap", &ev);
    k_tonemap<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>((float4*)rgba, n, type, ex, use_gamma);
    prof_end(c, "tonemap", ev, n);
    if (int r = hip_check(c, hipGetLastError(), "tonemap launch")) return r;
    return hip_check(c, hipStreamSynchronize(c->stream), "tonemap");
}

// in place on rgba (W x H float4); tmp: 2 W x H float4 scratch (ping-pong image, guide)
int denoise_dev(srt_ctx* c, float* rgba, float* tmp, int W, int H) {
    const dim3 grid((unsigned)((W + 15) / 16), (unsigned)((H + 15) / 16));
    const int64_t n = (int64_t)W * H;
    hipEvent_t ev;
    prof_begin(c, "denoise", &ev);
    float4* a = (float4*)rgba;
    float4* b = (float4*)tmp;
    float4* g = (float4*)tmp + n;
    for (int it = 0; it < kDenoisePasses; it++) {
        const float sigma = kDenoiseSigma * std::ldexp(1.0f, -it);
        const float inv_s2 = 1.0f / (sigma * sigma);
        k_denoise_guide<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(a, g, n);
        k_atrous<<<grid, 256, 0, c->stream>>>(a, g, b, W, H, 1 << it, inv_s2);
        std::swap(a, b);
    }
    // an odd pass count leaves the result in tmp
    if (a != (float4*)rgba)
        if (int r = hip_check(c, hipMemcpyAsync(rgba, a, (size_t)W * H * 16, hipMemcpyDeviceToDevice, c->stream), "denoise"))
            return r;
    prof_end(c, "denoise", ev, (int64_t)W * H);
    if (int r = hip_check(c, hipGetLastError(), "denoise launch")) return r;
    return hip_check(c, hipStreamSynchronize(c->stream), "denoise");
}

int accumulate_dev(srt_ctx* c, float* acc, float* rgba, int n, int frame_count) {
    hipEvent_t ev;
    prof_begin(c, "accumulate", &ev);
    k_accumulate<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>((float4*)acc, (float4*)rgba, n, frame_count);
    prof_end(c, "accumulate", ev, n);
    if (int r = hip_check(c, hipGetLastError(), "accumulate launch")) return r;
    return hip_check(c, hipStreamSynchronize(c->stream), "accumulate");
}

}  // namespace srt
