// Radiance MLP for gfx950: BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation, hand-written forward and backward kernels, Adam, and the
// online-training / inference hooks of the path tracer.
//
// Replaces the reference's NeuralRadiosity pass network
// (src/passes/npr/neural_radiosity.cu:6-52, tiny-cuda-nn HashGrid + CutlassMLP,
// inference only; training is a stub at :165-169) with:
//   input   9 raw features u (position normalised to the scene box, (n+1)/2,
//           (w_o+1)/2, as nr_lhs.cu:72-89 does for n and w), encoded as
//           [u, sin(2^k pi u), cos(2^k pi u), k = 0..2] -> 63 (+1 zero pad)
//   layers  63 -> 64 -> 64 -> 64 -> 64 (ReLU) -> 3 (linear)
//   loss    relative L2: mean (y - t)^2 / (y^2 + 0.01), y detached below
//   Adam    Kingma & Ba 2015, beta1 0.9, beta2 0.999, eps 1e-8, lr 1e-3
// Every activation and gradient that feeds an MFMA is rounded to BF16 (round
// to nearest even); sums are FP32.  Weight gradients are reduced over fixed
// sample chunks and then in chunk order, so training is bit-reproducible.
//
// MFMA fragment layout used (16x16x32, wave64): lane l = 16 g + r holds
// A[r][8g .. 8g+7] and B[8g .. 8g+7][r] (8 BF16 each), and receives
// D[4g .. 4g+3][r] (4 FP32).  Only the row/column/lane-group structure matters:
// A and B use the same K placement, so any K permutation inside the
// instruction cancels.
#include "srt_internal.h"
#include <cstring>
#include <cmath>
#include <vector>

namespace srt {

typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x4 __attribute__((__vector_size__(4 * sizeof(float))));

constexpr int MLP_IN = 64;        // 63 encoded features + 1 zero pad
constexpr int MLP_H = 64;
constexpr int MLP_OUTP = 16;      // 3 outputs padded to one MFMA tile
constexpr int MLP_LAYERS = 5;
constexpr int MLP_RAW = 9;
constexpr int MLP_OCT = 3;
constexpr int TILE = 16;          // samples per wave tile
constexpr int MLP_BLOCK = 256;    // 4 waves
constexpr int CHUNK = 4096;       // samples per weight-gradient partial

// rows (outputs) of layer l, columns (inputs) are always 64
__host__ __device__ constexpr int layer_rows(int l) { return l == 4 ? MLP_OUTP : MLP_H; }
// offsets into the padded parameter arrays (floats / bf16 elements)
__host__ __device__ constexpr int w_off(int l) { return l * MLP_H * 64; }
__host__ __device__ constexpr int b_off(int l) { return l * MLP_H; }
constexpr int W_TOTAL = 4 * 64 * 64 + MLP_OUTP * 64;    // 17408
constexpr int B_TOTAL = 4 * 64 + MLP_OUTP;              // 272

SRT_D uint16_t f2bf(float f) {      // round to nearest even (NaN kept quiet)
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
SRT_D float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

SRT_D bf16x8 ld_frag(const uint16_t* p) {          // 8 consecutive BF16 (16 B aligned)
    uint4 v = *(const uint4*)p;
    bf16x8 r;
    __builtin_memcpy(&r, &v, 16);
    return r;
}

struct MlpDev {
    const uint16_t* wf;     // forward weights, layer l: [rows(l)][64] row-major
    const uint16_t* wt;     // transposed weights, layer l: [64][rows(l)] row-major
    const float* bias;      // [B_TOTAL]
    float lo[3], inv_ext[3];   // position normalisation
};

// feature i of the encoding of raw input u[9]
SRT_D float enc_feature(const float* u, int i) {
    if (i < MLP_RAW) return u[i];
    if (i >= MLP_RAW + 2 * MLP_RAW * MLP_OCT) return 0.0f;
    int j = i - MLP_RAW;
    int oct = j / (2 * MLP_RAW);
    int r = j % (2 * MLP_RAW);
    int c = r % MLP_RAW;
    float x = u[c] * (float)(1 << oct) * kPi;
    float s, co;
    pt_sincosf(x, &s, &co);
    return r < MLP_RAW ? s : co;
}

// Per-wave tile machinery --------------------------------------------------
// act: LDS [TILE][64] BF16 for the current layer input (row-major)
// The wave computes out[TILE][rows] = act * W^T (+ bias), fragment by fragment.
template <int ROWS>
SRT_D void layer_mfma(const uint16_t* act, const uint16_t* wl, f32x4 acc[ROWS / 16], int lane) {
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int jb = 0; jb < ROWS / 16; jb++) acc[jb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int q = 0; q < 2; q++) {
        bf16x8 a = ld_frag(act + r * 64 + 32 * q + 8 * g);
#pragma unroll
        for (int jb = 0; jb < ROWS / 16; jb++) {
            bf16x8 b = ld_frag(wl + (16 * jb + r) * 64 + 32 * q + 8 * g);
            acc[jb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[jb], 0, 0, 0);
        }
    }
}

// ---------------------------------------------------------------------------
// forward: in9 [n][9] fp32 -> y [n][3] fp32; TRAIN also stores the BF16 layer
// inputs H_l transposed ([l][64][n]) for the weight gradients.
// ---------------------------------------------------------------------------
template <bool TRAIN>
__global__ void __launch_bounds__(MLP_BLOCK) k_mlp_fwd(const float* in9, int n_arg, const uint32_t* n_dev, MlpDev M, float* y,
                                                       uint16_t* Ht) {
    const int n = n_dev ? (int)*n_dev : n_arg;     // device-side count (integrator queries)
    __shared__ __attribute__((aligned(16))) uint16_t sw[W_TOTAL];
    __shared__ float sb[B_TOTAL];
    __shared__ __attribute__((aligned(16))) uint16_t sact[MLP_BLOCK / 64][2][TILE * 64];
    for (int i = threadIdx.x; i < W_TOTAL / 8; i += blockDim.x) ((uint4*)sw)[i] = ((const uint4*)M.wf)[i];
    for (int i = threadIdx.x; i < B_TOTAL; i += blockDim.x) sb[i] = M.bias[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int ntiles = (n + TILE - 1) / TILE;
    constexpr int WAVES = MLP_BLOCK / 64;
    // every wave of the block runs the same rounds (the layer barriers are
    // block-wide); a wave past the last tile computes on zeros and stores nothing
    for (int base = blockIdx.x * WAVES; base < ntiles; base += gridDim.x * WAVES) {
        const int tile = base + wv;
        const int s0 = tile * TILE;
        uint16_t* a0 = sact[wv][0];
        uint16_t* a1 = sact[wv][1];
        // encode: lane (r, g) writes features 16g..16g+15 of sample r
        {
            float u[MLP_RAW];
            const int s = s0 + r;
            for (int c = 0; c < MLP_RAW; c++) u[c] = (s < n) ? in9[(size_t)s * MLP_RAW + c] : 0.0f;
            for (int k = 0; k < 16; k++) {
                int f = 16 * g + k;
                uint16_t h = f2bf(s < n ? enc_feature(u, f) : 0.0f);
                a0[r * 64 + f] = h;
                if (TRAIN && s < n) Ht[((size_t)0 * 64 + f) * n + s] = h;
            }
        }
        __syncthreads();
        uint16_t* cur = a0;
        uint16_t* nxt = a1;
        for (int l = 0; l < 4; l++) {
            f32x4 acc[4];
            layer_mfma<64>(cur, sw + w_off(l), acc, lane);
            // bias + ReLU -> BF16 -> next activation rows 4g..4g+3, column 16jb + r
#pragma unroll
            for (int jb = 0; jb < 4; jb++) {
                int j = 16 * jb + r;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    int s = 4 * g + i;
                    float z = acc[jb][i] + sb[b_off(l) + j];
                    uint16_t h = f2bf(z > 0.0f ? z : 0.0f);
                    nxt[s * 64 + j] = h;
                    if (TRAIN && s0 + s < n) Ht[((size_t)(l + 1) * 64 + j) * n + (s0 + s)] = h;
                }
            }
            __syncthreads();
            uint16_t* t = cur; cur = nxt; nxt = t;
        }
        f32x4 acc4[1];
        layer_mfma<16>(cur, sw + w_off(4), acc4, lane);
        if (r < 3) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                int s = s0 + 4 * g + i;
                if (s < n) y[(size_t)s * 3 + r] = acc4[0][i] + sb[b_off(4) + r];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// loss gradient + backward through the layers.  Writes G_l^T ([l][rows][n],
// BF16) = dL/dz of layer l's output, l = 0..4, and per-block loss partials.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(MLP_BLOCK) k_mlp_bwd(const float* y, const float* target, int n, float inv_count,
                                                       MlpDev M, const uint16_t* Ht, uint16_t* Gt, float* loss_part) {
    __shared__ __attribute__((aligned(16))) uint16_t swt[W_TOTAL];
    __shared__ __attribute__((aligned(16))) uint16_t sg[MLP_BLOCK / 64][TILE * 64];
    __shared__ float sloss[MLP_BLOCK];
    for (int i = threadIdx.x; i < W_TOTAL / 8; i += blockDim.x) ((uint4*)swt)[i] = ((const uint4*)M.wt)[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int ntiles = (n + TILE - 1) / TILE;
    constexpr int WAVES = MLP_BLOCK / 64;
    float lsum = 0.0f;
    uint16_t* G = sg[wv];
    for (int base = blockIdx.x * WAVES; base < ntiles; base += gridDim.x * WAVES) {
        const int tile = base + wv;
        const int s0 = tile * TILE;
        // output gradient: sample 4g+i, channel r (< 3), padded to 16 columns
        for (int i = 0; i < 4; i++) {
            int s = 4 * g + i, gs = s0 + s;
            float gval = 0.0f;
            if (r < 3 && gs < n) {
                float yy = y[(size_t)gs * 3 + r], tt = target[(size_t)gs * 3 + r];
                float d = yy - tt, den = yy * yy + 0.01f;
                lsum += d * d / den;
                gval = 2.0f * d / den * inv_count;
            }
            uint16_t h = f2bf(gval);
            // G rows are samples, columns outputs (16 used, stride 64)
            G[s * 64 + r] = h;
            if (gs < n) Gt[((size_t)4 * 64 + r) * n + gs] = h;
        }
        __syncthreads();
        // layers 4..1: dh_l = G_l * W_l (W_l^T stored [64][rows]); z mask from H_l > 0
        for (int l = 4; l >= 1; l--) {
            const int rows = layer_rows(l);
            f32x4 acc[4];
#pragma unroll
            for (int kb = 0; kb < 4; kb++) acc[kb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            const uint16_t* wt = swt + w_off(l);
            for (int q = 0; q < rows / 32 + (rows % 32 ? 1 : 0); q++) {
                // A = G (rows = samples, K = outputs of layer l); pad K beyond rows with 0
                bf16x8 a;
                if (32 * q + 8 * g < rows) a = ld_frag(G + r * 64 + 32 * q + 8 * g);
                else { uint4 z = make_uint4(0, 0, 0, 0); __builtin_memcpy(&a, &z, 16); }
#pragma unroll
                for (int kb = 0; kb < 4; kb++) {
                    bf16x8 b;
                    if (32 * q + 8 * g < rows) {
                        // B[kappa = j][n = k] = W_l[j][k] = Wt_l[k][j]: need 8 consecutive j for fixed k
                        b = ld_frag(wt + (16 * kb + r) * rows + 32 * q + 8 * g);
                    } else { uint4 z = make_uint4(0, 0, 0, 0); __builtin_memcpy(&b, &z, 16); }
                    acc[kb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[kb], 0, 0, 0);
                }
            }
            __syncthreads();
            // dz_{l-1} = dh * [H_l > 0]; H_l is the input of layer l = output of layer l-1
#pragma unroll
            for (int kb = 0; kb < 4; kb++) {
                int k = 16 * kb + r;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    int s = 4 * g + i, gs = s0 + s;
                    float hv = (gs < n) ? bf2f(Ht[((size_t)l * 64 + k) * n + gs]) : 0.0f;
                    float dz = hv > 0.0f ? acc[kb][i] : 0.0f;
                    uint16_t h = f2bf(dz);
                    G[s * 64 + k] = h;
                    if (gs < n) Gt[((size_t)(l - 1) * 64 + k) * n + gs] = h;
                }
            }
            __syncthreads();
        }
    }
    sloss[threadIdx.x] = lsum;
    __syncthreads();
    for (int st = MLP_BLOCK / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sloss[threadIdx.x] += sloss[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = sloss[0];
}

// ---------------------------------------------------------------------------
// weight gradients: partial dW_l[j][k] = sum over a chunk of samples of
// G_l[j][s] * H_l[k][s] (both stored transposed, s contiguous).
// grid: (chunks, 16 output blocks (jb, kb), layers); one wave per block.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_mlp_dw(int n, const uint16_t* Ht, const uint16_t* Gt, float* part) {
    const int chunk = blockIdx.x, blk = blockIdx.y, l = blockIdx.z;
    const int rows = layer_rows(l);
    const int jb = blk >> 2, kb = blk & 3;
    if (16 * jb >= rows) return;
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    const int s_begin = chunk * CHUNK;
    const int s_end = min(n, s_begin + CHUNK);
    const uint16_t* G = Gt + ((size_t)l * 64 + 16 * jb + r) * n;   // A rows: j
    const uint16_t* H = Ht + ((size_t)l * 64 + 16 * kb + r) * n;   // B cols: k
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int s = s_begin; s < s_end; s += 32) {
        int base = s + 8 * g;
        bf16x8 a, b;
        if (base + 8 <= s_end && (n % 8) == 0) {
            a = ld_frag(G + base);
            b = ld_frag(H + base);
        } else {
            uint16_t ta[8], tb[8];
            for (int e = 0; e < 8; e++) {
                ta[e] = (base + e < s_end) ? G[base + e] : (uint16_t)0;
                tb[e] = (base + e < s_end) ? H[base + e] : (uint16_t)0;
            }
            __builtin_memcpy(&a, ta, 16);
            __builtin_memcpy(&b, tb, 16);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    // D[j = 16jb + 4g + i][k = 16kb + r]
    float* out = part + ((size_t)chunk * MLP_LAYERS + l) * (64 * 64);
#pragma unroll
    for (int i = 0; i < 4; i++) out[(16 * jb + 4 * g + i) * 64 + 16 * kb + r] = acc[i];
}

// bias gradients: db_l[j] partial over a chunk (fixed-order sums)
__global__ void __launch_bounds__(64) k_mlp_db(int n, const uint16_t* Gt, float* part) {
    const int chunk = blockIdx.x, l = blockIdx.z;
    const int j = threadIdx.x;
    if (j >= layer_rows(l)) return;
    const int s_begin = chunk * CHUNK, s_end = min(n, s_begin + CHUNK);
    const uint16_t* G = Gt + ((size_t)l * 64 + j) * n;
    float acc = 0.0f;
    for (int s = s_begin; s < s_end; s++) acc += bf2f(G[s]);
    part[((size_t)chunk * MLP_LAYERS + l) * 64 + j] = acc;
}

// reduce partials in chunk order, Adam update, refresh BF16 copies
__global__ void __launch_bounds__(256) k_mlp_adam(int nchunks, const float* wpart, const float* bpart, float* W, float* B,
                                                  float* mW, float* vW, float* mB, float* vB, float lr, float b1, float b2,
                                                  float eps, float c1, float c2, int apply, float* grad_out, uint16_t* wf,
                                                  uint16_t* wt) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < W_TOTAL) {
        int l = (p < 4 * 4096) ? p / 4096 : 4;
        int o = p - w_off(l);
        int j = o / 64, k = o % 64;
        bool live = !(l == 0 && k == 63) && !(l == 4 && j >= 3);
        float gsum = 0.0f;
        for (int c = 0; c < nchunks; c++) gsum += wpart[((size_t)c * MLP_LAYERS + l) * 4096 + j * 64 + k];
        if (!live) gsum = 0.0f;
        if (grad_out) grad_out[p] = gsum;
        float w = W[p];
        if (apply && live) {
            float m = b1 * mW[p] + (1.0f - b1) * gsum;
            float v = b2 * vW[p] + (1.0f - b2) * gsum * gsum;
            mW[p] = m;
            vW[p] = v;
            float mh = m / c1, vh = v / c2;
            w = w - lr * mh / (__builtin_sqrtf(vh) + eps);
            W[p] = w;
        }
        uint16_t h = f2bf(w);
        wf[p] = h;
        int rows = layer_rows(l);
        wt[w_off(l) + k * rows + j] = h;
    } else if (p < W_TOTAL + B_TOTAL) {
        int q = p - W_TOTAL;
        int l = (q < 256) ? q / 64 : 4;
        int j = q - b_off(l);
        bool live = !(l == 4 && j >= 3);
        float gsum = 0.0f;
        for (int c = 0; c < nchunks; c++) gsum += bpart[((size_t)c * MLP_LAYERS + l) * 64 + j];
        if (!live) gsum = 0.0f;
        if (grad_out) grad_out[p] = gsum;
        if (apply && live) {
            float m = b1 * mB[q] + (1.0f - b1) * gsum;
            float v = b2 * vB[q] + (1.0f - b2) * gsum * gsum;
            mB[q] = m;
            vB[q] = v;
            float mh = m / c1, vh = v / c2;
            B[q] = B[q] - lr * mh / (__builtin_sqrtf(vh) + eps);
        }
    }
}

__global__ void k_loss_sum(const float* part, int nparts, float* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float s = 0.0f;
        for (int i = 0; i < nparts; i++) s += part[i];
        out[0] = s;
    }
}

// ---------------------------------------------------------------------------
// host state
// ---------------------------------------------------------------------------
struct MlpState {
    DevBuf W, B, mW, vW, mB, vB, wf, wt, bias_dev;
    DevBuf Ht, Gt, y, wpart, bpart, loss_part, loss, grad, in_buf, tgt_buf, out_buf;
    int64_t cap = 0;        // training sample capacity of Ht/Gt/y
    int64_t infer_cap = 0;
    int step = 0;
    float lr = 1e-3f, b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float lo[3] = {0, 0, 0}, inv_ext[3] = {1, 1, 1};
};

static MlpState* mlp_of(srt_ctx* c) { return (MlpState*)c->mlp; }

static void mlp_free_all(srt_ctx* c) {
    MlpState* m = mlp_of(c);
    if (!m) return;
    DevBuf* all[] = {&m->W, &m->B, &m->mW, &m->vW, &m->mB, &m->vB, &m->wf, &m->wt, &m->bias_dev, &m->Ht, &m->Gt, &m->y,
                     &m->wpart, &m->bpart, &m->loss_part, &m->loss, &m->grad, &m->in_buf, &m->tgt_buf, &m->out_buf};
    for (DevBuf* b : all) dev_free(c, *b);
    delete m;
    c->mlp = nullptr;
}

static MlpDev mlp_dev(MlpState* m) {
    MlpDev d;
    d.wf = (const uint16_t*)m->wf.p;
    d.wt = (const uint16_t*)m->wt.p;
    d.bias = (const float*)m->B.p;
    for (int i = 0; i < 3; i++) { d.lo[i] = m->lo[i]; d.inv_ext[i] = m->inv_ext[i]; }
    return d;
}

// refresh BF16 copies without an update (apply = 0)
static int refresh_copies(srt_ctx* c, MlpState* m) {
    const int tot = W_TOTAL + B_TOTAL;
    k_mlp_adam<<<(tot + 255) / 256, 256, 0, c->stream>>>(0, nullptr, nullptr, (float*)m->W.p, (float*)m->B.p, nullptr,
                                                         nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 0,
                                                         nullptr, (uint16_t*)m->wf.p, (uint16_t*)m->wt.p);
    return hip_check(c, hipGetLastError(), "mlp refresh");
}

static int ensure_train(srt_ctx* c, MlpState* m, int64_t n) {
    if (m->cap >= n) return SRT_OK;
    DevBuf* bs[] = {&m->Ht, &m->Gt, &m->y, &m->wpart, &m->bpart, &m->tgt_buf};
    for (DevBuf* b : bs) dev_free(c, *b);
    int64_t chunks = (n + CHUNK - 1) / CHUNK;
    int r = dev_alloc(c, m->Ht, (size_t)MLP_LAYERS * 64 * n * 2);
    if (!r) r = dev_alloc(c, m->Gt, (size_t)MLP_LAYERS * 64 * n * 2);
    if (!r) r = dev_alloc(c, m->y, (size_t)n * 12);
    if (!r) r = dev_alloc(c, m->tgt_buf, (size_t)n * 12);
    if (!r) r = dev_alloc(c, m->wpart, (size_t)chunks * MLP_LAYERS * 4096 * 4);
    if (!r) r = dev_alloc(c, m->bpart, (size_t)chunks * MLP_LAYERS * 64 * 4);
    if (r) return r;
    m->cap = n;
    return SRT_OK;
}

static int ensure_in(srt_ctx* c, MlpState* m, int64_t n) {
    if (m->infer_cap >= n) return SRT_OK;
    dev_free(c, m->in_buf);
    dev_free(c, m->out_buf);
    int r = dev_alloc(c, m->in_buf, (size_t)n * 36);
    if (!r) r = dev_alloc(c, m->out_buf, (size_t)n * 12);
    if (r) return r;
    m->infer_cap = n;
    return SRT_OK;
}

static int fwd_grid(srt_ctx* c, int64_t n) {
    int64_t tiles = (n + TILE - 1) / TILE;
    int64_t blocks = (tiles + 3) / 4;
    hipDeviceProp_t prop;
    int cus = 256;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) cus = prop.multiProcessorCount;
    int64_t cap = (int64_t)cus * 3;
    return (int)std::max<int64_t>(1, std::min(blocks, cap));
}

// one training step on device arrays (in9 [n][9], target [n][3]); apply = 0
// computes gradients only (grad_out host pointer optional)
static int train_step_dev(srt_ctx* c, MlpState* m, const float* in9, const float* target, int n, int apply, float* loss_host,
                          float* grad_host) {
    if (int r = ensure_train(c, m, n)) return r;
    hipStream_t st = c->stream;
    MlpDev d = mlp_dev(m);
    const int grid = fwd_grid(c, n);
    if (!m->loss_part.p || m->loss_part.bytes < (size_t)grid * 4) {
        dev_free(c, m->loss_part);
        if (int r = dev_alloc(c, m->loss_part, (size_t)std::max(grid, 1024) * 4)) return r;
    }
    if (!m->loss.p) if (int r = dev_alloc(c, m->loss, 16)) return r;
    if (!m->grad.p) if (int r = dev_alloc(c, m->grad, (size_t)(W_TOTAL + B_TOTAL) * 4)) return r;
    hipEvent_t ev;
    prof_begin(c, "mlp_fwd", &ev);
    k_mlp_fwd<true><<<grid, MLP_BLOCK, 0, st>>>(in9, n, nullptr, d, (float*)m->y.p, (uint16_t*)m->Ht.p);
    prof_end(c, "mlp_fwd", ev, n);
    prof_begin(c, "mlp_bwd", &ev);
    k_mlp_bwd<<<grid, MLP_BLOCK, 0, st>>>((const float*)m->y.p, target, n, 1.0f / (3.0f * (float)n), d,
                                          (const uint16_t*)m->Ht.p, (uint16_t*)m->Gt.p, (float*)m->loss_part.p);
    prof_end(c, "mlp_bwd", ev, n);
    k_loss_sum<<<1, 64, 0, st>>>((const float*)m->loss_part.p, grid, (float*)m->loss.p);
    const int chunks = (n + CHUNK - 1) / CHUNK;
    prof_begin(c, "mlp_dw", &ev);
    k_mlp_dw<<<dim3(chunks, 16, MLP_LAYERS), 64, 0, st>>>(n, (const uint16_t*)m->Ht.p, (const uint16_t*)m->Gt.p,
                                                            (float*)m->wpart.p);
    k_mlp_db<<<dim3(chunks, 1, MLP_LAYERS), 64, 0, st>>>(n, (const uint16_t*)m->Gt.p, (float*)m->bpart.p);
    prof_end(c, "mlp_dw", ev, n);
    if (apply) m->step++;
    const float c1 = 1.0f - std::pow(m->b1, (float)std::max(m->step, 1));
    const float c2 = 1.0f - std::pow(m->b2, (float)std::max(m->step, 1));
    const int tot = W_TOTAL + B_TOTAL;
    prof_begin(c, "mlp_adam", &ev);
    k_mlp_adam<<<(tot + 255) / 256, 256, 0, st>>>(chunks, (const float*)m->wpart.p, (const float*)m->bpart.p, (float*)m->W.p,
                                                  (float*)m->B.p, (float*)m->mW.p, (float*)m->vW.p, (float*)m->mB.p,
                                                  (float*)m->vB.p, m->lr, m->b1, m->b2, m->eps, c1, c2, apply,
                                                  (float*)m->grad.p, (uint16_t*)m->wf.p, (uint16_t*)m->wt.p);
    prof_end(c, "mlp_adam", ev, tot);
    if (int r = hip_check(c, hipGetLastError(), "mlp train launch")) return r;
    if (loss_host) {
        float l = 0.0f;
        if (int r = hip_check(c, hipMemcpyAsync(&l, m->loss.p, 4, hipMemcpyDeviceToHost, st), "loss")) return r;
        if (int r = hip_check(c, hipStreamSynchronize(st), "mlp train")) return r;
        *loss_host = l / (3.0f * (float)n);
    }
    if (grad_host) {
        if (int r = hip_check(c, hipStreamSynchronize(st), "mlp train")) return r;
        if (int r = hip_check(c, hipMemcpy(grad_host, m->grad.p, (size_t)tot * 4, hipMemcpyDeviceToHost), "grads")) return r;
    }
    return SRT_OK;
}

int mlp_infer_dev(srt_ctx* c, const float* in9, int n, float* out3) {
    MlpState* m = mlp_of(c);
    if (!m) return set_err(c, SRT_E_STATE, "no MLP: call srt_mlp_create first");
    if (n == 0) return SRT_OK;
    hipEvent_t ev;
    prof_begin(c, "mlp_infer", &ev);
    k_mlp_fwd<false><<<fwd_grid(c, n), MLP_BLOCK, 0, c->stream>>>(in9, n, nullptr, mlp_dev(m), out3, nullptr);
    prof_end(c, "mlp_infer", ev, n);
    return hip_check(c, hipGetLastError(), "mlp infer launch");
}

// forward over a queue whose length lives in device memory (at most max_n)
int mlp_query_dev(srt_ctx* c, const float* in9, const uint32_t* count_dev, int max_n, float* out3) {
    MlpState* m = mlp_of(c);
    if (!m) return set_err(c, SRT_E_STATE, "no MLP: call srt_mlp_create first");
    hipEvent_t ev;
    prof_begin(c, "mlp_query", &ev);
    k_mlp_fwd<false><<<fwd_grid(c, max_n), MLP_BLOCK, 0, c->stream>>>(in9, 0, count_dev, mlp_dev(m), out3, nullptr);
    prof_end(c, "mlp_query", ev, 0);
    return hip_check(c, hipGetLastError(), "mlp query launch");
}

bool mlp_ready(srt_ctx* c) { return mlp_of(c) != nullptr; }

int mlp_train_dev(srt_ctx* c, const float* in9, const float* target, int n, float* loss_host) {
    MlpState* m = mlp_of(c);
    if (!m) return set_err(c, SRT_E_STATE, "no MLP: call srt_mlp_create first");
    return train_step_dev(c, m, in9, target, n, 1, loss_host, nullptr);
}

void mlp_bounds(srt_ctx* c, const float lo[3], const float hi[3]) {
    MlpState* m = mlp_of(c);
    if (!m) return;
    for (int i = 0; i < 3; i++) {
        float e = hi[i] - lo[i];
        m->lo[i] = lo[i];
        m->inv_ext[i] = e > 0.0f ? 1.0f / e : 1.0f;
    }
}
void mlp_get_bounds(srt_ctx* c, float lo[3], float inv_ext[3]) {
    MlpState* m = mlp_of(c);
    for (int i = 0; i < 3; i++) { lo[i] = m ? m->lo[i] : 0.0f; inv_ext[i] = m ? m->inv_ext[i] : 1.0f; }
}

// logical (unpadded) parameter order: W0 [64][63], b0 [64], W1..W3 [64][64] + b,
// W4 [3][64], b4 [3]
static int logical_count() { return 64 * 63 + 64 + 3 * (64 * 64 + 64) + 3 * 64 + 3; }

static void pad_from_logical(const float* in, std::vector<float>& W, std::vector<float>& B) {
    W.assign(W_TOTAL, 0.0f);
    B.assign(B_TOTAL, 0.0f);
    size_t p = 0;
    for (int l = 0; l < MLP_LAYERS; l++) {
        int rows = (l == 4) ? 3 : 64, cols = (l == 0) ? 63 : 64;
        for (int j = 0; j < rows; j++)
            for (int k = 0; k < cols; k++) W[w_off(l) + j * 64 + k] = in[p++];
        for (int j = 0; j < rows; j++) B[b_off(l) + j] = in[p++];
    }
}
static void logical_from_pad(const std::vector<float>& W, const std::vector<float>& B, float* out) {
    size_t p = 0;
    for (int l = 0; l < MLP_LAYERS; l++) {
        int rows = (l == 4) ? 3 : 64, cols = (l == 0) ? 63 : 64;
        for (int j = 0; j < rows; j++)
            for (int k = 0; k < cols; k++) out[p++] = W[w_off(l) + j * 64 + k];
        for (int j = 0; j < rows; j++) out[p++] = B[b_off(l) + j];
    }
}

}  // namespace srt

using namespace srt;

extern "C" {

int srt_mlp_num_params(void) { return logical_count(); }

int srt_mlp_create(srt_ctx* c, uint32_t seed, float lr) {
    SRT_IDLE(c);
    if (!c) return SRT_E_INVALID;
    if (!(lr > 0.0f)) return set_err(c, SRT_E_INVALID, "learning rate must be positive");
    if (c->mlp) mlp_free_all(c);
    MlpState* m = new MlpState();
    c->mlp = m;
    c->mlp_free = mlp_free_all;
    m->lr = lr;
    int r = SRT_OK;
    DevBuf* f32[] = {&m->W, &m->mW, &m->vW};
    for (DevBuf* b : f32) if (!r) r = dev_alloc(c, *b, (size_t)W_TOTAL * 4);
    DevBuf* f32b[] = {&m->B, &m->mB, &m->vB};
    for (DevBuf* b : f32b) if (!r) r = dev_alloc(c, *b, (size_t)B_TOTAL * 4);
    if (!r) r = dev_alloc(c, m->wf, (size_t)W_TOTAL * 2);
    if (!r) r = dev_alloc(c, m->wt, (size_t)W_TOTAL * 2);
    if (r) { mlp_free_all(c); return r; }
    // He-uniform init from a PCG32 stream; padded entries stay zero
    std::vector<float> W(W_TOTAL, 0.0f), B(B_TOTAL, 0.0f), z(W_TOTAL, 0.0f);
    uint64_t state = 0, inc = ((uint64_t)seed << 1u) | 1u;
    auto next = [&]() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    };
    next(); state += 0x853c49e6748fea9bULL; next();
    for (int l = 0; l < MLP_LAYERS; l++) {
        int rows = (l == 4) ? 3 : 64, cols = (l == 0) ? 63 : 64;
        float a = std::sqrt(6.0f / (float)cols);
        for (int j = 0; j < rows; j++)
            for (int k = 0; k < cols; k++) {
                float u = (float)(next() >> 8) * 5.9604644775390625e-8f;
                W[w_off(l) + j * 64 + k] = (2.0f * u - 1.0f) * a;
            }
    }
    r = hip_check(c, hipMemcpy(m->W.p, W.data(), W_TOTAL * 4, hipMemcpyHostToDevice), "upload");
    if (!r) r = hip_check(c, hipMemcpy(m->B.p, B.data(), B_TOTAL * 4, hipMemcpyHostToDevice), "upload");
    DevBuf* zs[] = {&m->mW, &m->vW};
    for (DevBuf* b : zs) if (!r) r = hip_check(c, hipMemcpy(b->p, z.data(), W_TOTAL * 4, hipMemcpyHostToDevice), "upload");
    DevBuf* zb[] = {&m->mB, &m->vB};
    for (DevBuf* b : zb) if (!r) r = hip_check(c, hipMemcpy(b->p, z.data(), B_TOTAL * 4, hipMemcpyHostToDevice), "upload");
    if (!r) r = refresh_copies(c, m);
    if (!r) r = hip_check(c, hipStreamSynchronize(c->stream), "mlp create");
    return r;
}

int srt_mlp_set_params(srt_ctx* c, const float* params, int32_t n) {
    SRT_IDLE(c);
    if (!c || !params) return SRT_E_INVALID;
    MlpState* m = mlp_of(c);
    if (!m) return set_err(c, SRT_E_STATE, "no MLP");
    if (n != logical_count()) return set_err(c, SRT_E_INVALID, "expected %d parameters", logical_count());
    std::vector<float> W, B;
    pad_from_logical(params, W, B);
    int r = hip_check(c, hipMemcpy(m->W.p, W.data(), W_TOTAL * 4, hipMemcpyHostToDevice), "upload");
    if (!r) r = hip_check(c, hipMemcpy(m->B.p, B.data(), B_TOTAL * 4, hipMemcpyHostToDevice), "upload");
    if (!r) r = refresh_copies(c, m);
    if (!r) r = hip_check(c, hipStreamSynchronize(c->stream), "set params");
    return r;
}

int srt_mlp_get_params(srt_ctx* c, float* params, int32_t n) {
    SRT_IDLE(c);
    if (!c || !params) return SRT_E_INVALID;
    MlpState* m = mlp_of(c);
    if (!m) return set_err(c, SRT_E_STATE, "no MLP");
    if (n != logical_count()) return set_err(c, SRT_E_INVALID, "expected %d parameters", logical_count());
    std::vector<float> W(W_TOTAL), B(B_TOTAL);
    if (int r = hip_check(c, hipStreamSynchronize(c->stream), "sync")) return r;
    int r = hip_check(c, hipMemcpy(W.data(), m->W.p, W_TOTAL * 4, hipMemcpyDeviceToHost), "download");
    if (!r) r = hip_check(c, hipMemcpy(B.data(), m->B.p, B_TOTAL * 4, hipMemcpyDeviceToHost), "download");
    if (!r) logical_from_pad(W, B, params);
    return r;
}

int srt_mlp_set_hparams(srt_ctx* c, float lr, float beta1, float beta2, float eps) {
    SRT_IDLE(c);
    if (!c) return SRT_E_INVALID;
    MlpState* m = mlp_of(c);
    if (!m) return set_err(c, SRT_E_STATE, "no MLP");
    if (!(lr > 0 && beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 && eps > 0)) return set_err(c, SRT_E_INVALID, "bad Adam hyper-parameters");
    m->lr = lr; m->b1 = beta1; m->b2 = beta2; m->eps = eps;
    return SRT_OK;
}

int srt_mlp_set_bounds(srt_ctx* c, const float lo[3], const float hi[3]) {
    SRT_IDLE(c);
    if (!c || !lo || !hi) return SRT_E_INVALID;
    if (!mlp_of(c)) return set_err(c, SRT_E_STATE, "no MLP");
    mlp_bounds(c, lo, hi);
    return SRT_OK;
}

// host arrays: in9 [n][9] raw features (already normalised to [0,1]), out3 [n][3]
int srt_mlp_infer(srt_ctx* c, int32_t n, const float* in9, float* out3) {
    SRT_IDLE(c);
    if (!c || n < 0 || (n && (!in9 || !out3))) return SRT_E_INVALID;
    MlpState* m = mlp_of(c);
    if (!m) return set_err(c, SRT_E_STATE, "no MLP");
    if (n == 0) return SRT_OK;
    if (int r = ensure_in(c, m, n)) return r;
    if (int r = hip_check(c, hipMemcpy(m->in_buf.p, in9, (size_t)n * 36, hipMemcpyHostToDevice), "upload")) return r;
    if (int r = mlp_infer_dev(c, (const float*)m->in_buf.p, n, (float*)m->out_buf.p)) return r;
    if (int r = hip_check(c, hipStreamSynchronize(c->stream), "infer")) return r;
    prof_collect(c);
    return hip_check(c, hipMemcpy(out3, m->out_buf.p, (size_t)n * 12, hipMemcpyDeviceToHost), "download");
}

// one Adam step (apply = 1) or gradients only (apply = 0); grads (optional) in
// the logical parameter order; loss = mean relative L2 before the step
int srt_mlp_train_step(srt_ctx* c, int32_t n, const float* in9, const float* target3, int32_t apply, float* loss,
                       float* grads) {
    SRT_IDLE(c);
    if (!c || n <= 0 || !in9 || !target3) return SRT_E_INVALID;
    MlpState* m = mlp_of(c);
    if (!m) return set_err(c, SRT_E_STATE, "no MLP");
    if (int r = ensure_in(c, m, n)) return r;
    if (int r = ensure_train(c, m, n)) return r;
    if (int r = hip_check(c, hipMemcpy(m->in_buf.p, in9, (size_t)n * 36, hipMemcpyHostToDevice), "upload")) return r;
    if (int r = hip_check(c, hipMemcpy(m->tgt_buf.p, target3, (size_t)n * 12, hipMemcpyHostToDevice), "upload")) return r;
    std::vector<float> g(grads ? (size_t)(W_TOTAL + B_TOTAL) : 0);
    int r = train_step_dev(c, m, (const float*)m->in_buf.p, (const float*)m->tgt_buf.p, n, apply, loss,
                           grads ? g.data() : nullptr);
    if (r) return r;
    prof_collect(c);
    if (grads) {
        std::vector<float> W(g.begin(), g.begin() + W_TOTAL), B(g.begin() + W_TOTAL, g.end());
        logical_from_pad(W, B, grads);
    }
    return SRT_OK;
}

int srt_mlp_save(srt_ctx* c, const char* path) {
    SRT_IDLE(c);
    if (!c || !path) return SRT_E_INVALID;
    MlpState* m = mlp_of(c);
    if (!m) return set_err(c, SRT_E_STATE, "no MLP");
    if (int r = hip_check(c, hipStreamSynchronize(c->stream), "sync")) return r;
    std::vector<float> buf(3 * (W_TOTAL + B_TOTAL));
    DevBuf* src[] = {&m->W, &m->mW, &m->vW};
    DevBuf* srcb[] = {&m->B, &m->mB, &m->vB};
    for (int i = 0; i < 3; i++) {
        if (int r = hip_check(c, hipMemcpy(buf.data() + i * W_TOTAL, src[i]->p, W_TOTAL * 4, hipMemcpyDeviceToHost), "save")) return r;
        if (int r = hip_check(c, hipMemcpy(buf.data() + 3 * W_TOTAL + i * B_TOTAL, srcb[i]->p, B_TOTAL * 4, hipMemcpyDeviceToHost), "save")) return r;
    }
    FILE* f = fopen(path, "wb");
    if (!f) return set_err(c, SRT_E_INVALID, "cannot open %s", path);
    const char magic[8] = {'S', 'R', 'T', 'M', 'L', 'P', '1', 0};
    int32_t hdr[2] = {m->step, W_TOTAL + B_TOTAL};
    float hp[4 + 6] = {m->lr, m->b1, m->b2, m->eps, m->lo[0], m->lo[1], m->lo[2], m->inv_ext[0], m->inv_ext[1], m->inv_ext[2]};
    bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(hdr, 4, 2, f) == 2 && fwrite(hp, 4, 10, f) == 10 &&
              fwrite(buf.data(), 4, buf.size(), f) == buf.size();
    fclose(f);
    return ok ? SRT_OK : set_err(c, SRT_E_INVALID, "write failed: %s", path);
}

int srt_mlp_load(srt_ctx* c, const char* path) {
    SRT_IDLE(c);
    if (!c || !path) return SRT_E_INVALID;
    MlpState* m = mlp_of(c);
    if (!m) return set_err(c, SRT_E_STATE, "no MLP: call srt_mlp_create first");
    FILE* f = fopen(path, "rb");
    if (!f) return set_err(c, SRT_E_INVALID, "cannot open %s", path);
    char magic[8];
    int32_t hdr[2];
    float hp[10];
    std::vector<float> buf(3 * (W_TOTAL + B_TOTAL));
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "SRTMLP1", 8) == 0 && fread(hdr, 4, 2, f) == 2 &&
              hdr[1] == W_TOTAL + B_TOTAL && fread(hp, 4, 10, f) == 10 && fread(buf.data(), 4, buf.size(), f) == buf.size();
    fclose(f);
    if (!ok) return set_err(c, SRT_E_INVALID, "not an MLP checkpoint: %s", path);
    m->step = hdr[0];
    m->lr = hp[0]; m->b1 = hp[1]; m->b2 = hp[2]; m->eps = hp[3];
    for (int i = 0; i < 3; i++) { m->lo[i] = hp[4 + i]; m->inv_ext[i] = hp[7 + i]; }
    DevBuf* dst[] = {&m->W, &m->mW, &m->vW};
    DevBuf* dstb[] = {&m->B, &m->mB, &m->vB};
    for (int i = 0; i < 3; i++) {
        if (int r = hip_check(c, hipMemcpy(dst[i]->p, buf.data() + i * W_TOTAL, W_TOTAL * 4, hipMemcpyHostToDevice), "load")) return r;
        if (int r = hip_check(c, hipMemcpy(dstb[i]->p, buf.data() + 3 * W_TOTAL + i * B_TOTAL, B_TOTAL * 4, hipMemcpyHostToDevice), "load")) return r;
    }
    if (int r = refresh_copies(c, m)) return r;
    return hip_check(c, hipStreamSynchronize(c->stream), "load");
}

}  // extern "C"
