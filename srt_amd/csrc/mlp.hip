// Radiance MLP for gfx950: BF16 MFMA (v_mfma_f32_32x32x16_bf16 and
// v_mfma_f32_16x16x32_bf16) with FP32 accumulation, hand-written forward and
// backward kernels, Adam, and the online-training / inference hooks of the
// path tracer.
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
// to nearest even); sums are FP32.  A training step is one kernel: weight
// gradients are reduced over fixed 1024-sample blocks and then in block
// order, so training is bit-reproducible.
//
// MFMA fragment layouts used (wave64; only the row/column/lane-group
// structure matters: A and B use the same K placement, so any K permutation
// inside the instruction cancels):
//   32x32x16: lane l = 32 g + r holds A[r][8g .. 8g+7] and B[8g .. 8g+7][r],
//             and receives D[8q + 4g + p][r] in element 4q + p (q, p < 4);
//   16x16x32: lane l = 16 g + r holds A[r][8g .. 8g+7] and B[8g .. 8g+7][r],
//             and receives D[4g + i][r] in element i (i < 4).
#include "srt_internal.h"
#include <cstring>
#include <cmath>
#include <vector>

namespace srt {

typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x4 __attribute__((__vector_size__(4 * sizeof(float))));
typedef float f32x16 __attribute__((__vector_size__(16 * sizeof(float))));

constexpr int MLP_IN = 64;        // 63 encoded features + 1 zero pad
constexpr int MLP_H = 64;
constexpr int MLP_OUTP = 16;      // 3 outputs padded to one MFMA tile
constexpr int MLP_LAYERS = 5;
constexpr int MLP_RAW = 9;
constexpr int MLP_OCT = 3;
constexpr int MLP_BLOCK = 256;    // 4 waves

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

// Encoding of the 9 raw inputs: [u, sin(2^k pi u), cos(2^k pi u), k = 0..2].
// sin/cos of pi u come from the hardware sine and cosine (v_sin_f32 /
// v_cos_f32 take the angle in turns: pi u = u/2 turns, u in [0, 1]); octaves
// 1 and 2 follow by the double-angle identities.  The features feed BF16
// operands (8 significant bits), far coarser than either step's error.
struct Enc {
    float u[MLP_RAW], s[MLP_OCT][MLP_RAW], c[MLP_OCT][MLP_RAW];
};
SRT_D void enc_init(Enc& e, const float* u) {
#pragma unroll
    for (int i = 0; i < MLP_RAW; i++) {
        e.u[i] = u[i];
        float s = __builtin_amdgcn_sinf(0.5f * u[i]), c = __builtin_amdgcn_cosf(0.5f * u[i]);
        e.s[0][i] = s; e.c[0][i] = c;
#pragma unroll
        for (int k = 1; k < MLP_OCT; k++) {
            const float s2 = (2.0f * s) * c, c2 = (c - s) * (c + s);
            s = s2; c = c2;
            e.s[k][i] = s; e.c[k][i] = c;
        }
    }
}
// feature f (compile-time after unrolling) of the 64-wide encoded input
SRT_D float enc_get(const Enc& e, int f) {
    if (f < MLP_RAW) return e.u[f];
    if (f >= MLP_RAW + 2 * MLP_RAW * MLP_OCT) return 0.0f;
    const int j = f - MLP_RAW, oct = j / (2 * MLP_RAW), r = j % (2 * MLP_RAW);
    return r < MLP_RAW ? e.s[oct][r] : e.c[oct][r - MLP_RAW];
}

typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef short i16x2v __attribute__((ext_vector_type(2)));

SRT_D uint16_t bf(float x) {                 // round to nearest even
    __bf16 h = (__bf16)x;
    uint16_t u;
    __builtin_memcpy(&u, &h, 2);
    return u;
}
// two values rounded to BF16 (one v_cvt_pk_bf16_f32), a in the low half (lower address)
SRT_D uint32_t bf2(float a, float b) {
    const bf16x2v h = __builtin_convertvector((f32x2v){a, b}, bf16x2v);
    uint32_t u;
    __builtin_memcpy(&u, &h, 4);
    return u;
}
// ReLU of two BF16 values: as 16-bit integers, negative values (sign bit set)
// are negative and non-negative values keep their order, so max(x, 0) on the
// integers is max(x, +0) on the floats (v_pk_max_i16)
SRT_D uint32_t relu2(uint32_t v) {
    i16x2v x;
    __builtin_memcpy(&x, &v, 4);
    x = __builtin_elementwise_max(x, (i16x2v){0, 0});
    __builtin_memcpy(&v, &x, 4);
    return v;
}
// g where h != 0, else 0, per 16-bit half; h holds non-negative BF16 values.
// Two packed integer instructions: min(h, 1) per half is 0 or 1, and the
// low-half product applies it (written out: the compiler would otherwise
// expand the pair into compares and selects)
SRT_D uint32_t mask2(uint32_t g, uint32_t h) {
    uint32_t m, out;
    asm("v_pk_min_u16 %0, %1, %2" : "=v"(m) : "v"(h), "v"(0x00010001u));
    asm("v_pk_mul_lo_u16 %0, %1, %2" : "=v"(out) : "v"(g), "v"(m));
    return out;
}
SRT_D uint64_t pack64(uint32_t lo, uint32_t hi) { return (uint64_t)lo | ((uint64_t)hi << 32); }

SRT_D f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
SRT_D f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// ---------------------------------------------------------------------------
// Wave tiles.  A wave owns 32 samples.  Its activation rows live in LDS as
// act[sample][feature] (BF16, row stride AST), the operand form of the next
// layer's product.  Layer outputs come out of v_mfma_f32_32x32x16_bf16 as
// D[sample][feature]: lane (r = lane & 31, g = lane >> 5) holds feature r of
// the 32-feature tile and samples 8q + 4g + p in element 4q + p.  That same
// register form, four consecutive samples per 64-bit word, is what the weight
// gradients need (sample-contiguous rows), so it is kept in registers (hpk)
// and staged as [feature][sample] rows.
// ---------------------------------------------------------------------------
constexpr int T_WAVES = 4;
constexpr int T_TILE = 32;                   // samples per wave
constexpr int T_SUB = T_WAVES * T_TILE;      // samples per block step
constexpr int T_CHUNK_MAX = 1024;            // samples per block (= per weight-gradient partial), at most
constexpr int AST = 72;                      // row stride (BF16) of 64-wide LDS rows (+16 B against bank conflicts)
constexpr int SST = T_SUB + 8;               // row stride of the staged [feature][sample] rows
constexpr int WT4ST = 24;                    // row stride of W_4^T ([64][16])
constexpr int WF_ELEMS = 4 * 64 * AST + MLP_OUTP * AST;
constexpr int WT_ELEMS = 3 * 64 * AST + 64 * WT4ST;
// samples per block for a step of n samples: a multiple of the block step,
// about n / 256 (one block per CU of the 256-CU part), 128 .. 1024; a
// function of n alone, so the reduction order does not depend on the device
__host__ __device__ constexpr int train_chunk(int n) {
    const int per = (n + 255) / 256;
    const int c = (per + T_SUB - 1) / T_SUB * T_SUB;
    return c < T_SUB ? T_SUB : (c > T_CHUNK_MAX ? T_CHUNK_MAX : c);
}

SRT_D void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the four BF16 values of one register group (samples 8q + 4g + p, p = 0..3,
// of one feature) to / from rows base, base + AST, base + 2 AST, base + 3 AST of act
SRT_D void put_rows(uint16_t* act, int base, uint32_t lo, uint32_t hi) {
    act[base] = (uint16_t)lo;
    act[base + AST] = (uint16_t)(lo >> 16);
    act[base + 2 * AST] = (uint16_t)hi;
    act[base + 3 * AST] = (uint16_t)(hi >> 16);
}
SRT_D uint64_t get_rows(const uint16_t* act, int base) {
    const uint32_t lo = (uint32_t)act[base] | ((uint32_t)act[base + AST] << 16);
    const uint32_t hi = (uint32_t)act[base + 2 * AST] | ((uint32_t)act[base + 3 * AST] << 16);
    return pack64(lo, hi);
}

// forward weights W_l[j][k] -> sWf[(64 l + j) AST + k]; bias -> sB
SRT_D void load_forward_weights(const MlpDev& M, uint16_t* sWf, float* sB) {
    for (int i = threadIdx.x; i < W_TOTAL / 8; i += blockDim.x) {
        const int e = i * 8, l = e < 4 * 4096 ? e / 4096 : 4, o = e - w_off(l), j = o / 64, k = o % 64;
        *(uint4*)&sWf[(l * 64 + j) * AST + k] = ((const uint4*)M.wf)[i];
    }
    for (int i = threadIdx.x; i < B_TOTAL; i += blockDim.x) sB[i] = M.bias[i];
}

// encode samples t0 .. t0+31 into act: lane (r, g) writes features 32g .. 32g+31 of sample t0 + r
SRT_D void wave_encode(uint16_t* act, const float* in9, int t0, int n, int r, int g) {
    const int s = t0 + r;
    float u[MLP_RAW];
#pragma unroll
    for (int c = 0; c < MLP_RAW; c++) u[c] = (s < n) ? in9[(size_t)s * MLP_RAW + c] : 0.0f;
    Enc e;
    enc_init(e, u);
    uint32_t pk[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float a0 = g ? enc_get(e, 32 + 2 * i) : enc_get(e, 2 * i);
        const float a1 = g ? enc_get(e, 33 + 2 * i) : enc_get(e, 2 * i + 1);
        pk[i] = bf2(a0, a1);
    }
#pragma unroll
    for (int m = 0; m < 4; m++)
        *(uint4*)&act[r * AST + 32 * g + 8 * m] = make_uint4(pk[4 * m], pk[4 * m + 1], pk[4 * m + 2], pk[4 * m + 3]);
}

// hidden layer l (0..3) on the wave's 32 samples: act (input rows) -> act
// (output rows); the output's register form goes to h[jt][q]
SRT_D void wave_hidden(uint16_t* act, const uint16_t* sWf, const float* sB, int l, uint64_t h[2][4], int r, int g) {
    f32x16 acc[2];
#pragma unroll
    for (int jt = 0; jt < 2; jt++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[jt][i] = 0.0f;
    const uint16_t* W = sWf + l * 64 * AST;
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
        const bf16x8 a = ld_frag(act + r * AST + 16 * ks + 8 * g);
#pragma unroll
        for (int jt = 0; jt < 2; jt++) acc[jt] = mfma32(a, ld_frag(W + (32 * jt + r) * AST + 16 * ks + 8 * g), acc[jt]);
    }
#pragma unroll
    for (int jt = 0; jt < 2; jt++) {
        const float bj = sB[b_off(l) + 32 * jt + r];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            // ReLU after rounding equals rounding after ReLU (rounding keeps the sign)
            const uint32_t lo = relu2(bf2(acc[jt][4 * q] + bj, acc[jt][4 * q + 1] + bj));
            const uint32_t hi = relu2(bf2(acc[jt][4 * q + 2] + bj, acc[jt][4 * q + 3] + bj));
            h[jt][q] = pack64(lo, hi);
            put_rows(act, (8 * q + 4 * g) * AST + 32 * jt + r, lo, hi);
        }
    }
}

// output layer on act (H_4 rows): lane (r16, g16) gets y[16 rt + 4 g16 + i][channel r16] (bias not added)
SRT_D void wave_output(const uint16_t* act, const uint16_t* sWf, f32x4 y[2], int lane) {
    const int r16 = lane & 15, g16 = lane >> 4;
    const uint16_t* W = sWf + 4 * 64 * AST;
#pragma unroll
    for (int rt = 0; rt < 2; rt++) {
        y[rt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
            y[rt] = mfma16(ld_frag(act + (16 * rt + r16) * AST + 32 * ks + 8 * g16),
                           ld_frag(W + r16 * AST + 32 * ks + 8 * g16), y[rt]);
    }
}

// ---------------------------------------------------------------------------
// inference: in9 [n][9] -> y [n][3]; n from n_dev (device-side queue count)
// when given.  Waves run independently over 32-sample tiles.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(MLP_BLOCK) k_mlp_infer(const float* in9, int n_arg, const uint32_t* n_dev, MlpDev M,
                                                         float* y) {
    const int n = n_dev ? (int)*n_dev : n_arg;
    __shared__ __attribute__((aligned(16))) uint16_t sWf[WF_ELEMS];
    __shared__ float sB[B_TOTAL];
    __shared__ __attribute__((aligned(16))) uint16_t sAct[MLP_BLOCK / 64][T_TILE * AST];
    load_forward_weights(M, sWf, sB);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 31, g = lane >> 5, r16 = lane & 15, g16 = lane >> 4;
    uint16_t* act = sAct[w];
    const int ntiles = (n + T_TILE - 1) / T_TILE;
    for (int tile = blockIdx.x * (MLP_BLOCK / 64) + w; tile < ntiles; tile += gridDim.x * (MLP_BLOCK / 64)) {
        const int t0 = tile * T_TILE;
        wave_encode(act, in9, t0, n, r, g);
        wave_lds_order();
        uint64_t h[2][4];
#pragma unroll
        for (int l = 0; l < 4; l++) {
            wave_hidden(act, sWf, sB, l, h, r, g);
            wave_lds_order();
        }
        f32x4 yv[2];
        wave_output(act, sWf, yv, lane);
        if (r16 < 3) {
#pragma unroll
            for (int rt = 0; rt < 2; rt++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int s = t0 + 16 * rt + 4 * g16 + i;
                    if (s < n) y[(size_t)s * 3 + r16] = yv[rt][i] + sB[b_off(4) + r16];
                }
        }
        wave_lds_order();
    }
}

// ---------------------------------------------------------------------------
// training step (forward, loss gradient, backward, weight/bias gradients) in
// one kernel.  Block b owns samples [b chunk, (b+1) chunk) and writes one
// partial of every weight and bias gradient; k_mlp_adam sums the partials in
// block order, so a step is bit-reproducible.  Per block step of 128 samples
// (4 waves x 32): each wave encodes, runs the forward pass, the loss gradient
// and the backward products on its own samples; the layer gradients G_l and
// inputs H_l of all 128 samples are staged in LDS as [feature][sample] rows
// and every wave accumulates one tile of dW_l = G_l^T H_l over them.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(MLP_BLOCK, 1) k_mlp_train(const float* in9, const float* target, int n, int chunk,
                                                             float inv_count, MlpDev M, float* wpart, float* bpart,
                                                             float* loss_part) {
    __shared__ __attribute__((aligned(16))) uint16_t sWf[WF_ELEMS];
    __shared__ __attribute__((aligned(16))) uint16_t sWt[WT_ELEMS];     // W_l^T[k][j], layers 1..3, then layer 4
    __shared__ __attribute__((aligned(16))) uint16_t sAct[T_WAVES][T_TILE * AST];
    __shared__ __attribute__((aligned(16))) uint16_t sGt[64 * SST];     // G_l^T[feature][sample of the block step]
    __shared__ __attribute__((aligned(16))) uint16_t sHt[64 * SST];     // H_l^T[feature][sample of the block step]
    __shared__ float sB[B_TOTAL];
    __shared__ float sRed[MLP_BLOCK];
    load_forward_weights(M, sWf, sB);
    for (int i = threadIdx.x; i < 3 * 64 * 8; i += MLP_BLOCK) {
        const int l = 1 + i / 512, rem = i % 512, k = rem / 8, j = (rem % 8) * 8;
        *(uint4*)&sWt[((l - 1) * 64 + k) * AST + j] = *(const uint4*)&M.wt[w_off(l) + k * 64 + j];
    }
    for (int i = threadIdx.x; i < 64 * 2; i += MLP_BLOCK) {
        const int k = i / 2, j = (i % 2) * 8;
        *(uint4*)&sWt[3 * 64 * AST + k * WT4ST + j] = *(const uint4*)&M.wt[w_off(4) + k * MLP_OUTP + j];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 31, g = lane >> 5, r16 = lane & 15, g16 = lane >> 4;
    const int jt_w = w >> 1, kt_w = w & 1;      // this wave's 32x32 tile of each 64x64 weight gradient
    uint16_t* act = sAct[w];
    f32x16 accW[4];
#pragma unroll
    for (int l = 0; l < 4; l++)
#pragma unroll
        for (int i = 0; i < 16; i++) accW[l][i] = 0.0f;
    f32x4 accW4 = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // bias gradients: rows 16w .. 16w + 15 of each layer (layer 4: wave 0) as a product with a ones column
    f32x4 dbm[MLP_LAYERS];
#pragma unroll
    for (int l = 0; l < MLP_LAYERS; l++) dbm[l] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 ones;
    {
        const uint4 o = make_uint4(0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u);     // BF16 1.0
        __builtin_memcpy(&ones, &o, 16);
    }
    float lsum = 0.0f;
    const int blk_begin = blockIdx.x * chunk;        // chunk: a multiple of T_SUB
    const int blk_end = min(n, blk_begin + chunk);

    // stage the register form of a 64-feature gradient or activation tile as [feature][sample] rows
    auto stage = [&](uint16_t* dst, const uint64_t v[2][4]) {
#pragma unroll
        for (int jt = 0; jt < 2; jt++)
#pragma unroll
            for (int q = 0; q < 4; q++) *(uint64_t*)&dst[(32 * jt + r) * SST + T_TILE * w + 8 * q + 4 * g] = v[jt][q];
    };
    // the staged gradient rows 16w .. 16w + 15 summed over the block step's samples
    auto bias_rows = [&](f32x4& acc) {
#pragma unroll
        for (int ks = 0; ks < T_SUB / 32; ks++) acc = mfma16(ld_frag(sGt + (16 * w + r16) * SST + 32 * ks + 8 * g16), ones, acc);
    };

    for (int sb = blk_begin; sb < blk_end; sb += T_SUB) {
        const int t0 = sb + T_TILE * w;
        uint64_t hpk[MLP_LAYERS][2][4];         // H_l (input of layer l), register form
        wave_encode(act, in9, t0, n, r, g);
        wave_lds_order();
#pragma unroll
        for (int jt = 0; jt < 2; jt++)
#pragma unroll
            for (int q = 0; q < 4; q++) hpk[0][jt][q] = get_rows(act, (8 * q + 4 * g) * AST + 32 * jt + r);
#pragma unroll
        for (int l = 0; l < 4; l++) {
            wave_hidden(act, sWf, sB, l, hpk[l + 1], r, g);
            wave_lds_order();
        }
        f32x4 yv[2];
        wave_output(act, sWf, yv, lane);
        wave_lds_order();
        // loss gradient G_4 (channels 0..2 of 16): rows of act (columns 0..15) and staged rows
#pragma unroll
        for (int rt = 0; rt < 2; rt++) {
            float gval[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int s = t0 + 16 * rt + 4 * g16 + i;
                gval[i] = 0.0f;
                if (r16 < 3 && s < n) {
                    const float yy = yv[rt][i] + sB[b_off(4) + r16], tt = target[(size_t)s * 3 + r16];
                    const float d = yy - tt, inv = 1.0f / (yy * yy + 0.01f);
                    lsum += (d * d) * inv;
                    gval[i] = ((2.0f * d) * inv) * inv_count;
                }
            }
            const uint32_t lo = bf2(gval[0], gval[1]), hi = bf2(gval[2], gval[3]);
            put_rows(act, (16 * rt + 4 * g16) * AST + r16, lo, hi);
            *(uint64_t*)&sGt[r16 * SST + T_TILE * w + 16 * rt + 4 * g16] = pack64(lo, hi);
        }
        stage(sHt, hpk[4]);
        wave_lds_order();
        // G_3 = (G_4 W_4) masked by H_4 > 0
        uint64_t gpk[2][4];
        {
            f32x16 acc[2];
#pragma unroll
            for (int kt = 0; kt < 2; kt++)
#pragma unroll
                for (int i = 0; i < 16; i++) acc[kt][i] = 0.0f;
            const bf16x8 a = ld_frag(act + r * AST + 8 * g);
#pragma unroll
            for (int kt = 0; kt < 2; kt++)
                acc[kt] = mfma32(a, ld_frag(sWt + 3 * 64 * AST + (32 * kt + r) * WT4ST + 8 * g), acc[kt]);
#pragma unroll
            for (int kt = 0; kt < 2; kt++)
#pragma unroll
                for (int q = 0; q < 4; q++)
                    gpk[kt][q] = pack64(mask2(bf2(acc[kt][4 * q], acc[kt][4 * q + 1]), (uint32_t)hpk[4][kt][q]),
                                        mask2(bf2(acc[kt][4 * q + 2], acc[kt][4 * q + 3]), (uint32_t)(hpk[4][kt][q] >> 32)));
        }
        __syncthreads();
        // dW_4 (rows 0..15, columns 16 w .. 16 w + 15) and db_4 over the block step
#pragma unroll
        for (int ks = 0; ks < T_SUB / 32; ks++) {
            const bf16x8 a = ld_frag(sGt + r16 * SST + 32 * ks + 8 * g16);
            accW4 = mfma16(a, ld_frag(sHt + (16 * w + r16) * SST + 32 * ks + 8 * g16), accW4);
            if (w == 0) dbm[4] = mfma16(a, ones, dbm[4]);
        }
        __syncthreads();
#pragma unroll
        for (int l = 3; l >= 0; l--) {
            stage(sGt, gpk);
            stage(sHt, hpk[l]);
            if (l >= 1) {
                // G_{l-1} = (G_l W_l) masked by H_l > 0; G_l rows into act as the product's operand
#pragma unroll
                for (int jt = 0; jt < 2; jt++)
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        put_rows(act, (8 * q + 4 * g) * AST + 32 * jt + r, (uint32_t)gpk[jt][q], (uint32_t)(gpk[jt][q] >> 32));
                wave_lds_order();
                f32x16 acc[2];
#pragma unroll
                for (int kt = 0; kt < 2; kt++)
#pragma unroll
                    for (int i = 0; i < 16; i++) acc[kt][i] = 0.0f;
                const uint16_t* Wt = sWt + (l - 1) * 64 * AST;
#pragma unroll
                for (int ks = 0; ks < 4; ks++) {
                    const bf16x8 a = ld_frag(act + r * AST + 16 * ks + 8 * g);
#pragma unroll
                    for (int kt = 0; kt < 2; kt++)
                        acc[kt] = mfma32(a, ld_frag(Wt + (32 * kt + r) * AST + 16 * ks + 8 * g), acc[kt]);
                }
#pragma unroll
                for (int kt = 0; kt < 2; kt++)
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        gpk[kt][q] = pack64(mask2(bf2(acc[kt][4 * q], acc[kt][4 * q + 1]), (uint32_t)hpk[l][kt][q]),
                                            mask2(bf2(acc[kt][4 * q + 2], acc[kt][4 * q + 3]),
                                                  (uint32_t)(hpk[l][kt][q] >> 32)));
                wave_lds_order();
            }
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < T_SUB / 16; ks++)
                accW[l] = mfma32(ld_frag(sGt + (32 * jt_w + r) * SST + 16 * ks + 8 * g),
                                 ld_frag(sHt + (32 * kt_w + r) * SST + 16 * ks + 8 * g), accW[l]);
            bias_rows(dbm[l]);
            __syncthreads();
        }
    }

    // partials of this block: dW_l tile (jt_w, kt_w): lane (r, g) holds D[32 jt_w + 8q + 4g + p][32 kt_w + r]
    float* wp = wpart + (size_t)blockIdx.x * MLP_LAYERS * 4096;
#pragma unroll
    for (int l = 0; l < 4; l++)
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int p = 0; p < 4; p++)
                wp[l * 4096 + (32 * jt_w + 8 * q + 4 * g + p) * 64 + 32 * kt_w + r] = accW[l][4 * q + p];
#pragma unroll
    for (int i = 0; i < 4; i++) wp[4 * 4096 + (4 * g16 + i) * 64 + 16 * w + r16] = accW4[i];
    // bias partials: lane (0, g16) of wave w holds rows 16w + 4 g16 + i (every column of the product is the same)
    if (r16 == 0) {
        float* bp = bpart + (size_t)blockIdx.x * MLP_LAYERS * 64;
#pragma unroll
        for (int l = 0; l < 4; l++)
#pragma unroll
            for (int i = 0; i < 4; i++) bp[l * 64 + 16 * w + 4 * g16 + i] = dbm[l][i];
        if (w == 0) {
#pragma unroll
            for (int i = 0; i < 4; i++) bp[4 * 64 + 4 * g16 + i] = dbm[4][i];
        }
    }
    sRed[threadIdx.x] = lsum;
    __syncthreads();
    for (int st = MLP_BLOCK / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sRed[threadIdx.x] += sRed[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = sRed[0];
}

// Sum the per-block partials, Adam update, refresh the BF16 copies.  Block b
// handles parameters p = 64 b + (t & 63); thread group t >> 6 (of 16) sums the
// partials t >> 6, t >> 6 + 16, ... in order and the 16 group sums are added
// in group order, so the reduction order depends on the partial count alone.
// Partial c of weight p is wpart[c * 5 * 4096 + p], of bias q bpart[c * 5 * 64 + q].
constexpr int ADAM_PARAMS = 64;
constexpr int ADAM_GROUPS = 16;
constexpr int ADAM_BLOCK = ADAM_PARAMS * ADAM_GROUPS;
__global__ void __launch_bounds__(ADAM_BLOCK) k_mlp_adam(int nchunks, const float* wpart, const float* bpart, float* W, float* B,
                                                  float* mW, float* vW, float* mB, float* vB, float lr, float b1, float b2,
                                                  float eps, float c1, float c2, int apply, float* grad_out, uint16_t* wf,
                                                  uint16_t* wt) {
    __shared__ float sg[ADAM_BLOCK];
    const int t = threadIdx.x, grp = t >> 6;
    const int p = blockIdx.x * ADAM_PARAMS + (t & 63);
    float part = 0.0f;
    if (p < W_TOTAL) {
        for (int c = grp; c < nchunks; c += ADAM_GROUPS) part += wpart[(size_t)c * (MLP_LAYERS * 4096) + p];
    } else if (p < W_TOTAL + B_TOTAL) {
        for (int c = grp; c < nchunks; c += ADAM_GROUPS) part += bpart[(size_t)c * (MLP_LAYERS * 64) + (p - W_TOTAL)];
    }
    sg[t] = part;
    __syncthreads();
    if (grp != 0) return;
    float gtot = sg[t];
#pragma unroll
    for (int k = 1; k < ADAM_GROUPS; k++) gtot += sg[t + 64 * k];
    if (p < W_TOTAL) {
        int l = (p < 4 * 4096) ? p / 4096 : 4;
        int o = p - w_off(l);
        int j = o / 64, k = o % 64;
        bool live = !(l == 0 && k == 63) && !(l == 4 && j >= 3);
        float gsum = live ? gtot : 0.0f;
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
        float gsum = live ? gtot : 0.0f;
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
    DevBuf wpart, bpart, loss_part, loss, grad, in_buf, tgt_buf, out_buf;
    int64_t cap = 0;        // training sample capacity of the partials and targets
    int64_t infer_cap = 0;
    int step = 0;
    float lr = 1e-3f, b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float lo[3] = {0, 0, 0}, inv_ext[3] = {1, 1, 1};
};

static MlpState* mlp_of(srt_ctx* c) { return (MlpState*)c->mlp; }

static void mlp_free_one(srt_ctx* c, MlpState* m) {
    if (!m) return;
    DevBuf* all[] = {&m->W, &m->B, &m->mW, &m->vW, &m->mB, &m->vB, &m->wf, &m->wt, &m->bias_dev, &m->wpart, &m->bpart, &m->loss_part, &m->loss, &m->grad, &m->in_buf, &m->tgt_buf, &m->out_buf};
    for (DevBuf* b : all) dev_free(c, *b);
    delete m;
}

static void mlp_free_all(srt_ctx* c) {
    for (auto& kv : c->mlps) mlp_free_one(c, (MlpState*)kv.second);
    c->mlps.clear();
    c->mlp = nullptr;
}

void* mlp_for_scene(srt_ctx* c, uint64_t h) {
    auto it = c->mlps.find(h);
    if (it == c->mlps.end()) it = c->mlps.find(0);
    return it == c->mlps.end() ? nullptr : it->second;
}

void mlp_release_scene(srt_ctx* c, uint64_t h) {
    auto it = c->mlps.find(h);
    if (h == 0 || it == c->mlps.end()) return;
    if (c->mlp == it->second) c->mlp = nullptr;
    mlp_free_one(c, (MlpState*)it->second);
    c->mlps.erase(it);
    if (c->mlp_sel == h) {
        c->mlp_sel = 0;
        auto d = c->mlps.find(0);
        c->mlp = d == c->mlps.end() ? nullptr : d->second;
    }
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
    k_mlp_adam<<<(tot + ADAM_PARAMS - 1) / ADAM_PARAMS, ADAM_BLOCK, 0, c->stream>>>(0, nullptr, nullptr, (float*)m->W.p, (float*)m->B.p, nullptr,
                                                         nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 0,
                                                         nullptr, (uint16_t*)m->wf.p, (uint16_t*)m->wt.p);
    return hip_check(c, hipGetLastError(), "mlp refresh");
}

static int ensure_train(srt_ctx* c, MlpState* m, int64_t n) {
    if (m->cap >= n) return SRT_OK;
    DevBuf* bs[] = {&m->wpart, &m->bpart, &m->loss_part, &m->tgt_buf};
    for (DevBuf* b : bs) dev_free(c, *b);
    // blocks of any step of at most n samples: ceil(n' / train_chunk(n')) <= max(256, ceil(n / 1024))
    const int64_t chunks = std::max<int64_t>(256, (n + T_CHUNK_MAX - 1) / T_CHUNK_MAX);
    int r = dev_alloc(c, m->tgt_buf, (size_t)n * 12);
    if (!r) r = dev_alloc(c, m->loss_part, (size_t)chunks * 4);
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

// inference grid: 32-sample tiles, 4 per block step, at most two blocks per CU (LDS)
static int infer_grid(srt_ctx* c, int64_t n) {
    int64_t tiles = (n + T_TILE - 1) / T_TILE;
    int64_t blocks = (tiles + MLP_BLOCK / 64 - 1) / (MLP_BLOCK / 64);
    hipDeviceProp_t prop;
    int cus = 256;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) cus = prop.multiProcessorCount;
    int64_t cap = (int64_t)cus * 2;
    return (int)std::max<int64_t>(1, std::min(blocks, cap));
}

// one training step on device arrays (in9 [n][9], target [n][3]); apply = 0
// computes gradients only (grad_out host pointer optional)
static int train_step_dev(srt_ctx* c, MlpState* m, const float* in9, const float* target, int n, int apply, float* loss_host,
                          float* grad_host) {
    if (int r = ensure_train(c, m, n)) return r;
    hipStream_t st = c->stream;
    MlpDev d = mlp_dev(m);
    const int chunk = train_chunk(n), chunks = (n + chunk - 1) / chunk;
    if (!m->loss.p) if (int r = dev_alloc(c, m->loss, 16)) return r;
    if (!m->grad.p) if (int r = dev_alloc(c, m->grad, (size_t)(W_TOTAL + B_TOTAL) * 4)) return r;
    hipEvent_t ev;
    prof_begin(c, "mlp_train", &ev);
    k_mlp_train<<<chunks, MLP_BLOCK, 0, st>>>(in9, target, n, chunk, 1.0f / (3.0f * (float)n), d, (float*)m->wpart.p,
                                              (float*)m->bpart.p, (float*)m->loss_part.p);
    prof_end(c, "mlp_train", ev, n);
    k_loss_sum<<<1, 64, 0, st>>>((const float*)m->loss_part.p, chunks, (float*)m->loss.p);
    if (apply) m->step++;
    const float c1 = 1.0f - std::pow(m->b1, (float)std::max(m->step, 1));
    const float c2 = 1.0f - std::pow(m->b2, (float)std::max(m->step, 1));
    const int tot = W_TOTAL + B_TOTAL;
    prof_begin(c, "mlp_adam", &ev);
    k_mlp_adam<<<(tot + ADAM_PARAMS - 1) / ADAM_PARAMS, ADAM_BLOCK, 0, st>>>(chunks, (const float*)m->wpart.p, (const float*)m->bpart.p, (float*)m->W.p,
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
    k_mlp_infer<<<infer_grid(c, n), MLP_BLOCK, 0, c->stream>>>(in9, n, nullptr, mlp_dev(m), out3);
    prof_end(c, "mlp_infer", ev, n);
    return hip_check(c, hipGetLastError(), "mlp infer launch");
}

// forward over a queue whose length lives in device memory (at most max_n)
int mlp_query_dev(srt_ctx* c, const float* in9, const uint32_t* count_dev, int max_n, float* out3) {
    MlpState* m = mlp_of(c);
    if (!m) return set_err(c, SRT_E_STATE, "no MLP: call srt_mlp_create first");
    hipEvent_t ev;
    prof_begin(c, "mlp_query", &ev);
    k_mlp_infer<<<infer_grid(c, max_n), MLP_BLOCK, 0, c->stream>>>(in9, 0, count_dev, mlp_dev(m), out3);
    uint32_t queries = 0;
    if (c->profile) {
        // profiling only: the query count, for the per-query rate in the statistics
        if (int r = hip_check(c, hipMemcpyAsync(&queries, count_dev, 4, hipMemcpyDeviceToHost, c->stream), "count")) return r;
        if (int r = hip_check(c, hipStreamSynchronize(c->stream), "count")) return r;
    }
    prof_end(c, "mlp_query", ev, queries);
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
    {
        auto it = c->mlps.find(c->mlp_sel);
        if (it != c->mlps.end()) {
            mlp_free_one(c, (MlpState*)it->second);
            c->mlps.erase(it);
        }
    }
    MlpState* m = new MlpState();
    c->mlps[c->mlp_sel] = m;
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
    if (r) {
        mlp_free_one(c, m);
        c->mlps.erase(c->mlp_sel);
        c->mlp = nullptr;
        return r;
    }
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

// Selects the MLP the other srt_mlp_* calls act on: scene h's own (0: the
// context's own).  Renders with mlp_mode 1 and training frames of a scene use
// its own MLP when it has one, else the context's own.
int srt_mlp_select(srt_ctx* c, srt_scene h) {
    SRT_IDLE(c);
    if (!c) return SRT_E_INVALID;
    if (h != 0 && !c->scenes.count(h)) return set_err(c, SRT_E_NOT_FOUND, "unknown scene %llu", (unsigned long long)h);
    c->mlp_sel = h;
    auto it = c->mlps.find(h);
    c->mlp = it == c->mlps.end() ? nullptr : it->second;
    return SRT_OK;
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
