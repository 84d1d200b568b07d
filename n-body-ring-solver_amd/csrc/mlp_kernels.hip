// mlp_kernels.hip -- BF16 MFMA kernels of the learned acceleration
// extrapolator: a 15 -> 128 -> 128 -> 128 -> 3 MLP with SiLU hidden
// activations, FP32 accumulation, trained with Adam on FP32 master weights.
//
// Every matrix product runs on MFMA (v_mfma_f32_16x16x16_bf16; the
// inference kernel's layers 2-4 on gfx950's v_mfma_f32_16x16x32_bf16) in
// "transposed" form: a layer computes D^T = W X^T with the weights as the A
// operand (rows = output features) and 16 particles as the columns of the B
// operand.  The accumulator layout of one layer -- lane (r, g), r = lane % 16,
// g = lane / 16, holds features 4g..4g+3 of particle r -- is exactly the
// B-operand layout the next layer needs, so activations never leave registers
// inside the forward or the backward sweep.  Weight gradients dW = delta^T H
// are a GEMM over the batch; the sweeps store delta and H feature-major (bf16)
// for it.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mlp_kernels.h"

namespace mlp {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round to nearest even: the conversion compiles to gfx950's
// v_cvt_pk_bf16_f32 (NaN stays NaN).
__device__ __forceinline__ unsigned short to_bf16(float f) {
    return __builtin_bit_cast(unsigned short, (__bf16)f);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 v;
    v[0] = (short)to_bf16(a);
    v[1] = (short)to_bf16(b);
    v[2] = (short)to_bf16(c);
    v[3] = (short)to_bf16(d);
    return v;
}

__device__ __forceinline__ bf16x4 load4(const unsigned short* __restrict__ p) {
    return __builtin_bit_cast(bf16x4, *(const uint2*)p);
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// v_exp_f32 and v_rcp_f32 (1 ulp) instead of an IEEE division: every value
// computed from it enters a matrix product as a bf16 operand (8 bits).
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_d(float z) {
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

// Feature-major bf16 store of one accumulator tile: rows f0 + 4g + j, column p.
__device__ __forceinline__ void store_fm(unsigned short* __restrict__ T, long ld, int f0, int g, long p, bf16x4 v) {
#pragma unroll
    for (int j = 0; j < 4; j++) T[(long)(f0 + 4 * g + j) * ld + p] = (unsigned short)v[j];
}

}  // namespace

// ---------------------------------------------------------------------------
// Training forward: Y = f(X) for rows [0, rows) with the pre-activations Z_l
// (row-major [rows][128] fp32) and the layer inputs H_l (feature-major bf16,
// H0 [16][ld], H1..H3 [128][ld]) kept for the backward pass.  Rows past
// `rows` (up to the 16-row tile) read zero inputs and are not written.
// A workgroup of kTrainWaves = 8 waves takes one 16-particle tile at a time:
// wave w computes the 16 output features 16 w .. 16 w + 15 of every hidden
// layer (the same MFMA chain, in the same order, as one wave computing all
// eight), and the layer's activations reach the other waves through LDS
// (two alternating buffers, one barrier per layer), so a tile's latency is
// three short chains instead of 144 dependent products, and a training batch
// of 4096 rows spreads over 256 workgroups.
// ---------------------------------------------------------------------------
constexpr int kTrainWaves = 8;

__global__ __launch_bounds__(64 * kTrainWaves) void k_mlp_fwd(Net net, const float* __restrict__ X, int rows,
                                                              float* __restrict__ Y, float* __restrict__ Z,
                                                              unsigned short* __restrict__ H, long ld) {
    __shared__ bf16x4 hs[2][8][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int tiles = (rows + 15) / 16;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const long p = (long)tile * 16 + r;
        const bool valid = p < rows;
        const float4 xin = valid ? *(const float4*)(X + p * 16 + 4 * g) : make_float4(0.f, 0.f, 0.f, 0.f);
        const bf16x4 h0 = pack4(xin.x, xin.y, xin.z, xin.w);
        if (w == 0) store_fm(H, ld, 0, g, p, h0);
        f32x4 acc = mfma(load4(net.W1 + (long)(16 * w + r) * 16 + 4 * g), h0, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
        for (int l = 0; l < 3; l++) {   // bias + SiLU of layer l+1, then (waves' share of) layer l+2's product
            const float* b = (l == 0) ? net.b1 : (l == 1) ? net.b2 : net.b3;
            f32x4 zz = acc;
#pragma unroll
            for (int j = 0; j < 4; j++) zz[j] += b[16 * w + 4 * g + j];
            if (valid) *(f32x4*)(Z + (long)l * ld * 128 + p * 128 + 16 * w + 4 * g) = zz;
            const bf16x4 hv = pack4(silu(zz[0]), silu(zz[1]), silu(zz[2]), silu(zz[3]));
            store_fm(H + 16 * ld + (long)l * 128 * ld, ld, 16 * w, g, p, hv);
            hs[l & 1][w][lane] = hv;
            __syncthreads();
            if (l < 2) {
                const unsigned short* W = (l == 0) ? net.W2 : net.W3;
                acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kt = 0; kt < 8; kt++) acc = mfma(load4(W + (long)(16 * w + r) * 128 + 16 * kt + 4 * g), hs[l & 1][kt][lane], acc);
            }
        }
        if (w == 0) {   // output layer from layer 3's activations (buffer 0)
            f32x4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 8; kt++) y = mfma(load4(net.W4 + (long)r * 128 + 16 * kt + 4 * g), hs[0][kt][lane], y);
            if (g == 0 && valid) *(float4*)(Y + p * 4) = make_float4(y[0] + net.b4[0], y[1] + net.b4[1], y[2] + net.b4[2], 0.f);
        }
        __syncthreads();   // buffer 0 is rewritten by the next tile's first layer
    }
}

// ---------------------------------------------------------------------------
// Inference forward (no saved activations): the arithmetic of k_mlp_fwd (bf16
// operands, FP32 accumulation) in fewer, wider instructions.  Layers 2-4 run
// on gfx950's v_mfma_f32_16x16x32_bf16: K = 32 per instruction, the lane's
// two 4-feature pieces of blocks 2 kb and 2 kb + 1 of the previous layer's
// output as one 8-value B operand, the weights stored in that order (Wtile),
// so activations still never leave registers: 76 MFMAs per 16-particle tile
// instead of 144.  Each accumulator starts from the layer's bias, so the
// bias costs no VALU instruction.  Every weight tile and bias is staged once
// per workgroup in LDS (72 KiB + 1.6 KiB, two workgroups per CU) and read
// back linearly, 8 or 16 bytes per lane.  An empty asm with a memory clobber
// at the top of the tile loop keeps the compiler from hoisting the weight
// tiles into registers (which would cost the occupancy).
// ---------------------------------------------------------------------------
constexpr int kInferWaves = 8;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(Net net, const float* __restrict__ X, int rows,
                                                                float* __restrict__ Y) {
    __shared__ uint4 wt[TILES * 32];
    __shared__ f32x4 bias[(3 * 128 + 16) / 4];
    {
        const uint4* src = (const uint4*)net.Wtile;
        for (int k = threadIdx.x; k < TILES * 32; k += 64 * kInferWaves) wt[k] = src[k];
        float* b = (float*)bias;
        for (int k = threadIdx.x; k < 3 * 128 + 16; k += 64 * kInferWaves)
            b[k] = k < 128 ? net.b1[k] : k < 256 ? net.b2[k - 128] : k < 384 ? net.b3[k - 256] : net.b4[k - 384];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const int tiles = (rows + 15) / 16;
    const int wave = blockIdx.x * kInferWaves + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * kInferWaves;
    const uint2* w1 = (const uint2*)wt;                 // W1 tiles: 8 bytes per lane
    const uint4* wide = wt + W1_TILES * 64 / 2;         // then 16 bytes per lane
    for (int tile = wave; tile < tiles; tile += nwaves) {
        asm volatile("" ::: "memory");
        const long p = (long)tile * 16 + r;
        const bool valid = p < rows;
        const float4 xin = valid ? *(const float4*)(X + p * 16 + 4 * g) : make_float4(0.f, 0.f, 0.f, 0.f);
        const bf16x4 h0 = pack4(xin.x, xin.y, xin.z, xin.w);
        bf16x4 h[8];
        f32x4 z[8];
#pragma unroll
        for (int m = 0; m < 8; m++) z[m] = mfma(__builtin_bit_cast(bf16x4, w1[m * 64 + lane]), h0, bias[4 * m + g]);
#pragma unroll
        for (int l = 0; l < 3; l++) {   // SiLU of layer l+1, then layer l+2's product
#pragma unroll
            for (int m = 0; m < 8; m++) h[m] = pack4(silu(z[m][0]), silu(z[m][1]), silu(z[m][2]), silu(z[m][3]));
            if (l < 2) {
                const int u0 = l == 0 ? U_W2 : U_W3;
#pragma unroll
                for (int m = 0; m < 8; m++) {
                    f32x4 acc = bias[32 * (l + 1) + 4 * m + g];
#pragma unroll
                    for (int kb = 0; kb < 4; kb++)
                        acc = mfma32(__builtin_bit_cast(s16x8, wide[(u0 + 4 * m + kb) * 64 + lane]),
                                     __builtin_shufflevector(h[2 * kb], h[2 * kb + 1], 0, 1, 2, 3, 4, 5, 6, 7), acc);
                    z[m] = acc;
                }
            }
        }
        f32x4 acc = bias[96 + g];   // b4 padded with zeros to 16
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
            acc = mfma32(__builtin_bit_cast(s16x8, wide[(U_W4 + kb) * 64 + lane]),
                         __builtin_shufflevector(h[2 * kb], h[2 * kb + 1], 0, 1, 2, 3, 4, 5, 6, 7), acc);
        if (g == 0 && valid) *(float4*)(Y + p * 4) = make_float4(acc[0], acc[1], acc[2], 0.f);
    }
}

// ---------------------------------------------------------------------------
// Backward of L = sum (y - t)^2 / (2 rows): deltas of every layer, stored
// feature-major bf16 (D4 [16][ld], D3..D1 [128][ld]); per-tile column sums of
// the deltas (bias gradients, fp32) and of the squared error.  Split like the
// training forward: wave w of a workgroup computes features 16 w .. 16 w + 15
// of each hidden layer's delta, the layer's deltas pass to the other waves
// through LDS (two alternating buffers, one barrier per layer).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64 * kTrainWaves) void k_mlp_bwd(Net net, const float* __restrict__ Y,
                                                              const float* __restrict__ T, int rows,
                                                              const float* __restrict__ Z, unsigned short* __restrict__ D,
                                                              long ld, float* __restrict__ dbpart,
                                                              float* __restrict__ losspart) {
    __shared__ bf16x4 dsh[2][8][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int tiles = (rows + 15) / 16;
    const float inv = 1.0f / (float)rows;
    // column sums over the 16 particles (lanes with equal g) by xor-shuffles
    auto colsum = [&](float v) {
        for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
        return v;
    };
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const long p = (long)tile * 16 + r;
        const bool valid = p < rows;
        // output delta: features 0..2 live in the g == 0 lanes
        float e[3] = {0.f, 0.f, 0.f};
        if (g == 0 && valid) {
            const float4 y = *(const float4*)(Y + p * 4);
            const float4 t = *(const float4*)(T + p * 4);
            e[0] = y.x - t.x; e[1] = y.y - t.y; e[2] = y.z - t.z;
        }
        const f32x4 d4 = {e[0] * inv, e[1] * inv, e[2] * inv, 0.f};
        const bf16x4 dl0 = pack4(d4[0], d4[1], d4[2], d4[3]);
        float* dbt = dbpart + (long)tile * (16 + 3 * 128);
        if (w == 0) {
            store_fm(D, ld, 0, g, p, dl0);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float s = colsum(d4[j]);
                if (r == 0) dbt[4 * g + j] = s;
            }
            const float sq = colsum(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            if (lane == 0) losspart[tile] = sq;
        }
        // hidden layers 3, 2, 1: this wave's 16 features of each
#pragma unroll
        for (int l = 2; l >= 0; l--) {
            f32x4 a = {0.f, 0.f, 0.f, 0.f};
            if (l == 2) {
                a = mfma(load4(net.W4T + (long)(16 * w + r) * 16 + 4 * g), dl0, a);
            } else {
                const unsigned short* WT = (l == 1) ? net.W3T : net.W2T;
#pragma unroll
                for (int kt = 0; kt < 8; kt++) a = mfma(load4(WT + (long)(16 * w + r) * 128 + 16 * kt + 4 * g), dsh[(l + 1) & 1][kt][lane], a);
            }
            const f32x4 zz = valid ? *(const f32x4*)(Z + (long)l * ld * 128 + p * 128 + 16 * w + 4 * g) : f32x4{0.f, 0.f, 0.f, 0.f};
            f32x4 d;
#pragma unroll
            for (int j = 0; j < 4; j++) d[j] = valid ? a[j] * silu_d(zz[j]) : 0.f;
            const bf16x4 dv = pack4(d[0], d[1], d[2], d[3]);
            store_fm(D + 16 * ld + (long)(2 - l) * 128 * ld, ld, 16 * w, g, p, dv);
            float* dbl = dbt + 16 + (long)l * 128;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float s = colsum(d[j]);
                if (r == 0) dbl[16 * w + 4 * g + j] = s;
            }
            if (l > 0) {
                dsh[l & 1][w][lane] = dv;
                __syncthreads();
            }
        }
        // no barrier at the end of a tile: buffer 0 was last read before this
        // tile's l = 1 barrier, and buffer 1 is rewritten only after the next
        // tile's l = 2 barrier, which every wave reaches after reading it
    }
}

// ---------------------------------------------------------------------------
// dW_l = delta_l^T H_{l-1}, K = ld (the padded batch): one workgroup of
// kDwWaves waves per 16x16 output tile.  Tiles: layer 1 (8 x 1), layer 2
// (8 x 8), layer 3 (8 x 8), layer 4 (1 x 8).  Wave w takes the batch steps
// k = 16 (w + kDwWaves i) and wave 0 adds the waves' accumulators in wave
// order (the batch split so that its sum, 256 MFMA steps for a batch of
// 4096, is not one latency-bound chain).  Writes fp32 into the packed
// gradient (W row-major [out][in], W1 with in = 15, W4 with out = 3).
// ---------------------------------------------------------------------------
constexpr int kDwWaves = 8;

__global__ __launch_bounds__(64 * kDwWaves) void k_mlp_dw(const unsigned short* __restrict__ H,
                                                          const unsigned short* __restrict__ D, long ld,
                                                          float* __restrict__ grad) {
    __shared__ f32x4 part[kDwWaves][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int t = blockIdx.x;   // 0 .. TILES - 1
    // layer, tile (mt, nt), operand bases
    int layer_id, mt, nt;
    if (t < 8) { layer_id = 0; mt = t; nt = 0; }
    else if (t < 72) { layer_id = 1; mt = (t - 8) / 8; nt = (t - 8) % 8; }
    else if (t < 136) { layer_id = 2; mt = (t - 72) / 8; nt = (t - 72) % 8; }
    else { layer_id = 3; mt = 0; nt = t - 136; }
    // delta of layer l: D layout [D4 16 rows][D3 128][D2 128][D1 128]
    const long drow = (layer_id == 3) ? 0 : 16 + (long)(2 - layer_id) * 128;
    const long hrow = (layer_id == 0) ? 0 : 16 + (long)(layer_id - 1) * 128;
    const unsigned short* A = D + (drow + 16 * mt + r) * ld;   // A[m][k] = delta[k][out]
    const unsigned short* B = H + (hrow + 16 * nt + r) * ld;   // B[k][n] = H[k][in]
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    // steps k = 16 (w + kDwWaves s), s < steps; eight steps' loads are issued
    // before their products (the sum stays one chain in k order)
    const int steps = (int)((ld - 16L * w + 16L * kDwWaves - 1) / (16L * kDwWaves));
    int s = 0;
    for (; s + 8 <= steps; s += 8) {
        bf16x4 a[8], b[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const long k = 16L * (w + (long)kDwWaves * (s + q));
            a[q] = load4(A + k + 4 * g);
            b[q] = load4(B + k + 4 * g);
        }
#pragma unroll
        for (int q = 0; q < 8; q++) acc = mfma(a[q], b[q], acc);
    }
    for (; s < steps; s++) {
        const long k = 16L * (w + (long)kDwWaves * s);
        acc = mfma(load4(A + k + 4 * g), load4(B + k + 4 * g), acc);
    }
    part[w][lane] = acc;
    __syncthreads();
    if (w != 0) return;
    acc = part[0][lane];
    for (int q = 1; q < kDwWaves; q++) {
        const f32x4 v = part[q][lane];
#pragma unroll
        for (int j = 0; j < 4; j++) acc[j] += v[j];
    }
    const int in_dim = (layer_id == 0) ? 15 : 128;
    const int out_dim = (layer_id == 3) ? 3 : 128;
    const long off = (layer_id == 0) ? 0 : (layer_id == 1) ? OFF_W2 : (layer_id == 2) ? OFF_W3 : OFF_W4;
    const int col = 16 * nt + r;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int row = 16 * mt + 4 * g + j;
        if (row < out_dim && col < in_dim) grad[off + (long)row * in_dim + col] = acc[j];
    }
}

// Bias gradients and loss: sums over the batch's tiles.  A workgroup of
// kDbWaves waves per 64 columns (column 400: the loss): wave w takes tiles
// w, w + kDbWaves, ..., and wave 0 adds the waves' sums in wave order.
constexpr int kDbWaves = 16;
constexpr int kDbCols = 16 + 3 * 128 + 1;   // dbpart rows, then the loss

__global__ __launch_bounds__(64 * kDbWaves) void k_mlp_db(const float* __restrict__ dbpart,
                                                          const float* __restrict__ losspart, int tiles,
                                                          float* __restrict__ grad, float* __restrict__ loss) {
    __shared__ float part[kDbWaves][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    float s = 0.f;
    if (i < kDbCols - 1) {
#pragma unroll 8
        for (int t = w; t < tiles; t += kDbWaves) s += dbpart[(long)t * (16 + 3 * 128) + i];
    } else if (i == kDbCols - 1) {
#pragma unroll 8
        for (int t = w; t < tiles; t += kDbWaves) s += losspart[t];
    }
    part[w][lane] = s;
    __syncthreads();
    if (w != 0 || i >= kDbCols) return;
    s = part[0][lane];
    for (int q = 1; q < kDbWaves; q++) s += part[q][lane];
    // dbpart rows: [0,16) output, [16,144) layer 1, [144,272) layer 2, [272,400) layer 3
    if (i < 3) grad[OFF_B4 + i] = s;
    else if (i >= 16 && i < 144) grad[OFF_B1 + (i - 16)] = s;
    else if (i >= 144 && i < 272) grad[OFF_B2 + (i - 144)] = s;
    else if (i >= 272 && i < 400) grad[OFF_B3 + (i - 272)] = s;
    else if (i == 400) *loss = s;
}

// Adam (Kingma & Ba 2015) with bias correction on the FP32 master weights.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies for the kernels: W1 [128][16] (column 15 zero), W2, W3 [128][128],
// W4 [16][128] (rows 3..15 zero) and the transposes W2T, W3T [128][128],
// W4T [128][16] (columns 3..15 zero); fp32 biases (b4 padded to 16).
__global__ __launch_bounds__(256) void k_publish(const float* __restrict__ p, Net net) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    unsigned short* W1 = (unsigned short*)net.W1;
    unsigned short* W2 = (unsigned short*)net.W2;
    unsigned short* W3 = (unsigned short*)net.W3;
    unsigned short* W4 = (unsigned short*)net.W4;
    unsigned short* W2T = (unsigned short*)net.W2T;
    unsigned short* W3T = (unsigned short*)net.W3T;
    unsigned short* W4T = (unsigned short*)net.W4T;
    if (i < 128 * 16) {
        const int o = i / 16, c = i % 16;
        W1[i] = (c < 15) ? to_bf16(p[OFF_W1 + o * 15 + c]) : 0;
        const int row = i / 16, k = i % 16;   // W4T[row = in][k = out]
        W4T[i] = (k < 3) ? to_bf16(p[OFF_W4 + k * 128 + row]) : 0;
        const int o4 = i / 128, c4 = i % 128;   // W4[o4][c4]
        W4[i] = (o4 < 3) ? to_bf16(p[OFF_W4 + o4 * 128 + c4]) : 0;
    }
    if (i < 128 * 128) {
        const int o = i / 128, c = i % 128;
        W2[i] = to_bf16(p[OFF_W2 + i]);
        W3[i] = to_bf16(p[OFF_W3 + i]);
        W2T[i] = to_bf16(p[OFF_W2 + c * 128 + o]);
        W3T[i] = to_bf16(p[OFF_W3 + c * 128 + o]);
    }
    if (i < 128) {
        ((float*)net.b1)[i] = p[OFF_B1 + i];
        ((float*)net.b2)[i] = p[OFF_B2 + i];
        ((float*)net.b3)[i] = p[OFF_B3 + i];
    }
    if (i < 16) ((float*)net.b4)[i] = (i < 3) ? p[OFF_B4 + i] : 0.f;
    if (i < TILES * 256) {   // MFMA operand tiles (see Net)
        float v;
        if (i < W1_TILES * 256) {
            const int m = i / 256, w = i % 256, ln = w / 4, j = w % 4, r = ln % 16, g = ln / 16;
            const int col = 4 * g + j;
            v = col < NIN ? p[OFF_W1 + (16 * m + r) * NIN + col] : 0.f;
        } else {
            const int k = i - W1_TILES * 256;
            const int u = k / 512, w = k % 512, ln = w / 8, e = w % 8, r = ln % 16, g = ln / 16;
            const int kb = u % 4, m = (u % 32) / 4;
            const int in = 32 * kb + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4));
            if (u < U_W4) v = p[(u < U_W3 ? OFF_W2 : OFF_W3) + (16 * m + r) * NH + in];
            else v = r < NOUT ? p[OFF_W4 + r * NH + in] : 0.f;
        }
        ((unsigned short*)net.Wtile)[i] = to_bf16(v);
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
// Training kernels: one workgroup per 16-particle tile, up to 1024 of them
// (each then takes every 1024th tile).
static int grid_for(int rows) {
    const int tiles = (rows + 15) / 16;
    return std::max(1, std::min(1024, tiles));
}

hipError_t forward(const Net& net, const float* X, int rows, float* Y, bool save, float* Z, unsigned short* H, long ld,
                   hipStream_t st) {
    if (rows <= 0) return hipSuccess;
    if (!save) {
        const int tiles = (rows + 15) / 16;
        const int blocks = std::max(1, std::min(512, (tiles + kInferWaves - 1) / kInferWaves));
        hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(64 * kInferWaves), 0, st, net, X, rows, Y);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_mlp_fwd, dim3(grid_for(rows)), dim3(64 * kTrainWaves), 0, st, net, X, rows, Y, Z, H, ld);
    return hipGetLastError();
}

hipError_t backward(const Net& net, const float* Y, const float* T, int rows, const float* Z, const unsigned short* H,
                    unsigned short* D, long ld, float* dbpart, float* losspart, float* grad, float* loss,
                    hipStream_t st) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_bwd, dim3(grid_for(rows)), dim3(64 * kTrainWaves), 0, st, net, Y, T, rows, Z, D, ld, dbpart,
                       losspart);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_dw, dim3(TILES), dim3(64 * kDwWaves), 0, st, H, (const unsigned short*)D, ld, grad);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_db, dim3((kDbCols + 63) / 64), dim3(64 * kDbWaves), 0, st, dbpart, losspart,
                       (rows + 15) / 16, grad, loss);
    return hipGetLastError();
}

hipError_t adam(float* p, const float* g, float* m, float* v, int n, float lr, int t, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t publish(const float* p, const Net& net, hipStream_t st) {
    hipLaunchKernelGGL(k_publish, dim3(TILES), dim3(256), 0, st, p, net);
    return hipGetLastError();
}

}  // namespace mlp
