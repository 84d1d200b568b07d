// mlp_kernels.h -- interface of the extrapolator network kernels (internal).
#pragma once

#include <hip/hip_runtime.h>

namespace mlp {

constexpr int NIN = 15, NH = 128, NOUT = 3;

// Packed FP32 parameters: W1 [128][15] b1 [128] W2 [128][128] b2 W3 [128][128] b3 W4 [3][128] b4 [3].
constexpr long OFF_W1 = 0;
constexpr long OFF_B1 = OFF_W1 + NH * NIN;
constexpr long OFF_W2 = OFF_B1 + NH;
constexpr long OFF_B2 = OFF_W2 + NH * NH;
constexpr long OFF_W3 = OFF_B2 + NH;
constexpr long OFF_B3 = OFF_W3 + NH * NH;
constexpr long OFF_W4 = OFF_B3 + NH;
constexpr long OFF_B4 = OFF_W4 + NOUT * NH;
constexpr long NPARAM = OFF_B4 + NOUT;

// Device copies the kernels read: bf16 weights (padded, plus transposes for
// the backward sweep) and fp32 biases.
// MFMA operand tiles of the forward weights (Wtile, bf16), read linearly by
// the inference kernel:
//  - W1: 8 tiles of 64 lanes x 4 values (16 x 16 blocks, K = 16): tile m
//    holds W1[16 m + r][4 g + j] at lane l = r + 16 g, element j;
//  - W2, W3, W4: 68 tiles of 64 lanes x 8 values (16 x 32 blocks, K = 32)
//    after the W1 tiles: tile u = 4 m + kb (W2, u < 32; W3 at 32 + 4 m + kb;
//    W4 at 64 + kb) holds, at lane r + 16 g, element e, the weight of output
//    16 m + r and input 32 kb + 4 g + e (e < 4) or 32 kb + 16 + 4 g + e - 4
//    (e >= 4): the two 16-feature blocks 2 kb, 2 kb + 1 that one lane holds of
//    the layer's input, in that order.
// TILES counts 16 x 16 blocks (W1 8, W2 64, W3 64, W4 8): also the weight
// gradient's tiles.
constexpr int TILES = 144;
constexpr int W1_TILES = 8, WIDE_TILES = 68, U_W2 = 0, U_W3 = 32, U_W4 = 64;
static_assert(W1_TILES * 64 * 4 + WIDE_TILES * 64 * 8 == TILES * 64 * 4, "tile storage");

struct Net {
    const unsigned short *W1, *W2, *W3, *W4;    // [128][16], [128][128] x2, [16][128]
    const unsigned short *W2T, *W3T, *W4T;      // [128][128] x2, [128][16]
    const float *b1, *b2, *b3, *b4;             // 128, 128, 128, 16
    const unsigned short* Wtile;                // TILES x 64 lanes x 4 bf16
};

// bytes of the bf16 + bias copies addressed by Net
constexpr long NET_BYTES = (128 * 16 + 2 * 128 * 128 + 16 * 128 + 2 * 128 * 128 + 128 * 16) * 2 + (3 * 128 + 16) * 4 +
                           TILES * 64 * 4 * 2;

// Forward on X [rows][16] (feature 15 zero) -> Y [rows][4].  With save, Z
// ([3][ld][128] fp32) and H ([16 + 3*128][ld] bf16, feature-major) are kept;
// ld >= rows rounded up to 16.
hipError_t forward(const Net& net, const float* X, int rows, float* Y, bool save, float* Z, unsigned short* H,
                   long ld, hipStream_t st);

// Backward of L = sum (y - t)^2 / (2 rows) after forward(save): writes the
// packed gradient and the loss (scratch: D [16 + 3*128][ld] bf16,
// dbpart [tiles][400], losspart [tiles]).
hipError_t backward(const Net& net, const float* Y, const float* T, int rows, const float* Z, const unsigned short* H,
                    unsigned short* D, long ld, float* dbpart, float* losspart, float* grad, float* loss,
                    hipStream_t st);

hipError_t adam(float* p, const float* g, float* m, float* v, int n, float lr, int t, hipStream_t st);

// Refresh the bf16 / bias copies of `net` from the fp32 master parameters p.
hipError_t publish(const float* p, const Net& net, hipStream_t st);

}  // namespace mlp
