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
// MFMA operand tiles of the forward weights: tile t holds, for lane
// l = r + 16 g, the 4 bf16 A-operand values W[16 m + r][16 kt + 4 g + j] of
// one 16 x 16 block (m, kt); tiles 0..7 are W1 (m), 8..71 W2 (8 + 8 m + kt),
// 72..135 W3, 136..143 W4 (kt).  Read linearly, 8 bytes per lane.
constexpr int TILES = 144, T_W1 = 0, T_W2 = 8, T_W3 = 72, T_W4 = 136;

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
