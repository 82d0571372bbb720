// Building blocks of the learned heat surrogate on gfx950: 3x3x3 convolutions as implicit GEMMs
// on the bf16 matrix cores (v_mfma_f32_16x16x16_bf16, fp32 accumulation), and the element-wise,
// pooling and optimizer kernels around them.
//
// Activations are channels-last bf16, [B][D][H][W][C] with C a multiple of 16; weights are
// [Cout][27][Cin] bf16 (tap t = (dz*3 + dy)*3 + dx, offsets dz-1, dy-1, dx-1), with fp32 master
// copies for the optimizer.  Padding is zero ("same" convolution).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace pm {
namespace nn {

using bf16 = uint16_t;

// Round-to-nearest-even fp32 -> bf16 and exact bf16 -> fp32 (finite values).
__host__ __device__ __forceinline__ bf16 to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(bf16 h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// Eight bf16 channels as one 16-byte word: element-wise kernels move channel groups of eight
// (every channel count and offset here is a multiple of 16).
__device__ __forceinline__ void unpack8(uint4 u, float f[8]) {
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        f[2 * q] = from_bf16((bf16)(w[q] & 0xFFFFu));
        f[2 * q + 1] = from_bf16((bf16)(w[q] >> 16));
    }
}
__device__ __forceinline__ uint4 pack8(const float f[8]) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) w[q] = (uint32_t)to_bf16(f[2 * q]) | ((uint32_t)to_bf16(f[2 * q + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// tanh-form GELU and its derivative, with tanh(u) written through the logistic function:
// 0.5 z (1 + tanh(u)) = z / (1 + e^(-2u)) and tanh(u) = 2 / (1 + e^(-2u)) - 1, using the hardware
// exp2 and reciprocal (v_exp_f32, v_rcp_f32, about 1 ulp each).  The results are rounded to bf16
// (2^-9) where they are stored, far above the difference from a correctly rounded tanh.
// The exponent -2u log2(e) = z (B + A z^2) takes one multiply and one fused multiply-add.
__device__ __forceinline__ float gelu_e2u(float z) {
    const float k2 = 2.0f * 0.7978845608028654f * 1.4426950408889634f;  // 2 sqrt(2/pi) log2(e)
    return __builtin_amdgcn_exp2f(z * __builtin_fmaf(z * z, -k2 * 0.044715f, -k2));
}
__device__ __forceinline__ float gelu(float z) {
    return z * __builtin_amdgcn_rcpf(1.0f + gelu_e2u(z));
}
__device__ __forceinline__ float gelu_grad(float z) {
    const float k = 0.7978845608028654f;
    const float t = 2.0f * __builtin_amdgcn_rcpf(1.0f + gelu_e2u(z)) - 1.0f;
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * k * (1.0f + 3.0f * 0.044715f * z * z);
}

// Shape of one convolution: the output has the input's spatial extent.  An M-tile is a box of
// TD x TH x TW output positions of one sample, at most 64 of them (fewer only when the whole
// volume is smaller); rows of the 64-row MFMA tile past the box are masked.
struct Conv {
    int B, D, H, W;
    int Cin, Cout;
    int TD, TH, TW;
    __host__ __device__ int64_t positions() const { return (int64_t)B * D * H * W; }
    __host__ __device__ int tile_positions() const { return TD * TH * TW; }
    __host__ __device__ int64_t tiles() const {
        return (int64_t)B * (D / TD) * (H / TH) * (W / TW);
    }
};

// Chooses the M-tile: as much of a row as fits, then rows, then planes.  Every extent here is a
// power of two, so the box divides the volume.
inline Conv make_conv(int B, int D, int H, int W, int Cin, int Cout) {
    Conv c{B, D, H, W, Cin, Cout, 1, 1, 64};
    c.TW = W < 64 ? W : 64;
    c.TH = H < 64 / c.TW ? H : 64 / c.TW;
    c.TD = D < 64 / (c.TW * c.TH) ? D : 64 / (c.TW * c.TH);
    return c;
}

// Host launchers (surrogate.hip).  All pointers are device pointers; `stream` orders them.
//
// y = conv(x, w) + bias; z (pre-activation) and a = act(z) stored as bf16 when non-null, with
// act = GELU when `gelu_out`, identity otherwise.  Output channel c is written at
// (position * ostride + ooff + c), so a convolution can write into a wider buffer.
//
// A grid too small to give 1024 workgroups with full-size MFMA blocks is split over the
// reduction (16-channel chunk x kernel plane): each split writes fp32 partials to `work`, and a
// reduction sums them in split order, so results do not depend on scheduling.  Without a
// `work` of at least conv_forward_workspace(g) floats the launch narrows its blocks instead.
//
// With `skip`, the input is the concatenation read in place: channels [0, cup) are x at half the
// resolution, upsampled by nearest neighbour, and channels [cup, Cin) are `skip` (full
// resolution, Cin - cup channels), so a decoder layer needs no concatenated buffer.
//
// `fuse` (optional) asks the epilogue for more from the activations it has staged: the 2x2x2
// average pool of a (channels-last, Cout channels, at half the resolution; the sum of the eight
// values in (dz, dy, dx) order times 0.125, rounded once), and the 1x1 head of a 16-channel
// layer (y[p] = head[16] + sum_c head[c] a[p][c], fp32, c in order).  The launch sets `fused`
// when it produced them, and then stores no `a` if `skip_a`; when it could not (a split launch,
// a box with odd extents, more than 16 channels for the head) it stores `a` and leaves them to
// the caller.
struct FwdFuse {
    bf16 *pool = nullptr;
    const float *head = nullptr;
    float *head_y = nullptr;
    bool skip_a = false;
    bool fused = false;
};
int conv_forward(const Conv &g, const bf16 *x, const bf16 *w, const float *bias, bf16 *z,
                 bf16 *a, bool gelu_out, int ostride, int ooff, hipStream_t stream,
                 float *work = nullptr, int64_t work_floats = 0, const bf16 *skip = nullptr,
                 int cup = 0, FwdFuse *fuse = nullptr);
// Floats of `work` conv_forward would use for this shape (0: it does not split).
int64_t conv_forward_workspace(const Conv &g);
// dx = conv(dy, flip(w)^T): the gradient of the input, fp32 accumulation, bf16 out.
int conv_backward_data(const Conv &g, const bf16 *dy, const bf16 *w, bf16 *w_flip_scratch,
                       bf16 *dx, hipStream_t stream, float *work = nullptr,
                       int64_t work_floats = 0);
// The transposed convolution conv_backward_data runs (Cin and Cout exchanged).
Conv transposed(const Conv &g);
// dw[co][t][ci] = sum_p dy[p][co] * x[p + off(t)][ci] (fp32), db[co] = sum_p dy[p][co];
// split over position tiles and reduced in a fixed order.  `scratch` holds
// splits * Cout * (27 * Cin + 1) floats.
int conv_backward_weight(const Conv &g, const bf16 *x, const bf16 *dy, float *dw, float *db,
                         float *scratch, int splits, hipStream_t stream);

}  // namespace nn
}  // namespace pm
