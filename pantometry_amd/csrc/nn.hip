// Implicit-GEMM 3x3x3 convolutions on the gfx950 bf16 matrix cores, and their gradients.
//
// Forward: a workgroup of four waves computes a 64-position M-tile (a TD x TH x TW box of one
// sample) against an N-block of up to 64 output channels.  For each 16-channel chunk of the
// input, the box's input neighbourhood ((TD+2) x (TH+2) x (TW+2) positions, zero outside the
// volume) is staged once in LDS, then every wave runs the 27 taps as MFMA k-steps: the A fragment
// (16 positions x 16 channels) comes from LDS at the tap's shift, the B fragment (16 channels x 16
// output channels) straight from the weight tensor.  One staged value feeds 27 taps.
//
// MFMA v_mfma_f32_16x16x16_bf16 lane layout (checked on the device): lane l holds A[l%16][4(l/16)
// + j], B[4(l/16) + j][l%16] for j < 4, and C[4(l/16) + r][l%16] for r < 4.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "nn.hpp"

namespace pm {
namespace nn {

typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// gfx950's 16x16x32 bf16 MFMA: lane l holds A[l%16][8(l/16)+j] and B[8(l/16)+j][l%16] (j < 8)
// and receives C[4(l/16)+r][l%16] (r < 4), the layout of the 16x16x16 form with K doubled.
__device__ __forceinline__ float4v mfma32(uint4 a, uint4 b, float4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

__device__ __forceinline__ float4v mfma16(short4v a, short4v b, float4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Decomposes an M-tile index into the box origin (b, d0, h0, w0).
__device__ __forceinline__ void tile_origin(const Conv &g, int64_t tile, int &b, int &d0, int &h0,
                                            int &w0) {
    const int tw = g.W / g.TW, th = g.H / g.TH, td = g.D / g.TD;
    int64_t t = tile;
    const int wt = (int)(t % tw);
    t /= tw;
    const int ht = (int)(t % th);
    t /= th;
    const int dt = (int)(t % td);
    b = (int)(t / td);
    d0 = dt * g.TD;
    h0 = ht * g.TH;
    w0 = wt * g.TW;
}

// Stages channels [c0, c0+16) of the box's neighbourhood in x (channel stride cs) into `s`:
// s[((pz * (TH+2) + py) * (TW+2) + px) * 16 + c], zero outside the volume.
__device__ __forceinline__ void stage_neighbourhood(const Conv &g, const bf16 *__restrict__ x,
                                                   int cs, int c0, int b, int d0, int h0, int w0,
                                                   bf16 *s, int tid, int nthreads) {
    const int SH = g.TH + 2, SW = g.TW + 2;
    const int npos = (g.TD + 2) * SH * SW;
    // Two 16-byte halves per position.
    for (int e = tid; e < npos * 2; e += nthreads) {
        const int half = e & 1, p = e >> 1;
        const int px = p % SW, py = (p / SW) % SH, pz = p / (SW * SH);
        const int z = d0 + pz - 1, y = h0 + py - 1, xx = w0 + px - 1;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (z >= 0 && z < g.D && y >= 0 && y < g.H && xx >= 0 && xx < g.W) {
            const int64_t pos = (((int64_t)b * g.D + z) * g.H + y) * g.W + xx;
            v = *(const uint4 *)(x + pos * cs + c0 + half * 8);
        }
        *(uint4 *)(s + p * 16 + half * 8) = v;
    }
}

// Box coordinates of tile position q (q < tile_positions()); the box's extents are powers of
// two (checked by the launchers), so these are shifts and masks.
__device__ __forceinline__ void box_coords(const Conv &g, int q, int &qd, int &qh, int &qw) {
    const int lw = __builtin_ctz(g.TW), lh = __builtin_ctz(g.TH);
    qw = q & (g.TW - 1);
    qh = (q >> lw) & (g.TH - 1);
    qd = q >> (lw + lh);
}

// q / d for 0 <= q < 2^20 and 0 < d < 2^10, from the float reciprocal `inv` = 1/d and one
// correction step (the float product is within one unit of the quotient).
__device__ __forceinline__ int div_small(int q, int d, float inv) {
    int r = (int)((float)q * inv);
    const int rem = q - r * d;
    r += rem >= d ? 1 : 0;
    r -= rem < 0 ? 1 : 0;
    return r;
}

// The forward kernel.  A workgroup of four waves computes a 64*RB-position M-tile (a TD x TH x TW
// box of one sample, smaller only when the volume is) against an N-block of 16*NT output
// channels; each wave owns 16*RB positions as RB 16-row MFMA blocks, so every B fragment (16
// channels x 16 output channels of one tap, read from the weights) feeds RB MFMAs and every A
// fragment (16 positions x 16 channels, read from LDS at the tap's shift) feeds NT.  Per
// 16-channel input chunk the box's neighbourhood is staged in LDS once, row by row; the 27 taps
// are unrolled so each tap's offset in the staged box is a constant plus the rows' strides.  The
// epilogue adds the bias, applies GELU, and writes bf16 through LDS so each position's channels
// leave in 16-byte stores.
//
// The input may be a concatenation read in place: with cup > 0, channels [0, cup) are x at half
// the resolution (nearest-neighbour upsampling) and channels [cup, Cin) are xs at full
// resolution; with cup = 0 all Cin channels are xs.
template <int NT, int RB, int SK>
__global__ __launch_bounds__(256, NT == 1 ? 4 : 2) void conv_fwd2(Conv g, const bf16 *__restrict__ x,
                                                 const bf16 *__restrict__ xs, int cup,
                                                 const bf16 *__restrict__ w,
                                                 const float *__restrict__ bias,
                                                 bf16 *__restrict__ zout,
                                                 bf16 *__restrict__ aout, int gelu_out,
                                                 int ostride, int ooff,
                                                 float *__restrict__ part,
                                                 bf16 *__restrict__ pool_out,
                                                 const float *__restrict__ head_w,
                                                 float *__restrict__ head_y) {
    constexpr int NB = 16 * NT;
    extern __shared__ __align__(16) bf16 smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int b, d0, h0, w0;
    tile_origin(g, blockIdx.x, b, d0, h0, w0);
    const int n0 = blockIdx.y * NB;
    const int SH = g.TH + 2, SW = g.TW + 2, SHW = SH * SW;
    const int TP = g.tile_positions();
    // The staged neighbourhood is two planes (channels 0-7 and 8-15 of the chunk) of 16-byte
    // positions, so the 16 lanes of a row group read 256 contiguous bytes.  Lane group
    // lg = lane/16 takes tap 2p + lg/2 of a tap pair p and channel half lg%2.
    const int lg = lane >> 4, qh = lg >> 1, qt = lg & 1;
    const int npos = (g.TD + 2) * SHW;
    const uint4 *sA = (const uint4 *)smem + qh * npos;

    // This lane's RB A rows: positions wave*16*RB + rb*16 + lane%16 of the box.
    int base[RB];
    bool live[RB];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int p = wave * 16 * RB + rb * 16 + (lane & 15);
        live[rb] = p < TP;
        int pd, ph, pw;
        box_coords(g, live[rb] ? p : TP - 1, pd, ph, pw);
        base[rb] = (pd * SH + ph) * SW + pw;
    }

    float4v acc[RB][NT];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int n = 0; n < NT; n++) acc[rb][n] = float4v{0.f, 0.f, 0.f, 0.f};

    const int64_t wrow = (int64_t)27 * g.Cin;
    const bf16 *wl = w + (int64_t)(n0 + (lane & 15)) * wrow + qh * 8;  // this lane's B column
    // The reduction runs over units (16-channel chunk, kernel plane dz); split z of gridDim.z
    // takes units [u0, u1).
    const int units = g.Cin / 16 * 3;
    const int u0 = (int)((int64_t)blockIdx.z * units / gridDim.z);
    const int u1 = (int)((int64_t)(blockIdx.z + 1) * units / gridDim.z);

    // Staging of a chunk: element e = tid + 256 k (k < SK) of the 2 npos 16-byte halves is half
    // e % 2 of staged position e / 2, so consecutive lanes read consecutive 16 bytes of the
    // channels-last input.  All of a thread's loads are issued before any is waited on, and the
    // next chunk's loads are issued before the current chunk's MFMAs.
    int dst[SK];  // LDS index of element k; past the neighbourhood, a spare slot after it
    int src[SK];  // its offset in xs (channel cup of the input); position 0 outside the volume
    int srcu[SK]; // its offset in the half-resolution x (channel 0), when cup > 0
    bool in[SK];  // inside the volume (else the loaded value is replaced by zero)
    const int cs = g.Cin - cup;  // channels of xs
    const float inv_sw = 1.0f / (float)SW, inv_sh = 1.0f / (float)SH;
    // Offsets are 32-bit (the launcher checks that every tensor's elements fit), and the tile's
    // own part of them is uniform: the box origin's position index, times the channel count, on
    // the scalar unit; each element adds its small signed displacement in 24-bit products.
    const int HW = g.H * g.W;
    const int pbase = ((b * g.D + d0) * g.H + h0) * g.W + w0;
    const int sbase = pbase * cs;
#pragma unroll
    for (int k = 0; k < SK; k++) {
        const int e = tid + 256 * k;
        const int p = e >> 1, half = e & 1;
        const int t = div_small(p, SW, inv_sw), pz = div_small(t, SH, inv_sh);
        const int px = p - t * SW, py = t - pz * SH;
        const int z = d0 + pz - 1, y = h0 + py - 1, xx = w0 + px - 1;
        const bool staged_here = e < 2 * npos;
        in[k] = staged_here && z >= 0 && z < g.D && y >= 0 && y < g.H && xx >= 0 && xx < g.W;
        dst[k] = staged_here ? half * npos + p : 2 * npos;
        const int rel = __mul24(pz - 1, HW) + __mul24(py - 1, g.W) + (px - 1);
        src[k] = in[k] ? sbase + __mul24(rel, cs) + half * 8 : 0;
        srcu[k] = 0;
    }
    if (cup > 0) {
        const unsigned Dh = g.D >> 1, Hh = g.H >> 1, Wh = g.W >> 1;
#pragma unroll
        for (int k = 0; k < SK; k++) {
            const int e = tid + 256 * k;
            const int p = e >> 1, half = e & 1;
            const int t = div_small(p, SW, inv_sw), pz = div_small(t, SH, inv_sh);
            const int px = p - t * SW, py = t - pz * SH;
            const unsigned z = (unsigned)(d0 + pz - 1) >> 1, y = (unsigned)(h0 + py - 1) >> 1;
            const unsigned xx = (unsigned)(w0 + px - 1) >> 1;
            const unsigned ph = __umul24(__umul24(__umul24((unsigned)b, Dh) + z, Hh) + y, Wh) + xx;
            srcu[k] = in[k] ? (int)(__umul24(ph, (unsigned)cup) + half * 8) : 0;
        }
    }
    // With one or two N-blocks the chunk's weights go through the same pipeline into LDS after
    // the neighbourhood: slot (tap * 2 + half) * NB + column holds input channels half*8 .. +8
    // of output channel n0 + column, so the loop reads B from LDS like A and waits on no global
    // load.  Element eb = tid + 256 k is that slot; past 54 NB the loads repeat slot 0's source.
    constexpr bool WB = NT <= 2;
    constexpr int SKB = WB ? (54 * NB + 255) / 256 : 1;
    const int aslots = 2 * npos + 1;
    // The prefetch registers: the neighbourhood's SK elements, then the weights' SKB.
    uint4 pre[SK + (WB ? SKB : 0)];
    // Source offset of weight element k (channel 0 of the chunk).
    auto bsrc = [&](int k) -> int {
        const int eb = tid + 256 * k;
        const int col = eb % NB, rest = eb / NB;
        return eb < 54 * NB ? (n0 + col) * (int)wrow + (rest >> 1) * g.Cin + (rest & 1) * 8 : 0;
    };
    const uint4 *sB = (const uint4 *)smem + aslots + qh * NB + (lane & 15);
    auto load_chunk = [&](int c0) {
        if (c0 < cup) {
#pragma unroll
            for (int k = 0; k < SK; k++) pre[k] = *(const uint4 *)(x + srcu[k] + c0);
        } else {
#pragma unroll
            for (int k = 0; k < SK; k++) pre[k] = *(const uint4 *)(xs + src[k] + c0 - cup);
        }
        if constexpr (WB) {
#pragma unroll
            for (int k = 0; k < SKB; k++) pre[SK + k] = *(const uint4 *)(w + bsrc(k) + c0);
        }
    };
    auto store_chunk = [&]() {
#pragma unroll
        for (int k = 0; k < SK; k++) {
            const uint4 v = pre[k];
            ((uint4 *)smem)[dst[k]] = make_uint4(in[k] ? v.x : 0u, in[k] ? v.y : 0u,
                                                 in[k] ? v.z : 0u, in[k] ? v.w : 0u);
        }
        if constexpr (WB) {
#pragma unroll
            for (int k = 0; k < SKB; k++) ((uint4 *)smem)[aslots + tid + 256 * k] = pre[SK + k];
        }
    };

    int staged = -1;
    if (u0 < u1) load_chunk((u0 / 3) * 16);
    for (int u = u0; u < u1; u++) {
        const int c0 = (u / 3) * 16, dz = u - (u / 3) * 3;
        if (c0 != staged) {
            staged = c0;
            __syncthreads();
            store_chunk();
            __syncthreads();
            // The next chunk this workgroup reduces over, in flight during this one's MFMAs.
            const int un = (u / 3 + 1) * 3;
            if (un < u1) load_chunk((un / 3) * 16);
        }
#pragma unroll
        for (int pr = 0; pr < 5; pr++) {
            // Taps 2pr and 2pr+1 of this kernel plane in one K=32 step; the tenth is padding: it
            // reads the ninth tap's positions against a zero B fragment, so every load is
            // unpredicated (rows past a small box read a clamped position and are never stored).
            const int mt = 2 * pr + qt;
            const bool tv = mt < 9;
            const int mc = tv ? mt : 8;
            const int toff = dz * SHW + (mc / 3) * SW + (mc % 3);
            uint4 bv[NT];
#pragma unroll
            for (int n = 0; n < NT; n++) {
                bv[n] = WB ? sB[((dz * 9 + mc) * 2) * NB + n * 16]
                           : *(const uint4 *)(wl + (int64_t)n * 16 * wrow + (dz * 9 + mc) * g.Cin + c0);
                if (!tv) bv[n] = make_uint4(0, 0, 0, 0);
            }
            uint4 a[RB];
#pragma unroll
            for (int rb = 0; rb < RB; rb++) a[rb] = sA[base[rb] + toff];
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
#pragma unroll
                for (int n = 0; n < NT; n++) acc[rb][n] = mfma32(a[rb], bv[n], acc[rb][n]);
        }
    }

    if (part) {
        // Split: fp32 partials [split][position][Cout], bias and activation left to the reduction.
        float *pz = part + (int64_t)blockIdx.z * g.positions() * g.Cout;
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int q = wave * 16 * RB + rb * 16 + (lane >> 4) * 4 + r;
                if (q >= TP) continue;
                int qd, qh, qw;
                box_coords(g, q, qd, qh, qw);
                const int pos = pbase + __mul24(qd, HW) + __mul24(qh, g.W) + qw;
#pragma unroll
                for (int n = 0; n < NT; n++) pz[pos * g.Cout + n0 + n * 16 + (lane & 15)] = acc[rb][n][r];
            }
        return;
    }

    // Epilogue through LDS: [position][NB + 8] bf16 (the pad spreads the four lane groups' rows
    // over all banks), then 16-byte stores per position.  A lane holds rows q0..q0+3 of one
    // channel; neighbouring lanes (channels c, c^1) swap packed pairs so that each writes two
    // 32-bit words of a channel pair, rows q0, q0+1 from the even lane and q0+2, q0+3 from the
    // odd one.
    constexpr int NS = NB + 8;
    float bb[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) bb[n] = bias ? bias[n0 + n * 16 + (lane & 15)] : 0.0f;
    bf16 *outs[2] = {zout, aout};
    const bool odd = lane & 1;
    const int cpair = (lane & 15) & ~1;
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
        bf16 *out = outs[pass];
        // The activations are staged for the pool and the head as well as for their own store.
        if (!out && (pass == 0 || (!pool_out && !head_y))) continue;
        __syncthreads();
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int n = 0; n < NT; n++) {
                uint32_t h[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float zz = acc[rb][n][r] + bb[n];
                    h[r] = to_bf16(pass == 1 && gelu_out ? gelu(zz) : zz);
                }
                const uint32_t mine = odd ? (h[2] | h[3] << 16) : (h[0] | h[1] << 16);
                const uint32_t send = odd ? (h[0] | h[1] << 16) : (h[2] | h[3] << 16);
                // quad_perm [1, 0, 3, 2]: each lane receives its neighbour's `send`.
                const uint32_t got = (uint32_t)__builtin_amdgcn_mov_dpp((int)send, 0xB1, 0xF, 0xF, false);
                const uint32_t lo = odd ? (got & 0xFFFFu) | (mine << 16) : (mine & 0xFFFFu) | (got << 16);
                const uint32_t hi = odd ? (got >> 16) | (mine & 0xFFFF0000u) : (mine >> 16) | (got & 0xFFFF0000u);
                const int q = wave * 16 * RB + rb * 16 + (lane >> 4) * 4 + (odd ? 2 : 0);
                uint32_t *row = (uint32_t *)(smem + q * NS + n * 16 + cpair);
                row[0] = lo;
                row[NS / 2] = hi;
            }
        __syncthreads();
        if (out) {
            for (int e = tid; e < TP * (NB / 8); e += 256) {
                const int q = e / (NB / 8), part = e % (NB / 8);
                int qd, qh, qw;
                box_coords(g, q, qd, qh, qw);
                const int pos = pbase + __mul24(qd, HW) + __mul24(qh, g.W) + qw;
                *(uint4 *)(out + pos * ostride + ooff + n0 + part * 8) = *(const uint4 *)(smem + q * NS + part * 8);
            }
        }
        if (pass == 1 && pool_out) {
            // 2x2x2 means of the staged box (even extents, checked by the launcher): item e is
            // channel group e % (NB/8) of coarse position e / (NB/8), the eight fine values
            // summed in (dz, dy, dx) order.
            const int PW = g.TW >> 1, PH = g.TH >> 1, PP = (g.TD >> 1) * PH * PW;
            const int lpw = __builtin_ctz(PW), lph = __builtin_ctz(PH);
            const int Dh = g.D >> 1, Hh = g.H >> 1, Wh = g.W >> 1;
            for (int e = tid; e < PP * (NB / 8); e += 256) {
                const int pq = e / (NB / 8), grp = e % (NB / 8);
                const int pw = pq & (PW - 1), ph = (pq >> lpw) & (PH - 1), pd = pq >> (lpw + lph);
                float sum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int q = ((2 * pd + (k >> 2)) * g.TH + 2 * ph + ((k >> 1) & 1)) * g.TW + 2 * pw + (k & 1);
                    float f[8];
                    unpack8(*(const uint4 *)(smem + q * NS + grp * 8), f);
#pragma unroll
                    for (int c = 0; c < 8; c++) sum[c] += f[c];
                }
#pragma unroll
                for (int c = 0; c < 8; c++) sum[c] *= 0.125f;
                const int ppos = ((b * Dh + (d0 >> 1) + pd) * Hh + (h0 >> 1) + ph) * Wh + (w0 >> 1) + pw;
                *(uint4 *)(pool_out + ppos * g.Cout + n0 + grp * 8) = pack8(sum);
            }
        }
        if (pass == 1 && head_y) {
            // The 1x1 head of a 16-channel layer: one position per thread, channels in order.
            for (int q = tid; q < TP; q += 256) {
                float sum = head_w[16];
#pragma unroll
                for (int c = 0; c < 16; c++) sum += head_w[c] * from_bf16(smem[q * NS + c]);
                int qd, qh, qw;
                box_coords(g, q, qd, qh, qw);
                head_y[pbase + __mul24(qd, HW) + __mul24(qh, g.W) + qw] = sum;
            }
        }
    }
}

// W'[ci][t][co] = W[co][26 - t][ci]: the data gradient is the forward convolution of dy with the
// flipped, transposed weights.
__global__ void flip_weights(const bf16 *__restrict__ w, bf16 *__restrict__ wf, int Cin, int Cout) {
    const int64_t n = (int64_t)Cout * 27 * Cin;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ci = (int)(i % Cin);
        const int t = (int)((i / Cin) % 27);
        const int co = (int)(i / ((int64_t)Cin * 27));
        wf[((int64_t)ci * 27 + (26 - t)) * Cout + co] = w[i];
    }
}

// Weight gradient for a 16 (co) x 16 (ci) block over all 27 taps: wave w owns taps w, w+4, ...
// The workgroup walks the M-tiles tile = split, split + splits, ...; per tile it stages the input
// neighbourhood (ci block) and dy (co block) in LDS, then each wave runs one k-step per 16
// positions of the box per tap.  Partials go to part[split][co][t][ci]; the workgroups of the
// first ci block also sum dy over their tiles for the bias, into bpart[split][co].
__global__ __launch_bounds__(256) void conv_wgrad_kernel(Conv g, const bf16 *__restrict__ x,
                                                         const bf16 *__restrict__ dy,
                                                         float *__restrict__ part,
                                                         float *__restrict__ bpart, int splits) {
    extern __shared__ __align__(16) bf16 smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ci0 = blockIdx.x * 16, co0 = blockIdx.y * 16, split = blockIdx.z;
    const int SH = g.TH + 2, SW = g.TW + 2;
    const int TP = g.tile_positions();
    const int KS = (TP + 15) / 16;
    bf16 *sx = smem;                                  // neighbourhood, 16 channels per position
    bf16 *sd = smem + (g.TD + 2) * SH * SW * 16;      // dy box, [64 positions][16 co]
    float4v acc[7];
#pragma unroll
    for (int i = 0; i < 7; i++) acc[i] = float4v{0.f, 0.f, 0.f, 0.f};
    const int kq = (lane >> 4) * 4;
    const int64_t ntiles = g.tiles();
    const bool bias_lane = blockIdx.x == 0 && tid < 16;
    float bsum = 0.0f;
    for (int64_t tile = split; tile < ntiles; tile += splits) {
        int b, d0, h0, w0;
        tile_origin(g, tile, b, d0, h0, w0);
        __syncthreads();
        stage_neighbourhood(g, x, g.Cin, ci0, b, d0, h0, w0, sx, tid, 256);
        for (int e = tid; e < 64 * 2; e += 256) {
            const int half = e & 1, q = e >> 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (q < TP) {
                int qd, qh, qw;
                box_coords(g, q, qd, qh, qw);
                const int64_t pos = (((int64_t)b * g.D + d0 + qd) * g.H + h0 + qh) * g.W + w0 + qw;
                v = *(const uint4 *)(dy + pos * g.Cout + co0 + half * 8);
            }
            *(uint4 *)(sd + q * 16 + half * 8) = v;
        }
        __syncthreads();
        if (bias_lane)
            for (int q = 0; q < TP; q++) bsum += from_bf16(sd[q * 16 + tid]);
#pragma unroll
        for (int i = 0; i < 7; i++) {
            const int t = wave + 4 * i;
            if (t < 27) {
                const int dz = t / 9, dyy = (t / 3) % 3, dx = t % 3;
                for (int ks = 0; ks < KS; ks++) {
                    short4v av, bv;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int q = ks * 16 + kq + j;  // position within the box
                        int qd, qh, qw;
                        box_coords(g, q < TP ? q : 0, qd, qh, qw);
                        av[j] = (short)sd[q * 16 + (lane & 15)];  // zero past the box
                        const int sp = ((qd + dz) * SH + qh + dyy) * SW + qw + dx;
                        bv[j] = (short)sx[sp * 16 + (lane & 15)];
                    }
                    acc[i] = mfma16(av, bv, acc[i]);
                }
            }
        }
    }
    // C[row = co][col = ci]: lane holds rows 4*(lane/16) + r, column lane%16.
#pragma unroll
    for (int i = 0; i < 7; i++) {
        const int t = wave + 4 * i;
        if (t >= 27) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int co = co0 + (lane >> 4) * 4 + r, ci = ci0 + (lane & 15);
            part[(((int64_t)split * g.Cout + co) * 27 + t) * g.Cin + ci] = acc[i][r];
        }
    }
    if (bias_lane) bpart[(int64_t)split * g.Cout + co0 + tid] = bsum;
}

// dw[i] = sum over splits of part[s][i], in split order.
__global__ void reduce_splits(const float *__restrict__ part, float *__restrict__ out, int64_t n,
                              int splits) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int k = 0; k < splits; k++) s += part[k * n + i];
        out[i] = s;
    }
}

static unsigned blocks_for(int64_t n, int threads) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + threads - 1) / threads, 1 << 20));
}

// The kernels' assumptions: channel counts in whole 16-channel blocks, a box that tiles the
// volume with at most 64 positions, and a box of 64 unless it is the whole sample.
static bool conv_shape_ok(const Conv &g) {
    const bool ok = g.B > 0 && g.Cin > 0 && g.Cout > 0 && g.Cin % 16 == 0 && g.Cout % 16 == 0 &&
                    g.TD > 0 && g.TH > 0 && g.TW > 0 && g.D % g.TD == 0 && g.H % g.TH == 0 &&
                    g.W % g.TW == 0 && g.tile_positions() <= 64 &&
                    (g.tile_positions() == 64 || (g.TD == g.D && g.TH == g.H && g.TW == g.W)) &&
                    (g.TW & (g.TW - 1)) == 0 && (g.TH & (g.TH - 1)) == 0;
    if (!ok) set_error("convolution shape %dx%dx%dx%d, %d->%d channels, box %dx%dx%d unsupported",
                       g.B, g.D, g.H, g.W, g.Cin, g.Cout, g.TD, g.TH, g.TW);
    return ok;
}

// The forward kernel's M-tile: a box of up to m positions (256 or 64), at most 16 wide and 4
// deep, as tall as fills it (the whole sample when the volume is smaller).
static Conv forward_box(const Conv &g, int m) {
    Conv f = g;
    f.TW = g.W < 16 ? g.W : 16;
    f.TD = g.D < 4 ? g.D : 4;
    if (f.TW * f.TD > m) f.TD = m / f.TW;
    f.TH = g.H < m / (f.TW * f.TD) ? g.H : m / (f.TW * f.TD);
    f.TD = g.D < m / (f.TW * f.TH) ? g.D : m / (f.TW * f.TH);
    return f;
}

template <int NT, int RB>
static void launch_fwd(const Conv &g, dim3 grid, size_t lds, hipStream_t stream, const bf16 *x,
                       const bf16 *xs, int cup,
                       const bf16 *w, const float *bias, bf16 *z, bf16 *a, bool gelu_out,
                       int ostride, int ooff, float *part, bf16 *pool, const float *head,
                       float *head_y) {
    // Staged 16-byte halves per thread: 2 (TD+2)(TH+2)(TW+2) over 256 threads.
    const int halves = 2 * (g.TD + 2) * (g.TH + 2) * (g.TW + 2);
    const int sk = (halves + 255) / 256;
    if (sk <= 2)
        conv_fwd2<NT, RB, 2><<<grid, 256, lds, stream>>>(g, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y);
    else if (sk <= 4)
        conv_fwd2<NT, RB, 4><<<grid, 256, lds, stream>>>(g, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y);
    else if (sk <= 6)
        conv_fwd2<NT, RB, 6><<<grid, 256, lds, stream>>>(g, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y);
    else
        conv_fwd2<NT, RB, 8><<<grid, 256, lds, stream>>>(g, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y);
}

// Sums the split partials of conv_fwd2 in split order, adds the bias and writes z and act(z) as
// bf16, eight channels (one 16-byte store) per thread.
__global__ void conv_split_reduce(const float *__restrict__ part, int splits, int64_t P, int Cout,
                                  const float *__restrict__ bias, bf16 *__restrict__ zout,
                                  bf16 *__restrict__ aout, int gelu_out, int ostride, int ooff) {
    const int G = Cout / 8;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P * G;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pos = i / G;
        const int c = (int)(i - pos * G) * 8;
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = 0.0f;
        for (int z = 0; z < splits; z++) {
            const float *src = part + ((int64_t)z * P + pos) * Cout + c;
            const float4 lo = *(const float4 *)src, hi = *(const float4 *)(src + 4);
            v[0] += lo.x, v[1] += lo.y, v[2] += lo.z, v[3] += lo.w;
            v[4] += hi.x, v[5] += hi.y, v[6] += hi.z, v[7] += hi.w;
        }
        if (bias) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] += bias[c + k];
        }
        bf16 zb[8], ab[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            zb[k] = to_bf16(v[k]);
            ab[k] = to_bf16(gelu_out ? gelu(v[k]) : v[k]);
        }
        if (zout) *(uint4 *)(zout + pos * ostride + ooff + c) = *(const uint4 *)zb;
        if (aout) *(uint4 *)(aout + pos * ostride + ooff + c) = *(const uint4 *)ab;
    }
}

struct FwdPlan {
    int rb, nt, splits;
};

// Most reuse first: 256-position boxes and the widest N-block.  While that launch would have
// fewer than 1024 workgroups, split it over the reduction (when a workspace may be used), else
// fall back to narrower N-blocks and 64-position boxes.
static FwdPlan plan_fwd(const Conv &g0, bool can_split) {
    const int nt_max = g0.Cout % 32 == 0 ? 2 : 1;
    const Conv f = forward_box(g0, 256);
    if (f.TW > 0 && f.TH > 0 && f.TD > 0) {
        const int64_t base = f.tiles() * (g0.Cout / (16 * nt_max));
        if (base >= 1024) return {4, nt_max, 1};
        const int units = g0.Cin / 16 * 3;
        const int splits = (int)std::min<int64_t>(units, (1024 + base - 1) / base);
        if (can_split && splits > 1) return {4, nt_max, splits};
    }
    const int opts[4][2] = {{4, nt_max}, {4, 1}, {1, nt_max}, {1, 1}};
    FwdPlan p{4, nt_max, 1};
    int64_t best = -1;
    for (const auto &o : opts) {
        const Conv c = forward_box(g0, 64 * o[0]);
        if (c.TW <= 0 || c.TH <= 0 || c.TD <= 0) continue;
        const int64_t wgs = c.tiles() * (g0.Cout / (16 * o[1]));
        if (wgs > best) best = wgs, p = {o[0], o[1], 1};
        if (wgs >= 1024) return {o[0], o[1], 1};
    }
    return p;
}

int64_t conv_forward_workspace(const Conv &g0) {
    const FwdPlan p = plan_fwd(g0, true);
    return p.splits > 1 ? (int64_t)p.splits * g0.positions() * g0.Cout : 0;
}

Conv transposed(const Conv &g) {
    Conv t = g;
    t.Cin = g.Cout;
    t.Cout = g.Cin;
    return t;
}

int conv_forward(const Conv &g0, const bf16 *x, const bf16 *w, const float *bias, bf16 *z, bf16 *a,
                 bool gelu_out, int ostride, int ooff, hipStream_t stream, float *work,
                 int64_t work_floats, const bf16 *skip, int cup, FwdFuse *fuse) {
    // Without a skip tensor the whole input is x; with one, x is the half-resolution part.
    const bf16 *xs = skip ? skip : x;
    if (!skip) cup = 0;
    if (skip && (cup <= 0 || cup % 16 || cup >= g0.Cin || g0.D % 2 || g0.H % 2 || g0.W % 2)) {
        set_error("conv_forward: an in-place concatenation needs cup in (0, Cin), a multiple of "
                  "16, and even extents");
        return PM_ERR_ARGUMENT;
    }
    FwdPlan plan = plan_fwd(g0, true);
    if (plan.splits > 1 && (!work || work_floats < (int64_t)plan.splits * g0.positions() * g0.Cout))
        plan = plan_fwd(g0, false);
    const int rb = plan.rb, nt = plan.nt;
    float *part = plan.splits > 1 ? work : nullptr;
    const int m = 64 * rb;
    const Conv g = forward_box(g0, m);
    const bool ok = g.B > 0 && g.Cin > 0 && g.Cout > 0 && g.Cin % 16 == 0 && g.Cout % 16 == 0 &&
                    g.TW > 0 && g.TH > 0 && g.TD > 0 && g.D % g.TD == 0 && g.H % g.TH == 0 &&
                    g.W % g.TW == 0 && g.tile_positions() <= m &&
                    (g.tile_positions() == m || (g.TD == g.D && g.TH == g.H && g.TW == g.W)) &&
                    ostride % 8 == 0 && ooff % 8 == 0 &&
                    2 * (g.TD + 2) * (g.TH + 2) * (g.TW + 2) <= 256 * 8 &&
                    (g.TW & (g.TW - 1)) == 0 && (g.TH & (g.TH - 1)) == 0 &&
                    g.positions() * std::max({g.Cin, g.Cout, ostride}) < ((int64_t)1 << 31);
    if (!ok) {
        set_error("conv_forward: shape %dx%dx%dx%d, %d->%d channels unsupported", g.B, g.D, g.H,
                  g.W, g.Cin, g.Cout);
        return PM_ERR_ARGUMENT;
    }
    const size_t stage = ((size_t)2 * (g.TD + 2) * (g.TH + 2) * (g.TW + 2) + 1) * 16 +
                         (nt <= 2 ? (size_t)((54 * 16 * nt + 255) / 256) * 256 * 16 : 0);
    const size_t epi = (size_t)m * (16 * nt + 8) * sizeof(bf16);
    const size_t lds = std::max(stage, epi);
    dim3 grid((unsigned)g.tiles(), (unsigned)(g.Cout / (16 * nt)), (unsigned)plan.splits);
    // The pool and the head in the epilogue: a launch that is not split, over boxes of even
    // extents (pool) or one 16-channel N-block after GELU (head).
    bf16 *pool = nullptr;
    const float *head = nullptr;
    float *head_y = nullptr;
    if (fuse) {
        fuse->fused = false;
        const bool pool_ok = !fuse->pool || (g.TD % 2 == 0 && g.TH % 2 == 0 && g.TW % 2 == 0 &&
                                             ooff == 0 && ostride == g.Cout);
        const bool head_ok = !fuse->head_y || (g.Cout == 16 && gelu_out && fuse->head);
        if (!part && pool_ok && head_ok && (fuse->pool || fuse->head_y)) {
            pool = fuse->pool, head = fuse->head, head_y = fuse->head_y;
            fuse->fused = true;
            if (fuse->skip_a) a = nullptr;
        }
    }
    const int key = nt * 10 + rb;
    switch (key) {
        case 14: launch_fwd<1, 4>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        case 24: launch_fwd<2, 4>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        case 34: launch_fwd<3, 4>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        case 44: launch_fwd<4, 4>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        case 11: launch_fwd<1, 1>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        case 21: launch_fwd<2, 1>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        case 31: launch_fwd<3, 1>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        default: launch_fwd<4, 1>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
    }
    PM_HIP(hipGetLastError());
    if (part) {
        const int64_t items = g.positions() * (g.Cout / 8);
        conv_split_reduce<<<blocks_for(items, 256), 256, 0, stream>>>(
            part, plan.splits, g.positions(), g.Cout, bias, z, a, gelu_out, ostride, ooff);
        PM_HIP(hipGetLastError());
    }
    return PM_OK;
}

int conv_backward_data(const Conv &g, const bf16 *dy, const bf16 *w, bf16 *wf, bf16 *dx,
                       hipStream_t stream, float *work, int64_t work_floats) {
    const int64_t n = (int64_t)g.Cout * 27 * g.Cin;
    flip_weights<<<blocks_for(n, 256), 256, 0, stream>>>(w, wf, g.Cin, g.Cout);
    PM_HIP(hipGetLastError());
    const Conv t = transposed(g);
    return conv_forward(t, dy, wf, nullptr, nullptr, dx, false, t.Cout, 0, stream, work, work_floats);
}

int conv_backward_weight(const Conv &g, const bf16 *x, const bf16 *dy, float *dw, float *db,
                         float *scratch, int splits, hipStream_t stream) {
    if (!conv_shape_ok(g) || splits < 1 || !db) return PM_ERR_ARGUMENT;
    const size_t lds = ((size_t)(g.TD + 2) * (g.TH + 2) * (g.TW + 2) * 16 + 64 * 16) * sizeof(bf16);
    const int64_t n = (int64_t)g.Cout * 27 * g.Cin;
    float *bpart = scratch + (int64_t)splits * n;
    dim3 grid((unsigned)(g.Cin / 16), (unsigned)(g.Cout / 16), (unsigned)splits);
    conv_wgrad_kernel<<<grid, 256, lds, stream>>>(g, x, dy, scratch, bpart, splits);
    PM_HIP(hipGetLastError());
    reduce_splits<<<blocks_for(n, 256), 256, 0, stream>>>(scratch, dw, n, splits);
    reduce_splits<<<blocks_for(g.Cout, 256), 256, 0, stream>>>(bpart, db, g.Cout, splits);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

}  // namespace nn
}  // namespace pm
