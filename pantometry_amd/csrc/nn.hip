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
// channels; each wave owns 16*RB positions as RB 16-position MFMA blocks.  The MFMA takes the
// weights as its A operand (16 output channels x 32 k) and the staged activations as B (32 k x 16
// positions), so every A fragment (read from LDS) feeds RB MFMAs, every B fragment (read from
// LDS at the tap's shift) feeds NT, and a lane receives four consecutive output channels of one
// position: the accumulators start from the bias, and the epilogue applies GELU, converts pairs
// to bf16 in one instruction and stores 8 bytes per lane straight from registers.  Per 16-channel input chunk the box's
// neighbourhood is staged in LDS once; the 27 taps run in pairs as K = 32 steps.  With FIXED the
// box is 16 x 4 x 4 at compile time, so tap offsets are immediates.
//
// The input may be a concatenation read in place: with cup > 0, channels [0, cup) are x at half
// the resolution (nearest-neighbour upsampling) and channels [cup, Cin) are xs at full
// resolution; with cup = 0 all Cin channels are xs.
//
// Staging loads are buffer loads: an element outside the volume (or past the neighbourhood) takes
// the offset 2^31, past the resource's 2^31-byte range, and reads zero, so no lane selects.  The
// launcher keeps every tensor below 2^30 elements, so valid offsets plus the uniform channel
// offset stay below 2^31.
constexpr uint32_t STAGE_OOB = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t stage_rsrc(const bf16 *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<bf16 *>(base), (short)0, (int)STAGE_OOB, 0x00020000);
}
__device__ __forceinline__ uint4 stage_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 v = {(__bf16)a, (__bf16)b};  // gfx950: one v_cvt_pk_bf16_f32, round to nearest even
    return __builtin_bit_cast(uint32_t, v);
}

template <int NT, int RB, int SK, int FIXED>
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
    constexpr bool FX = FIXED != 0;
    static_assert(!FX || RB == 4, "the fixed box is 256 positions");
    extern __shared__ __align__(16) bf16 smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int TW = FX ? 16 : g.TW, TH = FX ? 4 : g.TH, TD = FX ? 4 : g.TD;
    // Box origin of this workgroup's tile.
    int b, d0, h0, w0;
    {
        const int tw = g.W / TW, th = g.H / TH, td = g.D / TD;
        int t = blockIdx.x;
        w0 = (t % tw) * TW;
        t /= tw;
        h0 = (t % th) * TH;
        t /= th;
        d0 = (t % td) * TD;
        b = t / td;
    }
    const int n0 = blockIdx.y * NB;
    const int SH = TH + 2, SW = TW + 2, SHW = SH * SW;
    const int TP = TD * TH * TW;
    const int lw = FX ? 4 : __builtin_ctz(TW), lh = FX ? 2 : __builtin_ctz(TH);
    // The staged neighbourhood is two planes (channels 0-7 and 8-15 of the chunk) of 16-byte
    // positions, so the 16 lanes of a row group read 256 contiguous bytes.  Lane group
    // lg = lane/16 takes tap 2p + qt of a tap pair p and channel half qh: k = 16 qh + 8 qt + j.
    const int lg = lane >> 4, qh = lg >> 1, qt = lg & 1;
    const int npos = (TD + 2) * SHW;
    const uint4 *sA = (const uint4 *)smem + qh * npos;

    // This lane's RB rows: positions wave*16*RB + rb*16 + lane%16 of the box (the last one for
    // rows past a small box, which read in-range positions and are never stored), as offsets
    // in the staged box; `tap1` adds the second tap of a pair one column on, `tap16` the pair
    // that crosses a kernel row (taps 2 and 3: SW - 2 on).
    int base[RB];
    bool live[RB];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int p = wave * 16 * RB + rb * 16 + (lane & 15);
        live[rb] = FX || p < TP;
        const int q = live[rb] ? p : TP - 1;
        const int pw = q & (TW - 1), ph = (q >> lw) & (TH - 1), pd = q >> (lw + lh);
        base[rb] = (pd * SH + ph) * SW + pw;
    }

    // Accumulators start from the bias of the lane's four output channels (n0 + n*16 + 4(lane/16)
    // + r), or from zero in a split launch, whose reduction adds it.
    float4v acc[RB][NT];
#pragma unroll
    for (int n = 0; n < NT; n++) {
        float4v b0 = float4v{0.f, 0.f, 0.f, 0.f};
        if (bias && !part) {
#pragma unroll
            for (int r = 0; r < 4; r++) b0[r] = bias[n0 + n * 16 + (lane >> 4) * 4 + r];
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++) acc[rb][n] = b0;
    }

    const int wrow = 27 * g.Cin;
    // The reduction runs over units (16-channel chunk, kernel plane dz); split z of gridDim.z
    // takes units [u0, u1).
    const int units = g.Cin / 16 * 3;
    const int u0 = (int)((int64_t)blockIdx.z * units / gridDim.z);
    const int u1 = (int)((int64_t)(blockIdx.z + 1) * units / gridDim.z);

    // Staging of a chunk: element e = tid + 256 k (k < SK) of the 2 npos 16-byte halves is half
    // tid % 2 (the same for every k) of staged position tid / 2 + 128 k, so consecutive lanes read
    // consecutive 16 bytes of the channels-last input.  The position's coordinates are found once
    // and stepped by 128 positions per k.  All of a thread's loads are issued before any is
    // waited on, and the next chunk's loads are issued before the current chunk's MFMAs.
    uint32_t src[SK];   // byte offset in xs (channel cup of the input), or STAGE_OOB
    uint32_t srcu[SK];  // byte offset in the half-resolution x (channel 0), when cup > 0
    const int cs = g.Cin - cup;  // channels of xs
    // Offsets are 32-bit, and the tile's own part of them is uniform: the box origin's position
    // index, times the channel count, on the scalar unit; each element adds its small signed
    // displacement in 24-bit products.
    const int HW = g.H * g.W;
    const int pbase = ((b * g.D + d0) * g.H + h0) * g.W + w0;
    const int sbase = pbase * cs;
    const int half = tid & 1, p0 = tid >> 1;
    {
        int px, py, pz;
        if constexpr (FX) {
            const int t = p0 / 18;
            px = p0 - 18 * t, pz = t / 6, py = t - 6 * pz;
        } else {
            const int t = div_small(p0, SW, 1.0f / (float)SW);
            px = p0 - t * SW;
            pz = div_small(t, SH, 1.0f / (float)SH);
            py = t - pz * SH;
        }
        // 128 positions are a_rows rows of SW and b_cols more; a_rows rows are a_z planes of SH
        // rows and a_y more.
        const int a_rows = 128 / SW, b_cols = 128 - a_rows * SW;
        const int a_z = a_rows / SH, a_y = a_rows - a_z * SH;
        const unsigned Dh = g.D >> 1, Hh = g.H >> 1, Wh = g.W >> 1;
#pragma unroll
        for (int k = 0; k < SK; k++) {
            if (k > 0) {
                px += b_cols;
                const int cx = px >= SW ? 1 : 0;
                px -= cx * SW;
                py += a_y + cx;
                const int cy = py >= SH ? 1 : 0;
                py -= cy * SH;
                pz += a_z + cy;
            }
            // Only the last two of the SK elements can lie past the neighbourhood: SK is at most
            // one more than ceil(npos / 128).
            const bool staged = k < SK - 2 || p0 + 128 * k < npos;
            const int z = d0 + pz - 1, y = h0 + py - 1, xx = w0 + px - 1;
            const bool in = staged && (unsigned)z < (unsigned)g.D && (unsigned)y < (unsigned)g.H &&
                            (unsigned)xx < (unsigned)g.W;
            const int rel = __mul24(pz - 1, HW) + __mul24(py - 1, g.W) + (px - 1);
            src[k] = in ? (uint32_t)(sbase + __mul24(rel, cs) + half * 8) * 2u : STAGE_OOB;
            srcu[k] = STAGE_OOB;
            if (cup > 0 && in) {
                const unsigned ph = __umul24(__umul24(__umul24((unsigned)b, Dh) + ((unsigned)z >> 1), Hh) +
                                             ((unsigned)y >> 1), Wh) + ((unsigned)xx >> 1);
                srcu[k] = (__umul24(ph, (unsigned)cup) + half * 8) * 2u;
            }
        }
    }
    // With one or two N-blocks the chunk's weights go through the same pipeline into LDS after
    // the neighbourhood: slot ((dz * 10 + tap) * 2 + half) * NB + column holds input channels
    // half*8 .. +8 of output channel n0 + column for tap < 9 of kernel plane dz, and zeros for
    // tap 9 (the padding tap of a plane's fifth pair), so the loop reads A from LDS and waits on
    // no global load.  Element eb = tid + 256 k is that slot; past 60 NB the loads read zero.
    constexpr bool WB = NT <= 2;
    constexpr int SKB = WB ? (60 * NB + 255) / 256 : 1;
    const int aslots = 2 * npos + 1;
    uint32_t wsrc[WB ? SKB : 1];
    if constexpr (WB) {
#pragma unroll
        for (int k = 0; k < SKB; k++) {
            const int eb = tid + 256 * k;
            const int col = eb % NB, rest = eb / NB;  // rest = (dz * 10 + tap) * 2 + half
            const int hf = rest & 1, dt = rest >> 1, dz = dt / 10, tap = dt - dz * 10;
            wsrc[k] = (eb < 60 * NB && tap < 9)
                          ? (uint32_t)((n0 + col) * wrow + (dz * 9 + tap) * g.Cin + hf * 8) * 2u
                          : STAGE_OOB;
        }
    }
    // The prefetch registers: the neighbourhood's SK elements, then the weights' SKB.
    uint4 pre[SK + (WB ? SKB : 0)];
    const __amdgpu_buffer_rsrc_t RX = stage_rsrc(x), RS = stage_rsrc(xs), RW = stage_rsrc(w);
    auto load_chunk = [&](int c0) {
        if (c0 < cup) {
#pragma unroll
            for (int k = 0; k < SK; k++) pre[k] = stage_load(RX, srcu[k], (uint32_t)c0 * 2u);
        } else {
#pragma unroll
            for (int k = 0; k < SK; k++) pre[k] = stage_load(RS, src[k], (uint32_t)(c0 - cup) * 2u);
        }
        if constexpr (WB) {
#pragma unroll
            for (int k = 0; k < SKB; k++) pre[SK + k] = stage_load(RW, wsrc[k], (uint32_t)c0 * 2u);
        }
    };
    const int dst0 = half * npos + p0;
    auto store_chunk = [&]() {
#pragma unroll
        for (int k = 0; k < SK; k++) {
            const bool staged = k < SK - 2 || p0 + 128 * k < npos;
            ((uint4 *)smem)[staged ? dst0 + 128 * k : 2 * npos] = pre[k];
        }
        if constexpr (WB) {
#pragma unroll
            for (int k = 0; k < SKB; k++) ((uint4 *)smem)[aslots + tid + 256 * k] = pre[SK + k];
        }
    };
    // This lane's A column (weights): LDS slot of output channel n0 + lane%16, channel half qh,
    // tap qt of a pair; from global memory when the N-block is too wide to stage.
    const uint4 *sW = (const uint4 *)smem + aslots + (qt * 2 + qh) * NB + (lane & 15);
    const bf16 *wl = w + (int64_t)(n0 + (lane & 15)) * wrow + qh * 8;

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
        const uint4 *sAz = sA + dz * SHW;
        const uint4 *sWz = sW + dz * 20 * NB;
#pragma unroll
        for (int pr = 0; pr < 5; pr++) {
            // Taps 2pr and 2pr+1 of this kernel plane in one K=32 step; the tenth is padding: it
            // reads the ninth tap's positions against a zero A fragment.
            const int t0 = 2 * pr, t1 = pr < 4 ? 2 * pr + 1 : 8;
            const int o0 = (t0 / 3) * SW + t0 % 3, o1 = (t1 / 3) * SW + t1 % 3;
            const int toff = qt ? o1 : o0;
            uint4 av[NT];
#pragma unroll
            for (int n = 0; n < NT; n++)
                av[n] = WB ? sWz[pr * 4 * NB + n * 16]
                           : (qt && pr == 4) ? make_uint4(0, 0, 0, 0)
                                             : *(const uint4 *)(wl + (int64_t)n * 16 * wrow + (dz * 9 + (qt ? t1 : t0)) * g.Cin + c0);
            uint4 bvv[RB];
#pragma unroll
            for (int rb = 0; rb < RB; rb++) bvv[rb] = sAz[base[rb] + toff];
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
#pragma unroll
                for (int n = 0; n < NT; n++) acc[rb][n] = mfma32(av[n], bvv[rb], acc[rb][n]);
            // The pair's LDS reads issue together, then its MFMAs: one exposed LDS latency a pair.
            __builtin_amdgcn_sched_group_barrier(0x100, NT + RB, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, NT * RB, 0);
        }
    }

    // Lane's rows and channels: position q(rb) = wave*16*RB + rb*16 + lane%16 of the box, output
    // channels n0 + n*16 + c4 .. +3.
    const int c4 = (lane >> 4) * 4;
    int pos[RB];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int q = wave * 16 * RB + rb * 16 + (lane & 15);
        const int qw = q & (TW - 1), qh2 = (q >> lw) & (TH - 1), qd = q >> (lw + lh);
        pos[rb] = pbase + __mul24(qd, HW) + __mul24(qh2, g.W) + qw;
    }
    if (part) {
        // Split: fp32 partials [split][position][Cout], bias and activation left to the reduction.
        float *pz = part + (int64_t)blockIdx.z * g.positions() * g.Cout;
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            if (!live[rb]) continue;
#pragma unroll
            for (int n = 0; n < NT; n++)
                *(float4v *)(pz + (int64_t)pos[rb] * g.Cout + n0 + n * 16 + c4) = acc[rb][n];
        }
        return;
    }

    // The activations are staged in LDS ([position][NB + 8] bf16) only for the pool and the head.
    const bool fused = pool_out || head_y;
    constexpr int NS = NB + 8;
    if (fused) __syncthreads();  // every wave is done reading the staged chunk
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int q = wave * 16 * RB + rb * 16 + (lane & 15);
#pragma unroll
        for (int n = 0; n < NT; n++) {
            float z[4], a[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                z[r] = acc[rb][n][r];
                a[r] = gelu_out ? gelu(z[r]) : z[r];
            }
            const int64_t o = (int64_t)pos[rb] * ostride + ooff + n0 + n * 16 + c4;
            if (live[rb] && zout) *(uint2 *)(zout + o) = make_uint2(pack_bf16(z[0], z[1]), pack_bf16(z[2], z[3]));
            const uint2 av = make_uint2(pack_bf16(a[0], a[1]), pack_bf16(a[2], a[3]));
            if (live[rb] && aout) *(uint2 *)(aout + o) = av;
            if (fused) *(uint2 *)(smem + q * NS + n * 16 + c4) = av;
        }
    }
    if (!fused) return;
    __syncthreads();
    if (pool_out) {
        // 2x2x2 means of the staged box (even extents, checked by the launcher): item e is
        // channel group e % (NB/8) of coarse position e / (NB/8), the eight fine values summed in
        // (dz, dy, dx) order.
        const int PW = TW >> 1, PH = TH >> 1, PP = (TD >> 1) * PH * PW;
        const int lpw = lw - 1, lph = lh - 1;
        const int Dh = g.D >> 1, Hh = g.H >> 1, Wh = g.W >> 1;
        for (int e = tid; e < PP * (NB / 8); e += 256) {
            const int pq = e / (NB / 8), grp = e % (NB / 8);
            const int pw = pq & (PW - 1), ph = (pq >> lpw) & (PH - 1), pd = pq >> (lpw + lph);
            float sum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int q = ((2 * pd + (k >> 2)) * TH + 2 * ph + ((k >> 1) & 1)) * TW + 2 * pw + (k & 1);
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
    if (head_y) {
        // The 1x1 head of a 16-channel layer: one position per thread, channels in order.
        for (int q = tid; q < TP; q += 256) {
            float sum = head_w[16];
#pragma unroll
            for (int c = 0; c < 16; c++) sum += head_w[c] * from_bf16(smem[q * NS + c]);
            const int qw = q & (TW - 1), qh2 = (q >> lw) & (TH - 1), qd = q >> (lw + lh);
            head_y[pbase + __mul24(qd, HW) + __mul24(qh2, g.W) + qw] = sum;
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
    // Staged 16-byte halves per thread: 2 (TD+2)(TH+2)(TW+2) over 256 threads (6 for the
    // 16 x 4 x 4 box, which has a variant of its own).
    const int halves = 2 * (g.TD + 2) * (g.TH + 2) * (g.TW + 2);
    const int sk = (halves + 255) / 256;
#define PM_FWD(SK, FX) conv_fwd2<NT, RB, SK, FX><<<grid, 256, lds, stream>>>(g, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y)
    if constexpr (RB == 4) {
        if (g.TW == 16 && g.TH == 4 && g.TD == 4) {
            PM_FWD(6, 1);
            return;
        }
    }
    if (sk <= 2)
        PM_FWD(2, 0);
    else if (sk <= 4)
        PM_FWD(4, 0);
    else if (sk <= 6)
        PM_FWD(6, 0);
    else
        PM_FWD(8, 0);
#undef PM_FWD
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
        // With a short reduction (at most 64 input channels), 64-position boxes that fill the
        // chip without a split: no fp32 partials to write and reduce.  A long reduction keeps
        // the 256-position boxes' reuse and splits (32^3, 192 -> 64 channels: 41 against 63 µs).
        const Conv f2 = forward_box(g0, 128);
        if (g0.Cin <= 64 && f2.TW > 0 && f2.TH > 0 && f2.TD > 0 && f2.tile_positions() == 128 &&
            f2.tiles() * (g0.Cout / (16 * nt_max)) >= 512)
            return {2, nt_max, 1};
        const Conv f1 = forward_box(g0, 64);
        if (g0.Cin <= 64 && f1.TW > 0 && f1.TH > 0 && f1.TD > 0 &&
            f1.tiles() * (g0.Cout / (16 * nt_max)) >= 1024)
            return {1, nt_max, 1};
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
                    g.positions() * std::max({g.Cin, g.Cout, ostride}) < ((int64_t)1 << 30);
    if (!ok) {
        set_error("conv_forward: shape %dx%dx%dx%d, %d->%d channels unsupported", g.B, g.D, g.H,
                  g.W, g.Cin, g.Cout);
        return PM_ERR_ARGUMENT;
    }
    const size_t stage = ((size_t)2 * (g.TD + 2) * (g.TH + 2) * (g.TW + 2) + 1) * 16 +
                         (nt <= 2 ? (size_t)((60 * 16 * nt + 255) / 256) * 256 * 16 : 0);
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
        case 12: launch_fwd<1, 2>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        case 22: launch_fwd<2, 2>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        case 21: launch_fwd<2, 1>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        case 31: launch_fwd<3, 1>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        case 41: launch_fwd<4, 1>(g, grid, lds, stream, x, xs, cup, w, bias, z, a, gelu_out, ostride, ooff, part, pool, head, head_y); break;
        default:
            set_error("conv_forward: no kernel for %d N-tiles x %d row blocks", nt, rb);
            return PM_ERR_ARGUMENT;
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
