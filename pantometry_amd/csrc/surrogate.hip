// The learned heat surrogate: a four-level 3D U-Net (channels 16-32-64-128, two 3x3x3 convolutions
// with GELU per level, average-pool down, nearest-neighbour up with skip concatenation, a 1x1 head)
// mapping (T, log k) to the temperature k stencil steps later.  Trained and run on gfx950 by the
// implicit-GEMM bf16 MFMA kernels of nn.hip with fp32 accumulation, fp32 master weights and Adam.
//
// Inputs are normalised as ((T - 300 K) / 100 K, log k / 0.5); the network predicts the change
// (T_k - T_0) / 10 K, and a prediction is made conservative by a uniform shift that restores the
// block's total heat before it is handed back.
//
// The FP64 stencil stays the judge: see pm_surrogate_step_solid, which applies one surrogate step
// to a Solid3D in place and reports the audit drift; the rollout driver compares against the
// stencil and falls back to it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "nn.hpp"

namespace pm {

// Device view of a Solid3D handle's state (heat.hip).
int solid_device_state(pm_solid3d *s, double **cells, const double **kcell, int64_t counts[3],
                       hipStream_t *stream, bool *uniform_capacity, double *mobility);

namespace nn {

constexpr float T_SHIFT = 300.0f, T_SCALE = 100.0f, K_SCALE = 0.5f, OUT_SCALE = 10.0f;
constexpr int K_STEPS = 16;  // stencil steps per surrogate step

// ---- element-wise kernels ----------------------------------------------------------------

#define GRID_STRIDE(i, n) \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// The network's sixteen input channels at one cell of a cubic sample (edge^3, x fastest), from
// the temperatures, the conductivities and the per-step coupling a = dt * dx / C:
//   0      (T - 300 K) / 100 K
//   1      log k / 0.5
//   2..7   10 * a * k_f for the faces x-, x+, y-, y+, z-, z+ (k_f the harmonic mean; 0 on the
//          block's surface, which is insulated)
//   8..13  the face's share of the explicit step, a * k_f * (T_nb - T), times K_STEPS / 10 K:
//          the first-order estimate of the target, face by face
//   14     their sum
//   15     1 (marks the volume: the convolutions pad with zeros)
// The fourteen loads are issued together (a face off the block reads the cell itself and is
// masked afterwards), and the arithmetic is fp32 but for T_nb - T, taken in the fields' own
// precision: every feature is rounded to bf16, whose 2^-8 step is far above fp32's error.
template <class TT>
__device__ __forceinline__ void cell_features(const TT *__restrict__ t, const TT *__restrict__ k,
                                              int64_t base, int edge, int x, int y, int z,
                                              float coupling, bf16 *__restrict__ out) {
    const int64_t c = x + (int64_t)edge * (y + (int64_t)edge * z);
    const int64_t e = edge, ee = (int64_t)edge * edge;
    const int64_t off[6] = {-1, 1, -e, e, -ee, ee};
    const bool on[6] = {x > 0, x < edge - 1, y > 0, y < edge - 1, z > 0, z < edge - 1};
    const TT tc = t[base + c], kc = k[base + c];
    TT tn[6], kn[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int64_t nb = base + (on[q] ? c + off[q] : c);
        tn[q] = t[nb], kn[q] = k[nb];
    }
    float f[16];
    f[0] = (float)(tc - (TT)T_SHIFT) * (1.0f / T_SCALE);
    const float kcf = (float)kc;
    f[1] = __logf(kcf) * (1.0f / K_SCALE);
    float sum = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const float knf = (float)kn[q];
        const float kf = __fdividef(2.0f * kcf * knf, kcf + knf);
        const float af = on[q] ? coupling * kf : 0.0f;
        const float fl = on[q] ? af * (float)(tn[q] - tc) * ((float)K_STEPS / OUT_SCALE) : 0.0f;
        f[2 + q] = 10.0f * af;
        f[8 + q] = fl;
        sum += fl;
    }
    f[14] = sum;
    f[15] = 1.0f;
    uint32_t w[8];
#pragma unroll
    for (int q = 0; q < 8; q++) w[q] = (uint32_t)to_bf16(f[2 * q]) | ((uint32_t)to_bf16(f[2 * q + 1]) << 16);
    *(uint4 *)(out) = make_uint4(w[0], w[1], w[2], w[3]);
    *(uint4 *)(out + 8) = make_uint4(w[4], w[5], w[6], w[7]);
}

// Features of one block from fp64 device fields: workgroups of 64 x 4 cells over (x, y) and a
// grid over (x, y, z), so a cell's coordinates come from the launch, not from divisions.
__global__ __launch_bounds__(256) void pack_input(const double *__restrict__ t,
                                                  const double *__restrict__ k, int edge,
                                                  float coupling, bf16 *__restrict__ out) {
    const int x = (int)(blockIdx.x * 64 + threadIdx.x), y = (int)(blockIdx.y * 4 + threadIdx.y);
    const int z = (int)blockIdx.z;
    if (x >= edge || y >= edge) return;
    const int64_t c = x + (int64_t)edge * (y + (int64_t)edge * z);
    cell_features<double>(t, k, 0, edge, x, y, z, coupling, out + c * 16);
}
static dim3 pack_grid(int64_t edge) {
    return dim3((unsigned)((edge + 63) / 64), (unsigned)((edge + 3) / 4), (unsigned)edge);
}
static const dim3 PACK_BLOCK(64, 4);

// Features of a batch drawn from the training set (fp32 fields [sample][cell]).
__global__ void pack_batch(const float *__restrict__ t, const float *__restrict__ k,
                           const float *__restrict__ coupling, const int *__restrict__ idx,
                           int edge, int batch, bf16 *__restrict__ out) {
    const int64_t n = (int64_t)edge * edge * edge;
    GRID_STRIDE(i, n * batch) {
        const int64_t s = i / n, c = i % n;
        const int64_t base = (int64_t)idx[s] * n;
        const int cc = (int)c, x = cc % edge, y = (cc / edge) % edge, z = cc / (edge * edge);
        cell_features<float>(t, k, base, edge, x, y, z, coupling[idx[s]], out + i * 16);
    }
}

__global__ void gather_target(const float *__restrict__ set, const int *__restrict__ idx,
                              int64_t n, int batch, float *__restrict__ out) {
    GRID_STRIDE(i, n * batch) {
        const int64_t s = i / n, c = i % n;
        out[i] = set[(int64_t)idx[s] * n + c];
    }
}

// 2x2x2 average pool, channels-last: (sum of the eight fine values) * 0.125, rounded once.
__global__ void pool_fwd(const bf16 *__restrict__ x, bf16 *__restrict__ y, int B, int D, int H,
                         int W, int C) {
    const int d2 = D / 2, h2 = H / 2, w2 = W / 2, G = C / 8;
    GRID_STRIDE(i, (int64_t)B * d2 * h2 * w2 * G) {
        const int g = (int)(i % G);
        int64_t p = i / G;
        const int w = (int)(p % w2);
        p /= w2;
        const int h = (int)(p % h2);
        p /= h2;
        const int d = (int)(p % d2);
        const int b = (int)(p / d2);
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < 8; k++) {
            const int64_t q = (((int64_t)b * D + 2 * d + (k >> 2)) * H + 2 * h + ((k >> 1) & 1)) * W + 2 * w + (k & 1);
            float f[8];
            unpack8(*(const uint4 *)(x + q * C + g * 8), f);
#pragma unroll
            for (int c = 0; c < 8; c++) s[c] += f[c];
        }
#pragma unroll
        for (int c = 0; c < 8; c++) s[c] *= 0.125f;
        *(uint4 *)(y + (i / G) * C + g * 8) = pack8(s);
    }
}

// Gradient of the pool, added to dx (the skip path has already written its part there).
__global__ void pool_bwd_add(const bf16 *__restrict__ dy, bf16 *__restrict__ dx, int B, int D,
                             int H, int W, int C) {
    const int G = C / 8;
    GRID_STRIDE(i, (int64_t)B * D * H * W * G) {
        const int g = (int)(i % G);
        int64_t p = i / G;
        const int w = (int)(p % W);
        p /= W;
        const int h = (int)(p % H);
        p /= H;
        const int d = (int)(p % D);
        const int b = (int)(p / D);
        const int64_t q = (((int64_t)b * (D / 2) + d / 2) * (H / 2) + h / 2) * (W / 2) + w / 2;
        const int64_t o = (i / G) * C + g * 8;
        float a[8], c[8];
        unpack8(*(const uint4 *)(dx + o), a);
        unpack8(*(const uint4 *)(dy + q * C + g * 8), c);
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = a[k] + 0.125f * c[k];
        *(uint4 *)(dx + o) = pack8(a);
    }
}

// Nearest-neighbour upsampling by two into channels [0, C) of a buffer with Ctot channels.
__global__ void up_fwd(const bf16 *__restrict__ x, bf16 *__restrict__ y, int B, int D, int H,
                       int W, int C, int Ctot) {
    // x is [B][D][H][W][C]; y is [B][2D][2H][2W][Ctot]
    const int G = C / 8;
    GRID_STRIDE(i, (int64_t)B * 8 * D * H * W * G) {
        const int g = (int)(i % G);
        int64_t p = i / G;
        const int w = (int)(p % (2 * W));
        p /= 2 * W;
        const int h = (int)(p % (2 * H));
        p /= 2 * H;
        const int d = (int)(p % (2 * D));
        const int b = (int)(p / (2 * D));
        const int64_t src = (((int64_t)b * D + d / 2) * H + h / 2) * W + w / 2;
        *(uint4 *)(y + (i / G) * Ctot + g * 8) = *(const uint4 *)(x + src * C + g * 8);
    }
}

// Gradient of the upsampling: dx[p][c] = sum of the eight fine cells' dy[.][c], rounded once.
__global__ void up_bwd(const bf16 *__restrict__ dy, bf16 *__restrict__ dx, int B, int D, int H,
                       int W, int C, int Ctot) {
    const int G = C / 8;
    GRID_STRIDE(i, (int64_t)B * D * H * W * G) {
        const int g = (int)(i % G);
        int64_t p = i / G;
        const int w = (int)(p % W);
        p /= W;
        const int h = (int)(p % H);
        p /= H;
        const int d = (int)(p % D);
        const int b = (int)(p / D);
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < 8; k++) {
            const int64_t q = (((int64_t)b * 2 * D + 2 * d + (k >> 2)) * 2 * H + 2 * h + ((k >> 1) & 1)) * 2 * W + 2 * w + (k & 1);
            float f[8];
            unpack8(*(const uint4 *)(dy + q * Ctot + g * 8), f);
#pragma unroll
            for (int c = 0; c < 8; c++) s[c] += f[c];
        }
        *(uint4 *)(dx + (i / G) * C + g * 8) = pack8(s);
    }
}

// Copies x (C channels) into channels [off, off+C) of y (Ctot channels), or, with `add`, adds
// those channels of y into x (the skip's share of a concatenation gradient).
__global__ void channel_copy(const bf16 *__restrict__ src, bf16 *__restrict__ dst, int64_t npos,
                             int C, int sstride, int soff, int dstride, int doff, int add) {
    const int G = C / 8;
    GRID_STRIDE(i, npos * G) {
        const int g = (int)(i % G);
        const int64_t p = i / G;
        const uint4 v = *(const uint4 *)(src + p * sstride + soff + g * 8);
        uint4 *o = (uint4 *)(dst + p * dstride + doff + g * 8);
        if (add) {
            float a[8], b[8];
            unpack8(*o, a);
            unpack8(v, b);
#pragma unroll
            for (int k = 0; k < 8; k++) a[k] = a[k] + b[k];
            *o = pack8(a);
        } else {
            *o = v;
        }
    }
}

// dz = da * gelu'(z), in place into da, eight values per thread.
__global__ void gelu_bwd(const bf16 *__restrict__ z, bf16 *__restrict__ da, int64_t n) {
    GRID_STRIDE(i, n / 8) {
        float a[8], zz[8];
        unpack8(*(const uint4 *)(da + i * 8), a);
        unpack8(*(const uint4 *)(z + i * 8), zz);
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = a[k] * gelu_grad(zz[k]);
        *(uint4 *)(da + i * 8) = pack8(a);
    }
}

// The 1x1 head: y[p] = sum_c hw[c] a[p][c] + hb.
__global__ void head_fwd(const bf16 *__restrict__ a, const float *__restrict__ hw,
                         const float *__restrict__ hb, int64_t npos, float *__restrict__ y) {
    GRID_STRIDE(p, npos) {
        float s = hb[0];
#pragma unroll
        for (int c = 0; c < 16; c++) s += hw[c] * from_bf16(a[p * 16 + c]);
        y[p] = s;
    }
}

// Loss 0.5 * mean((y - t)^2) per block partial, and dy = (y - t) / n.
__global__ __launch_bounds__(256) void loss_fwd_bwd(const float *__restrict__ y,
                                                    const float *__restrict__ t, int64_t n,
                                                    float *__restrict__ dy,
                                                    float *__restrict__ part) {
    __shared__ float red[256];
    float s = 0.0f;
    GRID_STRIDE(i, n) {
        const float e = y[i] - t[i];
        s += e * e;
        dy[i] = e / (float)n;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Head gradients: da[p][c] = dy[p] hw[c]; per-block partials of dhw[c] and dhb.
__global__ __launch_bounds__(256) void head_bwd(const bf16 *__restrict__ a,
                                                const float *__restrict__ hw,
                                                const float *__restrict__ dy, int64_t npos,
                                                bf16 *__restrict__ da, float *__restrict__ part) {
    __shared__ float red[17][256];
    float g[17];
#pragma unroll
    for (int c = 0; c < 17; c++) g[c] = 0.0f;
    GRID_STRIDE(p, npos) {
        const float d = dy[p];
#pragma unroll
        for (int c = 0; c < 16; c++) {
            const float av = from_bf16(a[p * 16 + c]);
            da[p * 16 + c] = to_bf16(d * hw[c]);
            g[c] += d * av;
        }
        g[16] += d;
    }
#pragma unroll
    for (int c = 0; c < 17; c++) red[c][threadIdx.x] = g[c];
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k)
            for (int c = 0; c < 17; c++) red[c][threadIdx.x] += red[c][threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x < 17) part[blockIdx.x * 17 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void sum_parts(const float *__restrict__ part, int nparts, int width,
                          float *__restrict__ out) {
    const int c = threadIdx.x;
    if (c >= width) return;
    float s = 0.0f;
    for (int k = 0; k < nparts; k++) s += part[k * width + c];
    out[c] = s;
}

// Adam (bias-corrected) on fp32 master parameters; refreshes the bf16 copies of weights.
__global__ void adam_step(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, int64_t n, float lr, float b1, float b2,
                          float eps, float c1, float c2) {
    GRID_STRIDE(i, n) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

__global__ void to_bf16_kernel(const float *__restrict__ x, bf16 *__restrict__ y, int64_t n) {
    GRID_STRIDE(i, n) { y[i] = to_bf16(x[i]); }
}

// T += 10 K * y in place (fp64), with per-block partial sums of T before (part0) and after
// (part1): each thread folds its cells in index order, then the fixed block tree.
__global__ __launch_bounds__(256) void apply_prediction_sum(double *__restrict__ t,
                                                            const float *__restrict__ y, int64_t n,
                                                            double *__restrict__ part0,
                                                            double *__restrict__ part1) {
    __shared__ double red[4];
    double s0 = 0.0, s1 = 0.0;
    GRID_STRIDE(i, n) {
        const double u = t[i];
        const double v = u + (double)OUT_SCALE * (double)y[i];
        t[i] = v;
        s0 += u;
        s1 += v;
    }
    const double tot0 = block_sum<256>(s0, red);
    __syncthreads();
    const double tot1 = block_sum<256>(s1, red);
    if (threadIdx.x == 0) part0[blockIdx.x] = tot0, part1[blockIdx.x] = tot1;
}

// scal[0] = sum T before, scal[1] = after the prediction, scal[2] = the uniform shift that
// restores the total; fixed-order sums of the partials.
__global__ __launch_bounds__(1024) void shift_finalize(const double *__restrict__ part0,
                                                       const double *__restrict__ part1,
                                                       unsigned nparts, int64_t n,
                                                       double *__restrict__ scal) {
    __shared__ double red[16];
    const double before = fixed_order_sum_1024(part0, nparts, red);
    __syncthreads();
    const double after = fixed_order_sum_1024(part1, nparts, red);
    if (threadIdx.x == 0) {
        scal[0] = before;
        scal[1] = after;
        scal[2] = (before - after) / (double)n;
    }
}

__global__ __launch_bounds__(256) void shift_sum(double *__restrict__ t, int64_t n,
                                                 const double *__restrict__ scal,
                                                 double *__restrict__ part) {
    __shared__ double red[4];
    const double delta = scal[2];
    double s = 0.0;
    GRID_STRIDE(i, n) {
        const double v = t[i] + delta;
        t[i] = v;
        s += v;
    }
    const double tot = block_sum<256>(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// scal[3] = sum T after the shift.
__global__ __launch_bounds__(1024) void drift_finalize(const double *__restrict__ part,
                                                       unsigned nparts, double *__restrict__ scal) {
    __shared__ double red[16];
    const double fixed = fixed_order_sum_1024(part, nparts, red);
    if (threadIdx.x == 0) scal[3] = fixed;
}

__global__ __launch_bounds__(256) void sq_diff_f64(const double *__restrict__ a,
                                                   const double *__restrict__ b, double ref,
                                                   int64_t n, double *__restrict__ part) {
    __shared__ double red[4];
    double s = 0.0, q = 0.0;
    GRID_STRIDE(i, n) {
        const double e = a[i] - b[i], r = b[i] - ref;
        s += e * e;
        q += r * r;
    }
    const double ts = block_sum<256>(s, red);
    __syncthreads();
    const double tq = block_sum<256>(q, red);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = ts, part[2 * blockIdx.x + 1] = tq;
}

// One training pair straight from two device temperature fields and the conductivity: T_0 and k
// as fp32, the target (T_k - T_0) / 10 K as fp32.
__global__ void pack_pair(const double *__restrict__ t0, const double *__restrict__ k,
                          const double *__restrict__ tk, int64_t n, float *__restrict__ st,
                          float *__restrict__ sk, float *__restrict__ target) {
    GRID_STRIDE(i, n) {
        st[i] = (float)t0[i];
        sk[i] = (float)k[i];
        target[i] = (float)((tk[i] - t0[i]) / OUT_SCALE);
    }
}

__global__ void shift_f64(double *__restrict__ x, int64_t n, double delta) {
    GRID_STRIDE(i, n) { x[i] += delta; }
}

static unsigned grid_of(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

}  // namespace nn

// ---- the network -----------------------------------------------------------------------------

namespace {

using nn::bf16;

struct ConvLayer {
    int cin, cout, level;  // level 0..3: resolution R >> level
    int64_t w_off, b_off;  // offsets into the fp32 parameter array
    int64_t wbf_off;       // offset into the bf16 weight copy
};

struct Act {  // one activation tensor and its gradient
    nn::bf16 *z = nullptr, *a = nullptr, *da = nullptr;
    int level = 0, c = 0;
};

}  // namespace

}  // namespace pm

using namespace pm;
using namespace pm::nn;

struct pm_surrogate {
    int R = 64;          // resolution the buffers are sized for
    int batch = 1;
    std::vector<ConvLayer> layers;
    int64_t nparams = 0, nwbf = 0;
    DevBuf params, grads, m, v, wbf, wflip, wscratch, fsplit;
    int64_t wscratch_floats = 0, fsplit_floats = 0;
    int64_t head_off = 0;  // 16 weights then 1 bias
    int adam_t = 0;
    hipStream_t stream = nullptr;
    // activations
    std::vector<DevBuf> bufs;
    DevBuf xin, y, dy, target, parts, idx, hparts;
    bf16 *cat[3] = {nullptr, nullptr, nullptr}, *dcat[3] = {nullptr, nullptr, nullptr};
    bf16 *pool[3] = {nullptr, nullptr, nullptr}, *dpool[3] = {nullptr, nullptr, nullptr};
    Act act[14];
    // training set
    DevBuf set_t0, set_k, set_t, set_a;  // T_0, k, target per cell; coupling per sample
    int64_t set_n = 0, set_cells = 0;
    int set_edge = 0;
    // step scratch: reduction partials and scalars, and the step's timing events
    DevBuf red64;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t seed = 1;
};

namespace {

const int kLayerShape[14][3] = {
    // cin, cout, level
    {16, 16, 0}, {16, 16, 0},      // enc1
    {16, 32, 1}, {32, 32, 1},      // enc2
    {32, 64, 2}, {64, 64, 2},      // enc3
    {64, 128, 3}, {128, 128, 3},   // enc4
    {192, 64, 2}, {64, 64, 2},     // dec3
    {96, 32, 1}, {32, 32, 1},      // dec2
    {48, 16, 0}, {16, 16, 0},      // dec1
};

int64_t npos(const pm_surrogate *s, int level) {
    const int64_t r = s->R >> level;
    return (int64_t)s->batch * r * r * r;
}

nn::Conv conv_of(const pm_surrogate *s, int li) {
    const int r = s->R >> s->layers[li].level;
    return nn::make_conv(s->batch, r, r, r, s->layers[li].cin, s->layers[li].cout);
}

int alloc_bf(pm_surrogate *s, int64_t count, bf16 **out) {
    s->bufs.emplace_back();
    PM_CHECK(s->bufs.back().alloc((size_t)std::max<int64_t>(count, 1) * sizeof(bf16)));
    *out = s->bufs.back().as<bf16>();
    return PM_OK;
}

// (Re)allocates activation buffers for resolution R and batch B.
int size_buffers(pm_surrogate *s, int R, int B) {
    if (s->R == R && s->batch == B && !s->bufs.empty()) return PM_OK;
    if (R < 8 || (R & (R - 1))) {
        set_error("surrogate blocks must have a power-of-two edge of at least 8, not %d", R);
        return PM_ERR_ARGUMENT;
    }
    PM_HIP(hipStreamSynchronize(s->stream));
    s->bufs.clear();
    s->R = R, s->batch = B;
    for (int li = 0; li < 14; li++) {
        Act &a = s->act[li];
        a.level = s->layers[li].level, a.c = s->layers[li].cout;
        const int64_t n = npos(s, a.level) * a.c;
        PM_CHECK(alloc_bf(s, n, &a.z));
        PM_CHECK(alloc_bf(s, n, &a.a));
        PM_CHECK(alloc_bf(s, n, &a.da));
    }
    const int cat_c[3] = {48, 96, 192};  // at levels 0, 1, 2
    for (int l = 0; l < 3; l++) {
        PM_CHECK(alloc_bf(s, npos(s, l) * cat_c[l], &s->cat[l]));
        PM_CHECK(alloc_bf(s, npos(s, l) * cat_c[l], &s->dcat[l]));
    }
    const int pool_c[3] = {16, 32, 64};  // pooled outputs at levels 1, 2, 3
    for (int l = 0; l < 3; l++) {
        PM_CHECK(alloc_bf(s, npos(s, l + 1) * pool_c[l], &s->pool[l]));
        PM_CHECK(alloc_bf(s, npos(s, l + 1) * pool_c[l], &s->dpool[l]));
    }
    // The split-reduction workspace of the small-grid convolutions, forward and data gradient.
    int64_t need = 0;
    for (int li = 0; li < 14; li++) {
        const nn::Conv g = conv_of(s, li);
        need = std::max(need, nn::conv_forward_workspace(g));
        need = std::max(need, nn::conv_forward_workspace(nn::transposed(g)));
    }
    s->fsplit_floats = need;
    PM_CHECK(s->fsplit.alloc(sizeof(float) * (size_t)std::max<int64_t>(need, 1)));
    const int64_t n0 = npos(s, 0);
    PM_CHECK(s->xin.alloc((size_t)n0 * 16 * sizeof(bf16)));
    PM_CHECK(s->y.alloc((size_t)n0 * sizeof(float)));
    PM_CHECK(s->dy.alloc((size_t)n0 * sizeof(float)));
    PM_CHECK(s->target.alloc((size_t)n0 * sizeof(float)));
    PM_CHECK(s->parts.alloc(65536 * sizeof(float)));
    PM_CHECK(s->hparts.alloc(65536 * 17 * sizeof(float)));
    PM_CHECK(s->idx.alloc((size_t)B * sizeof(int)));
    return PM_OK;
}

// The forward pass over xin into y.  With `training` the pre-activations are kept for the
// backward pass; inference writes only the activations.
int run_forward(pm_surrogate *s, bool training) {
    hipStream_t st = s->stream;
    const float *P = s->params.as<float>();
    const bf16 *WB = s->wbf.as<bf16>();
    auto R = [&](int l) { return s->R >> l; };
    auto conv = [&](int li, const bf16 *in, Act &o, const bf16 *skip = nullptr, int cup = 0,
                    nn::FwdFuse *fuse = nullptr) -> int {
        const nn::Conv g = conv_of(s, li);
        return nn::conv_forward(g, in, WB + s->layers[li].wbf_off, P + s->layers[li].b_off,
                                training ? o.z : nullptr, o.a, true, g.Cout, 0, st,
                                s->fsplit.as<float>(), s->fsplit_floats, skip, cup, fuse);
    };
    // An encoder level's second convolution and its pooling: in one launch where the epilogue
    // can pool, else the pooling kernel after it.
    auto conv_pool = [&](int li, const bf16 *in, Act &o, bf16 *pooled, int level, int C) -> int {
        nn::FwdFuse f;
        f.pool = pooled;
        PM_CHECK(conv(li, in, o, nullptr, 0, &f));
        if (!f.fused) {
            nn::pool_fwd<<<grid_of(npos(s, level) * C / 8), 256, 0, st>>>(o.a, pooled, s->batch, R(level - 1), R(level - 1), R(level - 1), C);
            PM_HIP(hipGetLastError());
        }
        return PM_OK;
    };
    // A decoder level's input: up(low) ++ skip.  Training keeps the concatenation (its weight
    // gradient reads it); inference reads both parts in place.
    auto decoder = [&](int li, const bf16 *low, int clow, const bf16 *skip, int cskip, bf16 *cat,
                       int level) -> int {
        if (!training) return conv(li, low, s->act[li], skip, clow);
        nn::up_fwd<<<grid_of(npos(s, level) * clow / 8), 256, 0, st>>>(
            low, cat, s->batch, R(level + 1), R(level + 1), R(level + 1), clow, clow + cskip);
        nn::channel_copy<<<grid_of(npos(s, level) * cskip / 8), 256, 0, st>>>(
            skip, cat, npos(s, level), cskip, cskip, 0, clow + cskip, clow, 0);
        return conv(li, cat, s->act[li]);
    };
    PM_CHECK(conv(0, s->xin.as<bf16>(), s->act[0]));
    PM_CHECK(conv_pool(1, s->act[0].a, s->act[1], s->pool[0], 1, 16));
    PM_CHECK(conv(2, s->pool[0], s->act[2]));
    PM_CHECK(conv_pool(3, s->act[2].a, s->act[3], s->pool[1], 2, 32));
    PM_CHECK(conv(4, s->pool[1], s->act[4]));
    PM_CHECK(conv_pool(5, s->act[4].a, s->act[5], s->pool[2], 3, 64));
    PM_CHECK(conv(6, s->pool[2], s->act[6]));
    PM_CHECK(conv(7, s->act[6].a, s->act[7]));
    // decoder level 2: up(enc4) ++ enc3
    PM_CHECK(decoder(8, s->act[7].a, 128, s->act[5].a, 64, s->cat[2], 2));
    PM_CHECK(conv(9, s->act[8].a, s->act[9]));
    PM_CHECK(decoder(10, s->act[9].a, 64, s->act[3].a, 32, s->cat[1], 1));
    PM_CHECK(conv(11, s->act[10].a, s->act[11]));
    PM_CHECK(decoder(12, s->act[11].a, 32, s->act[1].a, 16, s->cat[0], 0));
    // The last convolution and the 1x1 head: in inference one launch, which stores only y.
    nn::FwdFuse f;
    f.head = P + s->head_off;
    f.head_y = s->y.as<float>();
    f.skip_a = !training;
    PM_CHECK(conv(13, s->act[12].a, s->act[13], nullptr, 0, &f));
    if (!f.fused) {
        nn::head_fwd<<<grid_of(npos(s, 0)), 256, 0, st>>>(s->act[13].a, P + s->head_off, P + s->head_off + 16, npos(s, 0), s->y.as<float>());
        PM_HIP(hipGetLastError());
    }
    return PM_OK;
}

// Backward through one conv layer: its output gradient is in act[li].da (bf16); writes the input
// gradient to `din` (when not null) and the parameter gradients into the grads array.
int conv_back(pm_surrogate *s, int li, const bf16 *in, bf16 *din) {
    hipStream_t st = s->stream;
    Act &o = s->act[li];
    const nn::Conv g = conv_of(s, li);
    const int64_t n = g.positions() * g.Cout;
    nn::gelu_bwd<<<grid_of(n / 8), 256, 0, st>>>(o.z, o.da, n);
    PM_HIP(hipGetLastError());
    float *G = s->grads.as<float>();
    const int64_t tiles = g.tiles();
    const int64_t blocks_ct = (int64_t)(g.Cin / 16) * (g.Cout / 16);
    int splits = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, std::max<int64_t>(1, 2048 / blocks_ct)));
    while ((int64_t)splits * g.Cout * (27 * g.Cin + 1) > s->wscratch_floats) splits /= 2;
    splits = std::max(splits, 1);
    PM_CHECK(nn::conv_backward_weight(g, in, o.da, G + s->layers[li].w_off, G + s->layers[li].b_off,
                                      s->wscratch.as<float>(), splits, st));
    if (din)
        PM_CHECK(nn::conv_backward_data(g, o.da, s->wbf.as<bf16>() + s->layers[li].wbf_off,
                                        s->wflip.as<bf16>(), din, st, s->fsplit.as<float>(),
                                        s->fsplit_floats));
    return PM_OK;
}

int run_backward(pm_surrogate *s) {
    hipStream_t st = s->stream;
    const float *P = s->params.as<float>();
    float *G = s->grads.as<float>();
    const int B = s->batch;
    auto R = [&](int l) { return s->R >> l; };
    // head
    const int64_t n0 = npos(s, 0);
    const unsigned hb = std::min<unsigned>(grid_of(n0), 4096);
    nn::head_bwd<<<hb, 256, 0, st>>>(s->act[13].a, P + s->head_off, s->dy.as<float>(), n0, s->act[13].da, s->hparts.as<float>());
    nn::sum_parts<<<1, 64, 0, st>>>(s->hparts.as<float>(), (int)hb, 17, G + s->head_off);
    PM_HIP(hipGetLastError());
    // decoder level 0
    PM_CHECK(conv_back(s, 13, s->act[12].a, s->act[12].da));
    PM_CHECK(conv_back(s, 12, s->cat[0], s->dcat[0]));
    nn::up_bwd<<<grid_of(npos(s, 1) * 32 / 8), 256, 0, st>>>(s->dcat[0], s->act[11].da, B, R(1), R(1), R(1), 32, 48);
    nn::channel_copy<<<grid_of(n0 * 16 / 8), 256, 0, st>>>(s->dcat[0], s->act[1].da, n0, 16, 48, 32, 16, 0, 0);
    // decoder level 1
    PM_CHECK(conv_back(s, 11, s->act[10].a, s->act[10].da));
    PM_CHECK(conv_back(s, 10, s->cat[1], s->dcat[1]));
    nn::up_bwd<<<grid_of(npos(s, 2) * 64 / 8), 256, 0, st>>>(s->dcat[1], s->act[9].da, B, R(2), R(2), R(2), 64, 96);
    nn::channel_copy<<<grid_of(npos(s, 1) * 32 / 8), 256, 0, st>>>(s->dcat[1], s->act[3].da, npos(s, 1), 32, 96, 64, 32, 0, 0);
    // decoder level 2
    PM_CHECK(conv_back(s, 9, s->act[8].a, s->act[8].da));
    PM_CHECK(conv_back(s, 8, s->cat[2], s->dcat[2]));
    nn::up_bwd<<<grid_of(npos(s, 3) * 128 / 8), 256, 0, st>>>(s->dcat[2], s->act[7].da, B, R(3), R(3), R(3), 128, 192);
    nn::channel_copy<<<grid_of(npos(s, 2) * 64 / 8), 256, 0, st>>>(s->dcat[2], s->act[5].da, npos(s, 2), 64, 192, 128, 64, 0, 0);
    // encoder level 3
    PM_CHECK(conv_back(s, 7, s->act[6].a, s->act[6].da));
    PM_CHECK(conv_back(s, 6, s->pool[2], s->dpool[2]));
    nn::pool_bwd_add<<<grid_of(npos(s, 2) * 64 / 8), 256, 0, st>>>(s->dpool[2], s->act[5].da, B, R(2), R(2), R(2), 64);
    // encoder level 2
    PM_CHECK(conv_back(s, 5, s->act[4].a, s->act[4].da));
    PM_CHECK(conv_back(s, 4, s->pool[1], s->dpool[1]));
    nn::pool_bwd_add<<<grid_of(npos(s, 1) * 32 / 8), 256, 0, st>>>(s->dpool[1], s->act[3].da, B, R(1), R(1), R(1), 32);
    // encoder level 1
    PM_CHECK(conv_back(s, 3, s->act[2].a, s->act[2].da));
    PM_CHECK(conv_back(s, 2, s->pool[0], s->dpool[0]));
    nn::pool_bwd_add<<<grid_of(n0 * 16 / 8), 256, 0, st>>>(s->dpool[0], s->act[1].da, B, R(0), R(0), R(0), 16);
    // encoder level 0
    PM_CHECK(conv_back(s, 1, s->act[0].a, s->act[0].da));
    PM_CHECK(conv_back(s, 0, s->xin.as<bf16>(), nullptr));
    PM_HIP(hipGetLastError());
    return PM_OK;
}

int refresh_bf16(pm_surrogate *s) {
    for (const ConvLayer &l : s->layers) {
        const int64_t n = (int64_t)l.cout * 27 * l.cin;
        nn::to_bf16_kernel<<<grid_of(n), 256, 0, s->stream>>>(s->params.as<float>() + l.w_off, s->wbf.as<bf16>() + l.wbf_off, n);
    }
    PM_HIP(hipGetLastError());
    return PM_OK;
}

// Indices of a batch must name pairs of the training set.
int check_indices(const pm_surrogate *s, const int *indices, int batch) {
    for (int i = 0; i < batch; i++)
        if (indices[i] < 0 || indices[i] >= s->set_n) {
            set_error("batch index %d out of the training set's %lld pairs", indices[i],
                      (long long)s->set_n);
            return PM_ERR_ARGUMENT;
        }
    return PM_OK;
}

// xorshift64* (Vigna, 2016) on the seed, for weight initialisation: uniform in [0, 1) from the top
// 53 bits of each output.
double init_unit(uint64_t *st) {
    uint64_t x = *st;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *st = x;
    return (double)((x * 0x2545F4914F6CDD1DULL) >> 11) / (double)(1ULL << 53);
}

}  // namespace

extern "C" {

int pm_surrogate_create(uint64_t seed, pm_surrogate **out) {
    if (!out) return PM_ERR_ARGUMENT;
    *out = nullptr;
    PM_CHECK(require_device());
    std::unique_ptr<pm_surrogate> s(new pm_surrogate());
    s->seed = seed;
    int64_t off = 0, woff = 0;
    for (int li = 0; li < 14; li++) {
        ConvLayer l{kLayerShape[li][0], kLayerShape[li][1], kLayerShape[li][2], 0, 0, 0};
        l.w_off = off;
        off += (int64_t)l.cout * 27 * l.cin;
        l.b_off = off;
        off += l.cout;
        l.wbf_off = woff;
        woff += (int64_t)l.cout * 27 * l.cin;
        s->layers.push_back(l);
    }
    s->head_off = off;
    off += 17;
    s->nparams = off;
    s->nwbf = woff;
    // He-style uniform initialisation, bound sqrt(6 / fan_in), from a seeded stream; biases and
    // the head start at zero, so the untrained network predicts no change.
    std::vector<float> h((size_t)off, 0.0f);
    uint64_t st = seed == 0 ? 0x9E3779B97F4A7C15ULL : seed;
    for (int li = 0; li < 14; li++) {
        const ConvLayer &l = s->layers[li];
        const double bound = std::sqrt(6.0 / (27.0 * l.cin));
        const int64_t nw = (int64_t)l.cout * 27 * l.cin;
        for (int64_t i = 0; i < nw; i++) {
            const double u = init_unit(&st) * 2.0 - 1.0;
            h[(size_t)(l.w_off + i)] = (float)(u * bound);
        }
    }
    PM_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    PM_CHECK(s->params.alloc(sizeof(float) * (size_t)off));
    PM_CHECK(s->grads.alloc(sizeof(float) * (size_t)off));
    PM_CHECK(s->m.alloc(sizeof(float) * (size_t)off));
    PM_CHECK(s->v.alloc(sizeof(float) * (size_t)off));
    PM_CHECK(s->wbf.alloc(sizeof(bf16) * (size_t)woff));
    int64_t maxw = 0;
    for (const ConvLayer &l : s->layers) maxw = std::max<int64_t>(maxw, (int64_t)l.cout * 27 * l.cin);
    PM_CHECK(s->wflip.alloc(sizeof(bf16) * (size_t)maxw));
    s->wscratch_floats = 64 * maxw;
    PM_CHECK(s->wscratch.alloc(sizeof(float) * (size_t)s->wscratch_floats));
    PM_HIP(hipMemcpy(s->params.d(), h.data(), sizeof(float) * (size_t)off, hipMemcpyHostToDevice));
    PM_HIP(hipMemset(s->m.d(), 0, sizeof(float) * (size_t)off));
    PM_HIP(hipMemset(s->v.d(), 0, sizeof(float) * (size_t)off));
    PM_CHECK(refresh_bf16(s.get()));
    PM_HIP(hipStreamSynchronize(s->stream));
    *out = s.release();
    return PM_OK;
}

void pm_surrogate_destroy(pm_surrogate *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    hipStream_t st = s->stream;
    delete s;
    if (st) (void)hipStreamDestroy(st);
}

int64_t pm_surrogate_parameter_count(const pm_surrogate *s) { return s ? s->nparams : -1; }

int pm_surrogate_get_parameters(pm_surrogate *s, float *out) {
    if (!s || !out) return PM_ERR_ARGUMENT;
    PM_HIP(hipMemcpy(out, s->params.d(), sizeof(float) * (size_t)s->nparams, hipMemcpyDeviceToHost));
    return PM_OK;
}

int pm_surrogate_set_parameters(pm_surrogate *s, const float *in) {
    if (!s || !in) return PM_ERR_ARGUMENT;
    PM_HIP(hipMemcpy(s->params.d(), in, sizeof(float) * (size_t)s->nparams, hipMemcpyHostToDevice));
    PM_CHECK(refresh_bf16(s));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

int pm_surrogate_reserve_set(pm_surrogate *s, int64_t samples, int64_t edge) {
    if (!s || samples <= 0 || edge <= 0 || samples > INT32_MAX) return PM_ERR_ARGUMENT;
    s->set_n = samples;
    s->set_edge = (int)edge;
    s->set_cells = edge * edge * edge;
    const size_t cells = (size_t)(samples * s->set_cells);
    PM_CHECK(s->set_t0.alloc(sizeof(float) * cells));
    PM_CHECK(s->set_k.alloc(sizeof(float) * cells));
    PM_CHECK(s->set_t.alloc(sizeof(float) * cells));
    PM_CHECK(s->set_a.alloc(sizeof(float) * (size_t)samples));
    return PM_OK;
}

// One training pair from fp64 host arrays: T_0, conductivity and T_k of edge^3 cells, and the
// coupling a = dt * dx / C of the stencil step that produced it.
int pm_surrogate_set_sample(pm_surrogate *s, int64_t index, const double *t0, const double *k,
                            const double *tk, double coupling) {
    if (!s || !t0 || !k || !tk || index < 0 || index >= s->set_n || !(coupling > 0.0))
        return PM_ERR_ARGUMENT;
    const int64_t n = s->set_cells;
    std::vector<float> a((size_t)n), b((size_t)n), t((size_t)n);
    for (int64_t c = 0; c < n; c++) {
        a[(size_t)c] = (float)t0[c];
        b[(size_t)c] = (float)k[c];
        t[(size_t)c] = (float)((tk[c] - t0[c]) / OUT_SCALE);
    }
    const float cf = (float)coupling;
    PM_HIP(hipMemcpy(s->set_t0.as<float>() + index * n, a.data(), sizeof(float) * a.size(), hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->set_k.as<float>() + index * n, b.data(), sizeof(float) * b.size(), hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->set_t.as<float>() + index * n, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->set_a.as<float>() + index, &cf, sizeof(float), hipMemcpyHostToDevice));
    return PM_OK;
}

int pm_surrogate_set_sample_from(pm_surrogate *s, int64_t index, pm_solid3d *before,
                                 pm_solid3d *after, double dt) {
    if (!s || !before || !after || index < 0 || index >= s->set_n || !(dt > 0.0))
        return PM_ERR_ARGUMENT;
    double *t0 = nullptr, *tk = nullptr;
    const double *k0 = nullptr, *kk = nullptr;
    int64_t c0[3], ck[3];
    hipStream_t s0, sk;
    bool uniform_c = false;
    double mob = 0.0;
    PM_CHECK(solid_device_state(before, &t0, &k0, c0, &s0, &uniform_c, &mob));
    PM_CHECK(solid_device_state(after, &tk, &kk, ck, &sk, nullptr, nullptr));
    if (!k0 || !uniform_c || c0[0] != s->set_edge || c0[1] != s->set_edge ||
        c0[2] != s->set_edge || ck[0] != c0[0] || ck[1] != c0[1] || ck[2] != c0[2]) {
        set_error("pm_surrogate_set_sample_from: blocks must match the set's edge and carry "
                  "per-cell conductivity with uniform capacity");
        return PM_ERR_ARGUMENT;
    }
    PM_HIP(hipStreamSynchronize(s0));
    PM_HIP(hipStreamSynchronize(sk));
    const int64_t n = s->set_cells;
    nn::pack_pair<<<grid_of(n), 256, 0, s->stream>>>(t0, k0, tk, n, s->set_t0.as<float>() + index * n,
                                                   s->set_k.as<float>() + index * n,
                                                   s->set_t.as<float>() + index * n);
    PM_HIP(hipGetLastError());
    const float cf = (float)(dt * mob);
    PM_HIP(hipMemcpyAsync(s->set_a.as<float>() + index, &cf, sizeof(float), hipMemcpyHostToDevice, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

}  // extern "C"

namespace {

// Input features and targets of a batch of the training set, then the forward pass and the loss
// gradient; returns the loss partials' count.
int batch_forward_loss(pm_surrogate *s, const int *indices, int batch, unsigned *lb_out) {
    PM_CHECK(check_indices(s, indices, batch));
    PM_CHECK(size_buffers(s, s->set_edge, batch));
    hipStream_t st = s->stream;
    const int64_t n = s->set_cells;
    PM_HIP(hipMemcpyAsync(s->idx.d(), indices, sizeof(int) * (size_t)batch, hipMemcpyHostToDevice, st));
    nn::pack_batch<<<grid_of(n * batch), 256, 0, st>>>(s->set_t0.as<float>(), s->set_k.as<float>(),
                                                     s->set_a.as<float>(), s->idx.as<int>(),
                                                     s->set_edge, batch, s->xin.as<bf16>());
    nn::gather_target<<<grid_of(n * batch), 256, 0, st>>>(s->set_t.as<float>(), s->idx.as<int>(), n, batch, s->target.as<float>());
    PM_HIP(hipGetLastError());
    PM_CHECK(run_forward(s, true));
    const int64_t n0 = npos(s, 0);
    const unsigned lb = std::min<unsigned>(grid_of(n0), 4096);
    nn::loss_fwd_bwd<<<lb, 256, 0, st>>>(s->y.as<float>(), s->target.as<float>(), n0, s->dy.as<float>(), s->parts.as<float>());
    PM_HIP(hipGetLastError());
    *lb_out = lb;
    return PM_OK;
}

double host_loss(pm_surrogate *s, const std::vector<float> &parts) {
    double tot = 0.0;
    for (float p : parts) tot += p;
    return 0.5 * tot / (double)npos(s, 0);
}

}  // namespace

extern "C" {

// One Adam step on the samples `indices[0..batch)` of the training set; `loss` receives
// 0.5 * mean squared error of the normalised change before the step.
int pm_surrogate_train_step(pm_surrogate *s, const int *indices, int batch, float lr, double *loss) {
    if (!s || !indices || batch <= 0 || s->set_n == 0 || !(lr > 0.0f)) return PM_ERR_ARGUMENT;
    hipStream_t st = s->stream;
    unsigned lb = 0;
    PM_CHECK(batch_forward_loss(s, indices, batch, &lb));
    PM_CHECK(run_backward(s));
    s->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)s->adam_t), c2 = 1.0f - std::pow(b2, (float)s->adam_t);
    nn::adam_step<<<grid_of(s->nparams), 256, 0, st>>>(s->params.as<float>(), s->grads.as<float>(), s->m.as<float>(), s->v.as<float>(), s->nparams, lr, b1, b2, 1e-8f, c1, c2);
    PM_HIP(hipGetLastError());
    PM_CHECK(refresh_bf16(s));
    std::vector<float> parts(lb);
    PM_HIP(hipMemcpyAsync(parts.data(), s->parts.d(), sizeof(float) * lb, hipMemcpyDeviceToHost, st));
    PM_HIP(hipStreamSynchronize(st));
    if (loss) *loss = host_loss(s, parts);
    return PM_OK;
}

// Gradients of 0.5 * mean((y - target)^2) for the given batch, without a step: the flat fp32
// gradient array (same layout as the parameters).
int pm_surrogate_gradients(pm_surrogate *s, const int *indices, int batch, float *grads, double *loss) {
    if (!s || !indices || !grads || batch <= 0 || s->set_n == 0) return PM_ERR_ARGUMENT;
    hipStream_t st = s->stream;
    unsigned lb = 0;
    PM_CHECK(batch_forward_loss(s, indices, batch, &lb));
    PM_CHECK(run_backward(s));
    std::vector<float> parts(lb);
    PM_HIP(hipMemcpyAsync(parts.data(), s->parts.d(), sizeof(float) * lb, hipMemcpyDeviceToHost, st));
    PM_HIP(hipMemcpyAsync(grads, s->grads.d(), sizeof(float) * (size_t)s->nparams, hipMemcpyDeviceToHost, st));
    PM_HIP(hipStreamSynchronize(st));
    if (loss) *loss = host_loss(s, parts);
    return PM_OK;
}

// The network output (normalised change, fp32) for fp64 host T and k of edge^3 cells and the
// coupling a = dt * dx / C.
int pm_surrogate_forward(pm_surrogate *s, int64_t edge, const double *t, const double *k,
                         double coupling, float *y) {
    if (!s || !t || !k || !y || edge <= 0 || edge > 4096 || !(coupling > 0.0))
        return PM_ERR_ARGUMENT;
    PM_CHECK(size_buffers(s, (int)edge, 1));
    const int64_t n = edge * edge * edge;
    DevBuf dt, dk;
    PM_CHECK(dt.alloc(sizeof(double) * (size_t)n));
    PM_CHECK(dk.alloc(sizeof(double) * (size_t)n));
    PM_HIP(hipMemcpyAsync(dt.d(), t, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, s->stream));
    PM_HIP(hipMemcpyAsync(dk.d(), k, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, s->stream));
    nn::pack_input<<<nn::pack_grid(edge), nn::PACK_BLOCK, 0, s->stream>>>(dt.d(), dk.d(), (int)edge, (float)coupling, s->xin.as<bf16>());
    PM_CHECK(run_forward(s, false));
    PM_HIP(hipMemcpyAsync(y, s->y.d(), sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

// One surrogate step applied to a cubic Solid3D with per-cell conductivity and uniform capacity,
// in place on the device: T <- T + 10 K * net(features(T, k, dt)), then shifted uniformly so
// sum(T) (the block's total heat, C uniform) is what it was; the shift is computed and applied on
// the device.  `drift` receives |sum T' - sum T| / |sum (T - T_ref)|, the audit's view of the
// step; `elapsed_ms` the device time of the whole step.
int pm_surrogate_step_solid(pm_surrogate *s, pm_solid3d *solid, double dt, double reference,
                            double *drift, double *elapsed_ms) {
    if (!s || !solid || !(dt > 0.0)) return PM_ERR_ARGUMENT;
    double *cells = nullptr;
    const double *kc = nullptr;
    int64_t counts[3];
    hipStream_t sst;
    bool uniform_c = false;
    double mob = 0.0;
    PM_CHECK(solid_device_state(solid, &cells, &kc, counts, &sst, &uniform_c, &mob));
    if (!kc || !uniform_c || counts[0] != counts[1] || counts[1] != counts[2]) {
        set_error("pm_surrogate_step_solid needs a cubic block with per-cell conductivity and "
                  "uniform capacity");
        return PM_ERR_ARGUMENT;
    }
    const int64_t edge = counts[0], n = edge * edge * edge;
    PM_HIP(hipStreamSynchronize(sst));
    PM_CHECK(size_buffers(s, (int)edge, 1));
    hipStream_t st = s->stream;
    constexpr unsigned MAXB = 4096;
    if (s->red64.bytes() < (3 * MAXB + 8) * sizeof(double))
        PM_CHECK(s->red64.alloc((3 * MAXB + 8) * sizeof(double)));
    double *part0 = s->red64.d(), *part1 = part0 + MAXB, *part2 = part1 + MAXB;
    double *scal = part2 + MAXB;
    if (!s->ev0) PM_HIP(hipEventCreate(&s->ev0));
    if (!s->ev1) PM_HIP(hipEventCreate(&s->ev1));
    const unsigned rb = std::min<unsigned>(grid_of(n), MAXB);
    PM_HIP(hipEventRecord(s->ev0, st));
    nn::pack_input<<<nn::pack_grid(edge), nn::PACK_BLOCK, 0, st>>>(cells, kc, (int)edge, (float)(dt * mob), s->xin.as<bf16>());
    PM_CHECK(run_forward(s, false));
    nn::apply_prediction_sum<<<rb, 256, 0, st>>>(cells, s->y.as<float>(), n, part0, part1);
    nn::shift_finalize<<<1, 1024, 0, st>>>(part0, part1, rb, n, scal);
    nn::shift_sum<<<rb, 256, 0, st>>>(cells, n, scal, part2);
    nn::drift_finalize<<<1, 1024, 0, st>>>(part2, rb, scal);
    PM_HIP(hipGetLastError());
    PM_HIP(hipEventRecord(s->ev1, st));
    double h[4];
    PM_HIP(hipMemcpyAsync(h, scal, sizeof(h), hipMemcpyDeviceToHost, st));
    PM_HIP(hipStreamSynchronize(st));
    const double before = h[0], fixed = h[3];
    const double stored = before - reference * (double)n;
    if (drift) *drift = std::fabs(fixed - before) / std::max(std::fabs(stored), 1e-300);
    float ms = 0.0f;
    PM_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    if (elapsed_ms) *elapsed_ms = ms;
    return PM_OK;
}

// ||T_a - T_b|| / ||T_b - reference|| between two blocks' temperatures, on the device.
int pm_solid3d_relative_l2(pm_solid3d *a, pm_solid3d *b, double reference, double *rel) {
    if (!rel) return PM_ERR_ARGUMENT;
    double diff = 0.0, norm = 0.0;
    PM_CHECK(pm_solid3d_l2_parts(a, b, reference, &diff, &norm));
    *rel = diff / norm;
    return PM_OK;
}

int pm_solid3d_l2_parts(pm_solid3d *a, pm_solid3d *b, double reference, double *diff, double *norm) {
    if (!a || !b || !diff || !norm) return PM_ERR_ARGUMENT;
    double *ca = nullptr, *cb = nullptr;
    const double *ka, *kb;
    int64_t na[3], nb[3];
    hipStream_t sa, sb;
    PM_CHECK(solid_device_state(a, &ca, &ka, na, &sa, nullptr, nullptr));
    PM_CHECK(solid_device_state(b, &cb, &kb, nb, &sb, nullptr, nullptr));
    if (na[0] != nb[0] || na[1] != nb[1] || na[2] != nb[2]) return PM_ERR_ARGUMENT;
    PM_HIP(hipStreamSynchronize(sa));
    PM_HIP(hipStreamSynchronize(sb));
    const int64_t n = na[0] * na[1] * na[2];
    const unsigned rb = std::min<unsigned>(grid_of(n), 4096);
    DevBuf part;
    PM_CHECK(part.alloc(sizeof(double) * 2 * rb));
    nn::sq_diff_f64<<<rb, 256, 0, sa>>>(ca, cb, reference, n, part.d());
    std::vector<double> h(2 * rb);
    PM_HIP(hipMemcpyAsync(h.data(), part.d(), sizeof(double) * 2 * rb, hipMemcpyDeviceToHost, sa));
    PM_HIP(hipStreamSynchronize(sa));
    double e = 0.0, q = 0.0;
    for (unsigned i = 0; i < rb; i++) e += h[2 * i], q += h[2 * i + 1];
    *diff = std::sqrt(e);
    *norm = std::sqrt(q);
    return PM_OK;
}

// Test entry points for one convolution on host fp32 arrays (converted to bf16 on the way in):
// x [B][D][H][W][Cin], w [Cout][27][Cin], bias [Cout], y [B][D][H][W][Cout] (pre-activation).
int pm_nn_conv3d(int B, int D, int H, int W, int Cin, int Cout, const float *x, const float *w,
                 const float *bias, float *y, int split) {
    PM_CHECK(require_device());
    const nn::Conv g = nn::make_conv(B, D, H, W, Cin, Cout);
    const int64_t nx = g.positions() * Cin, ny = g.positions() * Cout, nw = (int64_t)Cout * 27 * Cin;
    std::vector<bf16> hx((size_t)nx), hw((size_t)nw);
    for (int64_t i = 0; i < nx; i++) hx[(size_t)i] = to_bf16(x[i]);
    for (int64_t i = 0; i < nw; i++) hw[(size_t)i] = to_bf16(w[i]);
    DevBuf dx, dw, db, dz;
    PM_CHECK(dx.alloc(sizeof(bf16) * (size_t)nx));
    PM_CHECK(dw.alloc(sizeof(bf16) * (size_t)nw));
    PM_CHECK(db.alloc(sizeof(float) * (size_t)Cout));
    PM_CHECK(dz.alloc(sizeof(bf16) * (size_t)ny));
    PM_HIP(hipMemcpy(dx.d(), hx.data(), sizeof(bf16) * hx.size(), hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(dw.d(), hw.data(), sizeof(bf16) * hw.size(), hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(db.d(), bias, sizeof(float) * (size_t)Cout, hipMemcpyHostToDevice));
    const int64_t nwork = split ? nn::conv_forward_workspace(g) : 0;
    DevBuf work;
    PM_CHECK(work.alloc(sizeof(float) * (size_t)std::max<int64_t>(nwork, 1)));
    PM_CHECK(nn::conv_forward(g, dx.as<bf16>(), dw.as<bf16>(), db.as<float>(), dz.as<bf16>(), nullptr, false, Cout, 0, nullptr,
                              work.as<float>(), nwork));
    std::vector<bf16> hz((size_t)ny);
    PM_HIP(hipMemcpy(hz.data(), dz.d(), sizeof(bf16) * hz.size(), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < ny; i++) y[i] = from_bf16(hz[(size_t)i]);
    return PM_OK;
}

// Gradients of sum(y * dy) for the convolution above: dx [B][D][H][W][Cin], dw [Cout][27][Cin],
// db [Cout].
int pm_nn_conv3d_backward(int B, int D, int H, int W, int Cin, int Cout, const float *x,
                          const float *w, const float *dy, float *dx, float *dw, float *db,
                          int split) {
    PM_CHECK(require_device());
    const nn::Conv g = nn::make_conv(B, D, H, W, Cin, Cout);
    const int64_t nx = g.positions() * Cin, ny = g.positions() * Cout, nw = (int64_t)Cout * 27 * Cin;
    std::vector<bf16> hx((size_t)nx), hw((size_t)nw), hd((size_t)ny);
    for (int64_t i = 0; i < nx; i++) hx[(size_t)i] = to_bf16(x[i]);
    for (int64_t i = 0; i < nw; i++) hw[(size_t)i] = to_bf16(w[i]);
    for (int64_t i = 0; i < ny; i++) hd[(size_t)i] = to_bf16(dy[i]);
    DevBuf ddx, ddw, ddy, wf, gx, gw, gb, scratch;
    PM_CHECK(ddx.alloc(sizeof(bf16) * (size_t)nx));
    PM_CHECK(ddw.alloc(sizeof(bf16) * (size_t)nw));
    PM_CHECK(ddy.alloc(sizeof(bf16) * (size_t)ny));
    PM_CHECK(wf.alloc(sizeof(bf16) * (size_t)nw));
    PM_CHECK(gx.alloc(sizeof(bf16) * (size_t)nx));
    PM_CHECK(gw.alloc(sizeof(float) * (size_t)nw));
    PM_CHECK(gb.alloc(sizeof(float) * (size_t)Cout));
    const int splits = (int)std::min<int64_t>(16, g.tiles());
    PM_CHECK(scratch.alloc(sizeof(float) * (size_t)(splits * (nw + Cout))));
    PM_HIP(hipMemcpy(ddx.d(), hx.data(), sizeof(bf16) * hx.size(), hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(ddw.d(), hw.data(), sizeof(bf16) * hw.size(), hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(ddy.d(), hd.data(), sizeof(bf16) * hd.size(), hipMemcpyHostToDevice));
    const int64_t nwork = split ? nn::conv_forward_workspace(nn::transposed(g)) : 0;
    DevBuf work;
    PM_CHECK(work.alloc(sizeof(float) * (size_t)std::max<int64_t>(nwork, 1)));
    PM_CHECK(nn::conv_backward_data(g, ddy.as<bf16>(), ddw.as<bf16>(), wf.as<bf16>(), gx.as<bf16>(), nullptr,
                                    work.as<float>(), nwork));
    PM_CHECK(nn::conv_backward_weight(g, ddx.as<bf16>(), ddy.as<bf16>(), gw.as<float>(), gb.as<float>(), scratch.as<float>(), splits, nullptr));
    std::vector<bf16> hgx((size_t)nx);
    PM_HIP(hipMemcpy(hgx.data(), gx.d(), sizeof(bf16) * hgx.size(), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nx; i++) dx[i] = from_bf16(hgx[(size_t)i]);
    PM_HIP(hipMemcpy(dw, gw.d(), sizeof(float) * (size_t)nw, hipMemcpyDeviceToHost));
    PM_HIP(hipMemcpy(db, gb.d(), sizeof(float) * (size_t)Cout, hipMemcpyDeviceToHost));
    return PM_OK;
}

}  // extern "C"
