// Solid3D on gfx950: the explicit conduction sweep of crates/pantometry-thermal/src/solid.rs
// (plain branch), with the conservation audit fused into the sweep's store.
//
// Layout in HBM: x-fastest `i + nx*(j + ny*k)`, f64.  Two temperature buffers ping-pong (the
// reference's per-step `old = cells.clone()` becomes a pointer swap).  The operator is held in
// one of three forms, chosen by what the cells hold:
//   uniform      one conductivity, one capacity: face k and dt*dx/C are kernel scalars (16 B/cell)
//   conductive   per-cell conductivity, one capacity: the three new faces of each cell are
//                recomputed with the harmonic mean in the sweep, from the per-cell k (24 B/cell)
//   general      per-cell conductivity, mobility dx/C and capacity (40 B/cell)
// Every form is arithmetically the reference's: faces are series(k[c-1], k[c]) as resolve
// builds them (solid.rs:716-734), and each cell's update follows solid.rs:1776-1826 term by
// term with contraction off.
//
// The sweep marches in z.  A workgroup of 64 x TY threads owns a 64 x TY column tile and a
// chunk of z-planes; each thread keeps its own column's below/current/above temperatures and
// conductivities in registers, stages the current xy plane (with a one-cell halo) in LDS for
// its x and y neighbours, and carries the z face it computed for the plane above into the next
// plane.  The audit's C*(T'-T_ref) is summed per thread as the cell is stored, reduced per
// workgroup in a fixed order, and the per-workgroup partials are summed by a one-block finalize
// kernel in a fixed order, which also performs audit_with on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace pm {

struct HeatArgs {
    const double *__restrict__ told;
    double *__restrict__ tnew;
    const double *__restrict__ kcell;  // per-cell conductivity (conductive, general)
    const double *__restrict__ mob;    // per-cell dx/C (general)
    const double *__restrict__ cap;    // per-cell C (general)
    double *__restrict__ partials;     // one audit partial per workgroup
    const AuditState *audit;           // halted flag: copy instead of step
    double kface;                      // uniform face conductivity
    double f;                          // dt * dx/C (uniform capacity)
    double dts;                        // dt (general)
    double cap0;                       // uniform capacity
    double tref;                       // ledger reference temperature
    int nx, ny, nz, zc;
    int gx, gy, gz;                    // tiles along x, y and z
    int force;                         // run even if an audit breach froze the run
};

enum { K_UNIFORM = 0, K_PER_CELL = 1 };
enum { C_UNIFORM = 0, C_PER_CELL = 1 };

// Workgroups of one launch are spread over the XCDs round-robin (workgroup i runs on XCD i % 8).
// The sweep gives each XCD a contiguous run of tiles in x-fastest order, so a tile's x and y
// neighbours, whose rows it reads as halo, run on the same XCD and are found in its L2.
constexpr int XCDS = 8;


// The faces of one plane from the conductivities staged in LDS: each thread builds its own low x
// and y faces; the first TY threads add the tile's right-edge x faces and the last row of threads
// the top-edge y faces.  Face g along an axis of n cells lies between cells g-1 and g and is zero
// unless 0 < g < n (solid.rs:716-734).
// The harmonic face mean with the zero guard as a select: for two zero conductivities the
// division yields NaN and the select returns 0, as series() does.
__device__ __forceinline__ double series_nz(double a, double b) {
    const double sum = a + b;
    const double q = 2.0 * a * b / sum;
    return sum > 0.0 ? q : 0.0;
}

// n / d by the steps of the IEEE double division (reciprocal, two Newton steps, one residual
// correction) without its operand scaling and special-case fix-up.  Those two leave the operands
// and the quotient unchanged when d is a positive normal number, n a positive number of at least
// 2^DIV_N_MIN_EXP whose exponent exceeds d's by less than DIV_MAX_GAP, and n / d is normal, so
// there the result is the correctly rounded quotient bit for bit; for n = +0 and a positive
// normal d it is +0, as the division's is.
constexpr int DIV_N_MIN_EXP = -969;
constexpr int DIV_MAX_GAP = 768;
__device__ __forceinline__ double div_unscaled(double n, double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    const double q = n * r;
    e = __builtin_fma(-d, q, n);
    return __builtin_fma(e, r, q);
}

// Conductivities for which every harmonic face mean may use div_unscaled: +0, or in
// [2^K_SHORT_LO_EXP, 2^K_SHORT_HI_EXP].
constexpr int K_SHORT_LO_EXP = -190, K_SHORT_HI_EXP = 190;
constexpr double K_SHORT_LO = 0x1p-190, K_SHORT_HI = 0x1p190;
static_assert(K_SHORT_LO * 0x1p190 == 1.0 && K_SHORT_HI * 0x1p-190 == 1.0,
              "the bounds are 2^K_SHORT_LO_EXP and 2^K_SHORT_HI_EXP");
// For a, b in that range the numerator 2ab lies in [2^(1+2 lo), 2^(1+2 hi)], the denominator
// a + b in [2^(lo+1), 2^(hi+1)], and the quotient (the harmonic mean, between min(a, b) and
// 2 min(a, b)) in [2^lo, 2^(hi+1)]; so n's exponent exceeds d's by at most hi + 2.  These are the
// div_unscaled preconditions, and a = +0 gives 2ab = +0.
static_assert(1 + 2 * K_SHORT_LO_EXP >= DIV_N_MIN_EXP, "numerator below div_unscaled's range");
static_assert(1 + 2 * K_SHORT_HI_EXP < 1024, "numerator overflows");
static_assert(K_SHORT_LO_EXP + 1 >= -1022 && K_SHORT_HI_EXP + 1 < 1024, "denominator not normal");
static_assert(K_SHORT_LO_EXP >= -1022 && K_SHORT_HI_EXP + 1 < 1024, "quotient not normal");
static_assert(K_SHORT_HI_EXP + 2 < DIV_MAX_GAP, "exponent gap beyond div_unscaled's range");
__device__ __forceinline__ bool k_short_ok(double v) {
    return __double_as_longlong(v) == 0 || (v >= K_SHORT_LO && v <= K_SHORT_HI);
}

// series_nz(a, b) for a face that exists (`on`), else 0, with one select for both guards; with
// SHORT, for conductivities that all pass k_short_ok, the division is div_unscaled.
template <bool SHORT = false>
__device__ __forceinline__ double series_on(double a, double b, bool on) {
    const double sum = a + b;
    const double q = SHORT ? div_unscaled(2.0 * a * b, sum) : 2.0 * a * b / sum;
    return (on && sum > 0.0) ? q : 0.0;
}

// Sets *outside when some conductivity fails k_short_ok.
__global__ __launch_bounds__(256) void k_short_check(const double *k, int64_t n, int *outside) {
    bool bad = false;
    for (int64_t c = blockIdx.x * 256 + threadIdx.x; c < n; c += (int64_t)gridDim.x * 256)
        bad = bad || !k_short_ok(k[c]);
    if (bad) atomicOr(outside, 1);
}

// Clears the flag k_short_check sets (a kernel rather than a memset, so a captured graph holds
// no memset node).
__global__ void clear_flag(int *flag) { *flag = 0; }

template <int TY, bool SHORT = false>
__device__ __forceinline__ void build_faces(const double (*sk)[66], double (*fx)[65],
                                            double (*fy)[64], int tx, int ty, int tid,
                                            bool fxl_on, bool fyl_on, bool fxe_on, bool fye_on) {
    constexpr int TX = 64;
    const double c = sk[ty + 1][tx + 1];
    fx[ty][tx] = series_on<SHORT>(sk[ty + 1][tx], c, fxl_on);
    fy[ty][tx] = series_on<SHORT>(sk[ty][tx + 1], c, fyl_on);
    if (tid < TY) fx[tid][TX] = series_on<SHORT>(sk[tid + 1][TX], sk[tid + 1][TX + 1], fxe_on);
    if (ty == TY - 1) fy[TY][tx] = series_on<SHORT>(sk[TY][tx + 1], sk[TY + 1][tx + 1], fye_on);
}

// One sweep (solid.rs:1776-1826) over a 64 x TY column tile and a chunk of z-planes, with the
// audit's C*(T'-T_ref) summed as the cells are stored.
//
// Per plane z a thread stages T(z) and, in the per-cell-conductivity form, k(z+1) in LDS, then
// one barrier: the x/y faces of plane z were built from k(z) in the previous iteration and the
// faces of plane z+1 are built now, so each plane needs a single barrier.  The own column's
// T(z-1..z+2) and k(z..z+3) live in four-register rings and the halo values in two-register
// rings; the z loop is unrolled by four so every ring slot has a fixed register and a load
// issued for a later plane is not waited on until that plane's iteration uses it.
template <int KMODE, int CMODE, int TY, bool SHORT = false>
__global__ __launch_bounds__(64 * TY) void heat_sweep(HeatArgs a) {
    constexpr int TX = 64;
    constexpr int NT = TX * TY;
    constexpr bool KC = KMODE == K_PER_CELL;
    __shared__ double sT[2][TY + 2][TX + 2];
    __shared__ double sK[KC ? 2 : 1][KC ? TY + 2 : 1][TX + 2];
    __shared__ double fx[KC ? 2 : 1][KC ? TY : 1][TX + 1];
    __shared__ double fy[KC ? 2 : 1][KC ? TY + 1 : 1][TX];
    __shared__ double red[NT / 64];

    const int64_t ntiles = (int64_t)a.gx * a.gy * a.gz;
    const int64_t chunk = (ntiles + XCDS - 1) / XCDS;
    const int64_t tile = (int64_t)(blockIdx.x % XCDS) * chunk + blockIdx.x / XCDS;
    if (tile >= ntiles) return;
    const int64_t bx = tile % a.gx, by = (tile / a.gx) % a.gy, bz = tile / ((int64_t)a.gx * a.gy);

    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
    const int64_t nx = a.nx, ny = a.ny, plane = nx * ny;
    const int64_t x0 = bx * TX, y0 = by * TY;
    // Plane indices are uniform 32-bit values, so their clamps and tests stay on the scalar unit.
    const int nz = a.nz;
    const int z0 = (int)bz * a.zc, z1 = std::min(z0 + a.zc, nz);
    const int64_t gi = x0 + tx, gj = y0 + ty;
    const bool inside = gi < nx && gj < ny;
    const int64_t col = std::min(gi, nx - 1) + nx * std::min(gj, ny - 1);

    // A breach froze the run: later launches leave both buffers alone, and the host takes the
    // state from the launch that broke.
    if (a.audit->halted && !a.force) return;

    // Halo role: one halo cell per thread for the first 2*TX + 2*TY threads (bottom row, top row,
    // left column, right column), read from the clamped index so every LDS value is finite.
    int hr = -1, hc = -1;
    int64_t hidx = 0;
    if (tid < TX) {
        hr = 0, hc = tid + 1;
        hidx = std::min(x0 + tid, nx - 1) + nx * std::max<int64_t>(y0 - 1, 0);
    } else if (tid < 2 * TX) {
        hr = TY + 1, hc = tid - TX + 1;
        hidx = std::min(x0 + tid - TX, nx - 1) + nx * std::min(y0 + TY, ny - 1);
    } else if (tid < 2 * TX + TY) {
        hr = tid - 2 * TX + 1, hc = 0;
        hidx = std::max<int64_t>(x0 - 1, 0) + nx * std::min(y0 + hr - 1, ny - 1);
    } else if (tid < 2 * TX + 2 * TY) {
        hr = tid - 2 * TX - TY + 1, hc = TX + 1;
        hidx = std::min(x0 + TX, nx - 1) + nx * std::min(y0 + hr - 1, ny - 1);
    }
    const bool halo = hr >= 0;

    const double *__restrict__ T = a.told;
    const double *__restrict__ K = a.kcell;
    // Byte offsets within a plane (planes of this sweep hold fewer than 2^29 cells): each access
    // is a buffer access at the plane's scalar address plus one 32-bit lane offset.  A lane with
    // no halo cell loads its own column in the halo's place, so every lane issues every load and
    // the waits before their uses count exactly.
    const uint32_t ub = (uint32_t)col * 8u, hb = halo ? (uint32_t)hidx * 8u : ub;
    // Plane q clamped to [0, min(z1, nz-1)]: a plane past what the chunk uses is never fetched
    // from HBM, only re-read from a plane the chunk already holds.
    const int top = std::min(z1, nz - 1);
    auto at = [&](const double *p, int q) -> const double * {
        return p + plane * std::min(std::max(q, 0), top);
    };
    auto ld = [&](const double *p, int q, uint32_t off) { return buffer_load_f64(at(p, q), off); };

    // Faces on the block's outer boundary are zero (solid.rs:189-195); which ones a thread meets
    // does not change along z.
    const bool fxl_on = gi > 0 && gi < nx, fyl_on = gj > 0 && gj < ny;
    const bool fxe_on = x0 + TX < nx, fye_on = y0 + TY < ny;

    // Rings, indexed by (z - z0) modulo their length.
    double RT[4], RK[4], HT[2], HK[2], KZ[2], RM[2], RC[2];
    if constexpr (CMODE == C_PER_CELL) {
        RM[0] = ld(a.mob, z0, ub), RM[1] = ld(a.mob, z0 + 1, ub);
        RC[0] = ld(a.cap, z0, ub), RC[1] = ld(a.cap, z0 + 1, ub);
    }
    RT[3] = ld(T, z0 - 1, ub);
    RT[0] = ld(T, z0, ub);
    RT[1] = ld(T, z0 + 1, ub);
    RT[2] = ld(T, z0 + 2, ub);
    HT[0] = ld(T, z0, hb);
    HT[1] = ld(T, z0 + 1, hb);
    if constexpr (KC) {
        const double k_m = ld(K, z0 - 1, ub);
        RK[0] = ld(K, z0, ub);
        RK[1] = ld(K, z0 + 1, ub);
        RK[2] = ld(K, z0 + 2, ub);
        RK[3] = ld(K, z0 + 3, ub);
        const double h_k0 = ld(K, z0, hb);
        HK[0] = ld(K, z0 + 1, hb);
        HK[1] = ld(K, z0 + 2, hb);
        const double lo = series_nz(k_m, RK[0]), hi = series_nz(RK[0], RK[1]);
        KZ[0] = z0 > 0 ? lo : 0.0;
        KZ[1] = z0 + 1 < nz ? hi : 0.0;
        // The faces of plane z0, from k(z0) staged in the second buffer.
        sK[1][ty + 1][tx + 1] = RK[0];
        if (halo) sK[1][hr][hc] = h_k0;
        lds_barrier();
        build_faces<TY, SHORT>(sK[1], fx[0], fy[0], tx, ty, tid, fxl_on, fyl_on, fxe_on, fye_on);
    }

    // Uniform-form faces depend only on position.
    const double kxl_u = fxl_on ? a.kface : 0.0;
    const double kxh_u = (gi + 1 < nx) ? a.kface : 0.0;
    const double kyl_u = fyl_on ? a.kface : 0.0;
    const double kyh_u = (gj + 1 < ny) ? a.kface : 0.0;
    double *__restrict__ Tout = a.tnew;

    double acc = 0.0;
    for (int zb = z0; zb < z1; zb += 4) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            // A chunk whose length is not a multiple of four runs up to three idle planes at its
            // end (loads clamped, nothing stored) so the ring keeps its fixed registers.
            const int z = zb + i;
            const bool live = z < z1;
            const int b = i & 1;
            const double t_m = RT[(i + 3) & 3], t_c = RT[i], t_p = RT[(i + 1) & 3];

            sT[b][ty + 1][tx + 1] = t_c;
            if (halo) sT[b][hr][hc] = HT[b];
            HT[b] = ld(T, z + 2, hb);
            if constexpr (KC) {
                sK[b][ty + 1][tx + 1] = RK[(i + 1) & 3];
                if (halo) sK[b][hr][hc] = HK[b];
                HK[b] = ld(K, z + 3, hb);
            }
            lds_barrier();

            double kxl, kxh, kyl, kyh, kzl, kzh;
            if constexpr (KC) {
                build_faces<TY, SHORT>(sK[b], fx[b ^ 1], fy[b ^ 1], tx, ty, tid, fxl_on, fyl_on,
                                       fxe_on, fye_on);
                kxl = fx[b][ty][tx], kxh = fx[b][ty][tx + 1];
                kyl = fy[b][ty][tx], kyh = fy[b][ty + 1][tx];
                kzl = KZ[b], kzh = KZ[b ^ 1];
                // the face between z+1 and z+2
                KZ[b] = series_on<SHORT>(RK[(i + 1) & 3], RK[(i + 2) & 3], z + 2 < nz);
                RK[i] = ld(K, z + 4, ub);  // k(z) is no longer needed
            } else {
                kxl = kxl_u, kxh = kxh_u, kyl = kyl_u, kyh = kyh_u;
                kzl = z > 0 ? a.kface : 0.0;
                kzh = z + 1 < nz ? a.kface : 0.0;
            }

            const double tw = sT[b][ty + 1][tx], te = sT[b][ty + 1][tx + 2];
            const double ts = sT[b][ty][tx + 1], tn = sT[b][ty + 2][tx + 1];
            // solid.rs:1781-1798, in the reference's order: x-, x+, y-, y+, z-, z+.
            double flux = 0.0;
            flux = flux + kxl * (tw - t_c);
            flux = flux + kxh * (te - t_c);
            flux = flux + kyl * (ts - t_c);
            flux = flux + kyh * (tn - t_c);
            flux = flux + kzl * (t_m - t_c);
            flux = flux + kzh * (t_p - t_c);
            RT[(i + 3) & 3] = ld(T, z + 3, ub);  // T(z-1) is no longer needed
            double mob_z = 0.0, cap_z = a.cap0;
            if constexpr (CMODE == C_PER_CELL) {
                mob_z = RM[b], cap_z = RC[b];
                RM[b] = ld(a.mob, z + 2, ub);
                RC[b] = ld(a.cap, z + 2, ub);
            }
            if (inside && live) {
                double rise;
                if constexpr (CMODE == C_PER_CELL)
                    rise = (a.dts * mob_z) * flux;
                else
                    rise = a.f * flux;
                const double t_new = t_c + rise;
                buffer_store_f64_nt(t_new, Tout + plane * z, ub);
                acc += cap_z * (t_new - a.tref);
            }
        }
    }

    const double total = block_sum<NT>(acc, red);
    if (tid == 0) a.partials[tile] = total;
}

// ---- Two steps per pass -------------------------------------------------------------------
//
// The per-cell-conductivity sweep (capacity uniform or per cell) advanced two steps in one pass
// over HBM (temporal blocking).  Every cell-step is computed with exactly the arithmetic of
// heat_sweep, so the fields stay bit-identical; what changes is that T is read and written once
// per two steps (12 B per cell-step instead of 24 with uniform capacity; mobility and capacity,
// when per cell, are read once per two steps too), and that each face conductivity, built once
// per plane in LDS, serves both steps.
//
// A workgroup owns a 64 x TY column tile and a z-chunk.  Per plane z it computes the first step,
// T1(z), on the tile widened by one cell (E1, 66 x (TY+2)) less its four corners, which no
// second-step stencil reaches: the tile's cells by their own threads, the 144-cell ring around
// them by the last threads of the workgroup (role "ring 1", each keeping its own column's
// z-rings).  That needs T0 and k on the tile widened by two (E2, 68 x (TY+4)); the outer ring's
// needed cells are fetched by the first threads (role "ring 2").  With T1(z) in LDS and T1(z-2),
// T1(z-1) in registers, the second step T2(z-1) follows on the tile in the same iteration, with
// the faces of plane z-1 still in LDS.  The chunk's first
// step starts one plane early and ends one plane late (planes z0-1 .. z1), which the
// neighbouring chunks also compute: identical arithmetic on identical inputs, so the overlap
// agrees bit for bit.
//
// Per plane every wave performs four rounds of divisions: its own column's z face, and three
// rounds shared out over the 1388 in-plane faces of E1 and the ring's z faces, so no wave
// waits on another at the plane's one barrier.  LDS: T0(z) and k(z+1) on E2 and T1 on E1,
// double-buffered, and the faces of planes z-1, z and z+1 on E1.  The audit of each step is summed as its cells are
// produced: step one over the owned planes of T1, step two over the stored T2.
template <int TY, bool SHORT, int CMODE>
__global__ __launch_bounds__(64 * TY, 4) void heat_sweep2(HeatArgs a) {
    constexpr bool CC = CMODE == C_PER_CELL;
    constexpr int TX = 64, NT = TX * TY;
    constexpr int E1X = TX + 2, E1Y = TY + 2, E2X = TX + 4, E2Y = TY + 4;
    constexpr int R1 = 2 * (E1X - 2) + 2 * TY;         // ring-1 cells (E1 less the tile, corners)
    // Ring-2 cells: those of E2's outer ring the ring-1 stencils reach, and E1's four corners,
    // whose T0 and k the ring-1 stencils and faces next to them read.
    constexpr int R2 = 2 * (E2X - 4) + 2 * (E2Y - 4) + 4;
    constexpr int FXS = E1Y * (E1X + 1), FYS = (E1Y + 1) * E1X;  // face layout: x rows, y rows
    constexpr int NFX = FXS - 4, NFY = FYS - 4;        // faces needed: no corner cell's
    constexpr int NF = NFX + NFY;
    constexpr int F2 = NF - 2 * NT;                    // faces of the third round
    constexpr int RING1_FIRST = NT - R1;
    static_assert(R1 <= NT && R2 <= RING1_FIRST && F2 > 0 && F2 <= RING1_FIRST,
                  "roles do not fit");
    __shared__ double sT[2][E2Y][E2X];
    __shared__ double sK[2][E2Y][E2X];
    __shared__ double sF[4][FXS + FYS];
    __shared__ double s1[2][E1Y][E1X];
    __shared__ double red[NT / 64];

    const int64_t ntiles = (int64_t)a.gx * a.gy * a.gz;
    const int64_t chunk = (ntiles + XCDS - 1) / XCDS;
    const int64_t tile = (int64_t)(blockIdx.x % XCDS) * chunk + blockIdx.x / XCDS;
    if (tile >= ntiles) return;
    if (a.audit->halted && !a.force) return;  // a breach froze the run
    const int64_t bx = tile % a.gx, by = (tile / a.gx) % a.gy, bz = tile / ((int64_t)a.gx * a.gy);

    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
    const int64_t nx = a.nx, ny = a.ny, plane = nx * ny;
    const int64_t x0 = bx * TX, y0 = by * TY;
    // Plane indices are uniform 32-bit values, so their clamps and tests stay on the scalar unit.
    const int nz = a.nz;
    const int z0 = (int)bz * a.zc, z1 = std::min(z0 + a.zc, nz);
    auto colidx = [&](int64_t x, int64_t y) -> uint32_t {
        return (uint32_t)(std::min(std::max<int64_t>(x, 0), nx - 1) +
                          nx * std::min(std::max<int64_t>(y, 0), ny - 1));
    };

    // Tile cell, E1 coordinates (ty+1, tx+1).
    const int64_t gi = x0 + tx, gj = y0 + ty;
    const bool inside = gi < nx && gj < ny;
    const uint32_t ucol = colidx(gi, gj);
    // Ring 1 (the last R1 threads): E1 coordinates (r1r, r1c), corners excluded.
    int r1r = -1, r1c = -1;
    {
        const int q = tid - RING1_FIRST;
        if (q >= 0 && q < E1X - 2) r1r = 0, r1c = 1 + q;
        else if (q >= E1X - 2 && q < 2 * (E1X - 2)) r1r = E1Y - 1, r1c = 1 + q - (E1X - 2);
        else if (q >= 2 * (E1X - 2) && q < 2 * (E1X - 2) + TY) r1r = 1 + q - 2 * (E1X - 2), r1c = 0;
        else if (q >= 2 * (E1X - 2) + TY && q < R1) r1r = 1 + q - 2 * (E1X - 2) - TY, r1c = E1X - 1;
    }
    const bool ring1 = r1r >= 0;
    const uint32_t r1col = ring1 ? colidx(x0 - 1 + r1c, y0 - 1 + r1r) : 0u;
    // Ring 2 (the first R2 threads): E2 coordinates (r2r, r2c).
    int r2r = -1, r2c = -1;
    {
        constexpr int SIDES = 2 * (E2X - 4) + 2 * (E2Y - 4);
        if (tid < E2X - 4) r2r = 0, r2c = 2 + tid;
        else if (tid < 2 * (E2X - 4)) r2r = E2Y - 1, r2c = 2 + tid - (E2X - 4);
        else if (tid < 2 * (E2X - 4) + (E2Y - 4)) r2r = 2 + tid - 2 * (E2X - 4), r2c = 0;
        else if (tid < SIDES) r2r = 2 + tid - 2 * (E2X - 4) - (E2Y - 4), r2c = E2X - 1;
        else if (tid < R2) {
            const int q = tid - SIDES;  // E1's corners
            r2r = (q & 2) ? E2Y - 2 : 1;
            r2c = (q & 1) ? E2X - 2 : 1;
        }
    }
    const bool ring2 = r2r >= 0;
    const uint32_t r2col = ring2 ? colidx(x0 - 2 + r2c, y0 - 2 + r2r) : 0u;
    // Byte offsets within a plane: they fit 32 bits (the paired sweep takes planes of fewer than
    // 2^29 cells), so each access is a uniform plane address plus one 32-bit lane offset.
    // Every lane issues every load (a lane outside a ring re-reads its own column, which the
    // tile's loads bring into cache), so the count of loads in flight is the same on every path
    // and the waits before their uses stay exact.
    const uint32_t ub = ucol * 8u, r1b = ring1 ? r1col * 8u : ub, r2b = ring2 ? r2col * 8u : ub;

    // This thread's in-plane faces, one per round (the third only for the first F2 threads),
    // packed as: bits 0-11 the first conductivity's offset in a k plane on E2, bits 12-23 the
    // face's offset in a face buffer, bit 24 set for a y face (second conductivity one row on,
    // else one column on), bit 25 set when the face lies inside the block (faces on its surface
    // are zero).
    static_assert(E2X * E2Y < 4096 && FXS + FYS < 4096, "face descriptors need 12-bit offsets");
    uint32_t fdesc[3];
#pragma unroll
    for (int rnd = 0; rnd < 3; rnd++) {
        int j = rnd * NT + tid;
        if (j >= NF) j = NF - 1;  // idle lane of the third round: a valid face, never stored
        if (j < NFX) {
            // x faces in row-major order over E1Y x (E1X+1), skipping the corner cells' faces
            int ex = j + 1;
            if (ex >= E1X) ex++;
            if (ex >= FXS - (E1X + 1)) ex++;
            const int r = ex / (E1X + 1), c = ex % (E1X + 1);
            const int64_t x = x0 - 1 + c;
            fdesc[rnd] = (uint32_t)((r + 1) * E2X + c) | ((uint32_t)ex << 12) |
                         ((x > 0 && x < nx) ? (1u << 25) : 0u);
        } else {
            int ey = j - NFX + 1;
            if (ey >= E1X - 1) ey++;
            if (ey >= FYS - E1X) ey++;
            const int r = ey / E1X, c = ey % E1X;
            const int64_t y = y0 - 1 + r;
            fdesc[rnd] = (uint32_t)(r * E2X + c + 1) | ((uint32_t)(FXS + ey) << 12) | (1u << 24) |
                         ((y > 0 && y < ny) ? (1u << 25) : 0u);
        }
    }
    const bool third = tid < F2;
    auto fk = [](uint32_t d) -> int { return (int)(d & 0xFFFu); };
    auto fk2 = [](uint32_t d) -> int { return (int)(d & 0xFFFu) + ((d >> 24) & 1u ? E2X : 1); };
    auto fdst = [](uint32_t d) -> int { return (int)((d >> 12) & 0xFFFu); };
    auto fon = [](uint32_t d) -> bool { return (d >> 25) & 1u; };

    const double *__restrict__ T = a.told;
    const double *__restrict__ K = a.kcell;
    // The chunk reads planes z0-2 .. z1+1; later planes are clamped to what it already reads.
    const int top = std::min(z1 + 1, nz - 1);
    // Plane accesses go through a buffer resource built on the scalar unit from the plane's
    // address, with the 32-bit lane offset as the only per-lane operand (mobility and capacity).
    auto ld = [&](const double *p, int q, uint32_t off) -> double {
        return buffer_load_f64(p + plane * std::min(std::max(q, 0), top), off);
    };
    // T and k are addressed from the chunk's lowest plane through a resource built once, a plane
    // being a uniform byte offset (planes zlo .. top, fewer than 2^32 bytes: the host caps the
    // chunk); the output likewise from plane z0.
    const int zlo = std::max(z0 - 2, 0);
    const __amdgpu_buffer_rsrc_t RTs = buffer_rsrc(T + plane * zlo), RKs = buffer_rsrc(K + plane * zlo);
    const uint32_t pbytes = (uint32_t)plane * 8u;
    auto so = [&](int q) -> uint32_t { return (uint32_t)(std::min(std::max(q, 0), top) - zlo) * pbytes; };
    auto ldT = [&](int q, uint32_t off) -> double { return buffer_load_f64(RTs, off, so(q)); };
    auto ldK = [&](int q, uint32_t off) -> double { return buffer_load_f64(RKs, off, so(q)); };
    auto zon = [&](int q) -> bool { return q > 0 && q < nz; };  // face between q-1 and q
    // One explicit update (solid.rs:1781-1826 order) from the in-plane neighbours in `t` around
    // position p (row stride `stride`), the four in-plane faces and the own column.
    auto update = [&](const double *t, int stride, int p, double kxl, double kxh, double kyl,
                      double kyh, double t_m, double t_c, double t_p, double kzl,
                      double kzh, double rate) -> double {
        const double tw = t[p - 1], te = t[p + 1], ts = t[p - stride], tn = t[p + stride];
        double flux = 0.0;
        flux = flux + kxl * (tw - t_c);
        flux = flux + kxh * (te - t_c);
        flux = flux + kyl * (ts - t_c);
        flux = flux + kyh * (tn - t_c);
        flux = flux + kzl * (t_m - t_c);
        flux = flux + kzh * (t_p - t_c);
        return t_c + rate * flux;
    };
    // The factor of the flux: dt * dx/C, one for the block or (dt * mobility) per cell, as the
    // one-step sweep forms it.
    auto rate_of = [&](double mob) { return CC ? a.dts * mob : a.f; };
    // The x faces of E1 cell (r, c) are at r*(E1X+1) + c and + 1, the y faces at
    // FXS + r*E1X + c and FXS + (r+1)*E1X + c.
    const int fxo = (ty + 1) * (E1X + 1) + tx + 1, fyo = FXS + (ty + 1) * E1X + tx + 1;
    const int r1fx = ring1 ? r1r * (E1X + 1) + r1c : 0, r1fy = ring1 ? FXS + r1r * E1X + r1c : 0;

    // Rings, indexed by (plane - zs) mod 4.
    const int zs = z0 - 1;
    double RT[4], RK[4], KZ[4], U1[4];  // tile column: T0, k, z face below, T1
    // Ring-1 column: T0, k and z faces.  A ring-2 thread, never a ring-1 thread, keeps the T0(z)
    // and k(z+1) of its next staging in QT[0..1] and QK[0..1] (indexed by plane parity).
    double QT[4], QK[4], QZ[4];
    RT[3] = ldT(zs - 1, ub);
    RT[0] = ldT(zs, ub);
    RT[1] = ldT(zs + 1, ub);
    RT[2] = ldT(zs + 2, ub);
    const double k_m = ldK(zs - 1, ub);
    RK[0] = ldK(zs, ub);
    RK[1] = ldK(zs + 1, ub);
    RK[2] = ldK(zs + 2, ub);
    RK[3] = ldK(zs + 3, ub);
    {
        const double lo = series_nz(k_m, RK[0]), hi = series_nz(RK[0], RK[1]);
        KZ[0] = zon(zs) ? lo : 0.0;
        KZ[1] = zon(zs + 1) ? hi : 0.0;
    }
    KZ[2] = KZ[3] = 0.0;
    U1[0] = U1[1] = U1[2] = U1[3] = 0.0;
    QT[0] = QT[1] = QT[2] = QT[3] = 0.0;
    QK[0] = QK[1] = QK[2] = QK[3] = 0.0;
    QZ[0] = QZ[1] = QZ[2] = QZ[3] = 0.0;
    double hk0 = 0.0, PT[2] = {0.0, 0.0}, PK[2] = {0.0, 0.0};
    // Per-cell capacity: mobility and capacity of the tile column at planes z and z-1 (slot
    // (plane - zs) mod 2) and mobility of the ring-1 column at z and z+1.
    double RM[2] = {0.0, 0.0}, RC[2] = {0.0, 0.0}, QM[2] = {0.0, 0.0};
    if constexpr (CC) {
        RM[0] = ld(a.mob, zs, ub), RM[1] = ld(a.mob, zs - 1, ub);
        RC[0] = ld(a.cap, zs, ub), RC[1] = ld(a.cap, zs - 1, ub);
        QM[0] = ld(a.mob, zs, r1b), QM[1] = ld(a.mob, zs + 1, r1b);
    }
    if (ring1) {
        QT[3] = ldT(zs - 1, r1b);
        QT[0] = ldT(zs, r1b);
        QT[1] = ldT(zs + 1, r1b);
        QT[2] = ldT(zs + 2, r1b);
        const double q_m = ldK(zs - 1, r1b);
        QK[0] = ldK(zs, r1b);
        QK[1] = ldK(zs + 1, r1b);
        QK[2] = ldK(zs + 2, r1b);
        QK[3] = ldK(zs + 3, r1b);
        const double lo = series_nz(q_m, QK[0]), hi = series_nz(QK[0], QK[1]);
        QZ[0] = zon(zs) ? lo : 0.0;
        QZ[1] = zon(zs + 1) ? hi : 0.0;
    }
    if (ring2) {
        PT[0] = ldT(zs, r2b);
        PT[1] = ldT(zs + 1, r2b);
        hk0 = ldK(zs, r2b);
        PK[0] = ldK(zs + 1, r2b);
        PK[1] = ldK(zs + 2, r2b);
    }
    // The faces of plane zs, from k(zs) staged in the second buffer.
    sK[1][ty + 2][tx + 2] = RK[0];
    if (ring1) sK[1][r1r + 1][r1c + 1] = QK[0];
    if (ring2) sK[1][r2r][r2c] = hk0;
    lds_barrier();
    {
        const double *k = &sK[1][0][0];
        double *f = sF[0];
#pragma unroll
        for (int rnd = 0; rnd < 3; rnd++) {
            const uint32_t d = fdesc[rnd];
            const double v = series_nz(k[fk(d)], k[fk2(d)]);
            if (rnd < 2 || third) f[fdst(d)] = fon(d) ? v : 0.0;
        }
    }

    const __amdgpu_buffer_rsrc_t RW = buffer_rsrc(a.tnew + plane * z0);
    double acc1 = 0.0, acc2 = 0.0;
    // Plane z's faces are in buffer (z - zs) mod 4; planes z-1, z and z+1 are live at once.
    for (int zb = zs; zb <= z1; zb += 4) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int z = zb + i;
            const bool live = z <= z1;
            const int b = i & 1;
            const int fb = i, fbn = (i + 1) & 3, fbp = (i + 3) & 3;
            // The planes this iteration fetches, as offsets computed once on the scalar unit.
            const uint32_t T2 = so(z + 2), T3 = so(z + 3), K3 = so(z + 3), K4 = so(z + 4);

            // Stage T0(z) and k(z+1) on E2.
            sT[b][ty + 2][tx + 2] = RT[i];
            sK[b][ty + 2][tx + 2] = RK[(i + 1) & 3];
            if (ring1) {
                sT[b][r1r + 1][r1c + 1] = QT[i];
                sK[b][r1r + 1][r1c + 1] = QK[(i + 1) & 3];
            }
            if (ring2) {
                sT[b][r2r][r2c] = PT[b];
                sK[b][r2r][r2c] = PK[b];
            }
            PT[b] = buffer_load_f64(RTs, r2b, T2);
            PK[b] = buffer_load_f64(RKs, r2b, K3);
            lds_barrier();

            // Four rounds of divisions: the tile column's z face between planes z+1 and z+2, two
            // rounds of the faces of plane z+1, and a third shared between the remaining faces
            // and the ring-1 columns' z faces.
            {
                KZ[(i + 2) & 3] = series_on<SHORT>(RK[(i + 1) & 3], RK[(i + 2) & 3], zon(z + 2));
            }
            {
                const double *k = &sK[b][0][0];
                double *f = sF[fbn];
#pragma unroll
                for (int rnd = 0; rnd < 2; rnd++) {
                    const uint32_t d = fdesc[rnd];
                    f[fdst(d)] = series_on<SHORT>(k[fk(d)], k[fk2(d)], fon(d));
                }
                const uint32_t d = fdesc[2];
                const double ka = third ? k[fk(d)] : QK[(i + 1) & 3];
                const double kb = third ? k[fk2(d)] : QK[(i + 2) & 3];
                const double v = series_on<SHORT>(ka, kb, third ? fon(d) : zon(z + 2));
                if (third) f[fdst(d)] = v;
                if (ring1) QZ[(i + 2) & 3] = v;
            }
            RK[i] = buffer_load_f64(RKs, ub, K4);  // k(z) is no longer needed
            QK[i] = buffer_load_f64(RKs, r1b, K4);

            // First step: T1(z) on E1 (its corners aside).
            const double *f = sF[fb];
            const double t1 = update(&sT[b][0][0], E2X, (ty + 2) * E2X + tx + 2, f[fxo],
                                     f[fxo + 1], f[fyo], f[fyo + E1X], RT[(i + 3) & 3], RT[i],
                                     RT[(i + 1) & 3], KZ[i], KZ[(i + 1) & 3], rate_of(RM[b]));
            U1[i] = t1;
            s1[b][ty + 1][tx + 1] = t1;
            if (ring1)
                s1[b][r1r][r1c] = update(&sT[b][0][0], E2X, (r1r + 1) * E2X + r1c + 1, f[r1fx],
                                         f[r1fx + 1], f[r1fy], f[r1fy + E1X], QT[(i + 3) & 3],
                                         QT[i], QT[(i + 1) & 3], QZ[i], QZ[(i + 1) & 3],
                                         rate_of(QM[b]));
            RT[(i + 3) & 3] = buffer_load_f64(RTs, ub, T3);  // T0(z-1) is no longer needed
            QT[(i + 3) & 3] = buffer_load_f64(RTs, r1b, T3);
            if constexpr (CC) QM[b] = ld(a.mob, z + 2, r1b);
            if (inside && z >= z0 && z < z1) acc1 += (CC ? RC[b] : a.cap0) * (t1 - a.tref);

            // Second step: T2(z-1) on the tile, from T1(z-1) staged in the previous iteration.
            const double *g = sF[fbp];
            const double t2 = update(&s1[b ^ 1][0][0], E1X, (ty + 1) * E1X + tx + 1, g[fxo],
                                     g[fxo + 1], g[fyo], g[fyo + E1X], U1[(i + 2) & 3],
                                     U1[(i + 3) & 3], U1[i], KZ[(i + 3) & 3], KZ[i],
                                     rate_of(RM[b ^ 1]));
            const double cap_m = CC ? RC[b ^ 1] : a.cap0;
            if constexpr (CC) {
                // mobility and capacity of plane z+1, used from the next iteration on
                RM[b ^ 1] = ld(a.mob, z + 1, ub);
                RC[b ^ 1] = ld(a.cap, z + 1, ub);
            }
            if (inside && live && z - 1 >= z0) {
                buffer_store_f64_nt(t2, RW, ub, (uint32_t)(z - 1 - z0) * pbytes);
                acc2 += cap_m * (t2 - a.tref);
            }
        }
    }

    const double total1 = block_sum<NT>(acc1, red);
    __syncthreads();
    const double total2 = block_sum<NT>(acc2, red);
    if (tid == 0) {
        a.partials[tile] = total1;
        a.partials[ntiles + tile] = total2;
    }
}

// The ledger totals and audits of the two steps of a pass, in order: the first step's breach
// freezes the run before the second is audited.
__global__ __launch_bounds__(1024) void heat_finalize2(const double *partials, int64_t n,
                                                       AuditState *st, double lost,
                                                       double supplied, int64_t step,
                                                       double *totals) {
    __shared__ double scratch[16];
    if (st->halted) return;
    const double stored1 = fixed_order_sum_1024(partials, n, scratch);
    __syncthreads();
    const double stored2 = fixed_order_sum_1024(partials + n, n, scratch);
    if (threadIdx.x == 0) {
        double t1 = 0.0;
        t1 += stored1;
        t1 += lost;
        t1 += -supplied;
        const double s1 = rmax(rmax(rmax(0.0, std::fabs(stored1)), std::fabs(lost)), std::fabs(supplied));
        audit_step(st, t1, s1, step, totals);
        if (st->halted) return;
        double t2 = 0.0;
        t2 += stored2;
        t2 += lost;
        t2 += -supplied;
        const double s2 = rmax(rmax(rmax(0.0, std::fabs(stored2)), std::fabs(lost)), std::fabs(supplied));
        audit_step(st, t2, s2, step + 1, totals);
    }
}

// Ledger total and audit of the step that just ran: stored = sum of partials; the ledger is
// (0 + stored) + lost + (-supplied) with scale max(|stored|, |lost|, |supplied|).
__global__ __launch_bounds__(1024) void heat_finalize(const double *partials, int64_t n,
                                                      AuditState *st, double lost,
                                                      double supplied, int64_t step,
                                                      double *totals, double *stored_out) {
    __shared__ double scratch[16];
    if (step >= 0 && st->halted) return;
    const double stored = fixed_order_sum_1024(partials, n, scratch);
    if (threadIdx.x == 0) {
        double total = 0.0;
        total += stored;
        total += lost;
        total += -supplied;
        double scale = rmax(rmax(rmax(0.0, std::fabs(stored)), std::fabs(lost)), std::fabs(supplied));
        if (stored_out) *stored_out = stored;
        if (step >= 0) audit_step(st, total, scale, step, totals);
    }
}

// The per-cell rate mob*face_sum of resolve (solid.rs:757-776), with face_sum summed in the
// reference's order; per-block maximum (Rust f64::max) and a NaN count.
__global__ __launch_bounds__(256) void heat_rates(const double *kcell, const double *mob,
                                                  double kface, double mob0, int kmode,
                                                  int cmode, int nx, int ny, int nz,
                                                  double *block_max, int *block_nan) {
    __shared__ double smax[256];
    __shared__ int snan[256];
    const int64_t n = (int64_t)nx * ny * nz;
    double worst = 0.0;
    int nan = 0;
    for (int64_t c = blockIdx.x * 256 + threadIdx.x; c < n; c += (int64_t)gridDim.x * 256) {
        const int64_t i = c % nx, j = (c / nx) % ny, k = c / ((int64_t)nx * ny);
        double f[6];
        if (kmode == K_PER_CELL) {
            const double kc = kcell[c];
            f[0] = i > 0 ? series(kcell[c - 1], kc) : 0.0;
            f[1] = i + 1 < nx ? series(kc, kcell[c + 1]) : 0.0;
            f[2] = j > 0 ? series(kcell[c - nx], kc) : 0.0;
            f[3] = j + 1 < ny ? series(kc, kcell[c + nx]) : 0.0;
            f[4] = k > 0 ? series(kcell[c - (int64_t)nx * ny], kc) : 0.0;
            f[5] = k + 1 < nz ? series(kc, kcell[c + (int64_t)nx * ny]) : 0.0;
        } else {
            f[0] = i > 0 ? kface : 0.0;
            f[1] = i + 1 < nx ? kface : 0.0;
            f[2] = j > 0 ? kface : 0.0;
            f[3] = j + 1 < ny ? kface : 0.0;
            f[4] = k > 0 ? kface : 0.0;
            f[5] = k + 1 < nz ? kface : 0.0;
        }
        const double sum = f[0] + f[1] + f[2] + f[3] + f[4] + f[5];
        const double m = cmode == C_PER_CELL ? mob[c] : mob0;
        const double rate = m * sum;
        if (std::isnan(rate))
            nan++;
        else
            worst = rmax(worst, rate);
    }
    smax[threadIdx.x] = worst;
    snan[threadIdx.x] = nan;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smax[threadIdx.x] = rmax(smax[threadIdx.x], smax[threadIdx.x + s]);
            snan[threadIdx.x] += snan[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) block_max[blockIdx.x] = smax[0], block_nan[blockIdx.x] = snan[0];
}

// fill: cells under the mask take the substance; mobility = dx/C (solid.rs:764).
__global__ void heat_fill(const uint8_t *mask, int64_t n, double *kcell, double *cap, double *mob,
                          double k, double c, double dx) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (!mask[i]) continue;
        if (kcell) kcell[i] = k;
        if (cap) cap[i] = c;
        if (mob) mob[i] = dx / c;
    }
}

__global__ void fill_value(double *p, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// A uniform rise on every cell: the placeless bus heat of solid.rs:1752-1765.
__global__ void add_uniform(double *p, int64_t n, double rise) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] += rise;
}

// The bus heat with void cells, which take no share (solid.rs:1756-1764).
__global__ void add_uniform_void(double *p, const uint8_t *vd, int64_t n, double rise) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (!vd[i]) p[i] += rise;
}

// A void cell conducts and holds nothing: zero conductivity and capacity, and zero mobility
// rather than dx/0 (solid.rs:686-702, 753-756).
__global__ void void_zero(const uint8_t *vd, int64_t n, double *kcell, double *cap, double *mob) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (vd[i]) kcell[i] = 0.0, cap[i] = 0.0, mob[i] = 0.0;
}

__global__ void audit_reset(AuditState *st, double total, double scale, double tol) {
    st->prev_total = total;
    st->prev_scale = scale;
    st->tolerance = tol;
    st->halted = 0;
    st->breach_step = -1;
    st->before = st->after = st->scale = 0.0;
}

// ---- Phase change: the enthalpy method of solid.rs:1563-1648 ------------------------------
//
// A cell whose substance melts carries a melted fraction beside its temperature; its state is the
// one monotone number e (kelvin of sensible heat below the melting point, phi*L inside, L + above),
// and every heat that reaches it is added to e and inverted back.  `cap` is the capacity the rise
// was computed with (the mixed one of the last rebuild), cs and cl the solid and liquid ones.
struct PhaseArgs {
    double *melted;
    const double *point, *latent, *cs, *cl;
};

// Solid3D::enthalpy_joules, solid.rs:1578-1593, latent > 0.
__device__ __forceinline__ double enthalpy_of(const PhaseArgs &ph, int64_t c, double t) {
    const double phi = ph.melted[c];
    if (phi >= 1.0) return ph.latent[c] + ph.cl[c] * (t - ph.point[c]);
    if (phi <= 0.0) return ph.cs[c] * (t - ph.point[c]);
    return phi * ph.latent[c];
}

// Solid3D::add_kelvin and state_from_enthalpy, solid.rs:1620-1648: the cell's new temperature,
// its fraction written back.
__device__ __forceinline__ double add_kelvin(const PhaseArgs &ph, int64_t c, double t, double rise,
                                             double cap) {
    if (!ph.melted || !(ph.latent[c] > 0.0)) return t + rise;
    const double joules = rise * cap;
    const double e = enthalpy_of(ph, c, t) + joules;
    const double point = ph.point[c], latent = ph.latent[c];
    if (e <= 0.0) {
        ph.melted[c] = 0.0;
        return point + e / ph.cs[c];
    }
    if (e >= latent) {
        ph.melted[c] = 1.0;
        return point + (e - latent) / ph.cl[c];
    }
    ph.melted[c] = e / latent;
    return point;
}

// add_kelvin with the capacity read only where the cell melts (cap null: every cell holds cap0).
__device__ __forceinline__ double add_kelvin_lazy(const PhaseArgs &ph, int64_t c, double t, double rise,
                                                  const double *__restrict__ cap, double cap0) {
    if (!ph.melted || !(ph.latent[c] > 0.0)) return t + rise;
    return add_kelvin(ph, c, t, rise, cap ? cap[c] : cap0);
}

// The bus share with phases (and void): a uniform rise through add_kelvin, solid.rs:1752-1765.
__global__ void add_uniform_phase(double *p, const uint8_t *vd, int64_t n, double rise,
                                  const double *cap, PhaseArgs ph) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (!(vd && vd[i])) p[i] = add_kelvin(ph, i, p[i], rise, cap[i]);
}

// Solid3D::deposit, solid.rs:1364-1372, for a block that can melt.
__global__ void deposit_phase(double *T, int64_t c, double joules, const double *cap, PhaseArgs ph) {
    T[c] = add_kelvin(ph, c, T[c], joules / cap[c], cap[c]);
}

// A two-phase block's operator from its fractions, solid.rs:686-702 and 753-756: conductivity and
// capacity mixed arithmetically, void conducting and holding nothing, mobility dx/C.
__global__ void mix_phases(const double *melted, const double *ks, const double *kl,
                           const double *cs, const double *cl, const uint8_t *vd, int64_t n,
                           double dx, double *kcell, double *cap, double *mob) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        if (vd && vd[c]) {
            kcell[c] = 0.0, cap[c] = 0.0, mob[c] = 0.0;
            continue;
        }
        const double phi = melted[c];
        kcell[c] = ks[c] + phi * (kl[c] - ks[c]);
        cap[c] = cs[c] + phi * (cl[c] - cs[c]);
        mob[c] = dx / cap[c];
    }
}

// set_temperature over the cells of a mask (solid.rs:1346-1361), phase included.
__global__ void set_region(double *T, const uint8_t *mask, int64_t n, double kelvin,
                           const double *point, const double *latent, double *melted) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        if (!mask[c]) continue;
        T[c] = kelvin;
        if (melted && latent[c] > 0.0) melted[c] = kelvin > point[c] ? 1.0 : 0.0;
    }
}

// Naming a temperature names the phase of a cell that melts (solid.rs:1352-1360).
__global__ void phase_from_temperature(const double *T, const double *point, const double *latent,
                                       int64_t n, double *melted) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x)
        if (latent[c] > 0.0) melted[c] = T[c] > point[c] ? 1.0 : 0.0;
}

static inline int grid_for(int64_t n, int threads) {
    int64_t g = (n + threads - 1) / threads;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

}  // namespace pm

using namespace pm;

struct pm_solid3d {
    std::string name;
    int64_t nx = 1, ny = 1, nz = 1;
    double dx = 0.0;
    double reference = 0.0;  // the ledger's enthalpy reference, the constructor's temperature
    double absorbed = 0.0;
    double lost = 0.0, supplied = 0.0;          // zero on the plain branch, kept for the ledger
    double saved_lost = 0.0, saved_supplied = 0.0;
    // operator
    int kmode = K_UNIFORM, cmode = C_UNIFORM;
    double k0 = 0.0, cap0 = 0.0, mob0 = 0.0, kface = 0.0;
    double worst_rate = 0.0;
    double heat_capacity_sum = 0.0;
    bool heat_capacity_valid = false;
    // device state
    int device = 0;  // the device current at creation, which holds every buffer
    hipStream_t stream = nullptr;
    DevBuf t[2];
    int cur = 0;
    DevBuf kcell, cap, mob, saved;
    // Whether every conductivity passes k_short_ok, known since the field last changed.
    bool k_checked = false, k_short = false;
    DevBuf kflag;
    bool saved_valid = false;
    DevBuf partials, audit, scalars, stored_dev;
    int64_t partial_count = 0;
    int tile_y = 8, z_chunk = 0, prefetch = 2;  // z_chunk 0: chosen from the extents
    int steps_per_pass = 2;  // 2: pairs of steps in one pass where the form allows it
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Per-launch events of a profiled advance: [3*s] before the sweep, [3*s+1] between sweep and
    // audit, [3*s+2] after the audit.
    bool profiling = false;
    std::vector<hipEvent_t> prof;
    std::vector<int> prof_steps;  // steps each profiled launch advanced
    int64_t prof_used = 0;
    int prof_launch_steps = 1;    // steps per launch of the launches sweep_ms averages
    double sweep_ms = 0.0, audit_ms = 0.0;
    // Exposed faces and volumetric sources, the film and source branches of the sweep
    // (solid.rs:1192-1288, 1800-1864).  A block with either steps through step_ext.
    struct Film {
        int on = 0;
        double ambient = 0.0, h = 0.0, area = 0.0;
    };
    Film film[6];                      // Face order: XMin, XMax, YMin, YMax, ZMin, ZMax
    std::vector<pm_substance> subs;    // subs[0] is the constructor's
    std::vector<uint16_t> which;       // each cell's substance; empty while all are subs[0]
    std::map<int64_t, double> source;  // watts of each generating cell
    DevBuf bcell, bemis, bkcond, shed, src_w, src_of, ext_part;
    DevBuf film_part, ext_books;  // the film's rate maxima and (lost, supplied) of a device-resident run
    int64_t nboundary = 0, nsource = 0;
    bool tables_valid = false;
    // Void cells (solid.rs:238, 992-1005) and the pairs of solid cells facing each other across
    // them (solid.rs:239-245, 791-847).  Each gap cell lists the pairs it belongs to in pair
    // order: gent[gstart[q] .. gstart[q+1]) holds 2*pair + (0 for the pair's first cell, 1 for
    // its second).
    std::vector<uint8_t> voids;  // empty until a cell is emptied
    int64_t nvoid = 0;
    DevBuf dvoid;
    struct Gap {
        int64_t a, b;
        double coef;
    };
    std::vector<Gap> gaps;
    DevBuf gap_a, gap_b, gap_coef, gcell, gstart, gent, gap_part;
    int64_t ngcell = 0;
    // Phase change (solid.rs:586-680): once any substance melts, per-cell solid and liquid
    // conductivity and capacity, melting point, latent joules, and the melted fraction (with
    // the copy a checkpoint keeps).  Host copies of point and latent serve set_temperature.
    bool phase = false, two_phase = false;
    DevBuf ph_ks, ph_kl, ph_cs, ph_cl, ph_point, ph_latent, melted, saved_melted;
    std::vector<double> h_point, h_latent;
    int64_t nlatent = 0;
    PhaseArgs phase_args() const {
        PhaseArgs a{};
        if (phase) {
            a.melted = melted.d(), a.point = ph_point.d(), a.latent = ph_latent.d();
            a.cs = ph_cs.d(), a.cl = ph_cl.d();
        }
        return a;
    }

    int64_t n() const { return nx * ny * nz; }
    double *T() const { return t[cur].d(); }
    bool exposed() const {
        for (const Film &f : film)
            if (f.on) return true;
        return false;
    }
    bool extended() const { return exposed() || !source.empty() || !gaps.empty() || nlatent > 0; }
    bool is_void(int64_t c) const { return nvoid && voids[(size_t)c]; }
};

namespace pm {

// The z-planes one workgroup marches through.  Unless fixed by set_tile, chunks are as long as
// possible (at most 128 planes, so the two planes of halo a chunk re-reads stay a small share)
// while the launch still has about 1024 workgroups, four per compute unit; short chunks (at least
// 4 planes) are what fills the chip on small blocks.
// The paired sweep addresses T, k and its output from its chunk's lowest plane with 32-bit byte
// offsets: a chunk of zc planes reaches zc + 4 planes of an array, which must span fewer than
// 2^32 bytes (blocks whose planes are too large for one such chunk take the one-step sweep).
static int64_t pair_max_chunk(const pm_solid3d *s) {
    return (int64_t)(((uint64_t)1 << 32) - 1) / (s->nx * s->ny * 8) - 4;
}

static int z_chunk_of(const pm_solid3d *s, int tile_y, bool paired = false) {
    int64_t zc;
    if (s->z_chunk > 0) {
        zc = s->z_chunk;
    } else {
        const int64_t xy = ((s->nx + 63) / 64) * ((s->ny + tile_y - 1) / tile_y);
        const int64_t chunks = std::max<int64_t>(1, (1024 + xy - 1) / xy);
        zc = std::min<int64_t>(128, std::max<int64_t>(4, (s->nz + chunks - 1) / chunks));
    }
    if (paired) zc = std::max<int64_t>(1, std::min(zc, pair_max_chunk(s)));
    return (int)zc;
}

static int heat_grid(const pm_solid3d *s, dim3 *grid, dim3 *block, int tile_y = 0, bool paired = false) {
    const int ty = tile_y ? tile_y : s->tile_y;
    const int zc = z_chunk_of(s, ty, paired);
    *block = dim3(64, (unsigned)ty, 1);
    *grid = dim3((unsigned)((s->nx + 63) / 64), (unsigned)((s->ny + ty - 1) / ty),
                 (unsigned)((s->nz + zc - 1) / zc));
    return PM_OK;
}

// The two-steps-per-pass sweep exists for per-cell conductivity (capacity uniform or per cell),
// in 64 x 8 tiles.
constexpr int PAIR_TY = 8;
static bool pairs_enabled(const pm_solid3d *s) {
    return s->steps_per_pass == 2 && s->kmode == K_PER_CELL && pair_max_chunk(s) >= 1;
}

// Cells per partial of the stored-heat reduction: about 1024 partials, between 1024 and 16384
// cells each (a multiple of 256), so a block of 128^3 fills every compute unit and one of 512^3
// has 8192 partials.  Every path that sums the stored heat uses this partition, so they agree
// bit for bit.
constexpr int64_t STORED_PER_BLOCK = 16384;
static int64_t stored_per_block(int64_t n) {
    const int64_t want = ((n + 1023) / 1024 + 255) / 256 * 256;
    return std::min<int64_t>(STORED_PER_BLOCK, std::max<int64_t>(1024, want));
}

// Per-workgroup partials one launch of either sweep writes, or the blocks of the stored-heat
// reduction, whichever is more.
static int64_t partials_needed(const pm_solid3d *s) {
    dim3 g, b, g2;
    heat_grid(s, &g, &b);
    heat_grid(s, &g2, &b, PAIR_TY, true);
    const int64_t sweep = std::max<int64_t>((int64_t)g.x * g.y * g.z, 2 * (int64_t)g2.x * g2.y * g2.z);
    const int64_t per = stored_per_block(s->n());
    return std::max<int64_t>(sweep, (s->n() + per - 1) / per);
}

static int ensure_partials(pm_solid3d *s) {
    const int64_t need = partials_needed(s);
    if (s->partial_count < need || !s->partials) {
        PM_CHECK(s->partials.alloc(sizeof(double) * (size_t)need));
        s->partial_count = need;
    }
    return PM_OK;
}

// The device memory an audited run works in besides the cells and the operator: per-workgroup
// partials, the audit state, the conductivity-range flag, the stored-heat scalar and the per-step
// totals (null when the caller wants none).  Either the handle's own buffers or the library's
// static scratch (pm_solid3d_advance_static).
struct RunSpace {
    double *partials = nullptr;
    AuditState *audit = nullptr;
    int *kflag = nullptr;
    double *stored = nullptr;
    double *totals = nullptr;
};

static int handle_space(pm_solid3d *s, RunSpace *w) {
    PM_CHECK(ensure_partials(s));
    if (!s->kflag) PM_CHECK(s->kflag.alloc(sizeof(int)));
    w->partials = s->partials.d();
    w->audit = s->audit.as<AuditState>();
    w->kflag = s->kflag.as<int>();
    w->stored = s->stored_dev.d();
    w->totals = nullptr;
    return PM_OK;
}

template <int KM, int CM, int TY>
static void launch_sweep_pd(const pm_solid3d *s, const HeatArgs &a, dim3 g) {
    // One-dimensional launch: the kernel maps workgroup ids to tiles XCD by XCD.
    const int64_t ntiles = (int64_t)g.x * g.y * g.z;
    const unsigned blocks = (unsigned)(XCDS * ((ntiles + XCDS - 1) / XCDS));
    if (KM == K_PER_CELL && s->k_checked && s->k_short)
        heat_sweep<KM, CM, TY, true><<<blocks, dim3(64, TY), 0, s->stream>>>(a);
    else
        heat_sweep<KM, CM, TY, false><<<blocks, dim3(64, TY), 0, s->stream>>>(a);
}

template <int KM, int CM>
static void launch_sweep_t(const pm_solid3d *s, const HeatArgs &a, dim3 g) {
    switch (s->tile_y) {
        case 4: launch_sweep_pd<KM, CM, 4>(s, a, g); break;
        case 8: launch_sweep_pd<KM, CM, 8>(s, a, g); break;
        default: launch_sweep_pd<KM, CM, 16>(s, a, g); break;
    }
}

// Events around a launch of a profiled advance: [3l] before the sweep, [3l+1] between sweep and
// finalize, [3l+2] after the finalize, for launch l.
static bool prof_slot(pm_solid3d *s, int64_t launch) {
    return s->profiling && launch >= 0 && 3 * (launch + 1) <= (int64_t)s->prof.size();
}

// One sweep from t[cur] into t[cur^1] plus the finalize/audit; then swaps.  `launch` indexes the
// profiling events (-1: none); `force` runs the sweep even after a breach, without the audit.
static int sweep_once(pm_solid3d *s, double dt, int64_t step, const RunSpace &w,
                      int64_t launch = -1, bool force = false) {
    dim3 g, b;
    heat_grid(s, &g, &b);
    HeatArgs a{};
    a.told = s->t[s->cur].d();
    a.tnew = s->t[s->cur ^ 1].d();
    a.kcell = s->kcell.d();
    a.mob = s->mob.d();
    a.cap = s->cap.d();
    a.partials = w.partials;
    a.audit = w.audit;
    a.kface = s->kface;
    a.f = dt * s->mob0;
    a.dts = dt;
    a.cap0 = s->cap0;
    a.tref = s->reference;
    a.nx = (int)s->nx, a.ny = (int)s->ny, a.nz = (int)s->nz, a.zc = z_chunk_of(s, s->tile_y);
    a.gx = (int)g.x, a.gy = (int)g.y, a.gz = (int)g.z;
    a.force = force ? 1 : 0;
    const bool prof = prof_slot(s, launch);
    if (prof) PM_HIP(hipEventRecord(s->prof[3 * launch], s->stream));
    if (s->kmode == K_UNIFORM && s->cmode == C_UNIFORM)
        launch_sweep_t<K_UNIFORM, C_UNIFORM>(s, a, g);
    else if (s->cmode == C_UNIFORM)
        launch_sweep_t<K_PER_CELL, C_UNIFORM>(s, a, g);
    else
        launch_sweep_t<K_PER_CELL, C_PER_CELL>(s, a, g);
    PM_HIP(hipGetLastError());
    if (prof) PM_HIP(hipEventRecord(s->prof[3 * launch + 1], s->stream));
    if (!force) {
        heat_finalize<<<1, 1024, 0, s->stream>>>(w.partials, (int64_t)g.x * g.y * g.z, w.audit,
                                                 s->lost, s->supplied, step, w.totals, w.stored);
        PM_HIP(hipGetLastError());
    }
    if (prof) {
        PM_HIP(hipEventRecord(s->prof[3 * launch + 2], s->stream));
        s->prof_steps[(size_t)launch] = 1;
        s->prof_used = std::max<int64_t>(s->prof_used, launch + 1);
    }
    s->cur ^= 1;
    return PM_OK;
}

// Finds, once after each change of the conductivity field, whether every conductivity passes
// k_short_ok, so the paired sweep may build its faces with the unscaled division.
static int check_k_range(pm_solid3d *s, int *flag) {
    if (s->k_checked) return PM_OK;
    clear_flag<<<1, 1, 0, s->stream>>>(flag);
    k_short_check<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(s->kcell.d(), s->n(), flag);
    PM_HIP(hipGetLastError());
    int outside = 1;
    PM_HIP(hipMemcpyAsync(&outside, flag, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    s->k_short = outside == 0;
    s->k_checked = true;
    return PM_OK;
}

// Steps `step` and `step + 1` in one pass from t[cur] into t[cur^1], then both audits; swaps.
static int sweep_pair(pm_solid3d *s, double dt, int64_t step, const RunSpace &w, int64_t launch) {
    dim3 g, b;
    heat_grid(s, &g, &b, PAIR_TY, true);
    HeatArgs a{};
    a.told = s->t[s->cur].d();
    a.tnew = s->t[s->cur ^ 1].d();
    a.kcell = s->kcell.d();
    a.partials = w.partials;
    a.audit = w.audit;
    a.mob = s->cmode == C_PER_CELL ? s->mob.d() : nullptr;
    a.cap = s->cmode == C_PER_CELL ? s->cap.d() : nullptr;
    a.f = dt * s->mob0;
    a.dts = dt;
    a.cap0 = s->cap0;
    a.tref = s->reference;
    a.nx = (int)s->nx, a.ny = (int)s->ny, a.nz = (int)s->nz, a.zc = z_chunk_of(s, PAIR_TY, true);
    a.gx = (int)g.x, a.gy = (int)g.y, a.gz = (int)g.z;
    const int64_t ntiles = (int64_t)g.x * g.y * g.z;
    const bool prof = prof_slot(s, launch);
    if (prof) PM_HIP(hipEventRecord(s->prof[3 * launch], s->stream));
    const unsigned blocks = (unsigned)(XCDS * ((ntiles + XCDS - 1) / XCDS));
    const bool quick = s->k_checked && s->k_short;
    if (s->cmode == C_PER_CELL) {
        if (quick)
            heat_sweep2<PAIR_TY, true, C_PER_CELL><<<blocks, dim3(64, PAIR_TY), 0, s->stream>>>(a);
        else
            heat_sweep2<PAIR_TY, false, C_PER_CELL><<<blocks, dim3(64, PAIR_TY), 0, s->stream>>>(a);
    } else if (quick) {
        heat_sweep2<PAIR_TY, true, C_UNIFORM><<<blocks, dim3(64, PAIR_TY), 0, s->stream>>>(a);
    } else {
        heat_sweep2<PAIR_TY, false, C_UNIFORM><<<blocks, dim3(64, PAIR_TY), 0, s->stream>>>(a);
    }
    PM_HIP(hipGetLastError());
    if (prof) PM_HIP(hipEventRecord(s->prof[3 * launch + 1], s->stream));
    heat_finalize2<<<1, 1024, 0, s->stream>>>(w.partials, ntiles, w.audit, s->lost, s->supplied,
                                              step, w.totals);
    PM_HIP(hipGetLastError());
    if (prof) {
        PM_HIP(hipEventRecord(s->prof[3 * launch + 2], s->stream));
        s->prof_steps[(size_t)launch] = 2;
        s->prof_used = std::max<int64_t>(s->prof_used, launch + 1);
    }
    s->cur ^= 1;
    return PM_OK;
}

// Mean sweep and finalize durations per launch over the profiled launches of the last advance
// that advanced the most common number of steps (a trailing single step of a paired run is left
// out of the mean).
static int collect_profile(pm_solid3d *s) {
    int most = 1;
    for (int64_t n = 0; n < s->prof_used; n++) most = std::max(most, s->prof_steps[(size_t)n]);
    double sw = 0.0, au = 0.0;
    int64_t used = 0;
    for (int64_t n = 0; n < s->prof_used; n++) {
        if (s->prof_steps[(size_t)n] != most) continue;
        float a = 0.0f, b = 0.0f;
        PM_HIP(hipEventElapsedTime(&a, s->prof[3 * n], s->prof[3 * n + 1]));
        PM_HIP(hipEventElapsedTime(&b, s->prof[3 * n + 1], s->prof[3 * n + 2]));
        sw += a, au += b;
        used++;
    }
    s->sweep_ms = used ? sw / (double)used : 0.0;
    s->audit_ms = used ? au / (double)used : 0.0;
    s->prof_launch_steps = most;
    return PM_OK;
}

// Rebuild worst_rate after the operator changed (resolve, solid.rs:757-782).
static int resolve(pm_solid3d *s) {
    s->heat_capacity_valid = false;
    if (s->kmode == K_UNIFORM && s->cmode == C_UNIFORM) {
        // Every cell's rate is one of the 27 boundary classes; evaluate one cell of each.
        double worst = 0.0;
        bool nan = false;
        auto reps = [](int64_t n) {
            std::vector<int64_t> r{0};
            if (n > 2) r.push_back(1);
            if (n > 1) r.push_back(n - 1);
            return r;
        };
        for (int64_t k : reps(s->nz))
            for (int64_t j : reps(s->ny))
                for (int64_t i : reps(s->nx)) {
                    double f0 = i > 0 ? s->kface : 0.0, f1 = i + 1 < s->nx ? s->kface : 0.0;
                    double f2 = j > 0 ? s->kface : 0.0, f3 = j + 1 < s->ny ? s->kface : 0.0;
                    double f4 = k > 0 ? s->kface : 0.0, f5 = k + 1 < s->nz ? s->kface : 0.0;
                    double sum = f0 + f1 + f2 + f3 + f4 + f5;
                    double rate = s->mob0 * sum;
                    if (std::isnan(rate))
                        nan = true;
                    else
                        worst = rmax(worst, rate);
                }
        s->worst_rate = nan ? NAN : worst;
        return PM_OK;
    }
    const int blocks = 1024;
    DevBuf bm, bn;
    PM_CHECK(bm.alloc(sizeof(double) * blocks));
    PM_CHECK(bn.alloc(sizeof(int) * blocks));
    heat_rates<<<blocks, 256, 0, s->stream>>>(s->kcell.d(), s->mob.d(), s->kface, s->mob0,
                                              s->kmode, s->cmode, (int)s->nx, (int)s->ny,
                                              (int)s->nz, bm.d(), bn.as<int>());
    PM_HIP(hipGetLastError());
    std::vector<double> hm(blocks);
    std::vector<int> hn(blocks);
    PM_HIP(hipMemcpyAsync(hm.data(), bm.d(), sizeof(double) * blocks, hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipMemcpyAsync(hn.data(), bn.d(), sizeof(int) * blocks, hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    double worst = 0.0;
    bool nan = false;
    for (int b = 0; b < blocks; b++) {
        worst = rmax(worst, hm[b]);
        nan = nan || hn[b] > 0;
    }
    s->worst_rate = nan ? NAN : worst;
    return PM_OK;
}

static double capacity_of(const pm_substance *sub, double dx) {
    // Substance::heat_capacity: (rho * volume) * c_p, volume = (dx*dx)*dx.
    double volume = dx * dx * dx;
    return (sub->density * volume) * sub->specific_heat;
}

static bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }

static int promote_k(pm_solid3d *s) {
    if (s->kmode == K_PER_CELL) return PM_OK;
    PM_CHECK(s->kcell.alloc(sizeof(double) * (size_t)s->n()));
    fill_value<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(s->kcell.d(), s->n(), s->k0);
    s->k_checked = false;
    PM_HIP(hipGetLastError());
    s->kmode = K_PER_CELL;
    return PM_OK;
}

static int promote_c(pm_solid3d *s) {
    if (s->cmode == C_PER_CELL) return PM_OK;
    PM_CHECK(promote_k(s));
    PM_CHECK(s->cap.alloc(sizeof(double) * (size_t)s->n()));
    PM_CHECK(s->mob.alloc(sizeof(double) * (size_t)s->n()));
    fill_value<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(s->cap.d(), s->n(), s->cap0);
    fill_value<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(s->mob.d(), s->n(), s->mob0);
    PM_HIP(hipGetLastError());
    s->cmode = C_PER_CELL;
    return PM_OK;
}

// Heat capacity of one cell (host read of the device copy in the general form).
static int cell_capacity(pm_solid3d *s, int64_t idx, double *out) {
    if (s->cmode == C_UNIFORM) {
        *out = s->cap0;
        return PM_OK;
    }
    PM_HIP(hipMemcpyAsync(out, s->cap.d() + idx, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

static int total_capacity(pm_solid3d *s, double *out) {
    if (!s->heat_capacity_valid) {
        // `self.capacity.iter().sum::<f64>()`: a left fold in index order, on the host.
        double total = 0.0;
        if (s->cmode == C_UNIFORM) {
            for (int64_t c = 0; c < s->n(); c++) total += s->cap0;
        } else {
            std::vector<double> h((size_t)s->n());
            PM_HIP(hipMemcpyAsync(h.data(), s->cap.d(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->stream));
            PM_HIP(hipStreamSynchronize(s->stream));
            for (double v : h) total += v;
        }
        s->heat_capacity_sum = total;
        s->heat_capacity_valid = true;
    }
    *out = s->heat_capacity_sum;
    return PM_OK;
}

// The extended sweep's grid, a function of the block's size only, so its sums are reproducible:
// 256 workgroups of 1024 threads marching 64 x 16 columns over chunks of planes when the block
// has that many column tiles x planes (at 128^3 128 and 1024 workgroups of the earlier one-cell-a-
// thread form ran 37% and 8% slower than 256), else one cell a thread in as many workgroups as
// the cells need (at 24^3, 14 of them: 23.4 us a step against 26.7 with a workgroup a plane of
// column tiles).  Its per-workgroup loss partials are folded in two levels (loss_fold).  The
// rate kernels of the film and the gaps use at most EXT_BLOCKS workgroups of 256.
constexpr int EXT_BLOCKS = 1024, EXT_THREADS = 1024, EXT_SWEEP_MAX = 256, LOSS_CHUNK = 64;
struct ExtSweep {
    int blocks;
    bool march;
};
static ExtSweep ext_sweep_plan(int64_t nx, int64_t ny, int64_t nz) {
    const int64_t items = ((nx + 63) / 64) * ((ny + 15) / 16) * nz;
    if (items >= EXT_SWEEP_MAX) return {EXT_SWEEP_MAX, true};
    const int64_t n = nx * ny * nz;
    return {(int)std::min<int64_t>(EXT_SWEEP_MAX, std::max<int64_t>(1, (n + EXT_THREADS - 1) / EXT_THREADS)), false};
}
// The film's loss of a step from the sweep's partials: chunks of LOSS_CHUNK folded in index
// order, then the chunks' sums folded in order.  The same on the host (step_ext) and the device
// (ext_books).
static double loss_fold(const double *part, int n) {
    double total = 0.0;
    for (int c = 0; c < n; c += LOSS_CHUNK) {
        double chunk = 0.0;
        for (int i = c; i < std::min(n, c + LOSS_CHUNK); i++) chunk += part[i];
        total += chunk;
    }
    return total;
}

// ---- Exposed faces and volumetric sources ---------------------------------------------------
//
// The film acts only on the cells of exposed faces, so those cells are listed once (index order,
// with their substance's emissivity and conductivity) and a small kernel serves them: the
// state-dependent stability rate (solid.rs:1148-1185).  The sweep of a block with a film or a
// source is heat_step_ext, one thread per cell: a listed cell (found through a per-cell index
// map) takes the film's flux out of it from the step's input (solid.rs:1192-1221) inside its
// update, and a generating cell its source after it, as solid.rs:1767-1864 orders them.

struct FilmFaces {
    int on[6];
    double ambient[6], h[6], share[6];  // share: the face's area over its cell count
    double dx;
};

constexpr double STEFAN_BOLTZMANN = 5.670374419e-8;  // crates/pantometry-units/src/lib.rs:899

// t.powi(4) and t.powi(3) with a constant exponent: repeated squaring, as Rust lowers them.
__device__ __forceinline__ double pow4(double x) {
    const double x2 = x * x;
    return x2 * x2;
}
__device__ __forceinline__ double pow3(double x) { return x * (x * x); }

// A cell's coordinates from its index, with two 32-bit divisions when the block has fewer than
// 2^32 cells (64-bit division is a long software sequence on the GPU).
__device__ __forceinline__ void coords_of(int64_t c, int64_t nx, int64_t ny, int64_t nz, int64_t &i,
                                          int64_t &j, int64_t &k) {
    if (nx * ny * nz <= (int64_t)UINT32_MAX) {
        const uint32_t cu = (uint32_t)c, x = (uint32_t)nx;
        const uint32_t q = cu / x, kk = q / (uint32_t)ny;
        i = cu - q * x, j = q - kk * (uint32_t)ny, k = kk;
    } else {
        i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    }
}

// Environment::loss_from, crates/pantometry-thermal/src/lib.rs:104-110.
__device__ __forceinline__ double loss_from(double t, double ambient, double h, double area,
                                            double emissivity) {
    const double convective = h * area * (t - ambient);
    const double radiative = emissivity * STEFAN_BOLTZMANN * area * (pow4(t) - pow4(ambient));
    return convective + radiative;
}

// series_with_half_cell, solid.rs:110-122.
__device__ __forceinline__ double series_with_half_cell(double film, double conductivity,
                                                        double area, double dx) {
    if (film <= 0.0) return 0.0;
    const double half = 2.0 * conductivity * area / dx;
    if (!isfinite(half)) return film;
    if (half <= 0.0) return 0.0;
    return 1.0 / (1.0 / film + 1.0 / half);
}

// Face::holds, solid.rs:158-169.
__device__ __forceinline__ bool holds(int f, int64_t i, int64_t j, int64_t k, int64_t nx,
                                      int64_t ny, int64_t nz) {
    switch (f) {
        case 0: return i == 0;
        case 1: return i + 1 == nx;
        case 2: return j == 0;
        case 3: return j + 1 == ny;
        case 4: return k == 0;
        default: return k + 1 == nz;
    }
}

// The six faces of a cell in resolve's order and their left-fold sum (solid.rs:757-763).
__device__ __forceinline__ double face_sum_of(const double *kc, double kface, int64_t c, int64_t i,
                                              int64_t j, int64_t k, int64_t nx, int64_t ny,
                                              int64_t nz) {
    const int64_t plane = nx * ny;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0, f4 = 0.0, f5 = 0.0;
    if (kc) {
        // All six neighbours' conductivities load together (a face off the grid reads the cell
        // itself) and a missing face takes 0 by a select.
        const bool o0 = i > 0, o1 = i + 1 < nx, o2 = j > 0, o3 = j + 1 < ny, o4 = k > 0, o5 = k + 1 < nz;
        const double kk = kc[c];
        const double k0 = kc[o0 ? c - 1 : c], k1 = kc[o1 ? c + 1 : c], k2 = kc[o2 ? c - nx : c];
        const double k3 = kc[o3 ? c + nx : c], k4 = kc[o4 ? c - plane : c], k5 = kc[o5 ? c + plane : c];
        f0 = o0 ? series_nz(k0, kk) : 0.0, f1 = o1 ? series_nz(kk, k1) : 0.0;
        f2 = o2 ? series_nz(k2, kk) : 0.0, f3 = o3 ? series_nz(kk, k3) : 0.0;
        f4 = o4 ? series_nz(k4, kk) : 0.0, f5 = o5 ? series_nz(kk, k5) : 0.0;
    } else {
        f0 = i > 0 ? kface : 0.0, f1 = i + 1 < nx ? kface : 0.0;
        f2 = j > 0 ? kface : 0.0, f3 = j + 1 < ny ? kface : 0.0;
        f4 = k > 0 ? kface : 0.0, f5 = k + 1 < nz ? kface : 0.0;
    }
    return f0 + f1 + f2 + f3 + f4 + f5;
}

// A film's listed cells and what their rates read.
struct FilmList {
    FilmFaces F;
    const int64_t *bcell;
    const double *bemis, *bk;
    int64_t nb;
    const double *kc;
    double kface;
    const double *mob;
    double mob0;
    int64_t nx, ny, nz;
};

// The stability rate of listed cell b at temperatures T, mob * (face_sum +
// loss_conductance_at(T)/dx), or -INFINITY where the loss is zero.
__device__ __forceinline__ double film_rate(const FilmList &L, int64_t b, const double *__restrict__ T) {
    const FilmFaces &F = L.F;
    const int64_t nx = L.nx, ny = L.ny, nz = L.nz;
    const int64_t c = L.bcell[b];
    int64_t i, j, k;
    coords_of(c, nx, ny, nz, i, j, k);
    const double at = T[c];
    double loss = 0.0;
    for (int f = 0; f < 6; f++) {
        if (!F.on[f] || !holds(f, i, j, k, nx, ny, nz)) continue;
        const double hot = fmax(at, F.ambient[f]);
        const double h = F.h[f] + 4.0 * L.bemis[b] * STEFAN_BOLTZMANN * pow3(hot);
        loss += series_with_half_cell(h * F.share[f], L.bk[b], F.share[f], F.dx);
    }
    if (loss == 0.0) return -INFINITY;
    return (L.mob ? L.mob[c] : L.mob0) * (face_sum_of(L.kc, L.kface, c, i, j, k, nx, ny, nz) + loss / F.dx);
}

// The stability rate of each listed cell at its present temperature, skipped where the loss is
// zero; per-block maxima (NaN rates lose every comparison, as in f64::max).
__global__ __launch_bounds__(256) void film_rates(FilmFaces F, const int64_t *__restrict__ bcell,
                                                  const double *__restrict__ bemis,
                                                  const double *__restrict__ bk, int64_t nb,
                                                  const double *__restrict__ T,
                                                  const double *__restrict__ kc, double kface,
                                                  const double *__restrict__ mob, double mob0,
                                                  int64_t nx, int64_t ny, int64_t nz,
                                                  double *__restrict__ part) {
    __shared__ double red[256];
    const FilmList L{F, bcell, bemis, bk, nb, kc, kface, mob, mob0, nx, ny, nz};
    double worst = -INFINITY;
    for (int64_t b = blockIdx.x * 256 + threadIdx.x; b < nb; b += (int64_t)gridDim.x * 256)
        worst = fmax(worst, film_rate(L, b, T));
    red[threadIdx.x] = worst;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Solid3D::film_flux of listed cell b (cell c at (i, j, k), temperature t): each exposed face the
// cell lies on removes g (t - ambient) / dx, with g the face's loss over the gap in series with
// the half cell.
__device__ __forceinline__ double film_flux_at(const FilmFaces &F, const double *__restrict__ bemis,
                                               const double *__restrict__ bk, int64_t b, double t,
                                               int64_t i, int64_t j, int64_t k, int64_t nx,
                                               int64_t ny, int64_t nz) {
    double out = 0.0;
    for (int f = 0; f < 6; f++) {
        if (!F.on[f] || !holds(f, i, j, k, nx, ny, nz)) continue;
        const double share = F.share[f];
        const double gap = t - F.ambient[f];
        if (gap == 0.0) continue;
        const double bare = loss_from(t, F.ambient[f], F.h[f], share, bemis[b]);
        const double g = series_with_half_cell(bare / gap, bk[b], share, F.dx);
        out += g * gap / F.dx;
    }
    return out;
}

// film_rates and film_shed in one pass over the listed cells, for the device-resident advance:
// the rates' per-block maxima, and each cell's film flux from the step's input into `shed`
// (skipped once the run has stopped).  The shed of a step the limit then refuses is never read:
// the sweep of that step does not run.
__global__ __launch_bounds__(256) void film_rates_shed(FilmFaces F, const int64_t *__restrict__ bcell,
                                                       const double *__restrict__ bemis,
                                                       const double *__restrict__ bk, int64_t nb,
                                                       const double *__restrict__ T,
                                                       const double *__restrict__ kc, double kface,
                                                       const double *__restrict__ mob, double mob0,
                                                       int64_t nx, int64_t ny, int64_t nz,
                                                       double *__restrict__ part,
                                                       double *__restrict__ shed, const AuditState *st) {
    __shared__ double red[256];
    const FilmList L{F, bcell, bemis, bk, nb, kc, kface, mob, mob0, nx, ny, nz};
    const bool live = !(st && st->halted);
    double worst = -INFINITY;
    for (int64_t b = blockIdx.x * 256 + threadIdx.x; b < nb; b += (int64_t)gridDim.x * 256) {
        worst = fmax(worst, film_rate(L, b, T));
        if (live) {
            const int64_t c = bcell[b];
            int64_t i, j, k;
            coords_of(c, nx, ny, nz, i, j, k);
            shed[c] = film_flux_at(F, bemis, bk, b, T[c], i, j, k, nx, ny, nz);
        }
    }
    red[threadIdx.x] = worst;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// What the sweep needs of a film (faces: bit f set for an exposed face f; shed: the listed cells'
// film flux, computed by film_shed from the step's input; null without a film) and of the
// sources (src_of: each cell's index in the generating cells, -1 for none; null without
// sources).  Small, so the sweep's arguments fit its scalar registers.
struct ExtCells {
    uint32_t faces;
    const double *shed;
    const int32_t *src_of;
    const double *src_w;
};

// The film's flux out of every listed cell, from the step's input, into the dense `shed` the
// sweep reads at the cells of exposed faces.  A launch of its own: in the sweep, a wave holding
// one cell of an x face would evaluate it for that one lane.
__global__ __launch_bounds__(256) void film_shed(FilmFaces F, const int64_t *__restrict__ bcell,
                                                 const double *__restrict__ bemis,
                                                 const double *__restrict__ bk, int64_t nb,
                                                 const double *__restrict__ T,
                                                 double *__restrict__ shed, int64_t nx, int64_t ny,
                                                 int64_t nz, const AuditState *st) {
    if (st && st->halted) return;  // a device-resident run that stopped
    for (int64_t b = blockIdx.x * 256 + threadIdx.x; b < nb; b += (int64_t)gridDim.x * 256) {
        const int64_t c = bcell[b];
        int64_t i, j, k;
        coords_of(c, nx, ny, nz, i, j, k);
        shed[c] = film_flux_at(F, bemis, bk, b, T[c], i, j, k, nx, ny, nz);
    }
}

// One step of a block with a film or sources: the plain sweep term by term (faces recomputed
// from the per-cell conductivity with series, or the uniform face), the film's shed of a listed
// cell (film_flux from the step's input, by film_shed) inside the update, rise = (dt*mob)*(flux - shed)
// (solid.rs:1776-1818), then a generating cell's (W*dt)/C (solid.rs:1854-1864), as the
// reference adds it after the sweep.  What the film removed, ((dt*mob)*shed)*C per cell
// (solid.rs:1821-1823), is summed per block in a fixed order.
// The rest of a cell's step once its flux is known: the shed of a cell on an exposed face
// (`exposed`), rise = (dt*mob)*(flux - shed) through add_kelvin, the source after it, the store;
// returns what the film removed, ((dt*mob)*shed)*C.
__device__ __forceinline__ double ext_finish(const ExtCells &x, int64_t c, double t, double flux,
                                             bool exposed, double dts, const double *__restrict__ mob,
                                             double mob0, const double *__restrict__ cap, double cap0,
                                             const PhaseArgs &ph, double *__restrict__ tnew) {
    const double sh = x.shed && exposed ? x.shed[c] : 0.0;
    const double dtm = dts * (mob ? mob[c] : mob0);
    // The capacity is read only where it enters: a melting cell, the film's loss, a source.
    double out = add_kelvin_lazy(ph, c, t, dtm * (flux - sh), cap, cap0);
    const double lost = sh != 0.0 ? dtm * sh * (cap ? cap[c] : cap0) : 0.0;
    if (x.src_of) {
        const int32_t q = x.src_of[c];
        if (q >= 0) {
            const double cc = cap ? cap[c] : cap0;
            out = add_kelvin(ph, c, out, (x.src_w[q] * dts) / cc, cc);
        }
    }
    tnew[c] = out;
    return lost;
}

// One step of a block with a film or sources.  Every lane loads all six neighbours (a face off
// the grid reads the cell itself) and a skipped face adds +0.0 by a select: the flux starts at
// +0.0 and a sum of it with anything is never -0.0, so adding +0.0 leaves it as skipping the
// face would, and the loads issue together.  MARCH: work items of 64 x 16 columns by z-chunks
// (workgroups loop over them in order); a thread marches its column up the chunk with T and k
// of the planes below, at and above its cell in registers, and the face below a cell is the
// face above the one under it, series_nz(k below, k) either way, so it is reused.  Otherwise one
// cell a thread (small blocks, whose columns and planes give fewer workgroups than the chip
// holds).
template <bool MARCH>
__global__ __launch_bounds__(EXT_THREADS) void heat_step_ext(const double *__restrict__ told,
                                                     double *__restrict__ tnew,
                                                     const double *__restrict__ kc, double kface,
                                                     const double *__restrict__ mob, double mob0,
                                                     const double *__restrict__ cap, double cap0,
                                                     ExtCells x, double dts,
                                                     int64_t nx, int64_t ny, int64_t nz,
                                                     double *__restrict__ lost_part, PhaseArgs ph,
                                                     const AuditState *st) {
    __shared__ double red[EXT_THREADS / 64];
    if (st && st->halted) return;  // a device-resident run that stopped
    static_assert(EXT_THREADS == 1024, "64 x 16 columns a workgroup");
    const int64_t plane = nx * ny;
    const uint32_t fm = x.faces;
    double lost = 0.0;
    if constexpr (MARCH) {
        const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
        const int64_t tnx = (nx + 63) / 64, tiles = tnx * ((ny + 15) / 16);
        int64_t zch = std::max<int64_t>(1, std::min<int64_t>(nz, (int64_t)gridDim.x / tiles));
        const int64_t zc = (nz + zch - 1) / zch;
        zch = (nz + zc - 1) / zc;
        for (int64_t it = blockIdx.x; it < tiles * zch; it += gridDim.x) {
            const int64_t tile = it % tiles, z0 = (it / tiles) * zc, z1 = std::min(nz, z0 + zc);
            const int64_t i = (tile % tnx) * 64 + tx, j = (tile / tnx) * 16 + ty;
            if (i >= nx || j >= ny) continue;
            const bool ox0 = i > 0, ox1 = i + 1 < nx, oy0 = j > 0, oy1 = j + 1 < ny;
            const int64_t dxm = ox0 ? -1 : 0, dxp = ox1 ? 1 : 0, dym = oy0 ? -nx : 0, dyp = oy1 ? nx : 0;
            const bool side = (!ox0 && (fm & 1)) || (!ox1 && (fm & 2)) || (!oy0 && (fm & 4)) || (!oy1 && (fm & 8));
            int64_t c = i + nx * j + plane * z0;
            const int64_t cb = z0 > 0 ? c - plane : c;
            double tb = told[cb], t = told[c];
            double kk = 0.0, fzm = kface;
            if (kc) kk = kc[c], fzm = series_nz(kc[cb], kk);
            for (int64_t k = z0; k < z1; k++, c += plane) {
                const bool oz0 = k > 0, oz1 = k + 1 < nz;
                const int64_t ca = oz1 ? c + plane : c;
                const double ta = told[ca];
                const double txm = told[c + dxm], txp = told[c + dxp], tym = told[c + dym], typ = told[c + dyp];
                double f0 = kface, f1 = kface, f2 = kface, f3 = kface, f5 = kface, ka = 0.0;
                if (kc) {
                    ka = kc[ca];
                    const double kxm = kc[c + dxm], kxp = kc[c + dxp], kym = kc[c + dym], kyp = kc[c + dyp];
                    f0 = series_nz(kxm, kk), f1 = series_nz(kk, kxp);
                    f2 = series_nz(kym, kk), f3 = series_nz(kk, kyp);
                    f5 = series_nz(kk, ka);
                }
                double flux = 0.0;
                flux = flux + (ox0 ? f0 * (txm - t) : 0.0);
                flux = flux + (ox1 ? f1 * (txp - t) : 0.0);
                flux = flux + (oy0 ? f2 * (tym - t) : 0.0);
                flux = flux + (oy1 ? f3 * (typ - t) : 0.0);
                flux = flux + (oz0 ? fzm * (tb - t) : 0.0);
                flux = flux + (oz1 ? f5 * (ta - t) : 0.0);
                const bool exposed = side || (!oz0 && (fm & 16)) || (!oz1 && (fm & 32));
                lost += ext_finish(x, c, t, flux, exposed, dts, mob, mob0, cap, cap0, ph, tnew);
                tb = t, t = ta, kk = ka, fzm = f5;
            }
        }
    } else {
        const int64_t n = plane * nz;
        for (int64_t c = blockIdx.x * EXT_THREADS + threadIdx.x; c < n; c += (int64_t)gridDim.x * EXT_THREADS) {
            int64_t i, j, k;
            coords_of(c, nx, ny, nz, i, j, k);
            const bool on[6] = {i > 0, i + 1 < nx, j > 0, j + 1 < ny, k > 0, k + 1 < nz};
            const int64_t nb[6] = {on[0] ? c - 1 : c, on[1] ? c + 1 : c, on[2] ? c - nx : c,
                                   on[3] ? c + nx : c, on[4] ? c - plane : c, on[5] ? c + plane : c};
            const double t = told[c];
            double tn[6], f[6];
#pragma unroll
            for (int q = 0; q < 6; q++) tn[q] = told[nb[q]];
            if (kc) {
                const double kk = kc[c];
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    const double kn = kc[nb[q]];
                    f[q] = (q & 1) ? series_nz(kk, kn) : series_nz(kn, kk);
                }
            } else {
#pragma unroll
                for (int q = 0; q < 6; q++) f[q] = kface;
            }
            double flux = 0.0;
#pragma unroll
            for (int q = 0; q < 6; q++) flux = flux + (on[q] ? f[q] * (tn[q] - t) : 0.0);
            bool exposed = false;
#pragma unroll
            for (int q = 0; q < 6; q++) exposed = exposed || (!on[q] && (fm & (1u << q)));
            lost += ext_finish(x, c, t, flux, exposed, dts, mob, mob0, cap, cap0, ph, tnew);
        }
    }
    const double total = block_sum<EXT_THREADS>(lost, red);
    if (threadIdx.x == 0) lost_part[blockIdx.x] = total;
}

// ScalarField::at for Solid3D, solid.rs:2045-2102: trilinear between cell centres, clamped at
// the outer half cells, weights renormalised over the corners with weight.  With void cells
// (`vd` not null): NaN when the cell the point rounds into is void, and void corners carry no
// weight.
__global__ void heat_sample(const double *__restrict__ T, const uint8_t *__restrict__ vd,
                            int64_t nx, int64_t ny, int64_t nz, double dx,
                            const double *__restrict__ xyz, int64_t n, double *__restrict__ out) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const double q[3] = {xyz[3 * p] / dx - 0.5, xyz[3 * p + 1] / dx - 0.5, xyz[3 * p + 2] / dx - 0.5};
        if (isnan(q[0]) || isnan(q[1]) || isnan(q[2])) {
            out[p] = NAN;
            continue;
        }
        const int64_t len[3] = {nx, ny, nz};
        int64_t at[3];
        double fr[3];
        for (int a = 0; a < 3; a++) {
            const int64_t last = len[a] - 1;
            if (q[a] <= 0.0) {
                at[a] = 0, fr[a] = 0.0;
            } else if (q[a] >= (double)last) {
                at[a] = last, fr[a] = 0.0;
            } else {
                const double fl = floor(q[a]);
                at[a] = (int64_t)fl, fr[a] = q[a] - fl;
            }
        }
        if (vd) {
            // `(v.round().max(0.0) as usize).min(n - 1)`, solid.rs:2076: half away from zero.
            int64_t near[3];
            for (int a = 0; a < 3; a++) {
                double r = round(q[a]);
                if (!(r > 0.0)) r = 0.0;
                near[a] = min((int64_t)r, len[a] - 1);
            }
            if (vd[near[0] + nx * (near[1] + ny * near[2])]) {
                out[p] = NAN;
                continue;
            }
        }
        const int64_t ia[2] = {at[0], min(at[0] + 1, nx - 1)};
        const int64_t jb[2] = {at[1], min(at[1] + 1, ny - 1)};
        const int64_t kc[2] = {at[2], min(at[2] + 1, nz - 1)};
        const double wa[2] = {1.0 - fr[0], fr[0]}, wb[2] = {1.0 - fr[1], fr[1]};
        const double wc[2] = {1.0 - fr[2], fr[2]};
        double sum = 0.0, weight = 0.0;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < 2; c++) {
                    const double w = wa[a] * wb[b] * wc[c];
                    const int64_t at_ = ia[a] + nx * (jb[b] + ny * kc[c]);
                    if (w > 0.0 && !(vd && vd[at_])) {
                        sum += w * T[at_];
                        weight += w;
                    }
                }
        out[p] = weight > 0.0 ? sum / weight : NAN;
    }
}


// The gaps' exchange, solid.rs:1866-1878: each gap cell applies the pairs it belongs to in pair
// order, each pair's watts coef*(Ta^4 - Tb^4) from the step's input, so a cell in several pairs
// sums them in the reference's order; the pair's two cells evaluate the same expression.
__global__ __launch_bounds__(256) void gap_exchange(const int64_t *__restrict__ gcell,
                                                    const int64_t *__restrict__ gstart,
                                                    const int64_t *__restrict__ gent, int64_t ngcell,
                                                    const int64_t *__restrict__ ga,
                                                    const int64_t *__restrict__ gb,
                                                    const double *__restrict__ gcoef,
                                                    const double *__restrict__ told,
                                                    const double *__restrict__ cap, double dts,
                                                    double *__restrict__ tnew, PhaseArgs ph,
                                                    const AuditState *st) {
    if (st && st->halted) return;  // a device-resident run that stopped
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ngcell;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = gcell[q];
        double t = tnew[c];
        for (int64_t e = gstart[q]; e < gstart[q + 1]; e++) {
            const int64_t g = gent[e] >> 1;
            const double ta = told[ga[g]], tb = told[gb[g]];
            const double watts = gcoef[g] * (pow4(ta) - pow4(tb));
            if (watts == 0.0) continue;
            const double joules = watts * dts;
            t = add_kelvin(ph, c, t, (gent[e] & 1) ? joules / cap[c] : -joules / cap[c], cap[c]);
        }
        tnew[c] = t;
    }
}

// The gaps' part of worst_rate_now, solid.rs:1159-1171: 4 coef T_hot^3 over each cell's
// capacity, as a per-block maximum.
__global__ __launch_bounds__(256) void gap_rates(const int64_t *__restrict__ ga,
                                                 const int64_t *__restrict__ gb,
                                                 const double *__restrict__ gcoef, int64_t npairs,
                                                 const double *__restrict__ T,
                                                 const double *__restrict__ cap,
                                                 double *__restrict__ block_max) {
    __shared__ double smax[256];
    double worst = 0.0;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < npairs;
         g += (int64_t)gridDim.x * blockDim.x) {
        const double hotter = rmax(T[ga[g]], T[gb[g]]);
        const double gc = 4.0 * gcoef[g] * pow3(hotter);
        if (cap[ga[g]] > 0.0) worst = rmax(worst, gc / cap[ga[g]]);
        if (cap[gb[g]] > 0.0) worst = rmax(worst, gc / cap[gb[g]]);
    }
    smax[threadIdx.x] = worst;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) smax[threadIdx.x] = rmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) block_max[blockIdx.x] = smax[0];
}

// A device-resident run of the extended step halts on the device: audit breaches set 1 (as every
// audited run does); the limit sets 2 for a NaN ratio (a substance with no diffusivity) and 3 for
// a ratio over the limit, before the step runs.
constexpr int HALT_BREACH_FLAG = 1, HALT_NO_DIFFUSIVITY = 2, HALT_LIMIT = 3;

// check_ratio on the device for one step of a device-resident run: the block's static rate, the
// gaps' and the film's per-block maxima at the step's input, combined by f64::max (exact, so in
// any order), then dt * rate against 1 + 6e-12.  The ratio goes to `after`, the step to
// breach_step.
__global__ __launch_bounds__(256) void ext_limit(const double *__restrict__ gpart, int ng,
                                                 const double *__restrict__ fpart, int nf,
                                                 double worst_rate, double dt, AuditState *st,
                                                 int64_t step) {
    __shared__ double smax[256];
    if (st->halted) return;
    double m = -INFINITY;
    for (int i = threadIdx.x; i < ng; i += 256) m = rmax(m, gpart[i]);
    for (int i = threadIdx.x; i < nf; i += 256) m = rmax(m, fpart[i]);
    smax[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) smax[threadIdx.x] = rmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double rate = std::isnan(worst_rate) ? worst_rate : rmax(worst_rate, smax[0]);
        const double ratio = dt * rate;
        if (std::isnan(ratio)) {
            st->halted = HALT_NO_DIFFUSIVITY;
            st->breach_step = step;
        } else if (ratio > 1.0 + 6e-12) {
            st->halted = HALT_LIMIT;
            st->breach_step = step;
            st->after = ratio;
        }
    }
}

// The books of a device-resident run after each step, as the host keeps them for step_ext and
// advance_ext: the film's loss of the step as the left fold of the sweep's per-block sums, added
// to the running loss; the supply after the step (the left fold of W * dt over the generating
// cells, solid.rs:1856-1862, which the state does not enter, is folded on the host and read from
// `supplied_after`); the stored heat as the fixed-order sum of its partials; then the ledger
// (0 + stored) + lost + (-supplied) audited against the step before.  books = (lost, supplied).
static_assert(EXT_SWEEP_MAX / LOSS_CHUNK <= 1024, "ext_books folds a chunk of loss partials a thread");
__global__ __launch_bounds__(1024) void ext_books(const double *__restrict__ lost_part, int nlost,
                                                 const double *__restrict__ supplied_after,
                                                 const double *__restrict__ stored_part,
                                                 int64_t nstored, double *books, AuditState *st,
                                                 int64_t step, double *totals) {
    __shared__ double scratch[16];
    __shared__ double lp[EXT_SWEEP_MAX];
    __shared__ double chunk_sum[EXT_SWEEP_MAX / LOSS_CHUNK];
    if (st->halted) return;
    // loss_fold: the partials are loaded in parallel, each chunk folded in order by a thread of
    // its own, then the chunks' sums in order by thread 0.
    for (int i = threadIdx.x; i < nlost; i += 1024) lp[i] = lost_part[i];
    __syncthreads();
    const int chunks = (nlost + LOSS_CHUNK - 1) / LOSS_CHUNK;
    if ((int)threadIdx.x < chunks) {
        const int c = threadIdx.x * LOSS_CHUNK;
        double acc = 0.0;
        for (int i = c; i < min(nlost, c + LOSS_CHUNK); i++) acc += lp[i];
        chunk_sum[threadIdx.x] = acc;
    }
    const double stored = fixed_order_sum_1024(stored_part, nstored, scratch);  // syncs the block
    if (threadIdx.x == 0) {
        double lost = 0.0;
        for (int c = 0; c < chunks; c++) lost += chunk_sum[c];
        const double lost_total = books[0] + lost;
        const double supplied = supplied_after[step];
        books[0] = lost_total, books[1] = supplied;
        double total = 0.0;
        total += stored;
        total += lost_total;
        total += -supplied;
        const double scale = rmax(rmax(rmax(0.0, std::fabs(stored)), std::fabs(lost_total)), std::fabs(supplied));
        audit_step(st, total, scale, step, totals);
    }
}

static FilmFaces film_faces(const pm_solid3d *s) {
    FilmFaces F{};
    for (int f = 0; f < 6; f++) {
        const int64_t cells = f < 2 ? s->ny * s->nz : f < 4 ? s->nx * s->nz : s->nx * s->ny;
        F.on[f] = s->film[f].on;
        F.ambient[f] = s->film[f].ambient;
        F.h[f] = s->film[f].h;
        // env.area / cells_on(face).max(1) as f64, solid.rs:1201
        F.share[f] = s->film[f].area / (double)std::max<int64_t>(cells, 1);
    }
    F.dx = s->dx;
    return F;
}

// The listed cells of the exposed faces and the generating cells, on the device.
static int ensure_tables(pm_solid3d *s) {
    if (s->tables_valid) return PM_OK;
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz;
    std::vector<int64_t> cells;
    for (int f = 0; f < 6; f++) {
        if (!s->film[f].on) continue;
        const int axis = f / 2;
        const int64_t at = (f % 2 == 0) ? 0 : (axis == 0 ? nx : axis == 1 ? ny : nz) - 1;
        for (int64_t k = 0; k < nz; k++)
            for (int64_t j = 0; j < ny; j++)
                for (int64_t i = 0; i < nx; i++) {
                    const int64_t q = axis == 0 ? i : axis == 1 ? j : k;
                    if (q == at) cells.push_back(i + nx * (j + ny * k));
                }
    }
    std::sort(cells.begin(), cells.end());
    cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
    s->nboundary = (int64_t)cells.size();
    if (!cells.empty()) {
        std::vector<double> emis(cells.size()), kcond(cells.size());
        for (size_t b = 0; b < cells.size(); b++) {
            const pm_substance &sub = s->subs[s->which.empty() ? 0 : s->which[(size_t)cells[b]]];
            // thermal.map_or(0.0, emissivity) and thermal.map_or(INFINITY, conductivity),
            // solid.rs:1193-1195: a substance that does not say has neither.
            const bool thermal = !std::isnan(sub.conductivity) && !std::isnan(sub.specific_heat);
            emis[b] = thermal ? sub.emissivity : 0.0;
            kcond[b] = thermal ? sub.conductivity : INFINITY;
        }
        PM_CHECK(s->bcell.alloc(sizeof(int64_t) * cells.size()));
        PM_CHECK(s->bemis.alloc(sizeof(double) * cells.size()));
        PM_CHECK(s->bkcond.alloc(sizeof(double) * cells.size()));
        PM_HIP(hipMemcpy(s->bcell.d(), cells.data(), sizeof(int64_t) * cells.size(), hipMemcpyHostToDevice));
        PM_HIP(hipMemcpy(s->bemis.d(), emis.data(), sizeof(double) * cells.size(), hipMemcpyHostToDevice));
        PM_HIP(hipMemcpy(s->bkcond.d(), kcond.data(), sizeof(double) * cells.size(), hipMemcpyHostToDevice));
        if (!s->shed) PM_CHECK(s->shed.alloc(sizeof(double) * (size_t)s->n()));
    }
    s->nsource = 0;
    for (const auto &e : s->source) s->nsource += !s->is_void(e.first);
    if (s->nsource) {
        if (s->nsource > INT32_MAX) {
            set_error("%s: %lld generating cells, more than the source index map holds", s->name.c_str(),
                      (long long)s->nsource);
            return PM_ERR_ARGUMENT;
        }
        std::vector<int32_t> map((size_t)s->n(), -1);
        std::vector<double> sw;
        for (const auto &e : s->source)
            if (!s->is_void(e.first)) map[(size_t)e.first] = (int32_t)sw.size(), sw.push_back(e.second);
        PM_CHECK(s->src_of.alloc(sizeof(int32_t) * map.size()));
        PM_CHECK(s->src_w.alloc(sizeof(double) * sw.size()));
        PM_HIP(hipMemcpy(s->src_of.d(), map.data(), sizeof(int32_t) * map.size(), hipMemcpyHostToDevice));
        PM_HIP(hipMemcpy(s->src_w.d(), sw.data(), sizeof(double) * sw.size(), hipMemcpyHostToDevice));
    }
    // The sweep's loss partials, or the film rates' of max_stable_dt.
    if (!s->ext_part) PM_CHECK(s->ext_part.alloc(sizeof(double) * std::max(EXT_SWEEP_MAX, EXT_BLOCKS)));
    s->tables_valid = true;
    return PM_OK;
}

// Whether a substance melts (FusionProps present, with thermal properties to melt with) and
// whether its liquid differs from its solid (solid.rs:621-629, 640-646).
static bool melts(const pm_substance &m) {
    return !std::isnan(m.melting_point) && !std::isnan(m.latent_heat) && !std::isnan(m.conductivity) &&
           !std::isnan(m.specific_heat);
}
static bool two_phased(const pm_substance &m) {
    return !std::isnan(m.liquid_conductivity) && !std::isnan(m.conductivity) && !std::isnan(m.specific_heat) &&
           (m.liquid_conductivity != m.conductivity || m.liquid_specific_heat != m.specific_heat);
}

// The per-cell phase tables of resolve (solid.rs:599-646) from each cell's substance: solid and
// liquid conductivity and capacity, melting point (INFINITY for a substance that does not melt)
// and latent joules.  The first build gives every cell its fraction from its temperature when
// `fresh` (the constructor's), zero otherwise (solid.rs:648-675).
static int build_phase(pm_solid3d *s, bool fresh, double initial) {
    const int64_t n = s->n();
    const double volume = s->dx * s->dx * s->dx;
    PM_CHECK(promote_c(s));
    std::vector<double> ks((size_t)n), kl((size_t)n), cs((size_t)n), cl((size_t)n);
    s->h_point.assign((size_t)n, INFINITY);
    s->h_latent.assign((size_t)n, 0.0);
    s->nlatent = 0;
    for (int64_t c = 0; c < n; c++) {
        const pm_substance &m = s->subs[s->which.empty() ? 0 : s->which[(size_t)c]];
        const double k_solid = m.conductivity;           // NaN for a substance that does not say
        const double c_solid = (m.density * volume) * m.specific_heat;
        const bool liquid = !std::isnan(m.liquid_conductivity) && !std::isnan(m.liquid_specific_heat);
        ks[(size_t)c] = k_solid, cs[(size_t)c] = c_solid;
        kl[(size_t)c] = liquid ? m.liquid_conductivity : k_solid;
        cl[(size_t)c] = liquid ? (m.density * volume) * m.liquid_specific_heat : c_solid;
        if (melts(m)) {
            s->h_point[(size_t)c] = m.melting_point;
            s->h_latent[(size_t)c] = (m.density * volume) * m.latent_heat;
            s->nlatent += s->h_latent[(size_t)c] > 0.0;
        }
    }
    const size_t bytes = sizeof(double) * (size_t)n;
    if (!s->phase) {
        PM_CHECK(s->ph_ks.alloc(bytes));
        PM_CHECK(s->ph_kl.alloc(bytes));
        PM_CHECK(s->ph_cs.alloc(bytes));
        PM_CHECK(s->ph_cl.alloc(bytes));
        PM_CHECK(s->ph_point.alloc(bytes));
        PM_CHECK(s->ph_latent.alloc(bytes));
        PM_CHECK(s->melted.alloc(bytes));
        PM_CHECK(s->saved_melted.alloc(bytes));
        std::vector<double> phi((size_t)n, 0.0);
        if (fresh)
            for (int64_t c = 0; c < n; c++) phi[(size_t)c] = initial > s->h_point[(size_t)c] ? 1.0 : 0.0;
        PM_HIP(hipMemcpy(s->melted.d(), phi.data(), bytes, hipMemcpyHostToDevice));
        PM_HIP(hipMemcpy(s->saved_melted.d(), phi.data(), bytes, hipMemcpyHostToDevice));
        s->phase = true;
    }
    PM_HIP(hipMemcpy(s->ph_ks.d(), ks.data(), bytes, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->ph_kl.d(), kl.data(), bytes, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->ph_cs.d(), cs.data(), bytes, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->ph_cl.d(), cl.data(), bytes, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->ph_point.d(), s->h_point.data(), bytes, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->ph_latent.d(), s->h_latent.data(), bytes, hipMemcpyHostToDevice));
    s->tables_valid = false;
    return PM_OK;
}

// A two-phase block's operator from the fractions it carries (solid.rs:686-702), before resolve.
static int remix(pm_solid3d *s) {
    if (!s->two_phase) return PM_OK;
    PM_CHECK(promote_c(s));
    PM_CHECK(promote_k(s));
    mix_phases<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(
        s->melted.d(), s->ph_ks.d(), s->ph_kl.d(), s->ph_cs.d(), s->ph_cl.d(),
        s->nvoid ? s->dvoid.as<uint8_t>() : nullptr, s->n(), s->dx, s->kcell.d(), s->cap.d(), s->mob.d());
    PM_HIP(hipGetLastError());
    s->k_checked = false;
    s->heat_capacity_valid = false;
    return PM_OK;
}

// Void cells hold and conduct nothing: the general form with their k, C and mobility zeroed.
// Called after anything that writes the per-cell arrays (a void cell stays void, solid.rs:519).
static int apply_voids(pm_solid3d *s) {
    if (!s->nvoid) return PM_OK;
    PM_CHECK(promote_c(s));
    void_zero<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(s->dvoid.as<uint8_t>(), s->n(),
                                                            s->kcell.d(), s->cap.d(), s->mob.d());
    PM_HIP(hipGetLastError());
    s->k_checked = false;
    s->heat_capacity_valid = false;
    return PM_OK;
}

// Solid3D::find_gaps, solid.rs:791-847, on the host: along every x line, then y, then z, a run of
// void with solid at both ends pairs the two cells, coupled by sigma dx^2/(1/ea + 1/eb - 1) when
// both radiate.  Then the per-cell lists the exchange kernel walks.
static int find_gaps(pm_solid3d *s) {
    s->gaps.clear();
    s->ngcell = 0;
    s->tables_valid = false;
    if (!s->nvoid) return PM_OK;
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz;
    const double area = s->dx * s->dx;
    auto emissivity = [&](int64_t c) {
        const pm_substance &sub = s->subs[s->which.empty() ? 0 : s->which[(size_t)c]];
        const bool thermal = !std::isnan(sub.conductivity) && !std::isnan(sub.specific_heat);
        return thermal ? sub.emissivity : 0.0;  // thermal.map_or(0.0, emissivity)
    };
    auto walk = [&](int64_t first, int64_t stride, int64_t len) {
        int64_t last = -1;
        bool between = false;
        for (int64_t q = 0; q < len; q++) {
            const int64_t c = first + q * stride;
            if (s->voids[(size_t)c]) {
                between = last >= 0;
                continue;
            }
            if (last >= 0 && between) {
                const double ea = emissivity(last), eb = emissivity(c);
                if (ea > 0.0 && eb > 0.0) {
                    const double resistance = 1.0 / ea + 1.0 / eb - 1.0;
                    if (resistance > 0.0) s->gaps.push_back({last, c, STEFAN_BOLTZMANN * area / resistance});
                }
            }
            last = c;
            between = false;
        }
    };
    for (int64_t k = 0; k < nz; k++)
        for (int64_t j = 0; j < ny; j++) walk(nx * (j + ny * k), 1, nx);
    for (int64_t k = 0; k < nz; k++)
        for (int64_t i = 0; i < nx; i++) walk(i + nx * ny * k, nx, ny);
    for (int64_t j = 0; j < ny; j++)
        for (int64_t i = 0; i < nx; i++) walk(i + nx * j, nx * ny, nz);
    if (s->gaps.empty()) return PM_OK;
    const size_t np = s->gaps.size();
    std::vector<int64_t> ga(np), gb(np);
    std::vector<double> gc(np);
    std::map<int64_t, std::vector<int64_t>> per_cell;  // each cell's entries in pair order
    for (size_t g = 0; g < np; g++) {
        ga[g] = s->gaps[g].a, gb[g] = s->gaps[g].b, gc[g] = s->gaps[g].coef;
        per_cell[ga[g]].push_back(2 * (int64_t)g);
        per_cell[gb[g]].push_back(2 * (int64_t)g + 1);
    }
    std::vector<int64_t> cells, start{0}, ent;
    for (const auto &e : per_cell) {
        cells.push_back(e.first);
        ent.insert(ent.end(), e.second.begin(), e.second.end());
        start.push_back((int64_t)ent.size());
    }
    s->ngcell = (int64_t)cells.size();
    PM_CHECK(s->gap_a.alloc(sizeof(int64_t) * np));
    PM_CHECK(s->gap_b.alloc(sizeof(int64_t) * np));
    PM_CHECK(s->gap_coef.alloc(sizeof(double) * np));
    PM_CHECK(s->gcell.alloc(sizeof(int64_t) * cells.size()));
    PM_CHECK(s->gstart.alloc(sizeof(int64_t) * start.size()));
    PM_CHECK(s->gent.alloc(sizeof(int64_t) * ent.size()));
    PM_CHECK(s->gap_part.alloc(sizeof(double) * EXT_BLOCKS));
    PM_HIP(hipMemcpy(s->gap_a.d(), ga.data(), sizeof(int64_t) * np, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->gap_b.d(), gb.data(), sizeof(int64_t) * np, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->gap_coef.d(), gc.data(), sizeof(double) * np, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->gcell.d(), cells.data(), sizeof(int64_t) * cells.size(), hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->gstart.d(), start.data(), sizeof(int64_t) * start.size(), hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(s->gent.d(), ent.data(), sizeof(int64_t) * ent.size(), hipMemcpyHostToDevice));
    return PM_OK;
}

// Solid3D::worst_rate_now, solid.rs:1148-1185: the resolved limit, raised by each exposed cell's
// film at its present temperature.  Synchronises the handle's stream.
static int worst_rate_now(pm_solid3d *s, double *out) {
    *out = s->worst_rate;
    if ((!s->exposed() && s->gaps.empty()) || std::isnan(s->worst_rate)) return PM_OK;
    double gap_worst = 0.0;
    if (!s->gaps.empty()) {
        const int64_t np = (int64_t)s->gaps.size();
        const int gblocks = (int)std::min<int64_t>(EXT_BLOCKS, (np + 255) / 256);
        gap_rates<<<gblocks, 256, 0, s->stream>>>(s->gap_a.as<int64_t>(), s->gap_b.as<int64_t>(),
                                                  s->gap_coef.d(), np, s->T(), s->cap.d(),
                                                  s->gap_part.d());
        PM_HIP(hipGetLastError());
        std::vector<double> part((size_t)gblocks);
        PM_HIP(hipMemcpyAsync(part.data(), s->gap_part.d(), sizeof(double) * part.size(), hipMemcpyDeviceToHost, s->stream));
        PM_HIP(hipStreamSynchronize(s->stream));
        for (double w : part) gap_worst = std::fmax(gap_worst, w);
        *out = std::fmax(s->worst_rate, gap_worst);
    }
    if (!s->exposed()) return PM_OK;
    PM_CHECK(ensure_tables(s));
    if (s->nboundary == 0) return PM_OK;
    const int blocks = (int)std::min<int64_t>(EXT_BLOCKS, (s->nboundary + 255) / 256);
    film_rates<<<blocks, 256, 0, s->stream>>>(
        film_faces(s), s->bcell.as<int64_t>(), s->bemis.d(), s->bkcond.d(), s->nboundary, s->T(),
        s->kmode == K_PER_CELL ? s->kcell.d() : nullptr, s->kface,
        s->cmode == C_PER_CELL ? s->mob.d() : nullptr, s->mob0, s->nx, s->ny, s->nz,
        s->ext_part.d());
    PM_HIP(hipGetLastError());
    std::vector<double> part((size_t)blocks);
    PM_HIP(hipMemcpyAsync(part.data(), s->ext_part.d(), sizeof(double) * part.size(), hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    double worst = std::fmax(s->worst_rate, gap_worst);
    for (double p : part) worst = std::fmax(worst, p);
    *out = worst;
    return PM_OK;
}

// The film and source tables the extended sweep reads (ensure_tables has built them).
static ExtCells ext_cells(const pm_solid3d *s) {
    ExtCells x{};
    const bool film = s->exposed() && s->nboundary > 0;
    for (int f = 0; f < 6; f++)
        if (s->film[f].on) x.faces |= 1u << f;
    x.shed = film ? s->shed.d() : nullptr;
    x.src_of = s->nsource ? s->src_of.as<int32_t>() : nullptr;
    x.src_w = s->src_w.d();
    return x;
}

// One step of a block with a film or sources (the caller has checked the limit and applied the
// bus heat): the sweep into the other buffer with the film's shed and the sources, the gaps, and
// the books: `lost` by the fixed-order sum of what the film removed, `supplied` by the
// reference's own left fold over the generating cells (solid.rs:1856-1862).
static int step_ext(pm_solid3d *s, double dt) {
    PM_CHECK(ensure_tables(s));
    const double *kc = s->kmode == K_PER_CELL ? s->kcell.d() : nullptr;
    const double *mob = s->cmode == C_PER_CELL ? s->mob.d() : nullptr;
    const double *cap = s->cmode == C_PER_CELL ? s->cap.d() : nullptr;
    double *out = s->t[s->cur ^ 1].d();
    const ExtCells xc = ext_cells(s);
    if (xc.shed) {
        const int blocks = (int)std::min<int64_t>(1024, (s->nboundary + 255) / 256);
        film_shed<<<blocks, 256, 0, s->stream>>>(film_faces(s), s->bcell.as<int64_t>(), s->bemis.d(), s->bkcond.d(),
                                                 s->nboundary, s->T(), s->shed.d(), s->nx, s->ny, s->nz,
                                                 nullptr);
        PM_HIP(hipGetLastError());
    }
    const ExtSweep sw = ext_sweep_plan(s->nx, s->ny, s->nz);
    const int sweep_blocks = sw.blocks;
    (sw.march ? heat_step_ext<true> : heat_step_ext<false>)<<<sweep_blocks, EXT_THREADS, 0, s->stream>>>(s->T(), out, kc, s->kface, mob, s->mob0, cap,
                                                     s->cap0, xc, dt,
                                                     s->nx, s->ny, s->nz, s->ext_part.d(),
                                                     s->phase_args(), nullptr);
    PM_HIP(hipGetLastError());
    if (s->ngcell) {
        gap_exchange<<<grid_for(s->ngcell, 256), 256, 0, s->stream>>>(
            s->gcell.as<int64_t>(), s->gstart.as<int64_t>(), s->gent.as<int64_t>(), s->ngcell,
            s->gap_a.as<int64_t>(), s->gap_b.as<int64_t>(), s->gap_coef.d(), s->T(), s->cap.d(), dt, out,
            s->phase_args(), nullptr);
        PM_HIP(hipGetLastError());
    }
    std::vector<double> part((size_t)sweep_blocks);
    PM_HIP(hipMemcpyAsync(part.data(), s->ext_part.d(), sizeof(double) * part.size(), hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    s->cur ^= 1;
    s->lost += loss_fold(part.data(), sweep_blocks);
    for (const auto &e : s->source)
        if (!s->is_void(e.first)) s->supplied += e.second * dt;  // void makes nothing (solid.rs:1857)
    // A two-phase block re-derives its operator, because the fractions moved (solid.rs:1897-1899).
    if (s->two_phase) {
        PM_CHECK(remix(s));
        PM_CHECK(resolve(s));
    }
    return PM_OK;
}

static int check_ratio(const pm_solid3d *cs, double dt, pm_violation *v, int64_t step) {
    // The limit of a block with an exposed face depends on its state; reading it runs a kernel
    // on the handle's stream, which a const query may do (Domain::max_stable_dt is &self).
    pm_solid3d *s = const_cast<pm_solid3d *>(cs);
    double rate = s->worst_rate;
    PM_CHECK(worst_rate_now(s, &rate));
    const double ratio = dt * rate;
    std::string site = s->name + " (explicit 3D conduction)";
    if (std::isnan(ratio)) {
        violation_at(v, s->name.c_str(), "substance has no diffusivity", INFINITY, step);
        return PM_VIOLATION;
    }
    if (ratio > 1.0 + 6e-12) {
        const double F = 1.0 / 6.0;
        fill_violation(v, "Fourier number", site.c_str(), F, ratio * F, F, 1e-12, step);
        return PM_VIOLATION;
    }
    return PM_OK;
}

// Solid3D::stored_heat, solid.rs:1557-1614: each cell's enthalpy above the reference state, in
// enthalpy for a cell that melts, capacity*(t - reference) otherwise.
__global__ __launch_bounds__(256) void heat_stored_kernel(const double *T, const double *cap,
                                                          double cap0, double tref, int64_t n,
                                                          int64_t per_block, double *partials,
                                                          PhaseArgs ph) {
    __shared__ double red[4];
    const int64_t lo = blockIdx.x * per_block, hi = std::min(n, lo + per_block);
    double acc = 0.0;
    for (int64_t c = lo + threadIdx.x; c < hi; c += 256) {
        if (ph.melted && ph.latent[c] > 0.0) {
            const double point = ph.point[c], latent = ph.latent[c];
            const double r = tref > point ? latent + ph.cl[c] * (tref - point) : ph.cs[c] * (tref - point);
            acc += enthalpy_of(ph, c, T[c]) - r;
        } else {
            acc += (cap ? cap[c] : cap0) * (T[c] - tref);
        }
    }
    double total = block_sum<256>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

static int heat_stored_in(pm_solid3d *s, const RunSpace &w, double *stored) {
    const int64_t n = s->n();
    const int64_t per = stored_per_block(n), blocks = (n + per - 1) / per;
    heat_stored_kernel<<<(unsigned)blocks, 256, 0, s->stream>>>(
        s->T(), s->cmode == C_PER_CELL ? s->cap.d() : nullptr, s->cap0, s->reference, n, per,
        w.partials, s->phase_args());
    PM_HIP(hipGetLastError());
    heat_finalize<<<1, 1024, 0, s->stream>>>(w.partials, blocks, w.audit, 0.0, 0.0, -1, nullptr,
                                             w.stored);
    PM_HIP(hipGetLastError());
    PM_HIP(hipMemcpyAsync(stored, w.stored, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

int heat_stored(pm_solid3d *s, double *stored) {
    RunSpace w;
    PM_CHECK(handle_space(s, &w));
    return heat_stored_in(s, w, stored);
}

static int reset_audit(pm_solid3d *s, AuditState *audit, double total, double scale, double tol) {
    audit_reset<<<1, 1, 0, s->stream>>>(audit, total, scale, tol);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

static int reset_audit(pm_solid3d *s, double total, double scale, double tol) {
    return reset_audit(s, s->audit.as<AuditState>(), total, scale, tol);
}

// Fills n doubles with v (the totals of steps a breach stopped stay NaN).
__global__ void fill_nan(double *p, int64_t n) {
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        p[i] = __longlong_as_double(0x7FF8000000000000LL);
}

// The plain branch of advance: steps in launches of two (paired sweep) or one, each launch's
// audit on the device, in chunks of at most `chunk` steps whose totals are staged in w.totals
// and copied out after the chunk (a chunk is even unless it is the run's last, so the pairing of
// steps is the same for any chunk length).  Step indices the kernels see are relative to the
// chunk; a breach is reported at its index in the run.
static int advance_plain(pm_solid3d *s, double dt, uint64_t steps, double tolerance,
                         double *totals, double *elapsed_ms, pm_violation *v, const RunSpace &w,
                         uint64_t chunk, bool profiled) {
    PM_CHECK(check_ratio(s, dt, v, 0));
    // The opening ledger.
    double stored0 = 0.0;
    PM_CHECK(heat_stored_in(s, w, &stored0));
    double total0 = 0.0;
    total0 += stored0;
    total0 += s->lost;
    total0 += -s->supplied;
    const double scale0 = rmax(rmax(rmax(0.0, std::fabs(stored0)), std::fabs(s->lost)), std::fabs(s->supplied));
    PM_CHECK(reset_audit(s, w.audit, total0, scale0, tolerance));
    if (profiled) {
        while (s->prof.size() < 3 * (size_t)steps) {
            hipEvent_t ev;
            PM_HIP(hipEventCreate(&ev));
            s->prof.push_back(ev);
        }
        s->prof_steps.assign((size_t)steps, 0);
        s->prof_used = 0;
    }
    // Launches of a chunk in order: (first step in the chunk, steps, buffer read).
    struct Launch {
        int64_t first;
        int steps;
        int cur;
    };
    std::vector<Launch> launches;
    const bool pairs = pairs_enabled(s);
    // A two-phase block re-mixes its conductivities every step: it keeps the full division.
    if (s->kmode == K_PER_CELL && !s->two_phase && steps > 1) PM_CHECK(check_k_range(s, w.kflag));
    PM_HIP(hipEventRecord(s->ev0, s->stream));
    int64_t launch_no = 0;
    if (steps == 0) PM_HIP(hipEventRecord(s->ev1, s->stream));
    for (uint64_t base = 0; base < steps; base += chunk) {
        const uint64_t len = std::min<uint64_t>(chunk, steps - base);
        RunSpace cw = w;
        if (!totals) cw.totals = nullptr;
        if (cw.totals) {
            fill_nan<<<grid_for((int64_t)len, 256), 256, 0, s->stream>>>(cw.totals, (int64_t)len);
            PM_HIP(hipGetLastError());
        }
        launches.clear();
        for (uint64_t n = 0; n < len;) {
            const int64_t l = profiled ? launch_no : -1;
            launch_no++;
            if (pairs && n + 1 < len) {
                launches.push_back({(int64_t)n, 2, s->cur});
                PM_CHECK(sweep_pair(s, dt, (int64_t)n, cw, l));
                n += 2;
            } else {
                launches.push_back({(int64_t)n, 1, s->cur});
                PM_CHECK(sweep_once(s, dt, (int64_t)n, cw, l));
                n += 1;
            }
        }
        if (base + len == steps) PM_HIP(hipEventRecord(s->ev1, s->stream));
        AuditState st;
        PM_HIP(hipMemcpyAsync(&st, w.audit, sizeof st, hipMemcpyDeviceToHost, s->stream));
        if (cw.totals)
            PM_HIP(hipMemcpyAsync(totals + base, cw.totals, sizeof(double) * len, hipMemcpyDeviceToHost, s->stream));
        PM_HIP(hipStreamSynchronize(s->stream));
        if (st.halted) {
            // Launches after the breach did nothing: the state is what the breaking step made
            // it, in the buffer its launch wrote.  A breach on the first step of a pair is
            // recomputed alone from the pair's input, which nothing has overwritten.
            for (const Launch &l : launches) {
                if (st.breach_step < l.first || st.breach_step >= l.first + l.steps) continue;
                s->cur = l.cur;
                if (l.steps == 2 && st.breach_step == l.first) {
                    PM_CHECK(sweep_once(s, dt, -1, cw, -1, true));
                    PM_HIP(hipStreamSynchronize(s->stream));
                } else {
                    s->cur ^= 1;
                }
                break;
            }
            if (totals)  // steps after the breach were not taken
                for (uint64_t i = base + len; i < steps; i++) totals[i] = NAN;
            fill_violation(v, "energy", "simulation", st.before, st.after, st.scale, tolerance,
                           (int64_t)base + st.breach_step);
            return PM_VIOLATION;
        }
    }
    PM_HIP(hipStreamSynchronize(s->stream));
    if (elapsed_ms) {
        float ms = 0.0f;
        PM_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        *elapsed_ms = ms;
    }
    if (profiled) PM_CHECK(collect_profile(s));
    return PM_OK;
}

// The static scratch of pm_solid3d_advance_static: the audit state, the conductivity-range flag,
// the stored-heat scalar, per-workgroup partials and staged per-step totals, laid out at fixed
// offsets.  One copy per device for the life of the process.
constexpr int64_t STATIC_PARTIALS = 65536;  // partials of one launch (2 x tiles of a pair)
constexpr int64_t STATIC_TOTALS = 8192;     // steps per chunk of staged totals (even)
constexpr size_t STATIC_OFF_AUDIT = 0, STATIC_OFF_FLAG = 256, STATIC_OFF_STORED = 512,
                 STATIC_OFF_PARTIALS = 1024,
                 STATIC_OFF_TOTALS = STATIC_OFF_PARTIALS + sizeof(double) * STATIC_PARTIALS;
constexpr size_t STATIC_BYTES = STATIC_OFF_TOTALS + sizeof(double) * STATIC_TOTALS;
static_assert(sizeof(AuditState) <= STATIC_OFF_FLAG, "audit state overlaps the flag");
static_assert(STATIC_TOTALS % 2 == 0, "chunks of totals must keep the pairing of steps");

}  // namespace pm

static __device__ __align__(256) unsigned char scratch[pm::STATIC_BYTES];

namespace pm {

// The device address of `scratch` on each device, looked up once per device with that device
// current (the name is a host-side stand-in, not an address a kernel may use).
constexpr int STATIC_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[STATIC_DEVICES] = {};

static int static_scratch(int device, unsigned char **out) {
    if (device < 0 || device >= STATIC_DEVICES) {
        set_error("static scratch: device %d is beyond the %d the library tracks", device, STATIC_DEVICES);
        return PM_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        PM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("hipGetSymbolAddress returned no address for the static scratch");
            return PM_ERR_DEVICE;
        }
        g_scratch_ptr[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[device];
    return PM_OK;
}

}  // namespace pm

extern "C" {

static int pm_solid3d_readings_kelvin(pm_solid3d *s, double out[4]);

int pm_solid3d_create(const char *name, const pm_substance *substance, size_t nx, size_t ny,
                      size_t nz, double dx, double initial, pm_solid3d **out) {
    if (!out || !substance) {
        set_error("pm_solid3d_create: null substance or output");
        return PM_ERR_ARGUMENT;
    }
    *out = nullptr;
    PM_CHECK(require_device());
    nx = std::max<size_t>(nx, 1), ny = std::max<size_t>(ny, 1), nz = std::max<size_t>(nz, 1);
    // The sweeps address a plane with 32-bit byte offsets: fewer than 2^29 cells a plane.
    if (nx > (1u << 30) || ny > (1u << 30) || nz > (1u << 30) || nx * ny >= ((size_t)1 << 29)) {
        set_error("pm_solid3d_create: counts (%zu, %zu, %zu) are beyond the sweep's indexing", nx, ny, nz);
        return PM_ERR_ARGUMENT;
    }
    std::unique_ptr<pm_solid3d> s(new pm_solid3d());
    s->name = name ? name : "block";
    s->subs.push_back(*substance);
    s->nx = (int64_t)nx, s->ny = (int64_t)ny, s->nz = (int64_t)nz;
    s->dx = dx;
    s->reference = initial;
    // resolve, solid.rs:604-626 and 704-782, for one material.
    s->k0 = substance->conductivity;
    s->cap0 = capacity_of(substance, dx);
    s->mob0 = dx / s->cap0;
    {
        double sum = s->k0 + s->k0;
        s->kface = sum > 0.0 ? 2.0 * s->k0 * s->k0 / sum : 0.0;
    }
    PM_HIP(hipGetDevice(&s->device));
    PM_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    PM_HIP(hipEventCreate(&s->ev0));
    PM_HIP(hipEventCreate(&s->ev1));
    const size_t bytes = sizeof(double) * (size_t)s->n();
    PM_CHECK(s->t[0].alloc(bytes));
    PM_CHECK(s->t[1].alloc(bytes));
    PM_CHECK(s->audit.alloc(sizeof(AuditState)));
    PM_CHECK(s->stored_dev.alloc(sizeof(double)));
    fill_value<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(s->t[0].d(), s->n(), initial);
    PM_HIP(hipGetLastError());
    PM_CHECK(reset_audit(s.get(), 0.0, 0.0, 1e-9));
    if (melts(*substance) || two_phased(*substance)) {
        s->two_phase = two_phased(*substance);
        PM_CHECK(build_phase(s.get(), true, initial));
        PM_CHECK(remix(s.get()));
    }
    PM_CHECK(resolve(s.get()));
    PM_HIP(hipStreamSynchronize(s->stream));
    *out = s.release();
    return PM_OK;
}

void pm_solid3d_destroy(pm_solid3d *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    for (hipEvent_t ev : s->prof) (void)hipEventDestroy(ev);
    hipStream_t st = s->stream;
    delete s;
    if (st) (void)hipStreamDestroy(st);
}

int pm_solid3d_counts(const pm_solid3d *s, size_t counts[3]) {
    if (!s || !counts) return PM_ERR_ARGUMENT;
    counts[0] = (size_t)s->nx, counts[1] = (size_t)s->ny, counts[2] = (size_t)s->nz;
    return PM_OK;
}

int pm_solid3d_fill(pm_solid3d *s, const pm_substance *sub, const uint8_t *mask) {
    if (!s || !sub || !mask) {
        set_error("pm_solid3d_fill: null argument");
        return PM_ERR_ARGUMENT;
    }
    const double k = sub->conductivity;
    const double c = capacity_of(sub, s->dx);
    const int64_t n = s->n();
    // The block's list of materials grows with every fill, selected cells or not (solid.rs:523).
    if (two_phased(*sub)) s->two_phase = true;
    bool any = false;
    for (int64_t i = 0; i < n && !any; i++) any = mask[i] != 0;
    if (!any) return PM_OK;
    bool all = true;
    for (int64_t i = 0; i < n && all; i++) all = mask[i] != 0;
    // Which substance each cell holds, for the film's emissivity and conductivity.
    if (all) {
        s->subs.assign(1, *sub);
        s->which.clear();
    } else {
        size_t id = 0;
        while (id < s->subs.size() && std::memcmp(&s->subs[id], sub, sizeof *sub) != 0) id++;
        if (id == s->subs.size()) {
            if (id > 65535) {
                set_error("pm_solid3d_fill: more than 65536 substances in one block");
                return PM_ERR_ARGUMENT;
            }
            s->subs.push_back(*sub);
        }
        if (s->which.empty() && id != 0) s->which.assign((size_t)n, 0);
        if (!s->which.empty())
            for (int64_t i = 0; i < n; i++)
                if (mask[i]) s->which[(size_t)i] = (uint16_t)id;
    }
    s->tables_valid = false;
    const bool k_same = s->kmode == K_UNIFORM && same_bits(k, s->k0);
    const bool c_same = s->cmode == C_UNIFORM && same_bits(c, s->cap0);
    const bool phased = s->phase || melts(*sub) || s->two_phase;
    if (all && !phased && (s->kmode == K_UNIFORM) && (s->cmode == C_UNIFORM)) {
        // The whole block becomes the new substance: still one material.
        s->k0 = k, s->cap0 = c, s->mob0 = s->dx / c;
        double sum = k + k;
        s->kface = sum > 0.0 ? 2.0 * k * k / sum : 0.0;
        return resolve(s);
    }
    if (k_same && c_same && !phased) return s->nvoid ? find_gaps(s) : PM_OK;
    if (!c_same) PM_CHECK(promote_c(s));
    if (!k_same) PM_CHECK(promote_k(s));
    DevBuf dmask;
    PM_CHECK(dmask.alloc((size_t)n));
    PM_HIP(hipMemcpyAsync(dmask.d(), mask, (size_t)n, hipMemcpyHostToDevice, s->stream));
    heat_fill<<<grid_for(n, 256), 256, 0, s->stream>>>(
        dmask.as<uint8_t>(), n, s->kmode == K_PER_CELL ? s->kcell.d() : nullptr,
        s->cmode == C_PER_CELL ? s->cap.d() : nullptr,
        s->cmode == C_PER_CELL ? s->mob.d() : nullptr, k, c, s->dx);
    PM_HIP(hipGetLastError());
    s->k_checked = false;
    if (s->phase || melts(*sub) || s->two_phase) {
        PM_CHECK(build_phase(s, false, 0.0));
        PM_CHECK(remix(s));
    }
    PM_CHECK(apply_voids(s));
    PM_CHECK(resolve(s));
    PM_CHECK(find_gaps(s));  // a new substance brings its own emissivity
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

int pm_solid3d_set_conductivity(pm_solid3d *s, const double *k) {
    if (!s || !k) return PM_ERR_ARGUMENT;
    if (s->kmode != K_PER_CELL) {
        PM_CHECK(s->kcell.alloc(sizeof(double) * (size_t)s->n()));
        s->kmode = K_PER_CELL;
    }
    PM_HIP(hipMemcpyAsync(s->kcell.d(), k, sizeof(double) * (size_t)s->n(), hipMemcpyHostToDevice, s->stream));
    s->k_checked = false;
    if (s->phase) {
        // Both phases of every cell take the given conductivity.
        PM_HIP(hipMemcpyAsync(s->ph_ks.d(), k, sizeof(double) * (size_t)s->n(), hipMemcpyHostToDevice, s->stream));
        PM_HIP(hipMemcpyAsync(s->ph_kl.d(), k, sizeof(double) * (size_t)s->n(), hipMemcpyHostToDevice, s->stream));
        PM_CHECK(remix(s));
    }
    PM_CHECK(apply_voids(s));
    PM_CHECK(resolve(s));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

int pm_solid3d_set_cells(pm_solid3d *s, const double *cells) {
    if (!s || !cells) return PM_ERR_ARGUMENT;
    PM_HIP(hipMemcpyAsync(s->T(), cells, sizeof(double) * (size_t)s->n(), hipMemcpyHostToDevice, s->stream));
    if (s->phase) {
        phase_from_temperature<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(
            s->T(), s->ph_point.d(), s->ph_latent.d(), s->n(), s->melted.d());
        PM_HIP(hipGetLastError());
    }
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

int pm_solid3d_copy_cells(pm_solid3d *dst, pm_solid3d *src) {
    if (!dst || !src) return PM_ERR_ARGUMENT;
    if (dst->nx != src->nx || dst->ny != src->ny || dst->nz != src->nz) {
        set_error("pm_solid3d_copy_cells: extents differ");
        return PM_ERR_ARGUMENT;
    }
    if (dst == src) return PM_OK;
    PM_HIP(hipStreamSynchronize(src->stream));
    PM_HIP(hipMemcpyAsync(dst->T(), src->T(), sizeof(double) * (size_t)dst->n(), hipMemcpyDeviceToDevice, dst->stream));
    PM_HIP(hipStreamSynchronize(dst->stream));
    return PM_OK;
}

int pm_solid3d_get_cells(pm_solid3d *s, double *cells) {
    if (!s || !cells) return PM_ERR_ARGUMENT;
    PM_HIP(hipMemcpyAsync(cells, s->T(), sizeof(double) * (size_t)s->n(), hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

int pm_solid3d_set_temperature(pm_solid3d *s, size_t i, size_t j, size_t k, double kelvin) {
    if (!s) return PM_ERR_ARGUMENT;
    if ((int64_t)i >= s->nx || (int64_t)j >= s->ny || (int64_t)k >= s->nz) return PM_OK;  // ignored, solid.rs:1346
    const int64_t idx = (int64_t)i + s->nx * ((int64_t)j + s->ny * (int64_t)k);
    PM_HIP(hipMemcpyAsync(s->T() + idx, &kelvin, sizeof(double), hipMemcpyHostToDevice, s->stream));
    double phi = 0.0;
    if (s->phase && s->h_latent[(size_t)idx] > 0.0) {
        // A temperature names the phase of a cell that melts (solid.rs:1352-1360).
        phi = kelvin > s->h_point[(size_t)idx] ? 1.0 : 0.0;
        PM_HIP(hipMemcpyAsync(s->melted.d() + idx, &phi, sizeof(double), hipMemcpyHostToDevice, s->stream));
    }
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

int pm_solid3d_set_region(pm_solid3d *s, const uint8_t *mask, double kelvin) {
    if (!s || !mask) return PM_ERR_ARGUMENT;
    DevBuf dm;
    PM_CHECK(dm.alloc((size_t)s->n()));
    PM_HIP(hipMemcpyAsync(dm.d(), mask, (size_t)s->n(), hipMemcpyHostToDevice, s->stream));
    set_region<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(
        s->T(), dm.as<uint8_t>(), s->n(), kelvin, s->phase ? s->ph_point.d() : nullptr,
        s->phase ? s->ph_latent.d() : nullptr, s->phase ? s->melted.d() : nullptr);
    PM_HIP(hipGetLastError());
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

int pm_solid3d_set_melted_fraction(pm_solid3d *s, size_t i, size_t j, size_t k, double fraction) {
    if (!s) return PM_ERR_ARGUMENT;
    if ((int64_t)i >= s->nx || (int64_t)j >= s->ny || (int64_t)k >= s->nz) return PM_OK;
    const int64_t idx = (int64_t)i + s->nx * ((int64_t)j + s->ny * (int64_t)k);
    if (!s->phase || !(s->h_latent[(size_t)idx] > 0.0)) return PM_OK;  // solid.rs:452-454
    // f64::clamp: NaN stays NaN.
    const double phi = std::isnan(fraction) ? fraction : std::fmin(std::fmax(fraction, 0.0), 1.0);
    const double point = s->h_point[(size_t)idx];
    double t = 0.0;
    PM_HIP(hipMemcpyAsync(&t, s->T() + idx, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    if (phi >= 1.0)
        t = rmax(t, point);
    else if (phi <= 0.0)
        t = std::fmin(t, point);
    else
        t = point;
    PM_HIP(hipMemcpyAsync(s->melted.d() + idx, &phi, sizeof(double), hipMemcpyHostToDevice, s->stream));
    PM_HIP(hipMemcpyAsync(s->T() + idx, &t, sizeof(double), hipMemcpyHostToDevice, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

int pm_solid3d_melted(pm_solid3d *s, double *out, double *volume) {
    if (!s) return PM_ERR_ARGUMENT;
    std::vector<double> phi((size_t)s->n(), 0.0);
    if (s->phase) {
        PM_HIP(hipMemcpyAsync(phi.data(), s->melted.d(), sizeof(double) * phi.size(), hipMemcpyDeviceToHost, s->stream));
        PM_HIP(hipStreamSynchronize(s->stream));
    }
    if (out) std::memcpy(out, phi.data(), sizeof(double) * phi.size());
    if (volume) {
        double total = 0.0;  // melted.iter().sum() * cell_volume, solid.rs:476-478
        for (double v : phi) total += v;
        *volume = total * (s->dx * s->dx * s->dx);
    }
    return PM_OK;
}

size_t pm_solid3d_melting_cells(const pm_solid3d *s) { return s ? (size_t)s->nlatent : 0; }

int pm_solid3d_deposit(pm_solid3d *s, size_t i, size_t j, size_t k, double joules) {
    if (!s) return PM_ERR_ARGUMENT;
    if ((int64_t)i >= s->nx || (int64_t)j >= s->ny || (int64_t)k >= s->nz) return PM_OK;  // refused, solid.rs:1368
    const int64_t idx = (int64_t)i + s->nx * ((int64_t)j + s->ny * (int64_t)k);
    double cap = 0.0, t = 0.0;
    PM_CHECK(cell_capacity(s, idx, &cap));
    PM_HIP(hipMemcpyAsync(&t, s->T() + idx, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    if (s->phase) {
        deposit_phase<<<1, 1, 0, s->stream>>>(s->T(), idx, joules, s->cap.d(), s->phase_args());
        PM_HIP(hipGetLastError());
        PM_HIP(hipStreamSynchronize(s->stream));
        s->absorbed += joules;
        return PM_OK;
    }
    const double rise = joules / cap;
    t += rise;  // add_kelvin, latent <= 0 (solid.rs:1634)
    PM_HIP(hipMemcpyAsync(s->T() + idx, &t, sizeof(double), hipMemcpyHostToDevice, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    s->absorbed += joules;
    return PM_OK;
}

double pm_solid3d_max_stable_dt(const pm_solid3d *s) {
    if (!s) return NAN;
    double rate = s->worst_rate;
    if (worst_rate_now(const_cast<pm_solid3d *>(s), &rate) != PM_OK) return NAN;
    if (std::isnan(rate)) return INFINITY;
    return 1.0 / rate;
}

double pm_solid3d_stability_ratio(const pm_solid3d *s, double dt) {
    if (!s) return NAN;
    double rate = s->worst_rate;
    if (worst_rate_now(const_cast<pm_solid3d *>(s), &rate) != PM_OK) return NAN;
    return dt * rate;
}

int pm_solid3d_refuses(const pm_solid3d *s, double dt, pm_violation *v) {
    if (!s) return PM_ERR_ARGUMENT;
    return check_ratio(s, dt, v, -1);
}

int pm_solid3d_step(pm_solid3d *s, double dt, double gained, pm_violation *v) {
    if (!s) return PM_ERR_ARGUMENT;
    PM_CHECK(check_ratio(s, dt, v, -1));
    if (gained != 0.0) {
        s->absorbed += gained;
        double total = 0.0;
        PM_CHECK(total_capacity(s, &total));
        const double rise = gained / total;
        if (s->phase)
            add_uniform_phase<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(
                s->T(), s->nvoid ? s->dvoid.as<uint8_t>() : nullptr, s->n(), rise, s->cap.d(), s->phase_args());
        else if (s->nvoid)
            add_uniform_void<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(s->T(), s->dvoid.as<uint8_t>(), s->n(), rise);
        else
            add_uniform<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(s->T(), s->n(), rise);
        PM_HIP(hipGetLastError());
    }
    if (s->extended()) return step_ext(s, dt);
    PM_CHECK(reset_audit(s, 0.0, 0.0, INFINITY));
    RunSpace w;
    PM_CHECK(handle_space(s, &w));
    PM_CHECK(sweep_once(s, dt, -1, w));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

int pm_solid3d_ledger(pm_solid3d *s, pm_books *books) {
    if (!s || !books) return PM_ERR_ARGUMENT;
    double stored = 0.0;
    PM_CHECK(heat_stored(s, &stored));
    books->stored = stored;
    books->lost = s->lost;
    books->supplied = s->supplied;
    return PM_OK;
}

// advance for a block with a film, sources, gaps or latent heat, resident on the device: every
// step's limit (film_rates, gap_rates, ext_limit), its sweep with the film's shed and the sources
// (heat_step_ext), its gaps, then its stored heat and books (ext_books, which audits): four to six
// launches a step, enqueued on the handle's stream with no
// host synchronisation until the run ends.  The arithmetic and the order of every sum are those
// of step_ext and the host books, so fields, books and ledger totals are bit-identical to
// stepping one call at a time.  A refusal (halted before its step) or a breach (after it) makes
// the later launches return at once, so the state is the one the last step that ran left.  A
// block of two phases re-derives its operator on the host after every step and runs
// advance_ext instead.
static int advance_ext_device(pm_solid3d *s, double dt, uint64_t steps, double tolerance,
                              double *totals, double *elapsed_ms, pm_violation *v) {
    PM_CHECK(ensure_tables(s));
    RunSpace w;
    PM_CHECK(handle_space(s, &w));
    double stored = 0.0;
    PM_CHECK(heat_stored_in(s, w, &stored));
    double total0 = 0.0;
    total0 += stored;
    total0 += s->lost;
    total0 += -s->supplied;
    const double scale0 = rmax(rmax(rmax(0.0, std::fabs(stored)), std::fabs(s->lost)), std::fabs(s->supplied));
    PM_CHECK(reset_audit(s, w.audit, total0, scale0, tolerance));
    if (!s->ext_books) PM_CHECK(s->ext_books.alloc(2 * sizeof(double)));
    const double books0[2] = {s->lost, s->supplied};
    PM_HIP(hipMemcpy(s->ext_books.d(), books0, sizeof books0, hipMemcpyHostToDevice));
    DevBuf dtot, dsup;
    PM_CHECK(dtot.alloc(sizeof(double) * (size_t)std::max<uint64_t>(steps, 1)));
    PM_CHECK(dsup.alloc(sizeof(double) * (size_t)std::max<uint64_t>(steps, 1)));
    fill_nan<<<grid_for((int64_t)std::max<uint64_t>(steps, 1), 256), 256, 0, s->stream>>>(dtot.d(), (int64_t)std::max<uint64_t>(steps, 1));
    PM_HIP(hipGetLastError());
    // The generating cells' W * dt in the order the books fold them (void cells make nothing).
    std::vector<double> joules;
    for (const auto &e : s->source)
        if (!s->is_void(e.first)) joules.push_back(e.second * dt);
    double supplied = s->supplied;
    std::vector<double> sup;

    const bool film = s->exposed() && s->nboundary > 0;
    const int64_t npairs = (int64_t)s->gaps.size();
    const int fblocks = film ? (int)std::min<int64_t>(EXT_BLOCKS, (s->nboundary + 255) / 256) : 0;
    const int gblocks = npairs ? (int)std::min<int64_t>(EXT_BLOCKS, (npairs + 255) / 256) : 0;
    if (!s->film_part) PM_CHECK(s->film_part.alloc(sizeof(double) * EXT_BLOCKS));
    const double *kc = s->kmode == K_PER_CELL ? s->kcell.d() : nullptr;
    const double *mob = s->cmode == C_PER_CELL ? s->mob.d() : nullptr;
    const double *cap = s->cmode == C_PER_CELL ? s->cap.d() : nullptr;
    const int64_t n = s->n();
    const int64_t per_stored = stored_per_block(n), nstored = (n + per_stored - 1) / per_stored;
    const FilmFaces faces = film_faces(s);
    const PhaseArgs ph = s->phase_args();
    AuditState *st = w.audit;
    const ExtCells xc = ext_cells(s);
    const ExtSweep sw = ext_sweep_plan(s->nx, s->ny, s->nz);
    const int sweep_blocks = sw.blocks;

    PM_HIP(hipEventRecord(s->ev0, s->stream));
    constexpr uint64_t CHUNK = 64;  // steps whose supply is folded and sent ahead of their launches
    for (uint64_t q = 0; q < steps; q++) {
        if (q % CHUNK == 0) {
            // Folded here while the device runs the steps before; the copy is ordered before the
            // chunk's launches on the stream (a copy from pageable memory is staged at the call).
            const uint64_t len = std::min(CHUNK, steps - q);
            sup.resize(len);
            for (uint64_t r = 0; r < len; r++) {
                for (double j : joules) supplied += j;
                sup[r] = supplied;
            }
            PM_HIP(hipMemcpyAsync(dsup.d() + q, sup.data(), sizeof(double) * len, hipMemcpyHostToDevice, s->stream));
        }
        const double *tin = s->t[(s->cur + q) & 1].d();
        double *tout = s->t[(s->cur + q + 1) & 1].d();
        if (gblocks)
            gap_rates<<<gblocks, 256, 0, s->stream>>>(s->gap_a.as<int64_t>(), s->gap_b.as<int64_t>(),
                                                      s->gap_coef.d(), npairs, tin, s->cap.d(), s->gap_part.d());
        if (film)
            film_rates_shed<<<fblocks, 256, 0, s->stream>>>(faces, s->bcell.as<int64_t>(), s->bemis.d(),
                                                            s->bkcond.d(), s->nboundary, tin, kc, s->kface,
                                                            mob, s->mob0, s->nx, s->ny, s->nz,
                                                            s->film_part.d(), s->shed.d(), st);
        ext_limit<<<1, 256, 0, s->stream>>>(s->gap_part.d(), gblocks, s->film_part.d(), fblocks,
                                            s->worst_rate, dt, st, (int64_t)q);
        (sw.march ? heat_step_ext<true> : heat_step_ext<false>)<<<sweep_blocks, EXT_THREADS, 0, s->stream>>>(tin, tout, kc, s->kface, mob, s->mob0, cap,
                                                         s->cap0, xc, dt, s->nx, s->ny, s->nz,
                                                         s->ext_part.d(), ph, st);
        if (s->ngcell)
            gap_exchange<<<grid_for(s->ngcell, 256), 256, 0, s->stream>>>(
                s->gcell.as<int64_t>(), s->gstart.as<int64_t>(), s->gent.as<int64_t>(), s->ngcell,
                s->gap_a.as<int64_t>(), s->gap_b.as<int64_t>(), s->gap_coef.d(), tin, s->cap.d(), dt,
                tout, ph, st);
        heat_stored_kernel<<<(unsigned)nstored, 256, 0, s->stream>>>(tout, cap, s->cap0, s->reference, n,
                                                                     per_stored, w.partials, ph);
        ext_books<<<1, 1024, 0, s->stream>>>(s->ext_part.d(), sweep_blocks, dsup.d(), w.partials, nstored,
                                              s->ext_books.d(), st, (int64_t)q, dtot.d());
        PM_HIP(hipGetLastError());
    }
    PM_HIP(hipEventRecord(s->ev1, s->stream));
    AuditState a;
    double books[2];
    PM_HIP(hipMemcpyAsync(&a, st, sizeof a, hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipMemcpyAsync(books, s->ext_books.d(), sizeof books, hipMemcpyDeviceToHost, s->stream));
    if (totals && steps)
        PM_HIP(hipMemcpyAsync(totals, dtot.d(), sizeof(double) * steps, hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    if (elapsed_ms) {
        float ms = 0.0f;
        PM_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        *elapsed_ms = ms;
    }
    // The steps that ran: up to a refusal's step, through a breach's.
    const uint64_t ran = !a.halted ? steps : a.halted == HALT_BREACH_FLAG ? (uint64_t)a.breach_step + 1
                                                                          : (uint64_t)a.breach_step;
    s->cur = (int)((s->cur + ran) & 1);
    s->lost = books[0], s->supplied = books[1];
    if (!a.halted) return PM_OK;
    if (a.halted == HALT_NO_DIFFUSIVITY) {
        violation_at(v, s->name.c_str(), "substance has no diffusivity", INFINITY, a.breach_step);
    } else if (a.halted == HALT_LIMIT) {
        const double F = 1.0 / 6.0;
        const std::string site = s->name + " (explicit 3D conduction)";
        fill_violation(v, "Fourier number", site.c_str(), F, a.after * F, F, 1e-12, a.breach_step);
    } else {
        fill_violation(v, "energy", "simulation", a.before, a.after, a.scale, tolerance, a.breach_step);
    }
    return PM_VIOLATION;
}

// advance for a block with a film or sources: the limit is checked before every step (it moves
// with the state), each step runs through step_ext, and the books are audited on the host after
// every step, stored + lost - supplied against the step before (conserved.rs:323-363).  A
// refusal or a breach stops the run with the state the last step left.
static int advance_ext(pm_solid3d *s, double dt, uint64_t steps, double tolerance, double *totals,
                       double *elapsed_ms, pm_violation *v) {
    double stored = 0.0;
    PM_CHECK(heat_stored(s, &stored));
    auto books = [&](double st, double *total, double *scale) {
        *total = 0.0;
        *total += st;
        *total += s->lost;
        *total += -s->supplied;
        *scale = rmax(rmax(rmax(0.0, std::fabs(st)), std::fabs(s->lost)), std::fabs(s->supplied));
    };
    double before = 0.0, sb = 0.0;
    books(stored, &before, &sb);
    PM_HIP(hipEventRecord(s->ev0, s->stream));
    for (uint64_t n = 0; n < steps; n++) {
        PM_CHECK(check_ratio(s, dt, v, (int64_t)n));
        PM_CHECK(step_ext(s, dt));
        PM_CHECK(heat_stored(s, &stored));
        double after = 0.0, sa = 0.0;
        books(stored, &after, &sa);
        if (totals) totals[n] = after;
        const double scale = rmax(rmax(rmax(std::fabs(before), std::fabs(after)), sb), sa);
        if (scale >= 1e-300 && std::fabs(after - before) / scale > tolerance) {
            fill_violation(v, "energy", "simulation", before, after, scale, tolerance, (int64_t)n);
            return PM_VIOLATION;
        }
        before = after, sb = sa;
    }
    PM_HIP(hipEventRecord(s->ev1, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    if (elapsed_ms) {
        float ms = 0.0f;
        PM_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        *elapsed_ms = ms;
    }
    return PM_OK;
}

int pm_solid3d_advance(pm_solid3d *s, double dt, uint64_t steps, double tolerance,
                       double *totals, double *elapsed_ms, pm_violation *v) {
    if (!s) return PM_ERR_ARGUMENT;
    if (s->extended())
        return s->two_phase ? advance_ext(s, dt, steps, tolerance, totals, elapsed_ms, v)
                            : advance_ext_device(s, dt, steps, tolerance, totals, elapsed_ms, v);
    RunSpace w;
    PM_CHECK(handle_space(s, &w));
    DevBuf dtot;
    if (totals && steps) {
        PM_CHECK(dtot.alloc(sizeof(double) * (size_t)steps));
        w.totals = dtot.d();
    }
    return advance_plain(s, dt, steps, tolerance, totals, elapsed_ms, v, w,
                         std::max<uint64_t>(steps, 1), s->profiling);
}

int pm_solid3d_advance_static(pm_solid3d *s, double dt, uint64_t steps, double tolerance,
                              double *totals, double *elapsed_ms, pm_violation *v) {
    if (!s) return PM_ERR_ARGUMENT;
    if (s->extended()) {
        set_error("pm_solid3d_advance_static: a block with a film, sources, gaps or a phase "
                  "change steps through the extended path, which needs per-call buffers; use "
                  "pm_solid3d_advance");
        return PM_ERR_ARGUMENT;
    }
    const int64_t need = partials_needed(s);
    if (need > STATIC_PARTIALS) {
        set_error("pm_solid3d_advance_static: the block needs %lld audit partials per launch, "
                  "beyond the static scratch's %lld; use pm_solid3d_advance",
                  (long long)need, (long long)STATIC_PARTIALS);
        return PM_ERR_ARGUMENT;
    }
    int current = 0;
    PM_HIP(hipGetDevice(&current));
    if (current != s->device) PM_HIP(hipSetDevice(s->device));
    unsigned char *base = nullptr;
    int rc = static_scratch(s->device, &base);
    if (rc == PM_OK) {
        RunSpace w;
        w.audit = reinterpret_cast<AuditState *>(base + STATIC_OFF_AUDIT);
        w.kflag = reinterpret_cast<int *>(base + STATIC_OFF_FLAG);
        w.stored = reinterpret_cast<double *>(base + STATIC_OFF_STORED);
        w.partials = reinterpret_cast<double *>(base + STATIC_OFF_PARTIALS);
        w.totals = reinterpret_cast<double *>(base + STATIC_OFF_TOTALS);
        rc = advance_plain(s, dt, steps, tolerance, totals, elapsed_ms, v, w,
                           (uint64_t)STATIC_TOTALS, false);
    }
    if (current != s->device) (void)hipSetDevice(current);
    return rc;
}

int pm_solid3d_losing_from(pm_solid3d *s, int face, const pm_environment *env) {
    if (!s || !env || face < 0 || face > 5) {
        set_error("pm_solid3d_losing_from: face must be 0..5 (XMin, XMax, YMin, YMax, ZMin, ZMax)");
        return PM_ERR_ARGUMENT;
    }
    s->film[face].on = 1;
    s->film[face].ambient = env->ambient;
    s->film[face].h = env->convection_w_per_m2_k;
    s->film[face].area = env->area;
    s->tables_valid = false;
    return PM_OK;
}

int pm_solid3d_dissipating(pm_solid3d *s, double watts, const uint8_t *mask) {
    if (!s || !mask) {
        set_error("pm_solid3d_dissipating: null argument");
        return PM_ERR_ARGUMENT;
    }
    // Over the selected cells that are there (solid.rs:1066-1075).
    int64_t chosen = 0;
    for (int64_t c = 0; c < s->n(); c++) chosen += mask[c] != 0 && !s->is_void(c);
    if (chosen == 0) return PM_OK;
    const double each = watts / (double)chosen;
    for (int64_t c = 0; c < s->n(); c++) {
        if (!mask[c] || s->is_void(c)) continue;
        if (each == 0.0)
            s->source.erase(c);
        else
            s->source[c] = each;
    }
    s->tables_valid = false;
    return PM_OK;
}

int pm_solid3d_empty(pm_solid3d *s, const uint8_t *mask) {
    if (!s || !mask) {
        set_error("pm_solid3d_empty: null argument");
        return PM_ERR_ARGUMENT;
    }
    const int64_t n = s->n();
    if (s->voids.empty()) s->voids.assign((size_t)n, 0);
    int64_t count = 0;
    for (int64_t c = 0; c < n; c++) {
        if (mask[c]) s->voids[(size_t)c] = 1;
        count += s->voids[(size_t)c];
    }
    if (count == 0) {
        s->voids.clear();
        return PM_OK;
    }
    s->nvoid = count;
    if (!s->dvoid) PM_CHECK(s->dvoid.alloc((size_t)n));
    PM_HIP(hipMemcpyAsync(s->dvoid.d(), s->voids.data(), (size_t)n, hipMemcpyHostToDevice, s->stream));
    PM_CHECK(remix(s));  // resolve mixes a two-phase block from the fractions it carries now
    PM_CHECK(apply_voids(s));
    PM_CHECK(resolve(s));
    PM_CHECK(find_gaps(s));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

size_t pm_solid3d_void_cells(const pm_solid3d *s) { return s ? (size_t)s->nvoid : 0; }

int pm_solid3d_is_void(const pm_solid3d *s, size_t i, size_t j, size_t k) {
    if (!s || (int64_t)i >= s->nx || (int64_t)j >= s->ny || (int64_t)k >= s->nz) return 0;
    return s->is_void((int64_t)i + s->nx * ((int64_t)j + s->ny * (int64_t)k)) ? 1 : 0;
}

size_t pm_solid3d_gap_pairs(const pm_solid3d *s, int64_t *a, int64_t *b, double *coef, size_t capacity) {
    if (!s) return 0;
    for (size_t g = 0; g < s->gaps.size() && g < capacity; g++) {
        if (a) a[g] = s->gaps[g].a;
        if (b) b[g] = s->gaps[g].b;
        if (coef) coef[g] = s->gaps[g].coef;
    }
    return s->gaps.size();
}

int pm_solid3d_sample(pm_solid3d *s, const double *xyz, size_t count, double *out) {
    if (!s || (count && (!xyz || !out))) return PM_ERR_ARGUMENT;
    if (count == 0) return PM_OK;
    DevBuf dp, dv;
    PM_CHECK(dp.alloc(sizeof(double) * 3 * count));
    PM_CHECK(dv.alloc(sizeof(double) * count));
    PM_HIP(hipMemcpyAsync(dp.d(), xyz, sizeof(double) * 3 * count, hipMemcpyHostToDevice, s->stream));
    heat_sample<<<grid_for((int64_t)count, 256), 256, 0, s->stream>>>(
        s->T(), s->nvoid ? s->dvoid.as<uint8_t>() : nullptr, s->nx, s->ny, s->nz, s->dx, dp.d(),
        (int64_t)count, dv.d());
    PM_HIP(hipGetLastError());
    PM_HIP(hipMemcpyAsync(out, dv.d(), sizeof(double) * count, hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
}

int pm_solid3d_generated_power(const pm_solid3d *s, double *watts) {
    if (!s || !watts) return PM_ERR_ARGUMENT;
    // `self.source.iter().sum()` over every cell in index order; zeros add nothing.
    double total = 0.0;
    for (const auto &e : s->source) total += e.second;
    *watts = total;
    return PM_OK;
}

int pm_solid3d_checkpoint(pm_solid3d *s) {
    if (!s) return PM_ERR_ARGUMENT;
    if (!s->saved) PM_CHECK(s->saved.alloc(sizeof(double) * (size_t)s->n()));
    PM_HIP(hipMemcpyAsync(s->saved.d(), s->T(), sizeof(double) * (size_t)s->n(), hipMemcpyDeviceToDevice, s->stream));
    if (s->phase)
        PM_HIP(hipMemcpyAsync(s->saved_melted.d(), s->melted.d(), sizeof(double) * (size_t)s->n(), hipMemcpyDeviceToDevice, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    s->saved_valid = true;
    s->saved_lost = s->lost;
    s->saved_supplied = s->supplied;
    return PM_OK;
}

int pm_solid3d_restore(pm_solid3d *s) {
    if (!s) return PM_ERR_ARGUMENT;
    if (!s->saved_valid) {
        // `saved` starts as the constructor's cells (solid.rs:401): every cell at the reference.
        fill_value<<<grid_for(s->n(), 256), 256, 0, s->stream>>>(s->T(), s->n(), s->reference);
        PM_HIP(hipGetLastError());
    } else {
        PM_HIP(hipMemcpyAsync(s->T(), s->saved.d(), sizeof(double) * (size_t)s->n(), hipMemcpyDeviceToDevice, s->stream));
    }
    // The fractions go back too; the operator is not rebuilt (solid.rs:1945-1950).
    if (s->phase)
        PM_HIP(hipMemcpyAsync(s->melted.d(), s->saved_melted.d(), sizeof(double) * (size_t)s->n(), hipMemcpyDeviceToDevice, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    s->lost = s->saved_lost;
    s->supplied = s->saved_supplied;
    return PM_OK;
}

int pm_solid3d_statistics(pm_solid3d *s, double out[3]) {
    if (!s || !out) return PM_ERR_ARGUMENT;
    double r[4];
    PM_CHECK(pm_solid3d_readings_kelvin(s, r));
    out[0] = r[0], out[1] = r[1], out[2] = r[2];
    return PM_OK;
}

int pm_solid3d_readings(pm_solid3d *s, double out[4]) {
    if (!s || !out) return PM_ERR_ARGUMENT;
    PM_CHECK(pm_solid3d_readings_kelvin(s, out));
    out[0] -= 273.15, out[1] -= 273.15, out[2] -= 273.15;
    return PM_OK;
}

static int pm_solid3d_readings_kelvin(pm_solid3d *s, double out[4]) {
    if (!s || !out) return PM_ERR_ARGUMENT;
    std::vector<double> h((size_t)s->n());
    PM_CHECK(pm_solid3d_get_cells(s, h.data()));
    // peak_temperature / mean_temperature / coldest_temperature, solid.rs:1378-1405, in celsius.
    // Over what is there: void cells have no vote (solid.rs:1377-1405).
    double peak = -1.7976931348623157e308, cold = 1.7976931348623157e308, sum = 0.0;
    int64_t count = 0;
    for (int64_t c = 0; c < s->n(); c++) {
        if (s->is_void(c)) continue;
        const double t = h[(size_t)c];
        peak = rmax(peak, t);
        cold = std::fmin(cold, t);
        sum += t;
        count++;
    }
    out[0] = peak;
    out[1] = count == 0 ? NAN : sum / (double)count;
    out[2] = cold;
    out[3] = s->absorbed;
    return PM_OK;
}

double pm_solid3d_heat_capacity(pm_solid3d *s) {
    if (!s) return NAN;
    double total = NAN;
    if (total_capacity(s, &total) != PM_OK) return NAN;
    return total;
}

int pm_solid3d_set_tile(pm_solid3d *s, int tile_y, int z_chunk, int prefetch) {
    if (!s) return PM_ERR_ARGUMENT;
    if (tile_y && tile_y != 4 && tile_y != 8 && tile_y != 16) {
        set_error("tile_y must be 4, 8 or 16");
        return PM_ERR_ARGUMENT;
    }
    if (z_chunk < 0 || prefetch < 0 || prefetch > 3) {
        set_error("z_chunk must be positive and prefetch 1, 2 or 3");
        return PM_ERR_ARGUMENT;
    }
    if (tile_y) s->tile_y = tile_y;
    if (z_chunk) s->z_chunk = z_chunk;
    if (prefetch) s->prefetch = prefetch;
    s->partial_count = 0;
    return PM_OK;
}

int pm_solid3d_set_steps_per_pass(pm_solid3d *s, int steps) {
    if (!s) return PM_ERR_ARGUMENT;
    if (steps != 1 && steps != 2) {
        set_error("steps per pass must be 1 or 2");
        return PM_ERR_ARGUMENT;
    }
    s->steps_per_pass = steps;
    return PM_OK;
}

int pm_solid3d_steps_per_pass(const pm_solid3d *s) {
    if (!s) return -1;
    return pairs_enabled(s) ? 2 : 1;
}

int pm_solid3d_set_profiling(pm_solid3d *s, int on) {
    if (!s) return PM_ERR_ARGUMENT;
    s->profiling = on != 0;
    return PM_OK;
}

int pm_solid3d_kernel_times(pm_solid3d *s, double *sweep_ms, double *audit_ms,
                            int *steps_per_launch) {
    if (!s) return PM_ERR_ARGUMENT;
    if (sweep_ms) *sweep_ms = s->sweep_ms;
    if (audit_ms) *audit_ms = s->audit_ms;
    if (steps_per_launch) *steps_per_launch = s->prof_launch_steps;
    return PM_OK;
}

}  // extern "C"

namespace pm {

// Device view of a block for the surrogate: the current temperatures, the per-cell conductivity
// (null when uniform), the extents, the block's stream, whether the capacity is uniform and, if
// it is, the mobility dx / C.
int solid_device_state(pm_solid3d *s, double **cells, const double **kcell, int64_t counts[3],
                       hipStream_t *stream, bool *uniform_capacity, double *mobility) {
    if (!s) return PM_ERR_ARGUMENT;
    *cells = s->T();
    *kcell = s->kmode == K_PER_CELL ? s->kcell.d() : nullptr;
    counts[0] = s->nx, counts[1] = s->ny, counts[2] = s->nz;
    *stream = s->stream;
    if (uniform_capacity) *uniform_capacity = s->cmode == C_UNIFORM;
    if (mobility) *mobility = s->mob0;
    return PM_OK;
}

}  // namespace pm
