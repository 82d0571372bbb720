// Hall on gfx950: the staggered pressure/velocity leapfrog of
// crates/pantometry-acoustic/src/hall.rs, with the energy invariant (hall.rs:334-384) fused into
// the step.
//
// Layout in HBM: node pressure [nx ny nz], face velocities vx [(nx-1) ny nz], vy [nx (ny-1) nz],
// vz [nx ny (nz-1)], all x-fastest f64.  Each of the four arrays ping-pongs between two buffers,
// so `pressure_prev` is simply the other pressure buffer (hall.rs:463 becomes free) and a step
// reads p, vx, vy, vz once and writes them once: 64 B per node.
//
// One fused kernel per step.  The thread of node n computes the new velocities on the three
// faces above n (the ones n owns) and, redundantly, on the three faces below it (owned by the
// neighbours), from the old pressure, exactly as hall.rs:433-457 does; then the new pressure from
// those six new velocities (hall.rs:465-505).  A face's new velocity is the same expression
// whichever thread evaluates it, so the redundancy is bit-identical.  The thread then adds its
// node's share of the invariant and the shares of the faces it owns.  The kernel marches in z
// over column tiles, so each array is read from HBM once per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace pm {

struct HallArgs {
    const double *__restrict__ p;
    const double *__restrict__ vx;
    const double *__restrict__ vy;
    const double *__restrict__ vz;
    double *__restrict__ p_out;
    double *__restrict__ vx_out;
    double *__restrict__ vy_out;
    double *__restrict__ vz_out;
    double *__restrict__ partials;
    double *__restrict__ supplied_step;  // the source's supplied energy this step
    const AuditState *audit;
    double gain, scale, two_rc2, volume, half_rho;
    double pot, kin;  // volume / (2 rho c^2) and half_rho x volume (the paired step's audit)
    int64_t src;         // source node or -1
    double source;       // its increment this step
    double source_prev;  // its increment the step before
    int nx, ny, nz;
    // Two steps per pass: the pressure after the first step (the second step's pressure_prev),
    // the source of the second step (node or -1, increment) and its supplied energy.
    double *__restrict__ p_mid;
    int64_t src2;
    double source2;
    double *__restrict__ supplied_step2;
    int force;  // run even if an audit breach froze the run
    int64_t strip_tiles;  // the paired step's last-column strips (hall_strip2), 0 without
};

__device__ __forceinline__ double wall(int64_t k, int64_t n) { return (k == 0 || k + 1 == n) ? 2.0 : 1.0; }

// Workgroups of one launch go round-robin over the eight XCDs; the marching kernel gives each XCD
// a contiguous run of tiles so a tile's neighbours (whose rows it reads as halo) share its L2.
constexpr int XCDS_HALL = 8;

// The z-marching form of the same step.  A workgroup owns a 64 x TY column tile of nodes and a
// chunk of z-planes; each thread keeps its own column's p(k-1..k+2) and vz(k-1..k+1) in
// registers, so every array is read from HBM once per step.  Per plane the workgroup stages p(k)
// with a one-node halo, vx(k) with the face left of the tile and vy(k) with the face below it in
// LDS (double-buffered, one barrier per plane) for the in-plane neighbours.  The arithmetic
// follows hall.rs:433-505 expression for expression.
template <int TY>
__global__ __launch_bounds__(64 * TY) void hall_march(HallArgs a, int gx, int gy, int gz, int zc) {
    constexpr int TX = 64, NT = TX * TY;
    __shared__ double sP[2][TY + 2][TX + 2];
    __shared__ double sVX[2][TY][TX + 1];  // faces x0-1 .. x0+63 of each row
    __shared__ double sVY[2][TY + 1][TX];  // faces y0-1 .. y0+TY-1 of each column
    __shared__ double red[NT / 64];

    const int64_t ntiles = (int64_t)gx * gy * gz;
    const int64_t chunk = (ntiles + XCDS_HALL - 1) / XCDS_HALL;
    const int64_t tile = (int64_t)(blockIdx.x % XCDS_HALL) * chunk + blockIdx.x / XCDS_HALL;
    if (tile >= ntiles) return;
    const int64_t bx = tile % gx, by = (tile / gx) % gy, bz = tile / ((int64_t)gx * gy);

    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz, plane = nx * ny;
    const int64_t x0 = bx * TX, y0 = by * TY;
    const int64_t z0 = bz * zc, z1 = std::min<int64_t>(z0 + zc, nz);
    const int64_t i = x0 + tx, j = y0 + ty;
    const bool inside = i < nx && j < ny;
    const int64_t ci = std::min(i, nx - 1), cj = std::min(j, ny - 1);
    const int64_t col = ci + nx * cj;
    const bool has_vx = nx > 1, has_vy = ny > 1, has_vz = nz > 1;
    // Face offsets within a plane of each velocity array (clamped to a face that exists).
    const int64_t vxp = (nx - 1) * ny, vyp = nx * (ny - 1);
    const int64_t vxcol = std::min(ci, nx - 2) + (nx - 1) * cj;
    const int64_t vycol = ci + nx * std::min(cj, ny - 2);
    auto kc = [&](int64_t q) { return std::min(std::max<int64_t>(q, 0), nz - 1); };
    auto kz = [&](int64_t q) { return std::min(std::max<int64_t>(q, 0), nz - 2); };

    // A breach froze the run: later launches leave every buffer alone, and the host takes the
    // state from the launch that broke.
    if (a.audit->halted && !a.force) return;

    // Halo roles: p one node around the tile (bottom row, top row, left and right columns), the
    // vx face left of each row (first TY threads), the vy face below each column (first row).
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
    const bool vxh = tid < TY && x0 > 0 && has_vx;  // face x0-1 of row y0+tid
    const int64_t vxhidx = (x0 - 1 >= 0 ? x0 - 1 : 0) + (nx - 1) * std::min(y0 + tid, ny - 1);
    const bool vyh = ty == 0 && y0 > 0 && has_vy;   // face y0-1 of column x0+tx
    const int64_t vyhidx = ci + nx * (y0 - 1 >= 0 ? y0 - 1 : 0);

    const double *__restrict__ P = a.p;
    auto ldp = [&](int64_t idx, int64_t k) { return P[idx + plane * kc(k)]; };
    auto ldvx = [&](bool ok, int64_t idx, int64_t k) { return ok ? a.vx[idx + vxp * kc(k)] : 0.0; };
    auto ldvy = [&](bool ok, int64_t idx, int64_t k) { return ok ? a.vy[idx + vyp * kc(k)] : 0.0; };
    auto ldvz = [&](int64_t k) { return has_vz ? a.vz[col + plane * kz(k)] : 0.0; };

    // Own column: p(k-1), p(k), p(k+1), p(k+2); vz(k-1), vz(k), vz(k+1); vx(k), vx(k+1);
    // vy(k), vy(k+1).  Halos: the values of planes k and k+1.
    double pm_ = ldp(col, z0 - 1), pc_ = ldp(col, z0), pp_ = ldp(col, z0 + 1), pq_ = ldp(col, z0 + 2);
    double zm_ = ldvz(z0 - 1), zc_ = ldvz(z0), zp_ = ldvz(z0 + 1);
    double xc_ = ldvx(has_vx, vxcol, z0), xp_ = ldvx(has_vx, vxcol, z0 + 1);
    double yc_ = ldvy(has_vy, vycol, z0), yp_ = ldvy(has_vy, vycol, z0 + 1);
    double hp0 = halo ? ldp(hidx, z0) : 0.0, hp1 = halo ? ldp(hidx, z0 + 1) : 0.0;
    double hx0 = ldvx(vxh, vxhidx, z0), hx1 = ldvx(vxh, vxhidx, z0 + 1);
    double hy0 = ldvy(vyh, vyhidx, z0), hy1 = ldvy(vyh, vyhidx, z0 + 1);

    const double g = a.gain;
    const double wi = wall(i, nx), wj = wall(j, ny);
    double acc = 0.0;
    for (int64_t k = z0; k < z1; k++) {
        const int b = (int)((k - z0) & 1);
        sP[b][ty + 1][tx + 1] = pc_;
        sVX[b][ty][tx + 1] = xc_;
        sVY[b][ty + 1][tx] = yc_;
        if (halo) sP[b][hr][hc] = hp0;
        if (tid < TY) sVX[b][tid][0] = hx0;
        if (ty == 0) sVY[b][0][tx] = hy0;
        lds_barrier();

        const double pc = pc_;
        double hi_x = 0.0, lo_x = 0.0, hi_y = 0.0, lo_y = 0.0, hi_z = 0.0, lo_z = 0.0;
        if (i + 1 < nx) hi_x = sVX[b][ty][tx + 1] - g * (sP[b][ty + 1][tx + 2] - pc);
        if (i > 0) lo_x = sVX[b][ty][tx] - g * (pc - sP[b][ty + 1][tx]);
        if (j + 1 < ny) hi_y = sVY[b][ty + 1][tx] - g * (sP[b][ty + 2][tx + 1] - pc);
        if (j > 0) lo_y = sVY[b][ty][tx] - g * (pc - sP[b][ty][tx + 1]);
        if (k + 1 < nz) hi_z = zc_ - g * (pp_ - pc);
        if (k > 0) lo_z = zm_ - g * (pc - pm_);
        const double wk = wall(k, nz);
        double pn = pc - a.scale * (wi * (hi_x - lo_x) + wj * (hi_y - lo_y) + wk * (hi_z - lo_z));
        const int64_t n = col + plane * k;
        if (inside) {
            if (n == a.src) {
                pn = pn + a.source;
                const double share = 1.0 / (wi * wj * wk);
                *a.supplied_step = share * (a.source + a.source_prev) * pc / a.two_rc2 * a.volume;
            }
            a.p_out[n] = pn;
            if (i + 1 < nx) a.vx_out[vxcol + vxp * k] = hi_x;
            if (j + 1 < ny) a.vy_out[vycol + vyp * k] = hi_y;
            if (k + 1 < nz) a.vz_out[n] = hi_z;
            const double share = 1.0 / (wi * wj * wk);
            acc += share * pn * pc / a.two_rc2 * a.volume;
            if (i + 1 < nx) acc += 1.0 / (wj * wk) * a.half_rho * hi_x * hi_x * a.volume;
            if (j + 1 < ny) acc += 1.0 / (wi * wk) * a.half_rho * hi_y * hi_y * a.volume;
            if (k + 1 < nz) acc += 1.0 / (wi * wj) * a.half_rho * hi_z * hi_z * a.volume;
        }

        // Advance the rings and fetch the planes two ahead.
        pm_ = pc_, pc_ = pp_, pp_ = pq_, pq_ = ldp(col, k + 3);
        zm_ = zc_, zc_ = zp_, zp_ = ldvz(k + 2);
        xc_ = xp_, xp_ = ldvx(has_vx, vxcol, k + 2);
        yc_ = yp_, yp_ = ldvy(has_vy, vycol, k + 2);
        hp0 = hp1, hp1 = halo ? ldp(hidx, k + 2) : 0.0;
        hx0 = hx1, hx1 = ldvx(vxh, vxhidx, k + 2);
        hy0 = hy1, hy1 = ldvy(vyh, vyhidx, k + 2);
    }
    const double total = block_sum<NT>(acc, red);
    if (tid == 0) a.partials[tile] = total;
}

// The LDS a TX x TY tile of the paired step uses, in doubles; the same for its transpose.
template <int TX, int TY>
constexpr int pair_lds_doubles() {
    return 2 * ((TY + 4) * (TX + 4) + (TY + 2) * (TX + 3) + (TY + 3) * (TX + 2) + 3 * (TY + 2) * (TX + 2)) +
           TX * TY / 64;
}

// Two steps per pass (temporal blocking) of the marching step, on TX x TY column tiles (TX 64,
// 56 or 48; 64 unless set).  Per plane k the workgroup
// computes the first step on the tile widened by one node (E1, (TX+2) x (TY+2), its four corners
// aside, which no second-step stencil reaches): each node's six new face velocities v' and its
// new pressure p' (hall.rs:433-505), from p, vx, vy staged on E2 (the tile widened by two) and its
// own column's p and vz in registers.  The tile's threads keep their nodes' v' in registers and
// p' in LDS; in the same iteration they take the second step on plane k-1 from those.  A face's
// v' is computed by both nodes it joins and a ring node's p' by two workgroups: the same
// expression on the same inputs each time, so every copy agrees bit for bit.  One pass reads p,
// vx, vy, vz once and writes p'' (the new pressure), p' (its pressure_prev) and the v'' faces:
// 72 B per two node-steps instead of 128.  The first step's audit sums the owned nodes' p'p and
// the owned faces' v'^2 as they are produced, the second step's those of p''p' and v''^2.
//
// This is one tile of it: nodes x0 .. x0+TX-1, y0 .. y0+TY-1 of planes z0 .. z0+zc-1, all of
// which it stores and audits, or (LAST) only those of the hall's last column.  Its audit sums go
// to partials[tile] and partials[ntiles + tile].  `lds` holds pair_lds_doubles<TX, TY>().
template <int TX, int TY, bool LAST>
__device__ __forceinline__ void pair_tile(const HallArgs &a, double *lds, int tx, int ty, int64_t tile,
                                          int64_t ntiles, int x0, int y0, int z0, int zc) {
    constexpr int NT = TX * TY;
    static_assert(NT % 64 == 0, "whole waves");
    constexpr int E1X = TX + 2, E1Y = TY + 2, E2X = TX + 4, E2Y = TY + 4;
    constexpr int R1 = 2 * (E1X - 2) + 2 * TY;             // E1 ring nodes, corners aside
    constexpr int R2 = 2 * (E2X - 4) + 2 * (E2Y - 4) + 4;  // E2 ring nodes reached, E1 corners
    constexpr int RING1_FIRST = NT - R1;
    static_assert(R1 <= NT && R2 <= RING1_FIRST, "roles do not fit");
    // p(k) on E2; vx(k) of E1 rows (face c between E2 columns c, c+1); vy(k) of E1 columns (face
    // r between E2 rows r, r+1); p'(k) on E1 with the v'x(k), v'y(k) of each E1 node's +x, +y
    // faces; and one double per wave for the audit's sum.  Each is double-buffered by plane.
    double(*sP)[E2Y][E2X] = reinterpret_cast<double(*)[E2Y][E2X]>(lds);
    double(*sVX)[E1Y][E2X - 1] = reinterpret_cast<double(*)[E1Y][E2X - 1]>(sP + 2);
    double(*sVY)[E2Y - 1][E1X] = reinterpret_cast<double(*)[E2Y - 1][E1X]>(sVX + 2);
    double(*sQ)[E1Y][E1X] = reinterpret_cast<double(*)[E1Y][E1X]>(sVY + 2);
    double(*sHX)[E1Y][E1X] = sQ + 2;
    double(*sHY)[E1Y][E1X] = sHX + 2;
    double *red = &sHY[2][0][0];

    const int tid = tx + TX * ty;
    const int nx = a.nx, ny = a.ny;
    const int64_t plane = (int64_t)nx * ny;
    // Plane indices are uniform 32-bit values, so their clamps and tests stay on the scalar unit.
    // The paired step is launched only for halls at least two nodes along every axis, so every
    // velocity array has faces.
    const int nz = a.nz;
    const int z1 = std::min(z0 + zc, nz);
    const int64_t vxp = (int64_t)(nx - 1) * ny, vyp = (int64_t)nx * (ny - 1);
    auto cx = [&](int x) { return std::min(std::max(x, 0), nx - 1); };
    auto cy = [&](int y) { return std::min(std::max(y, 0), ny - 1); };
    auto kc = [&](int q) { return std::min(std::max(q, 0), nz - 1); };
    auto kz = [&](int q) { return std::min(std::max(q, 0), nz - 2); };
    // A node's byte offsets within a plane of p (and vz), vx (its +x face) and vy (its +y face),
    // clamped to ones that exist (they fit 32 bits: planes of the paired step hold at most
    // 107,374,182 nodes), and its in-plane index, or ~0 off the hall.
    struct Col {
        int x, y;
        uint32_t p, vx, vy, n;
    };
    auto column = [&](int x, int y) {
        Col c;
        c.x = x, c.y = y;
        c.p = (uint32_t)(cx(x) + nx * cy(y)) * 8u;
        c.vx = (uint32_t)(std::min(cx(x), nx - 2) + (nx - 1) * cy(y)) * 8u;
        c.vy = (uint32_t)(cx(x) + nx * std::min(cy(y), ny - 2)) * 8u;
        c.n = (x >= 0 && x < nx && y >= 0 && y < ny) ? (uint32_t)(x + nx * y) : ~0u;
        return c;
    };
    // The sources of the two steps as (in-plane index, plane); none matches no node.
    const uint32_t sn1 = a.src >= 0 ? (uint32_t)(a.src % plane) : ~0u - 1;
    const int sk1 = a.src >= 0 ? (int)(a.src / plane) : INT32_MIN;
    const uint32_t sn2 = a.src2 >= 0 ? (uint32_t)(a.src2 % plane) : ~0u - 1;
    const int sk2 = a.src2 >= 0 ? (int)(a.src2 / plane) : INT32_MIN;

    // Roles.  Every thread: its tile node, E1 (ty+1, tx+1).  The last R1 threads: an E1 ring node
    // (r1r, r1c).  The first R2 threads: an E2 node whose p (and, on the left column and bottom
    // row and at E1's corners, whose +x or +y face) the ring stencils read.
    const Col tc = column(x0 + tx, y0 + ty);
    const bool inside = (LAST ? tc.x == nx - 1 : tc.x < nx) && tc.y < ny;
    int r1r = -1, r1c = -1;
    {
        const int q = tid - RING1_FIRST;
        if (q >= 0 && q < E1X - 2) r1r = 0, r1c = 1 + q;
        else if (q >= E1X - 2 && q < 2 * (E1X - 2)) r1r = E1Y - 1, r1c = 1 + q - (E1X - 2);
        else if (q >= 2 * (E1X - 2) && q < 2 * (E1X - 2) + TY) r1r = 1 + q - 2 * (E1X - 2), r1c = 0;
        else if (q >= 2 * (E1X - 2) + TY && q < R1) r1r = 1 + q - 2 * (E1X - 2) - TY, r1c = E1X - 1;
    }
    const bool ring1 = r1r >= 0;
    // A lane outside a role loads its own tile column in that role's place: every lane issues
    // every load, so the count of loads in flight is the same on every path and the waits before
    // their uses stay exact.
    const Col rc = ring1 ? column(x0 - 1 + r1c, y0 - 1 + r1r) : tc;
    int r2r = -1, r2c = -1;
    {
        constexpr int SIDES = 2 * (E2X - 4) + 2 * (E2Y - 4);
        if (tid < E2X - 4) r2r = 0, r2c = 2 + tid;
        else if (tid < 2 * (E2X - 4)) r2r = E2Y - 1, r2c = 2 + tid - (E2X - 4);
        else if (tid < 2 * (E2X - 4) + (E2Y - 4)) r2r = 2 + tid - 2 * (E2X - 4), r2c = 0;
        else if (tid < SIDES) r2r = 2 + tid - 2 * (E2X - 4) - (E2Y - 4), r2c = E2X - 1;
        else if (tid < R2) {
            const int q = tid - SIDES;
            r2r = (q & 2) ? E2Y - 2 : 1;
            r2c = (q & 1) ? E2X - 2 : 1;
        }
    }
    const bool ring2 = r2r >= 0;
    const bool corner = ring2 && (r2r == 1 || r2r == E2Y - 2) && (r2c == 1 || r2c == E2X - 2);
    const Col hc = ring2 ? column(x0 - 2 + r2c, y0 - 2 + r2r) : tc;
    // which faces a ring-2 thread stages: the +x face on E2's left column and at E1's corners,
    // the +y face on E2's bottom row and at E1's corners (0 where those faces do not exist)
    const bool s_vx = ring2 && (r2c == 0 || corner), s_vy = ring2 && (r2r == 0 || corner);
    const bool h_vx = s_vx && hc.x >= 0 && hc.x + 1 < nx;
    const bool h_vy = s_vy && hc.y >= 0 && hc.y + 1 < ny;

    // Each array is addressed from the chunk's lowest plane through a resource built once; a
    // plane is a uniform byte offset from there.  Planes zb .. z1 + 2 are read and z0 .. z1 - 1
    // written, and the host caps the chunk so that (zc + 4) planes span fewer than 2^32 bytes.
    const int zb = std::max(z0 - 2, 0), zbz = std::min(zb, nz - 2);
    const __amdgpu_buffer_rsrc_t RP = buffer_rsrc(a.p + plane * zb), RX = buffer_rsrc(a.vx + vxp * zb);
    const __amdgpu_buffer_rsrc_t RY = buffer_rsrc(a.vy + vyp * zb), RZ = buffer_rsrc(a.vz + plane * zbz);
    const uint32_t pb = (uint32_t)plane * 8u, xb = (uint32_t)vxp * 8u, yb = (uint32_t)vyp * 8u;
    auto oP = [&](int q) { return (uint32_t)(kc(q) - zb) * pb; };
    auto oX = [&](int q) { return (uint32_t)(kc(q) - zb) * xb; };
    auto oY = [&](int q) { return (uint32_t)(kc(q) - zb) * yb; };
    auto oZ = [&](int q) { return (uint32_t)(kz(q) - zbz) * pb; };
    const __amdgpu_buffer_rsrc_t WP = buffer_rsrc(a.p_out + plane * z0), WM = buffer_rsrc(a.p_mid + plane * z0);
    const __amdgpu_buffer_rsrc_t WX = buffer_rsrc(a.vx_out + vxp * z0), WY = buffer_rsrc(a.vy_out + vyp * z0);
    const __amdgpu_buffer_rsrc_t WZ = buffer_rsrc(a.vz_out + plane * std::min(z0, nz - 2));

    const double g = a.gain, sc = a.scale;
    // Wall weights are 1 or 2, so 1 / (a product of them) is exactly a product of 1 or 0.5.
    auto inv_wall = [](int64_t k, int64_t n) { return (k == 0 || k + 1 == n) ? 0.5 : 1.0; };
    // The first step at E1 node (r, c) of plane k: new face velocities and pressure.
    struct Step {
        double hx, lx, hy, ly, hz, lz, p;
    };
    auto first = [&](int b, int r, int c, const Col &cc, int k, double pm, double pcur,
                     double pp, double vzm, double vzc) {
        Step o;
        const int x = cc.x, y = cc.y;
        o.hx = 0.0, o.lx = 0.0, o.hy = 0.0, o.ly = 0.0, o.hz = 0.0, o.lz = 0.0;
        if (x + 1 < nx) o.hx = sVX[b][r][c + 1] - g * (sP[b][r + 1][c + 2] - pcur);
        if (x > 0) o.lx = sVX[b][r][c] - g * (pcur - sP[b][r + 1][c]);
        if (y + 1 < ny) o.hy = sVY[b][r + 1][c] - g * (sP[b][r + 2][c + 1] - pcur);
        if (y > 0) o.ly = sVY[b][r][c] - g * (pcur - sP[b][r][c + 1]);
        if (k + 1 < nz) o.hz = vzc - g * (pp - pcur);
        if (k > 0) o.lz = vzm - g * (pcur - pm);
        const double wi = wall(x, nx), wj = wall(y, ny), wk = wall(k, nz);
        double pn = pcur - sc * (wi * (o.hx - o.lx) + wj * (o.hy - o.ly) + wk * (o.hz - o.lz));
        if (cc.n == sn1 && k == sk1) pn = pn + a.source;
        o.p = pn;
        return o;
    };

    const int zs = z0 - 1;
    // Own columns: p(k-1..k+2), vz(k-1..k+1), vx(k), vy(k); the tile node also p'(k-2..k) and
    // the previous plane's v'.  Ring 2: p, vx, vy of plane k.  Loads for plane k+1 (k+2 for
    // p and vz ahead of use) are issued right after plane k is staged.
    double tpm = buffer_load_f64(RP, tc.p, oP(zs - 1)), tpc = buffer_load_f64(RP, tc.p, oP(zs));
    double tpp = buffer_load_f64(RP, tc.p, oP(zs + 1)), tpq = 0.0;
    double tzm = buffer_load_f64(RZ, tc.p, oZ(zs - 1)), tzc = buffer_load_f64(RZ, tc.p, oZ(zs)), tzp = 0.0;
    double txc = buffer_load_f64(RX, tc.vx, oX(zs)), tyc = buffer_load_f64(RY, tc.vy, oY(zs));
    double rpm = buffer_load_f64(RP, rc.p, oP(zs - 1)), rpc = buffer_load_f64(RP, rc.p, oP(zs));
    double rpp = buffer_load_f64(RP, rc.p, oP(zs + 1)), rpq = 0.0;
    double rzm = buffer_load_f64(RZ, rc.p, oZ(zs - 1)), rzc = buffer_load_f64(RZ, rc.p, oZ(zs)), rzp = 0.0;
    double rxc = buffer_load_f64(RX, rc.vx, oX(zs)), ryc = buffer_load_f64(RY, rc.vy, oY(zs));
    double hp0 = buffer_load_f64(RP, hc.p, oP(zs)), hx0 = buffer_load_f64(RX, hc.vx, oX(zs));
    double hy0 = buffer_load_f64(RY, hc.vy, oY(zs));
    double qm = 0.0, qc = 0.0;  // the tile node's p'(k-2), p'(k-1)
    double plz = 0.0;           // its v'z below plane k-1 (the one above is this plane's lz)

    const double iwi = inv_wall(tc.x, nx), iwj = inv_wall(tc.y, ny);
    const double wi = wall(tc.x, nx), wj = wall(tc.y, ny);
    // The invariant's terms (hall.rs:334-384) with its constant factors folded: volume / (2 rho
    // c^2) and rho / 2 x volume, each rounded once on the host, times the node's wall shares
    // (exact powers of two): p'p and the +x, +y faces' v'^2 weighted by this lane's constants and
    // then by the plane's share, the +z face's by its own, summed with fused multiply-adds.  A
    // term moves by an ulp or two from the reference's; the totals are held to 1e-12 of the
    // ledger's scale, not bit for bit.  A face the node does not own has v' = 0 (as has a
    // node's +z face on the top plane), and its weight is 0 as well.
    const double w_p = iwi * iwj * a.pot;
    const double w_x = tc.x + 1 < nx ? iwj * a.kin : 0.0, w_y = tc.y + 1 < ny ? iwi * a.kin : 0.0;
    const double w_z = iwi * iwj * a.kin;
    double acc1 = 0.0, acc2 = 0.0;
    for (int k = zs; k <= z1; k++) {
        const int b = (k - zs) & 1;
        // Stage p(k), vx(k), vy(k) on E2, then fetch the next planes into the freed registers.
        sP[b][ty + 2][tx + 2] = tpc;
        sVX[b][ty + 1][tx + 2] = txc;
        sVY[b][ty + 2][tx + 1] = tyc;
        if (ring1) {
            sP[b][r1r + 1][r1c + 1] = rpc;
            sVX[b][r1r][r1c + 1] = rxc;
            sVY[b][r1r + 1][r1c] = ryc;
        }
        if (ring2) {
            sP[b][r2r][r2c] = hp0;
            if (s_vx) sVX[b][r2r - 1][r2c] = h_vx ? hx0 : 0.0;
            if (s_vy) sVY[b][r2r][r2c - 1] = h_vy ? hy0 : 0.0;
        }
        {
            const uint32_t P2 = oP(k + 2), P1 = oP(k + 1), X1 = oX(k + 1), Y1 = oY(k + 1), Z1 = oZ(k + 1);
            tpq = buffer_load_f64(RP, tc.p, P2);
            tzp = buffer_load_f64(RZ, tc.p, Z1);
            txc = buffer_load_f64(RX, tc.vx, X1);
            tyc = buffer_load_f64(RY, tc.vy, Y1);
            rpq = buffer_load_f64(RP, rc.p, P2);
            rzp = buffer_load_f64(RZ, rc.p, Z1);
            rxc = buffer_load_f64(RX, rc.vx, X1);
            ryc = buffer_load_f64(RY, rc.vy, Y1);
            hp0 = buffer_load_f64(RP, hc.p, P1);
            hx0 = buffer_load_f64(RX, hc.vx, X1);
            hy0 = buffer_load_f64(RY, hc.vy, Y1);
        }
        lds_barrier();

        // First step on plane k for the tile node and the ring node (every lane computes it, a
        // lane outside the ring on its own node again), and the second step on plane k-1 from p'
        // of planes k-2, k-1 (registers, LDS) and k and the v' faces of plane k-1 (in-plane ones
        // staged by their nodes, z ones in registers): three independent chains in one block,
        // which the compiler interleaves.  Then the writes, stores and audits.
        const Step t1 = first(b, ty + 1, tx + 1, tc, k, tpm, tpc, tpp, tzm, tzc);
        const Step r1 = first(b, ring1 ? r1r : ty + 1, ring1 ? r1c : tx + 1, rc, k, rpm, rpc, rpp, rzm, rzc);
        const int k2 = k - 1;
        const double pc = qc;
        double hx = 0.0, lx = 0.0, hy = 0.0, ly = 0.0, hz = 0.0, lz = 0.0;
        {
            const int r = ty + 1, c = tx + 1;
            if (tc.x + 1 < nx) hx = sHX[b ^ 1][r][c] - g * (sQ[b ^ 1][r][c + 1] - pc);
            if (tc.x > 0) lx = sHX[b ^ 1][r][c - 1] - g * (pc - sQ[b ^ 1][r][c - 1]);
            if (tc.y + 1 < ny) hy = sHY[b ^ 1][r][c] - g * (sQ[b ^ 1][r + 1][c] - pc);
            if (tc.y > 0) ly = sHY[b ^ 1][r - 1][c] - g * (pc - sQ[b ^ 1][r - 1][c]);
            if (k2 + 1 < nz) hz = t1.lz - g * (t1.p - pc);
            if (k2 > 0) lz = plz - g * (pc - qm);
        }
        const double wk2 = wall(k2, nz);
        double pn = pc - sc * (wi * (hx - lx) + wj * (hy - ly) + wk2 * (hz - lz));
        const bool at_src = tc.n == sn2 && k2 == sk2;
        if (at_src) pn = pn + a.source2;

        sQ[b][ty + 1][tx + 1] = t1.p;
        sHX[b][ty + 1][tx + 1] = t1.hx;
        sHY[b][ty + 1][tx + 1] = t1.hy;
        if (ring1) {
            sQ[b][r1r][r1c] = r1.p;
            sHX[b][r1r][r1c] = r1.hx;
            sHY[b][r1r][r1c] = r1.hy;
        }
        if (inside && k2 >= z0 && k2 < z1) {
            const double iwk = inv_wall(k2, nz);
            if (at_src) {
                const double share = iwi * iwj * iwk;
                *a.supplied_step2 = share * (a.source2 + a.source) * pc / a.two_rc2 * a.volume;
            }
            const uint32_t o = (uint32_t)(k2 - z0);
            buffer_store_f64_nt(pn, WP, tc.p, o * pb);
            buffer_store_f64_nt(pc, WM, tc.p, o * pb);
            if (tc.x + 1 < nx) buffer_store_f64_nt(hx, WX, tc.vx, o * xb);
            if (tc.y + 1 < ny) buffer_store_f64_nt(hy, WY, tc.vy, o * yb);
            if (k2 + 1 < nz) buffer_store_f64_nt(hz, WZ, tc.p, o * pb);
            const double e = std::fma(w_y, hy * hy, std::fma(w_x, hx * hx, w_p * (pn * pc)));
            acc2 = std::fma(w_z, hz * hz, std::fma(iwk, e, acc2));
        }
        if (inside && k >= z0 && k < z1) {
            const double iwk = inv_wall(k, nz);
            if (tc.n == sn1 && k == sk1) {
                const double share = iwi * iwj * iwk;
                *a.supplied_step = share * (a.source + a.source_prev) * tpc / a.two_rc2 * a.volume;
            }
            const double e = std::fma(w_y, t1.hy * t1.hy, std::fma(w_x, t1.hx * t1.hx, w_p * (t1.p * tpc)));
            acc1 = std::fma(w_z, t1.hz * t1.hz, std::fma(iwk, e, acc1));
        }
        qm = qc, qc = t1.p;
        plz = t1.lz;

        // Advance the rings.
        tpm = tpc, tpc = tpp, tpp = tpq;
        tzm = tzc, tzc = tzp;
        rpm = rpc, rpc = rpp, rpp = rpq;
        rzm = rzc, rzc = rzp;
    }
    const double total1 = block_sum<NT>(acc1, red);
    __syncthreads();
    const double total2 = block_sum<NT>(acc2, red);
    if (tid == 0) {
        a.partials[tile] = total1;
        a.partials[ntiles + tile] = total2;
    }
}

// The paired step's launch: gx x gy x gz tiles of TX x TY columns.
template <int TX, int TY>
__global__ __launch_bounds__(TX * TY, 4) void hall_march2(HallArgs a, int gx, int gy, int gz, int zc) {
    __shared__ double lds[pair_lds_doubles<TX, TY>()];
    const int64_t ntiles = (int64_t)gx * gy * gz;
    const int64_t chunk = (ntiles + XCDS_HALL - 1) / XCDS_HALL;
    const int64_t tile = (int64_t)(blockIdx.x % XCDS_HALL) * chunk + blockIdx.x / XCDS_HALL;
    if (tile >= ntiles) return;
    if (a.audit->halted && !a.force) return;  // a breach froze the run
    const int64_t bx = tile % gx, by = (tile / gx) % gy, bz = tile / ((int64_t)gx * gy);
    pair_tile<TX, TY, false>(a, lds, threadIdx.x, threadIdx.y, tile, ntiles + a.strip_tiles,
                             (int)bx * TX, (int)by * TY, (int)bz * zc, zc);
}

// The last column of a hall gx * TX + 1 nodes across, beside hall_march2's gx x-tiles (a
// seventh x-tile at 385 = 6 x 64 + 1 nodes would march every plane for one live column): strips
// of the transposed tile, TY wide at x0 = nx - TY and TX tall, that own only that column, over
// short z-chunks so the launch is one round of short-lived workgroups.  Their audit sums follow
// the tiles' (partials from `first`).
template <int TX, int TY>
__global__ __launch_bounds__(TX * TY, 4) void hall_strip2(HallArgs a, int gs, int gz, int zc, int64_t first) {
    static_assert(pair_lds_doubles<TX, TY>() == pair_lds_doubles<TY, TX>(), "the tile's LDS");
    __shared__ double lds[pair_lds_doubles<TX, TY>()];
    const int64_t strip = blockIdx.x;
    if (strip >= (int64_t)gs * gz) return;
    if (a.audit->halted && !a.force) return;  // a breach froze the run
    const int sy = (int)(strip % gs), bz = (int)(strip / gs);
    const int tid = threadIdx.x + TX * threadIdx.y;
    pair_tile<TY, TX, true>(a, lds, tid % TY, tid / TY, first + strip, first + a.strip_tiles,
                            a.nx - TY, sy * TX, bz * zc, zc);
}

// The audits of the two steps of a pass, in order: supplied energy and ledger of the first step,
// then of the second (a breach on the first freezes the run before the second is booked).
__global__ __launch_bounds__(1024) void hall_finalize2(const double *partials, int64_t n,
                                                       AuditState *st, double offset,
                                                       double *supplied_total,
                                                       const double *supplied1,
                                                       const double *supplied2, int64_t step,
                                                       double *totals) {
    __shared__ double scratch[16];
    if (st->halted) return;
    const double inv1 = fixed_order_sum_1024(partials, n, scratch);
    __syncthreads();
    const double inv2 = fixed_order_sum_1024(partials + n, n, scratch);
    if (threadIdx.x == 0) {
        double supplied = *supplied_total;
        if (supplied1) supplied += *supplied1;
        double energy = inv1 + offset;
        double total = 0.0;
        total += energy;
        total += -supplied;
        audit_step(st, total, rmax(rmax(0.0, std::fabs(energy)), std::fabs(supplied)), step, totals);
        *supplied_total = supplied;
        if (st->halted) return;
        if (supplied2) supplied += *supplied2;
        energy = inv2 + offset;
        total = 0.0;
        total += energy;
        total += -supplied;
        audit_step(st, total, rmax(rmax(0.0, std::fabs(energy)), std::fabs(supplied)), step + 1, totals);
        *supplied_total = supplied;
    }
}

// The invariant of the current state without stepping (the release datum, ledger reads).
__global__ __launch_bounds__(256) void hall_invariant_kernel(const double *p, const double *pp,
                                                             const double *vx, const double *vy,
                                                             const double *vz, int nx, int ny,
                                                             int nz, double two_rc2,
                                                             double volume, double half_rho,
                                                             double *partials) {
    __shared__ double red[4];
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t c = blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (c < n) {
        const int64_t i = c % nx, j = (c / nx) % ny, k = c / ((int64_t)nx * ny);
        const double wi = wall(i, nx), wj = wall(j, ny), wk = wall(k, nz);
        acc += 1.0 / (wi * wj * wk) * p[c] * pp[c] / two_rc2 * volume;
        if (i + 1 < nx) {
            const double v = vx[i + (int64_t)(nx - 1) * (j + (int64_t)ny * k)];
            acc += 1.0 / (wj * wk) * half_rho * v * v * volume;
        }
        if (j + 1 < ny) {
            const double v = vy[i + (int64_t)nx * (j + (int64_t)(ny - 1) * k)];
            acc += 1.0 / (wi * wk) * half_rho * v * v * volume;
        }
        if (k + 1 < nz) {
            const double v = vz[c];
            acc += 1.0 / (wi * wj) * half_rho * v * v * volume;
        }
    }
    const double total = block_sum<256>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Room::invariant_si, room.rs:243-281: the potential share of every node, then the kinetic share
// of every vx face and every vy face; blocks [0, bp) potential, [bp, bp+bx) vx, the rest vy.
__global__ __launch_bounds__(256) void room_invariant_kernel(const double *p, const double *pp,
                                                             const double *vx, const double *vy,
                                                             int nx, int ny, double rc2,
                                                             double volume, double density,
                                                             double *partials) {
    __shared__ double red[4];
    const int64_t np = (int64_t)nx * ny, nfx = (int64_t)(nx - 1) * ny, nfy = (int64_t)nx * (ny - 1);
    const int64_t bp = (np + 255) / 256, bx = (nfx + 255) / 256;
    const int64_t b = blockIdx.x;
    double acc = 0.0;
    if (b < bp) {
        const int64_t k = b * 256 + threadIdx.x;
        if (k < np) {
            const double share = 1.0 / (wall(k % nx, nx) * wall(k / nx, ny));
            acc = share * p[k] * pp[k] / (2.0 * rc2) * volume;
        }
    } else if (b < bp + bx) {
        const int64_t k = (b - bp) * 256 + threadIdx.x;
        if (k < nfx) acc = density * vx[k] * vx[k] / 2.0 * volume / wall(k / (nx - 1), ny);
    } else {
        const int64_t k = (b - bp - bx) * 256 + threadIdx.x;
        if (k < nfy) acc = density * vy[k] * vy[k] / 2.0 * volume / wall(k % nx, nx);
    }
    const double total = block_sum<256>(acc, red);
    if (threadIdx.x == 0) partials[b] = total;
}

// The room's invariant from room_invariant_kernel's partials: each group summed in a fixed
// order, added as potential + kinetic_x + kinetic_y (room.rs:280) into out[0]; with an audit
// state, the ledger (energy = invariant + offset) is audited as step `step`.  Nothing runs once
// an audit has halted the run.
__global__ __launch_bounds__(1024) void room_total(const double *part, int64_t bp, int64_t bx,
                                                   int64_t by, AuditState *st, double offset,
                                                   int64_t step, double *totals, double *out) {
    __shared__ double red[16];
    if (st && st->halted) return;
    const double potential = fixed_order_sum_1024(part, bp, red);
    __syncthreads();
    const double kinetic_x = fixed_order_sum_1024(part + bp, bx, red);
    __syncthreads();
    const double kinetic_y = fixed_order_sum_1024(part + bp + bx, by, red);
    if (threadIdx.x == 0) {
        const double inv = potential + kinetic_x + kinetic_y;
        out[0] = inv;
        if (st) {
            const double energy = inv + offset;
            audit_step(st, energy, std::fabs(energy), step, totals);
        }
    }
}

// Sums the step's partials into the invariant and audits energy + offset - supplied.
__global__ __launch_bounds__(1024) void hall_finalize(const double *partials, int64_t n,
                                                      AuditState *st, double offset,
                                                      double *supplied_total,
                                                      const double *supplied_step,
                                                      int64_t step, double *totals,
                                                      double *invariant_out) {
    __shared__ double scratch[16];
    if (step >= 0 && st->halted) return;
    const double inv = fixed_order_sum_1024(partials, n, scratch);
    if (threadIdx.x == 0) {
        if (invariant_out) *invariant_out = inv;
        if (step >= 0) {
            double supplied = *supplied_total;
            if (supplied_step) supplied += *supplied_step;
            *supplied_total = supplied;
            const double energy = inv + offset;
            // Ledger::new().with(ENERGY, energy).with(ENERGY, -supplied)
            double total = 0.0;
            total += energy;
            total += -supplied;
            const double scale = rmax(rmax(0.0, std::fabs(energy)), std::fabs(supplied));
            audit_step(st, total, scale, step, totals);
        }
    }
}

__global__ void hall_zero(double *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

__global__ void hall_audit_reset(AuditState *st, double total, double scale, double tol) {
    st->prev_total = total;
    st->prev_scale = scale;
    st->tolerance = tol;
    st->halted = 0;
    st->breach_step = -1;
    st->before = st->after = st->scale = 0.0;
}

}  // namespace pm

using namespace pm;

namespace pm {

// clamp_index, hall.rs:684-694.
__device__ __forceinline__ void clamp_index(double q, int64_t n, int64_t *i, double *f) {
    if (isnan(q) || q <= 0.0) {
        *i = 0, *f = 0.0;
        return;
    }
    const double last = (double)(n - 1);
    if (q >= last) {
        *i = n - 1, *f = 0.0;
        return;
    }
    const double fl = floor(q);
    *i = (int64_t)fl, *f = q - fl;
}

__device__ __forceinline__ double lerp(double lo, double hi, double t) { return lo * (1.0 - t) + hi * t; }

// ScalarField::at for Hall, hall.rs:590-617: trilinear between nodes (x, then y, then z),
// clamped at the walls; p[0] for a NaN point.
__global__ void hall_sample(const double *__restrict__ p, int64_t nx, int64_t ny, int64_t nz,
                            double dx, const double *__restrict__ xyz, int64_t n,
                            double *__restrict__ out) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        const double qx = xyz[3 * q] / dx, qy = xyz[3 * q + 1] / dx, qz = xyz[3 * q + 2] / dx;
        if (isnan(qx) || isnan(qy) || isnan(qz)) {
            out[q] = p[0];
            continue;
        }
        int64_t i, j, k;
        double fx, fy, fz;
        clamp_index(qx, nx, &i, &fx);
        clamp_index(qy, ny, &j, &fy);
        clamp_index(qz, nz, &k, &fz);
        const int64_t i1 = min(i + 1, nx - 1), j1 = min(j + 1, ny - 1), k1 = min(k + 1, nz - 1);
        auto g = [&](int64_t a, int64_t b, int64_t c) { return p[a + nx * (b + ny * c)]; };
        const double z0 = lerp(lerp(g(i, j, k), g(i1, j, k), fx), lerp(g(i, j1, k), g(i1, j1, k), fx), fy);
        const double z1 = lerp(lerp(g(i, j, k1), g(i1, j, k1), fx), lerp(g(i, j1, k1), g(i1, j1, k1), fx), fy);
        out[q] = lerp(z0, z1, fz);
    }
}

}  // namespace pm

struct pm_hall {
    std::string name;
    int64_t nx = 2, ny = 1, nz = 1;
    double dx = 0.0, speed = 0.0, density = 0.0;
    bool staggered = false;
    double energy_offset = 0.0, energy_datum = 0.0;
    // A Room (crates/pantometry-acoustic/src/room.rs) is a hall one node deep: the same
    // arithmetic, under the two-dimensional Courant limit, with the Room's own invariant for a
    // room `room_depth` metres deep.
    int dims = 3;
    double room_depth = 0.0;
    DevBuf room_part;  // the room invariant's per-block partials
    int64_t room_groups[3] = {0, 0, 0};
    hipStream_t stream = nullptr;
    // Pressure rotates through three buffers: current, previous (pressure_prev) and a spare the
    // two-steps-per-pass kernel writes its intermediate pressure into.  Velocities ping-pong.
    DevBuf p[3], vx[2], vy[2], vz[2];
    int ip = 0, iq = 1, is = 2, iv = 0;
    int steps_per_pass = 2;
    DevBuf saved_p, saved_vx, saved_vy, saved_vz;
    bool saved_valid = false, saved_staggered = false;
    double saved_supplied = 0.0;  // the source's booked energy and cursor at the checkpoint
    uint64_t saved_cursor = 0;
    DevBuf partials, audit, scalars;  // scalars: [0] invariant, [1] supplied total, [2] supplied step
    int64_t partial_count = 0;
    DevBuf src_inc;
    int64_t src = -1;
    uint64_t src_count = 0, src_cursor = 0;
    std::vector<double> src_host;
    double supplied_host = 0.0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int tile_y = 8, z_chunk = 0;  // marching tiles; z_chunk 0: chosen from the extents
    int tile_x = 0;               // the paired step's tile width (64, 56, 48); 0: from nx

    int64_t nodes() const { return nx * ny * nz; }
    int64_t nfx() const { return (nx - 1) * ny * nz; }
    int64_t nfy() const { return nx * (ny - 1) * nz; }
    int64_t nfz() const { return nx * ny * (nz - 1); }
};

namespace pm {

// Tiles of the marching step: tile_x x tile_y columns and, unless fixed, z-chunks of at most 128
// planes, short enough that a launch has the workgroups hall_grid asks for, at least 4 planes.
struct HallGrid {
    int gx, gy, gz, zc;
    // The paired step's last-column strips (hall_strip2): gs along y by gzs z-chunks of zcs.
    int gs = 0, gzs = 0, zcs = 0;
    int64_t tiles() const { return (int64_t)gx * gy * gz; }
    int64_t strips() const { return (int64_t)gs * gzs; }
};

// The paired step's tile width: the hall's own setting, else 64.  Narrower tiles leave fewer
// idle lanes in the last x-tile but run fewer waves per workgroup, and the step is bound by the
// latency of its per-plane chain, not by issue: at 385 nodes 56-wide tiles (2% idle lanes) ran
// 4.6% slower than 64-wide ones (14% idle) and 48-wide ones 15% slower.
static int hall_tile_x(const pm_hall *h) { return h->tile_x ? h->tile_x : 64; }

// The paired step addresses each array from its chunk's lowest plane with 32-bit byte offsets:
// a chunk of zc planes reaches zc + 4 planes of an array, which must span fewer than 2^32 bytes.
static int64_t hall_pair_max_chunk(const pm_hall *h) {
    return (int64_t)(((uint64_t)1 << 32) - 1) / (h->nx * h->ny * 8) - 4;
}

// Tiles of a step.  For the paired step (`paired`), a hall of gx * tile_x + 1 nodes along x
// takes gx x-tiles and strips for its last column (hall_strip2) instead of gx + 1 x-tiles.
static HallGrid hall_grid(const pm_hall *h, int tile_y = 0, int tile_x = 64, int64_t max_zc = 0,
                          bool paired = false) {
    const int ty = tile_y ? tile_y : h->tile_y;
    HallGrid g;
    g.gx = (int)((h->nx + tile_x - 1) / tile_x);
    const bool last = paired && h->nx > tile_x && h->nx % tile_x == 1;
    if (last) g.gx = (int)(h->nx / tile_x);
    g.gy = (int)((h->ny + ty - 1) / ty);
    const int64_t xy = (int64_t)g.gx * g.gy;
    // At least `least` workgroups, so the last, partly filled round of workgroups is a small part
    // of the launch.  The single step: 2560, five rounds of two workgroups per CU over 256 CUs
    // (at 385^3 z-chunks of 49, 4% faster than 97).  The paired step: 1536, three rounds (at
    // 385^3 with the last column's strips, z-chunks of 65: 0.618 ms a step against 0.627-0.633 at
    // 49 and 77, 0.647 at 97 and 0.665 at 129).
    const int64_t least = paired ? 1536 : 2560;
    const int64_t chunks = std::max<int64_t>((least + xy - 1) / xy, (h->nz + 127) / 128);
    g.zc = h->z_chunk > 0 ? h->z_chunk
                          : (int)std::max<int64_t>(4, (h->nz + chunks - 1) / chunks);
    if (max_zc > 0 && g.zc > max_zc) g.zc = (int)max_zc;
    g.gz = (int)((h->nz + g.zc - 1) / g.zc);
    if (last) {
        // z-chunks short enough for one round of two strips per CU (at 385^3: 7 x 65 strips of
        // 6 planes)
        g.gs = (int)((h->ny + tile_x - 1) / tile_x);
        const int64_t want = std::max<int64_t>(1, 512 / g.gs);
        g.zcs = (int)std::max<int64_t>(4, (h->nz + want - 1) / want);
        if (max_zc > 0 && g.zcs > max_zc) g.zcs = (int)max_zc;
        g.gzs = (int)((h->nz + g.zcs - 1) / g.zcs);
    }
    return g;
}

// The room's invariant, per block of 256 terms, into room_part (potential blocks, then vx, then
// vy; the three counts in room_groups).
static int room_partials(pm_hall *h, const double *p, const double *pp, const double *vx,
                         const double *vy) {
    const int64_t bp = (h->nodes() + 255) / 256, bx = (h->nfx() + 255) / 256;
    const int64_t by = (h->nfy() + 255) / 256;
    if (!h->room_part) PM_CHECK(h->room_part.alloc(sizeof(double) * (size_t)(bp + bx + by)));
    h->room_groups[0] = bp, h->room_groups[1] = bx, h->room_groups[2] = by;
    const double rc2 = h->density * h->speed * h->speed;
    room_invariant_kernel<<<(unsigned)(bp + bx + by), 256, 0, h->stream>>>(
        p, pp, vx, vy, (int)h->nx, (int)h->ny, rc2, h->dx * h->dx * h->room_depth, h->density,
        h->room_part.d());
    PM_HIP(hipGetLastError());
    return PM_OK;
}

static int hall_invariant(pm_hall *h, const double *p, const double *pp, const double *vx,
                          const double *vy, const double *vz, double *out) {
    const double rc2 = h->density * h->speed * h->speed;
    if (rc2 <= 0.0) {
        *out = 0.0;
        return PM_OK;
    }
    if (h->dims == 2) {
        PM_CHECK(room_partials(h, p, pp, vx, vy));
        room_total<<<1, 1024, 0, h->stream>>>(h->room_part.d(), h->room_groups[0], h->room_groups[1],
                                              h->room_groups[2], nullptr, 0.0, -1, nullptr, h->scalars.d());
        PM_HIP(hipGetLastError());
        PM_HIP(hipMemcpyAsync(out, h->scalars.d(), sizeof(double), hipMemcpyDeviceToHost, h->stream));
        PM_HIP(hipStreamSynchronize(h->stream));
        return PM_OK;
    }
    const int64_t n = h->nodes();
    const int64_t blocks = (n + 255) / 256;
    DevBuf part;
    PM_CHECK(part.alloc(sizeof(double) * (size_t)blocks));
    const double volume = h->dx * h->dx * h->dx;
    hall_invariant_kernel<<<(unsigned)blocks, 256, 0, h->stream>>>(
        p, pp, vx, vy, vz, (int)h->nx, (int)h->ny, (int)h->nz, 2.0 * rc2, volume, 0.5 * h->density, part.d());
    PM_HIP(hipGetLastError());
    hall_finalize<<<1, 1024, 0, h->stream>>>(part.d(), blocks, h->audit.as<AuditState>(), 0.0,
                                             nullptr, nullptr, -1, nullptr, h->scalars.d());
    PM_HIP(hipGetLastError());
    PM_HIP(hipMemcpyAsync(out, h->scalars.d(), sizeof(double), hipMemcpyDeviceToHost, h->stream));
    PM_HIP(hipStreamSynchronize(h->stream));
    return PM_OK;
}

static int hall_invariant_now(pm_hall *h, double *out) {
    return hall_invariant(h, h->p[h->ip].d(), h->p[h->iq].d(), h->vx[h->iv].d(),
                          h->vy[h->iv].d(), h->vz[h->iv].d(), out);
}

static int hall_check(const pm_hall *h, double dt, pm_violation *v, int64_t step) {
    const double courant = h->speed * dt * std::sqrt((double)h->dims) / h->dx;
    if (courant > 1.0 + 1e-12) {
        std::string site = h->name + (h->dims == 2 ? " (two-dimensional wave equation)"
                                                   : " (three-dimensional wave equation)");
        fill_violation(v, "Courant number", site.c_str(), 1.0, courant, 1.0, 1e-12, step);
        return PM_VIOLATION;
    }
    return PM_OK;
}

// The source terms of the step the cursor is at: whether the source acts, its increment and
// the previous increment (hall.rs has no source; the extension books it as supplied energy).
struct SourceStep {
    int64_t node;
    double inc, prev;
};
static SourceStep source_at(const pm_hall *h, uint64_t cursor) {
    SourceStep s{-1, 0.0, 0.0};
    if (h->src >= 0 && cursor <= h->src_count) {
        s.node = h->src;
        if (cursor < h->src_count) s.inc = h->src_host[cursor];
        if (cursor > 0) s.prev = h->src_host[cursor - 1];
    }
    return s;
}

static HallArgs hall_args(pm_hall *h, double hstep, double vh) {
    const double rc2 = h->density * h->speed * h->speed;
    HallArgs a{};
    a.partials = h->partials.d();
    a.supplied_step = h->scalars.d() + 2;
    a.supplied_step2 = h->scalars.d() + 3;
    a.audit = h->audit.as<AuditState>();
    a.gain = vh / (h->density * h->dx);
    a.scale = hstep * rc2 / h->dx;
    a.two_rc2 = 2.0 * rc2;
    a.volume = h->dx * h->dx * h->dx;
    a.half_rho = 0.5 * h->density;
    a.pot = a.volume / a.two_rc2;
    a.kin = a.half_rho * a.volume;
    a.nx = (int)h->nx, a.ny = (int)h->ny, a.nz = (int)h->nz;
    a.src = a.src2 = -1;
    return a;
}

static int hall_partials(pm_hall *h, int64_t count) {
    if (h->partial_count < count) {
        PM_CHECK(h->partials.alloc(sizeof(double) * (size_t)count));
        h->partial_count = count;
    }
    return PM_OK;
}

// One fused step from the current buffers into the others, then the finalize (audit when
// step >= 0); `force` runs it even after a breach, without the finalize.
static int hall_sweep(pm_hall *h, double hstep, double vh, int64_t step, double *totals_dev,
                      double offset, bool force = false) {
    const HallGrid g = hall_grid(h);
    const int64_t nparts = g.tiles();
    PM_CHECK(hall_partials(h, nparts));
    HallArgs a = hall_args(h, hstep, vh);
    const int v = h->iv, vo = h->iv ^ 1;
    a.p = h->p[h->ip].d(), a.vx = h->vx[v].d(), a.vy = h->vy[v].d(), a.vz = h->vz[v].d();
    a.p_out = h->p[h->iq].d(), a.vx_out = h->vx[vo].d(), a.vy_out = h->vy[vo].d(), a.vz_out = h->vz[vo].d();
    a.force = force ? 1 : 0;
    const SourceStep src = source_at(h, h->src_cursor);
    a.src = src.node, a.source = src.inc, a.source_prev = src.prev;
    const bool sourced = src.node >= 0;
    h->src_cursor++;
    if (sourced) {
        hall_zero<<<1, 64, 0, h->stream>>>(h->scalars.d() + 2, 1);
        PM_HIP(hipGetLastError());
    }
    const unsigned blocks = (unsigned)(XCDS_HALL * ((nparts + XCDS_HALL - 1) / XCDS_HALL));
    if (h->tile_y == 16)
        hall_march<16><<<blocks, dim3(64, 16), 0, h->stream>>>(a, g.gx, g.gy, g.gz, g.zc);
    else
        hall_march<8><<<blocks, dim3(64, 8), 0, h->stream>>>(a, g.gx, g.gy, g.gz, g.zc);
    PM_HIP(hipGetLastError());
    if (!force && h->dims == 3) {  // a room audits its own invariant (room_total)
        hall_finalize<<<1, 1024, 0, h->stream>>>(h->partials.d(), nparts, h->audit.as<AuditState>(), offset,
                                                 h->scalars.d() + 1, sourced ? h->scalars.d() + 2 : nullptr, step,
                                                 totals_dev, h->scalars.d());
        PM_HIP(hipGetLastError());
    }
    std::swap(h->ip, h->iq);
    h->iv = vo;
    return PM_OK;
}

// Steps `step` and `step + 1` (both full leapfrog steps) in one pass, then both audits.
static int hall_sweep_pair(pm_hall *h, double hstep, int64_t step, double *totals_dev, double offset) {
    const int ty = h->tile_y == 16 ? 16 : 8, tx = hall_tile_x(h);
    const HallGrid g = hall_grid(h, ty, tx, hall_pair_max_chunk(h), true);
    const int64_t nparts = g.tiles() + g.strips();
    PM_CHECK(hall_partials(h, 2 * nparts));
    HallArgs a = hall_args(h, hstep, hstep);
    a.strip_tiles = g.strips();
    const int v = h->iv, vo = h->iv ^ 1;
    a.p = h->p[h->ip].d(), a.vx = h->vx[v].d(), a.vy = h->vy[v].d(), a.vz = h->vz[v].d();
    a.p_out = h->p[h->iq].d(), a.p_mid = h->p[h->is].d();
    a.vx_out = h->vx[vo].d(), a.vy_out = h->vy[vo].d(), a.vz_out = h->vz[vo].d();
    const SourceStep s1 = source_at(h, h->src_cursor), s2 = source_at(h, h->src_cursor + 1);
    a.src = s1.node, a.source = s1.inc, a.source_prev = s1.prev;
    a.src2 = s2.node, a.source2 = s2.inc;
    h->src_cursor += 2;
    if (s1.node >= 0 || s2.node >= 0) {
        hall_zero<<<1, 64, 0, h->stream>>>(h->scalars.d() + 2, 2);
        PM_HIP(hipGetLastError());
    }
    const unsigned blocks = (unsigned)(XCDS_HALL * ((g.tiles() + XCDS_HALL - 1) / XCDS_HALL));
    const unsigned sblocks = (unsigned)g.strips();
    const dim3 wg(tx, ty);
    // The strips run after the tiles on the same stream: at 385^3 the tiles take 1151 us a pass and
    // the strips 69 us.  Forked onto a second stream they ran beside the tiles from the start and
    // stretched the pair to 1270 us.
#define PM_HALL_PAIR(TX, TY)                                                                                \
    do {                                                                                                    \
        hall_march2<TX, TY><<<blocks, wg, 0, h->stream>>>(a, g.gx, g.gy, g.gz, g.zc);                      \
        if (sblocks) hall_strip2<TX, TY><<<sblocks, wg, 0, h->stream>>>(a, g.gs, g.gzs, g.zcs, g.tiles()); \
    } while (0)
    if (ty == 16) {
        if (tx == 56) PM_HALL_PAIR(56, 16);
        else if (tx == 48) PM_HALL_PAIR(48, 16);
        else PM_HALL_PAIR(64, 16);
    } else {
        if (tx == 56) PM_HALL_PAIR(56, 8);
        else if (tx == 48) PM_HALL_PAIR(48, 8);
        else PM_HALL_PAIR(64, 8);
    }
#undef PM_HALL_PAIR
    PM_HIP(hipGetLastError());
    hall_finalize2<<<1, 1024, 0, h->stream>>>(h->partials.d(), nparts, h->audit.as<AuditState>(), offset,
                                              h->scalars.d() + 1, s1.node >= 0 ? h->scalars.d() + 2 : nullptr,
                                              s2.node >= 0 ? h->scalars.d() + 3 : nullptr, step, totals_dev);
    PM_HIP(hipGetLastError());
    const int ip = h->ip, iq = h->iq, is = h->is;
    h->ip = iq, h->iq = is, h->is = ip;
    h->iv = vo;
    return PM_OK;
}

static int hall_reset_audit(pm_hall *h, double total, double scale, double tol) {
    hall_audit_reset<<<1, 1, 0, h->stream>>>(h->audit.as<AuditState>(), total, scale, tol);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

static int hall_copy(pm_hall *h, DevBuf &dst, const DevBuf &src, int64_t count) {
    if (count <= 0) return PM_OK;
    PM_HIP(hipMemcpyAsync(dst.d(), src.d(), sizeof(double) * (size_t)count, hipMemcpyDeviceToDevice, h->stream));
    return PM_OK;
}

// The first step: half a velocity step, then the startup check (hall.rs:422-428, 513-528).
static int hall_first_step(pm_hall *h, double hstep, pm_violation *v, int64_t step_index) {
    PM_CHECK(hall_reset_audit(h, 0.0, 0.0, INFINITY));
    PM_CHECK(hall_sweep(h, hstep, 0.5 * hstep, 0, nullptr, 0.0));
    h->staggered = true;
    double inv = 0.0;
    PM_HIP(hipMemcpyAsync(&inv, h->scalars.d(), sizeof(double), hipMemcpyDeviceToHost, h->stream));
    double supplied_step = 0.0;
    PM_HIP(hipMemcpyAsync(&supplied_step, h->scalars.d() + 1, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    PM_HIP(hipStreamSynchronize(h->stream));
    h->supplied_host = supplied_step;
    // The fused sum is the hall's invariant; a room measures its own.
    if (h->dims == 2) PM_CHECK(hall_invariant_now(h, &inv));
    // The startup offset is measured on the leapfrog alone: a source's first increment is
    // booked as supplied, not folded into the offset.
    const double inv_scheme = inv - supplied_step;
    const double offset = h->energy_datum - inv_scheme;
    const double scale = rmax(std::fabs(h->energy_datum), std::fabs(inv_scheme));
    if (scale > 0.0 && std::fabs(offset) / scale > 0.25) {
        std::string site = h->name + " (starting the leapfrog)";
        fill_violation(v, "energy", site.c_str(), h->energy_datum, inv_scheme, scale, 0.25, step_index);
        return PM_VIOLATION;
    }
    h->energy_offset = offset;
    return PM_OK;
}

}  // namespace pm

extern "C" {

int pm_hall_create(const char *name, double width, double height, double depth, size_t across,
                   double speed, double density, pm_hall **out) {
    if (!out) return PM_ERR_ARGUMENT;
    *out = nullptr;
    PM_CHECK(require_device());
    std::unique_ptr<pm_hall> h(new pm_hall());
    h->name = name ? name : "hall";
    // Hall::new, hall.rs:108-134.
    const int64_t nx = (int64_t)std::max<size_t>(across, 2);
    const double dx = std::fmax(width / (double)(nx - 1), 2.2250738585072014e-308);
    auto quantise = [&](double l) { return (int64_t)std::fmax(std::round(l / dx), 0.0) + 1; };
    h->nx = nx, h->ny = quantise(height), h->nz = quantise(depth);
    h->dx = dx, h->speed = speed, h->density = density;
    if (h->nx * h->ny > ((int64_t)1 << 31) || h->nz > (1 << 30)) {
        set_error("pm_hall_create: node counts beyond the kernel's indexing");
        return PM_ERR_ARGUMENT;
    }
    PM_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    PM_HIP(hipEventCreate(&h->ev0));
    PM_HIP(hipEventCreate(&h->ev1));
    PM_CHECK(h->p[2].alloc(sizeof(double) * (size_t)h->nodes()));
    PM_HIP(hipMemsetAsync(h->p[2].d(), 0, h->p[2].bytes(), h->stream));
    for (int b = 0; b < 2; b++) {
        PM_CHECK(h->p[b].alloc(sizeof(double) * (size_t)h->nodes()));
        PM_CHECK(h->vx[b].alloc(sizeof(double) * (size_t)std::max<int64_t>(h->nfx(), 1)));
        PM_CHECK(h->vy[b].alloc(sizeof(double) * (size_t)std::max<int64_t>(h->nfy(), 1)));
        PM_CHECK(h->vz[b].alloc(sizeof(double) * (size_t)std::max<int64_t>(h->nfz(), 1)));
        PM_HIP(hipMemsetAsync(h->p[b].d(), 0, h->p[b].bytes(), h->stream));
        PM_HIP(hipMemsetAsync(h->vx[b].d(), 0, h->vx[b].bytes(), h->stream));
        PM_HIP(hipMemsetAsync(h->vy[b].d(), 0, h->vy[b].bytes(), h->stream));
        PM_HIP(hipMemsetAsync(h->vz[b].d(), 0, h->vz[b].bytes(), h->stream));
    }
    PM_CHECK(h->audit.alloc(sizeof(AuditState)));
    PM_CHECK(h->scalars.alloc(sizeof(double) * 4));
    PM_HIP(hipMemsetAsync(h->scalars.d(), 0, sizeof(double) * 4, h->stream));
    PM_CHECK(hall_reset_audit(h.get(), 0.0, 0.0, 1e-9));
    PM_HIP(hipStreamSynchronize(h->stream));
    *out = h.release();
    return PM_OK;
}

int pm_room_create(const char *name, double width, double height, size_t cells_across,
                   double depth, double density, double speed, pm_hall **out) {
    if (!out) return PM_ERR_ARGUMENT;
    *out = nullptr;
    // Room::new, room.rs:90-120: at least three nodes each way, square cells from the width.
    const int64_t nx = (int64_t)std::max<size_t>(cells_across, 3);
    const double dx = width / (double)(nx - 1);
    const double rows = std::round(height / dx);
    const int64_t ny = std::max<int64_t>((rows > 0.0 ? (int64_t)rows : 0) + 1, 3);
    // The hall's constructor quantises a height to the same rows; the node counts are then
    // set to the room's own, whatever the hall's rules would have made of them.
    pm_hall *h = nullptr;
    PM_CHECK(pm_hall_create(name ? name : "room", width, (double)(ny - 1) * dx, 0.0, (size_t)nx, speed,
                            density, &h));
    if (h->nx != nx || h->ny != ny || h->nz != 1 || h->dx != dx) {
        pm_hall_destroy(h);
        set_error("pm_room_create: node counts (%lld, %lld) do not map onto one plane",
                  (long long)nx, (long long)ny);
        return PM_ERR_ARGUMENT;
    }
    h->dims = 2;
    h->room_depth = depth;
    *out = h;
    return PM_OK;
}

int pm_hall_dimensions(const pm_hall *h) { return h ? h->dims : 0; }

void pm_hall_destroy(pm_hall *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    hipStream_t st = h->stream;
    delete h;
    if (st) (void)hipStreamDestroy(st);
}

int pm_hall_nodes(const pm_hall *h, size_t nodes[3]) {
    if (!h || !nodes) return PM_ERR_ARGUMENT;
    nodes[0] = (size_t)h->nx, nodes[1] = (size_t)h->ny, nodes[2] = (size_t)h->nz;
    return PM_OK;
}

double pm_hall_spacing(const pm_hall *h) { return h ? h->dx : NAN; }

int pm_hall_release(pm_hall *h, const double *pressure) {
    if (!h || !pressure) return PM_ERR_ARGUMENT;
    h->ip = 0, h->iq = 1, h->is = 2, h->iv = 0;
    PM_HIP(hipMemcpyAsync(h->p[2].d(), pressure, sizeof(double) * (size_t)h->nodes(), hipMemcpyHostToDevice, h->stream));
    for (int b = 0; b < 2; b++) {
        PM_HIP(hipMemcpyAsync(h->p[b].d(), pressure, sizeof(double) * (size_t)h->nodes(), hipMemcpyHostToDevice, h->stream));
        PM_HIP(hipMemsetAsync(h->vx[b].d(), 0, h->vx[b].bytes(), h->stream));
        PM_HIP(hipMemsetAsync(h->vy[b].d(), 0, h->vy[b].bytes(), h->stream));
        PM_HIP(hipMemsetAsync(h->vz[b].d(), 0, h->vz[b].bytes(), h->stream));
    }
    PM_HIP(hipMemsetAsync(h->scalars.d(), 0, sizeof(double) * 4, h->stream));
    h->staggered = false;
    h->energy_offset = 0.0;
    h->supplied_host = 0.0;
    h->src_cursor = 0;
    PM_CHECK(hall_invariant_now(h, &h->energy_datum));
    return PM_OK;
}

int pm_hall_set_source(pm_hall *h, size_t node, const double *increments, uint64_t count) {
    if (!h) return PM_ERR_ARGUMENT;
    if (count == 0 || !increments) {
        h->src = -1, h->src_count = 0;
        h->src_host.clear();
        return PM_OK;
    }
    if (h->dims == 2) {
        set_error("pm_hall_set_source: a room has no source (room.rs)");
        return PM_ERR_ARGUMENT;
    }
    if ((int64_t)node >= h->nodes()) {
        set_error("pm_hall_set_source: node %zu is outside the hall", node);
        return PM_ERR_ARGUMENT;
    }
    h->src = (int64_t)node;
    h->src_host.assign(increments, increments + count);
    h->src_count = count;
    h->src_cursor = 0;
    return PM_OK;
}

double pm_hall_max_stable_dt(const pm_hall *h) {
    if (!h) return NAN;
    if (h->speed <= 0.0) return INFINITY;
    return h->dx / (h->speed * std::sqrt((double)h->dims));
}

double pm_hall_courant(const pm_hall *h, double dt) {
    if (!h) return NAN;
    return h->speed * dt * std::sqrt((double)h->dims) / h->dx;
}

int pm_hall_step(pm_hall *h, double dt, pm_violation *v) {
    if (!h) return PM_ERR_ARGUMENT;
    PM_CHECK(hall_check(h, dt, v, -1));
    if (dt <= 0.0) return PM_OK;
    if (!h->staggered) return hall_first_step(h, dt, v, -1);
    PM_CHECK(hall_reset_audit(h, 0.0, 0.0, INFINITY));
    PM_CHECK(hall_sweep(h, dt, dt, 0, nullptr, h->energy_offset));
    double supplied = 0.0;
    PM_HIP(hipMemcpyAsync(&supplied, h->scalars.d() + 1, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    PM_HIP(hipStreamSynchronize(h->stream));
    h->supplied_host = supplied;
    return PM_OK;
}

int pm_hall_sample(pm_hall *h, const double *xyz, size_t count, double *out) {
    if (!h || (count && (!xyz || !out))) return PM_ERR_ARGUMENT;
    if (count == 0) return PM_OK;
    DevBuf dp, dv;
    PM_CHECK(dp.alloc(sizeof(double) * 3 * count));
    PM_CHECK(dv.alloc(sizeof(double) * count));
    PM_HIP(hipMemcpyAsync(dp.d(), xyz, sizeof(double) * 3 * count, hipMemcpyHostToDevice, h->stream));
    hall_sample<<<(unsigned)std::min<size_t>((count + 255) / 256, 4096), 256, 0, h->stream>>>(
        h->p[h->ip].d(), h->nx, h->ny, h->nz, h->dx, dp.d(), (int64_t)count, dv.d());
    PM_HIP(hipGetLastError());
    PM_HIP(hipMemcpyAsync(out, dv.d(), sizeof(double) * count, hipMemcpyDeviceToHost, h->stream));
    PM_HIP(hipStreamSynchronize(h->stream));
    return PM_OK;
}

int pm_hall_energy(pm_hall *h, double *invariant, double *offset, double *supplied) {
    if (!h) return PM_ERR_ARGUMENT;
    double inv = 0.0;
    PM_CHECK(hall_invariant_now(h, &inv));
    if (invariant) *invariant = inv;
    if (offset) *offset = h->energy_offset;
    if (supplied) *supplied = h->supplied_host;
    return PM_OK;
}

namespace pm {

// A room's steps: the hall's sweep (its fused three-dimensional audit left out), then the
// room's own invariant and the audit of its ledger (energy only: a room has no source), all on
// the device.  The marching kernel returns at once after a breach, so the buffers that step
// wrote hold the state the breaking step made; the host finds them from the step index.
static int room_advance(pm_hall *h, double dt, uint64_t steps, double tolerance, double *totals,
                        double *elapsed_ms, pm_violation *v) {
    PM_HIP(hipEventRecord(h->ev0, h->stream));
    uint64_t first = 0;
    if (!h->staggered) {
        int st = hall_first_step(h, dt, v, 0);
        if (st != PM_OK) return st;
        first = 1;
    }
    double inv = 0.0;
    PM_CHECK(hall_invariant_now(h, &inv));
    const double total0 = inv + h->energy_offset;
    if (first == 1 && totals) totals[0] = total0;
    PM_CHECK(hall_reset_audit(h, total0, std::fabs(total0), tolerance));
    DevBuf dtot;
    PM_CHECK(dtot.alloc(sizeof(double) * (size_t)steps));
    struct Buffers {
        int ip, iq, is, iv;
    };
    std::vector<Buffers> after;  // the buffers holding the state after each step
    for (uint64_t n = first; n < steps; n++) {
        PM_CHECK(hall_sweep(h, dt, dt, (int64_t)n, nullptr, h->energy_offset));
        PM_CHECK(room_partials(h, h->p[h->ip].d(), h->p[h->iq].d(), h->vx[h->iv].d(), h->vy[h->iv].d()));
        room_total<<<1, 1024, 0, h->stream>>>(h->room_part.d(), h->room_groups[0], h->room_groups[1],
                                              h->room_groups[2], h->audit.as<AuditState>(), h->energy_offset,
                                              (int64_t)n, dtot.d(), h->scalars.d());
        PM_HIP(hipGetLastError());
        after.push_back({h->ip, h->iq, h->is, h->iv});
    }
    PM_HIP(hipEventRecord(h->ev1, h->stream));
    AuditState st;
    PM_HIP(hipMemcpyAsync(&st, h->audit.d(), sizeof st, hipMemcpyDeviceToHost, h->stream));
    PM_HIP(hipStreamSynchronize(h->stream));
    const uint64_t done = st.halted ? (uint64_t)st.breach_step + 1 : steps;
    if (totals && done > first)
        PM_HIP(hipMemcpy(totals + first, dtot.d() + first, sizeof(double) * (done - first), hipMemcpyDeviceToHost));
    if (elapsed_ms) {
        float ms = 0.0f;
        PM_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *elapsed_ms = ms;
    }
    if (st.halted) {
        const Buffers &b = after[(size_t)(st.breach_step - (int64_t)first)];
        h->ip = b.ip, h->iq = b.iq, h->is = b.is, h->iv = b.iv;
        fill_violation(v, "energy", "simulation", st.before, st.after, st.scale, tolerance, st.breach_step);
        return PM_VIOLATION;
    }
    return PM_OK;
}

}  // namespace pm

int pm_hall_advance(pm_hall *h, double dt, uint64_t steps, double tolerance, double *totals,
                    double *elapsed_ms, pm_violation *v) {
    if (!h) return PM_ERR_ARGUMENT;
    PM_CHECK(hall_check(h, dt, v, 0));
    if (dt <= 0.0 || steps == 0) return PM_OK;
    if (h->dims == 2) return room_advance(h, dt, steps, tolerance, totals, elapsed_ms, v);
    DevBuf dtot;
    PM_CHECK(dtot.alloc(sizeof(double) * (size_t)steps));
    PM_HIP(hipMemsetAsync(dtot.d(), 0xFF, sizeof(double) * (size_t)steps, h->stream));
    uint64_t first = 0;
    PM_HIP(hipEventRecord(h->ev0, h->stream));
    if (!h->staggered) {
        // The half step and its startup check run host-driven; the audit of that step against
        // the released ledger holds by construction of the offset.
        int st = hall_first_step(h, dt, v, 0);
        if (st != PM_OK) return st;
        first = 1;
    }
    // The opening ledger of the device loop: energy (invariant + offset) minus what was supplied.
    double inv = 0.0;
    PM_CHECK(hall_invariant_now(h, &inv));
    const double energy = inv + h->energy_offset;
    double total0 = 0.0;
    total0 += energy;
    total0 += -h->supplied_host;
    const double scale0 = rmax(rmax(0.0, std::fabs(energy)), std::fabs(h->supplied_host));
    if (first == 1 && totals) totals[0] = total0;
    PM_CHECK(hall_reset_audit(h, total0, scale0, tolerance));
    // Launches in order, with the buffers and source cursor each started from.
    struct Launch {
        int64_t first;
        int steps;
        int ip, iq, is, iv;
        uint64_t cursor;
    };
    std::vector<Launch> launches;
    // The paired step addresses five planes of an array or more with 32-bit byte offsets
    // (planes of at most 107,374,182 nodes) and assumes faces along every axis: other halls step one
    // at a time.
    const bool pairs = h->steps_per_pass == 2 && h->nx >= 2 && h->ny >= 2 && h->nz >= 2 &&
                       hall_pair_max_chunk(h) >= 1;
    for (uint64_t n = first; n < steps;) {
        launches.push_back({(int64_t)n, 1, h->ip, h->iq, h->is, h->iv, h->src_cursor});
        if (pairs && n + 1 < steps) {
            launches.back().steps = 2;
            PM_CHECK(hall_sweep_pair(h, dt, (int64_t)n, dtot.d(), h->energy_offset));
            n += 2;
        } else {
            PM_CHECK(hall_sweep(h, dt, dt, (int64_t)n, dtot.d(), h->energy_offset));
            n += 1;
        }
    }
    PM_HIP(hipEventRecord(h->ev1, h->stream));
    AuditState st;
    PM_HIP(hipMemcpyAsync(&st, h->audit.d(), sizeof st, hipMemcpyDeviceToHost, h->stream));
    if (totals && steps > first)
        PM_HIP(hipMemcpyAsync(totals + first, dtot.d() + first, sizeof(double) * (steps - first), hipMemcpyDeviceToHost, h->stream));
    double supplied = 0.0;
    PM_HIP(hipMemcpyAsync(&supplied, h->scalars.d() + 1, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    PM_HIP(hipStreamSynchronize(h->stream));
    h->supplied_host = supplied;
    if (elapsed_ms) {
        float ms = 0.0f;
        PM_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *elapsed_ms = ms;
    }
    if (st.halted) {
        // Launches after the breach did nothing: the state is what the breaking step made it.
        // A breach on the first step of a pair is recomputed alone from the pair's input, which
        // nothing has overwritten.
        for (const Launch &l : launches) {
            if (st.breach_step < l.first || st.breach_step >= l.first + l.steps) continue;
            h->ip = l.ip, h->iq = l.iq, h->is = l.is, h->iv = l.iv;
            h->src_cursor = l.cursor;
            if (l.steps == 2 && st.breach_step == l.first) {
                PM_CHECK(hall_sweep(h, dt, dt, -1, nullptr, h->energy_offset, true));
                PM_HIP(hipStreamSynchronize(h->stream));
            } else if (l.steps == 2) {
                const int ip = h->ip, iq = h->iq, is = h->is;
                h->ip = iq, h->iq = is, h->is = ip;
                h->iv ^= 1;
                h->src_cursor += 2;
            } else {
                std::swap(h->ip, h->iq);
                h->iv ^= 1;
                h->src_cursor += 1;
            }
            break;
        }
        fill_violation(v, "energy", "simulation", st.before, st.after, st.scale, tolerance, st.breach_step);
        return PM_VIOLATION;
    }
    return PM_OK;
}

int pm_hall_set_steps_per_pass(pm_hall *h, int steps) {
    if (!h) return PM_ERR_ARGUMENT;
    if (steps != 1 && steps != 2) {
        set_error("steps per pass must be 1 or 2");
        return PM_ERR_ARGUMENT;
    }
    h->steps_per_pass = steps;
    return PM_OK;
}

int pm_hall_get_fields(pm_hall *h, double *p, double *pp, double *vx, double *vy, double *vz) {
    if (!h) return PM_ERR_ARGUMENT;
    const int c = h->iv;
    if (p) PM_HIP(hipMemcpyAsync(p, h->p[h->ip].d(), sizeof(double) * (size_t)h->nodes(), hipMemcpyDeviceToHost, h->stream));
    if (pp) PM_HIP(hipMemcpyAsync(pp, h->p[h->iq].d(), sizeof(double) * (size_t)h->nodes(), hipMemcpyDeviceToHost, h->stream));
    if (vx && h->nfx()) PM_HIP(hipMemcpyAsync(vx, h->vx[c].d(), sizeof(double) * (size_t)h->nfx(), hipMemcpyDeviceToHost, h->stream));
    if (vy && h->nfy()) PM_HIP(hipMemcpyAsync(vy, h->vy[c].d(), sizeof(double) * (size_t)h->nfy(), hipMemcpyDeviceToHost, h->stream));
    if (vz && h->nfz()) PM_HIP(hipMemcpyAsync(vz, h->vz[c].d(), sizeof(double) * (size_t)h->nfz(), hipMemcpyDeviceToHost, h->stream));
    PM_HIP(hipStreamSynchronize(h->stream));
    return PM_OK;
}

int pm_hall_checkpoint(pm_hall *h) {
    if (!h) return PM_ERR_ARGUMENT;
    if (!h->saved_p) {
        PM_CHECK(h->saved_p.alloc(h->p[0].bytes()));
        PM_CHECK(h->saved_vx.alloc(h->vx[0].bytes()));
        PM_CHECK(h->saved_vy.alloc(h->vy[0].bytes()));
        PM_CHECK(h->saved_vz.alloc(h->vz[0].bytes()));
    }
    const int c = h->iv;
    PM_CHECK(hall_copy(h, h->saved_p, h->p[h->ip], h->nodes()));
    PM_CHECK(hall_copy(h, h->saved_vx, h->vx[c], h->nfx()));
    PM_CHECK(hall_copy(h, h->saved_vy, h->vy[c], h->nfy()));
    PM_CHECK(hall_copy(h, h->saved_vz, h->vz[c], h->nfz()));
    PM_HIP(hipStreamSynchronize(h->stream));
    h->saved_valid = true;
    h->saved_staggered = h->staggered;
    h->saved_supplied = h->supplied_host;
    h->saved_cursor = h->src_cursor;
    return PM_OK;
}

int pm_hall_restore(pm_hall *h) {
    if (!h) return PM_ERR_ARGUMENT;
    if (!h->saved_valid) return PM_OK;  // `if let Some(..) = self.saved` (hall.rs:547)
    const int c = h->iv;
    PM_CHECK(hall_copy(h, h->p[h->ip], h->saved_p, h->nodes()));
    PM_CHECK(hall_copy(h, h->p[h->iq], h->saved_p, h->nodes()));  // pressure_prev = p
    PM_CHECK(hall_copy(h, h->vx[c], h->saved_vx, h->nfx()));
    PM_CHECK(hall_copy(h, h->vy[c], h->saved_vy, h->nfy()));
    PM_CHECK(hall_copy(h, h->vz[c], h->saved_vz, h->nfz()));
    // The source's books go back with the fields.
    h->supplied_host = h->saved_supplied;
    h->src_cursor = h->saved_cursor;
    PM_HIP(hipMemcpyAsync(h->scalars.d() + 1, &h->saved_supplied, sizeof(double), hipMemcpyHostToDevice, h->stream));
    PM_HIP(hipStreamSynchronize(h->stream));
    h->staggered = h->saved_staggered;
    return PM_OK;
}

int pm_hall_set_tile(pm_hall *h, int tile_y, int z_chunk) {
    if (!h) return PM_ERR_ARGUMENT;
    if ((tile_y && tile_y != 8 && tile_y != 16) || z_chunk < 0) {
        set_error("hall tiles: tile_y must be 8 or 16 and z_chunk positive");
        return PM_ERR_ARGUMENT;
    }
    if (tile_y) h->tile_y = tile_y;
    h->z_chunk = z_chunk;
    return PM_OK;
}

int pm_hall_set_tile_width(pm_hall *h, int tile_x) {
    if (!h) return PM_ERR_ARGUMENT;
    if (tile_x && tile_x != 64 && tile_x != 56 && tile_x != 48) {
        set_error("hall tiles: tile_x must be 64, 56 or 48 (0: chosen from the extent)");
        return PM_ERR_ARGUMENT;
    }
    h->tile_x = tile_x;
    return PM_OK;
}

double pm_hall_peak_pressure(pm_hall *h) {
    if (!h) return NAN;
    std::vector<double> p((size_t)h->nodes());
    if (pm_hall_get_fields(h, p.data(), nullptr, nullptr, nullptr, nullptr) != PM_OK) return NAN;
    double m = 0.0;
    for (double v : p) m = rmax(m, std::fabs(v));
    return m;
}

}  // extern "C"
