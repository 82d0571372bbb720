// An ensemble of independent uniform heat blocks of uneven sizes, stepped together on gfx950.
//
// The reference runs such a study with Ensemble::run (crates/pantometry-core/src/ensemble.rs:
// 102-147): independent samples, contiguous chunks per thread, each sample a single-threaded
// Solid3D.  Here every member is a uniform Solid3D (plain branch) whose cells are concatenated
// into one device array per ping-pong side, and two schedules share the work:
//
//   resident   a member small enough for both of its temperature copies to fit in one
//              workgroup's LDS is stepped entirely in LDS: one workgroup per member runs every
//              step of the run in a single launch and touches HBM only to load and store.
//   streamed   every other member is cut into tiles (whole rows of the xy plane, or 64 x 8
//              for rows over 128 cells) by a z-chunk; one launch covers every tile of every
//              streamed member and advances it two steps (temporal blocking in LDS: T crosses
//              HBM once per two steps), and the launches of a run are captured once into a
//              hipGraph and replayed.
//
// The two schedules run on two streams joined inside the captured graph.  Which members are
// resident and how deep the z-chunks are is chosen by pm_ensemble_tune, which times the
// candidates on the ensemble itself.  Per-member arithmetic is Solid3D's plain sweep term by
// term (solid.rs:1776-1826), so every member is bit-identical to a standalone Solid3D.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "common.hpp"

namespace pm {

struct Member {
    int64_t off;          // first cell in the concatenated arrays
    int nx, ny, nz;
    int tx, ty;           // streamed tile: tx columns by ty rows
    int pad;
    double kface;         // series(k, k)
    double f;             // dt * dx / C
    double cap;           // C
    double tref;          // ledger reference
};

// A streamed work item: one tile of one member over the planes [z0, z1).  member < 0 pads the
// grid so that every XCD gets the same number of blocks.
struct Tile {
    int member;
    int x0, y0;
    int z0, z1;
    int pad;
};

constexpr int ENS_TILE = 512;  // most cells of a streamed tile
constexpr int ENS_W0 = 1024;   // most cells of a staged plane (tile + two-cell halo)
constexpr int ENS_NT = 1024;    // threads per streamed workgroup
constexpr int ENS_R = ENS_W0 / ENS_NT;  // staged cells per thread
constexpr int ENS_WAVES = 8;   // waves per SIMD the register budget is set for
constexpr int ENS_XCDS = 8;
constexpr int ENS_D = 2;       // planes prefetched ahead in registers

// One cell's update in the uniform form; neighbours off the block carry a zero face.
__device__ __forceinline__ double uniform_cell(double t, double tw, double te, double ts, double tn,
                                               double td, double tu, double kxl, double kxh,
                                               double kyl, double kyh, double kzl, double kzh,
                                               double f) {
    double flux = 0.0;
    flux = flux + kxl * (tw - t);
    flux = flux + kxh * (te - t);
    flux = flux + kyl * (ts - t);
    flux = flux + kyh * (tn - t);
    flux = flux + kzl * (td - t);
    flux = flux + kzh * (tu - t);
    return t + f * flux;
}

// The uniform update of a cell whose own value t and z neighbours td/tu are in registers and
// whose in-plane neighbours sit at index c of a plane staged in LDS with row stride w; `mask`
// says which in-plane neighbours exist (bit 0 x-1, 1 x+1, 2 y-1, 3 y+1; a missing one reads
// the cell itself, as in the uniform sweep).
__device__ __forceinline__ double staged_cell(const double *pc, int c, int w, unsigned mask,
                                              double t, double td, double tu, double kface,
                                              double kzl, double kzh, double f) {
    const double tw = pc[c - (int)(mask & 1u)];
    const double te = pc[c + (int)((mask >> 1) & 1u)];
    const double ts = pc[(mask & 4u) ? c - w : c];
    const double tn = pc[(mask & 8u) ? c + w : c];
    return uniform_cell(t, tw, te, ts, tn, td, tu, (mask & 1u) ? kface : 0.0,
                        (mask & 2u) ? kface : 0.0, (mask & 4u) ? kface : 0.0,
                        (mask & 8u) ? kface : 0.0, kzl, kzh, f);
}

__device__ __forceinline__ unsigned face_mask(int x, int y, int nx, int ny) {
    return (x > 0 ? 1u : 0u) | (x + 1 < nx ? 2u : 0u) | (y > 0 ? 4u : 0u) | (y + 1 < ny ? 8u : 0u);
}

constexpr unsigned IN_FIRST = 16u;  // the cell takes the first step (tile + S-1 halo)
constexpr unsigned IN_TILE = 32u;   // the cell is the workgroup's own

// Streamed schedule: S (1 or 2) steps of every work item per launch.  A workgroup marches its
// tile in z.  Each thread owns up to two cells of the staged plane (tile + S-cell halo) and
// keeps their z columns in registers: T prefetched D planes ahead, then T(i-2..i), and for
// S = 2 the first-step values T1(i-3..i-1).  Per plane i it stages T(i) in LDS (for the
// in-plane neighbours; double-buffered), computes the first step at plane i-1 on the cells of
// the tile widened by S-1 (into LDS for the second step's neighbours) and, for S = 2, the
// second step at plane i-2 on the tile's own cells; one LDS barrier per plane.  The planes next
// to a z-chunk are recomputed by both chunks (same inputs, same operations, the same bits).
// T crosses HBM once per S steps.
template <int S, int D>
__global__ __launch_bounds__(ENS_NT, ENS_WAVES) void ens_tile(const Member *__restrict__ members,
                                                      const Tile *__restrict__ tiles,
                                                      const double *__restrict__ told,
                                                      double *__restrict__ tnew) {
    __shared__ double sT[2][ENS_W0];
    __shared__ double s1[S == 2 ? 2 : 1][S == 2 ? ENS_W0 : 1];
    const Tile it = tiles[blockIdx.x];
    if (it.member < 0) return;  // the whole workgroup leaves before any barrier
    const Member m = members[it.member];
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    const int64_t plane = (int64_t)nx * ny;
    const int ox0 = it.x0, ox1 = min(it.x0 + m.tx, nx), oy0 = it.y0, oy1 = min(it.y0 + m.ty, ny);
    const int lx0 = max(ox0 - S, 0), lx1 = min(ox1 + S, nx);
    const int ly0 = max(oy0 - S, 0), ly1 = min(oy1 + S, ny);
    const int w0 = lx1 - lx0, n0 = w0 * (ly1 - ly0);
    const int t = threadIdx.x;
    const double kf = m.kface, f = m.f;
    // The member's T and T' through buffer resources: the lane's in-plane byte offset plus the
    // plane as a uniform offset (members hold fewer than 2^29 cells, so both fit 32 bits).
    const __amdgpu_buffer_rsrc_t RT = buffer_rsrc(told + m.off), RN = buffer_rsrc(tnew + m.off);
    const uint32_t pb = (uint32_t)plane * 8u;

    // This thread's cells: in-plane offset (-1 for none) and role bits.
    int g[ENS_R];
    unsigned info[ENS_R];
#pragma unroll
    for (int r = 0; r < ENS_R; r++) {
        const int q = t + r * ENS_NT;
        g[r] = -1, info[r] = 0;
        if (q < n0) {
            const int x = lx0 + q % w0, y = ly0 + q / w0;
            g[r] = y * nx + x;
            info[r] = face_mask(x, y, nx, ny);
            if (x >= ox0 - (S - 1) && x < ox1 + (S - 1) && y >= oy0 - (S - 1) && y < oy1 + (S - 1))
                info[r] |= IN_FIRST;
            if (x >= ox0 && x < ox1 && y >= oy0 && y < oy1) info[r] |= IN_TILE;
        }
    }

    const int z0 = it.z0, z1 = it.z1;
    const int lo_load = max(z0 - S, 0), hi_load = min(z1 - 1 + S, nz - 1);
    const int lo1 = max(z0 - (S - 1), 0), hi1 = min(z1 - 1 + (S - 1), nz - 1);
    const int iend = z1 - 1 + S;

    double pre[D][ENS_R];
#pragma unroll
    for (int u = 0; u < D; u++)
#pragma unroll
        for (int r = 0; r < ENS_R; r++)
            pre[u][r] = (lo_load + u <= hi_load && g[r] >= 0)
                            ? buffer_load_f64(RT, (uint32_t)g[r] * 8u, (uint32_t)(lo_load + u) * pb)
                            : 0.0;
    double zm[ENS_R] = {}, zc[ENS_R] = {}, zp[ENS_R] = {};  // T(i-2), T(i-1), T(i)
    double ym[ENS_R] = {}, yc[ENS_R] = {}, yp[ENS_R] = {};  // T1(i-3), T1(i-2), T1(i-1)

    auto phase = [&](int i, double (&reg)[ENS_R]) {
#pragma unroll
        for (int r = 0; r < ENS_R; r++) zm[r] = zc[r], zc[r] = zp[r], zp[r] = reg[r];
        if (i <= hi_load) {
#pragma unroll
            for (int r = 0; r < ENS_R; r++)
                if (g[r] >= 0) sT[i & 1][t + r * ENS_NT] = zp[r];
            if (i + D <= hi_load) {
#pragma unroll
                for (int r = 0; r < ENS_R; r++)
                    if (g[r] >= 0) reg[r] = buffer_load_f64(RT, (uint32_t)g[r] * 8u, (uint32_t)(i + D) * pb);
            }
        }
        const int p = i - 1;
        if (p >= lo1 && p <= hi1) {
            const double *pc = sT[p & 1];
            const double kzl = p > 0 ? kf : 0.0, kzh = p + 1 < nz ? kf : 0.0;
#pragma unroll
            for (int r = 0; r < ENS_R; r++) {
                if (!(info[r] & IN_FIRST)) continue;
                const double v = staged_cell(pc, t + r * ENS_NT, w0, info[r], zc[r],
                                             p > 0 ? zm[r] : zc[r], p + 1 < nz ? zp[r] : zc[r],
                                             kf, kzl, kzh, f);
                if constexpr (S == 2) {
                    s1[p & 1][t + r * ENS_NT] = v;
                    ym[r] = yc[r], yc[r] = yp[r], yp[r] = v;
                } else {
                    buffer_store_f64_nt(v, RN, (uint32_t)g[r] * 8u, (uint32_t)p * pb);
                }
            }
        } else if constexpr (S == 2) {
#pragma unroll
            for (int r = 0; r < ENS_R; r++) ym[r] = yc[r], yc[r] = yp[r];
        }
        if constexpr (S == 2) {
            const int p2 = i - 2;
            if (p2 >= z0 && p2 < z1) {
                const double *pc = s1[p2 & 1];
                const double kzl = p2 > 0 ? kf : 0.0, kzh = p2 + 1 < nz ? kf : 0.0;
#pragma unroll
                for (int r = 0; r < ENS_R; r++) {
                    if (!(info[r] & IN_TILE)) continue;
                    const double v = staged_cell(pc, t + r * ENS_NT, w0, info[r], yc[r],
                                                 p2 > 0 ? ym[r] : yc[r],
                                                 p2 + 1 < nz ? yp[r] : yc[r], kf, kzl, kzh, f);
                    buffer_store_f64_nt(v, RN, (uint32_t)g[r] * 8u, (uint32_t)p2 * pb);
                }
            }
        }
        lds_barrier();
    };

    for (int i = lo_load; i <= iend; i += D) {
#pragma unroll
        for (int u = 0; u < D; u++)
            if (i + u <= iend) phase(i + u, pre[u]);
    }
}

// Resident schedule: one workgroup per member runs `steps` steps in LDS.  Reads from `told`,
// leaves the result in `tout`.
__global__ __launch_bounds__(1024) void ens_resident(const Member *__restrict__ members,
                                                     const int *__restrict__ which,
                                                     const double *__restrict__ told,
                                                     double *__restrict__ tout, int steps) {
    extern __shared__ double lds[];
    const Member m = members[which[blockIdx.x]];
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    const int n = nx * ny * nz, plane = nx * ny;
    double *a = lds, *b = lds + n;
    for (int c = threadIdx.x; c < n; c += 1024) a[c] = told[m.off + c];
    __syncthreads();
    for (int s = 0; s < steps; s++) {
        for (int c = threadIdx.x; c < n; c += 1024) {
            const int i = c % nx, j = (c / nx) % ny, k = c / plane;
            const double t = a[c];
            b[c] = uniform_cell(t, a[i > 0 ? c - 1 : c], a[i + 1 < nx ? c + 1 : c],
                                a[j > 0 ? c - nx : c], a[j + 1 < ny ? c + nx : c],
                                a[k > 0 ? c - plane : c], a[k + 1 < nz ? c + plane : c],
                                i > 0 ? m.kface : 0.0, i + 1 < nx ? m.kface : 0.0,
                                j > 0 ? m.kface : 0.0, j + 1 < ny ? m.kface : 0.0,
                                k > 0 ? m.kface : 0.0, k + 1 < nz ? m.kface : 0.0, m.f);
        }
        __syncthreads();
        double *tmp = a;
        a = b, b = tmp;
    }
    for (int c = threadIdx.x; c < n; c += 1024) tout[m.off + c] = a[c];
}

// Per-member stored heat: one workgroup per member folds its cells in a fixed order.
__global__ __launch_bounds__(256) void ens_stored(const Member *__restrict__ members,
                                                  const double *__restrict__ T, double *out) {
    __shared__ double red[4];
    const Member m = members[blockIdx.x];
    const int64_t n = (int64_t)m.nx * m.ny * m.nz;
    double acc = 0.0;
    for (int64_t c = threadIdx.x; c < n; c += 256) acc += m.cap * (T[m.off + c] - m.tref);
    const double total = block_sum<256>(acc, red);
    if (threadIdx.x == 0) out[blockIdx.x] = total;
}

}  // namespace pm

using namespace pm;

struct pm_ensemble {
    std::vector<pm_member> spec;
    std::vector<Member> members;
    std::vector<double> worst_rate;
    int64_t cells = 0;
    hipStream_t stream = nullptr, side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, ev0 = nullptr, ev1 = nullptr;
    DevBuf t[2];
    int cur = 0;
    DevBuf dmembers, dtiles, dresident, dstored;
    int resident_edge = 0;  // members with max edge <= this run resident
    int z_chunk = 32;       // target planes per streamed tile
    int tile_policy = 0;    // streamed tile shapes (tile_shape)
    int steps_per_pass = 2;
    int64_t tile_blocks = 0, tile_count = 0, resident_count = 0;
    size_t lds_limit = 0;
    hipGraphExec_t graph = nullptr;
    uint64_t graph_steps = 0;
    int graph_cur = -1;
};

namespace pm {

static size_t resident_bytes(const pm_member &m) { return 2 * sizeof(double) * m.nx * m.ny * m.nz; }

// The streamed tile of a member under a tile policy; each keeps the tile to at most ENS_TILE
// cells and its plane with a two-cell halo to at most ENS_W0 (tile_fits), else the member falls
// back to policy 0.
//   0 whole rows when a row is at most 128 cells (no x halo), as many as fit; else 64 x 8
//   1 half as many whole rows (twice the workgroups, more halo rows); else 64 x 4
//   2 whole rows up to 64 cells; wider rows in 64 x 8 tiles
//   3 32-column tiles, as many rows as fit
constexpr int ENS_POLICIES = 4;
static void tile_shape(int nx, int ny, int policy, int *tx, int *ty) {
    switch (policy) {
        case 1:
            if (nx <= 128) {
                *tx = nx;
                *ty = std::max(1, std::min(ny, ENS_TILE / nx / 2));
            } else {
                *tx = 64;
                *ty = std::min(ny, 4);
            }
            return;
        case 2:
            if (nx <= 64) {
                *tx = nx;
                *ty = std::max(1, std::min(ny, ENS_TILE / nx));
            } else {
                *tx = 64;
                *ty = std::min(ny, 8);
            }
            return;
        case 3:
            *tx = std::min(nx, 32);
            *ty = std::max(1, std::min(ny, ENS_TILE / *tx));
            return;
        default:
            if (nx <= 128) {
                *tx = nx;
                *ty = std::max(1, std::min(ny, ENS_TILE / nx));
            } else {
                *tx = 64;
                *ty = std::min(ny, 8);
            }
    }
}

static bool tile_fits(const Member &m) {
    const int w0 = std::min(m.tx + 4, m.nx), h0 = std::min(m.ty + 4, m.ny);
    return m.tx * m.ty <= ENS_TILE && w0 * h0 <= ENS_W0;
}

// Sets every member's streamed tile from the ensemble's tile policy (policy 0 where the
// policy's tile does not fit) and uploads the member table.
static int ens_apply_tiles(pm_ensemble *e) {
    for (Member &d : e->members) {
        Member t = d;
        tile_shape(d.nx, d.ny, e->tile_policy, &t.tx, &t.ty);
        if (!tile_fits(t)) tile_shape(d.nx, d.ny, 0, &t.tx, &t.ty);
        d.tx = t.tx, d.ty = t.ty;
    }
    if (!e->members.empty())
        PM_HIP(hipMemcpy(e->dmembers.d(), e->members.data(), sizeof(Member) * e->members.size(),
                         hipMemcpyHostToDevice));
    return PM_OK;
}

static int ens_plan(pm_ensemble *e) {
    PM_CHECK(ens_apply_tiles(e));
    std::vector<int> resident;
    // Streamed tiles are grouped by (member, z-chunk); the groups go, largest first, to the XCD
    // with the least work so far, so the tiles that share halo rows run on one XCD's L2 and
    // the XCDs finish together.  Launch block b runs on XCD b mod 8.
    struct Group { int member, z0, z1; double cost; };
    std::vector<Group> groups;
    for (size_t idx = 0; idx < e->spec.size(); idx++) {
        const pm_member &s = e->spec[idx];
        const size_t edge = std::max(s.nx, std::max(s.ny, s.nz));
        if ((int)edge <= e->resident_edge && resident_bytes(s) <= e->lds_limit) {
            resident.push_back((int)idx);
            continue;
        }
        const int nz = (int)s.nz;
        const int chunks = (nz + e->z_chunk - 1) / e->z_chunk;
        const int zc = (nz + chunks - 1) / chunks;
        for (int z0 = 0; z0 < nz; z0 += zc)
            groups.push_back({(int)idx, z0, std::min(z0 + zc, nz), (double)s.nx * s.ny * std::min(zc, nz - z0)});
    }
    std::stable_sort(groups.begin(), groups.end(), [](const Group &a, const Group &b) { return a.cost > b.cost; });
    std::vector<std::vector<Tile>> per(ENS_XCDS);
    std::vector<double> load(ENS_XCDS, 0.0);
    for (const Group &g : groups) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        const Member &m = e->members[g.member];
        for (int y0 = 0; y0 < m.ny; y0 += m.ty)
            for (int x0 = 0; x0 < m.nx; x0 += m.tx) per[x].push_back({g.member, x0, y0, g.z0, g.z1, 0});
        load[x] += g.cost;
    }
    size_t depth = 0, count = 0;
    for (const auto &v : per) depth = std::max(depth, v.size()), count += v.size();
    std::vector<Tile> tiles(depth * ENS_XCDS, Tile{-1, 0, 0, 0, 0, 0});
    for (int x = 0; x < ENS_XCDS; x++)
        for (size_t j = 0; j < per[x].size(); j++) tiles[j * ENS_XCDS + x] = per[x][j];
    e->tile_blocks = (int64_t)tiles.size();
    e->tile_count = (int64_t)count;
    e->resident_count = (int64_t)resident.size();
    if (!tiles.empty()) {
        PM_CHECK(e->dtiles.alloc(sizeof(Tile) * tiles.size()));
        PM_HIP(hipMemcpy(e->dtiles.d(), tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice));
    }
    if (!resident.empty()) {
        PM_CHECK(e->dresident.alloc(sizeof(int) * resident.size()));
        PM_HIP(hipMemcpy(e->dresident.d(), resident.data(), sizeof(int) * resident.size(), hipMemcpyHostToDevice));
    }
    if (e->graph) {
        (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    return PM_OK;
}

// The launches of `steps` steps from buffer c0, recorded into the capture of e->stream: resident
// members on the side stream (forked and joined by events), tiles on the main stream.
static int ens_record(pm_ensemble *e, uint64_t steps, size_t max_lds, int c0) {
    PM_HIP(hipEventRecord(e->fork, e->stream));
    PM_HIP(hipStreamWaitEvent(e->side, e->fork, 0));
    const int cend = (int)((c0 + steps) & 1);
    if (e->resident_count > 0)
        ens_resident<<<(unsigned)e->resident_count, 1024, max_lds, e->side>>>(
            e->dmembers.as<Member>(), e->dresident.as<int>(), e->t[c0].d(), e->t[cend].d(), (int)steps);
    // Two-step passes in an even number, so that after `steps` steps the state sits in the same
    // buffer as after `steps` single steps; the rest (at most three) as single steps.
    const uint64_t pairs = e->steps_per_pass == 2 ? (steps / 2) & ~(uint64_t)1 : 0;
    const uint64_t singles = steps - 2 * pairs;
    int c = c0;
    for (uint64_t s = 0; s < pairs + singles && e->tile_blocks > 0; s++) {
        if (s < pairs)
            ens_tile<2, ENS_D><<<(unsigned)e->tile_blocks, ENS_NT, 0, e->stream>>>(
                e->dmembers.as<Member>(), e->dtiles.as<Tile>(), e->t[c].d(), e->t[c ^ 1].d());
        else
            ens_tile<1, ENS_D><<<(unsigned)e->tile_blocks, ENS_NT, 0, e->stream>>>(
                e->dmembers.as<Member>(), e->dtiles.as<Tile>(), e->t[c].d(), e->t[c ^ 1].d());
        c ^= 1;
    }
    PM_HIP(hipEventRecord(e->join, e->side));
    PM_HIP(hipStreamWaitEvent(e->stream, e->join, 0));
    return PM_OK;
}

// Captures `steps` steps starting from buffer `cur` into a graph.
static int ens_capture(pm_ensemble *e, uint64_t steps) {
    if (e->graph && e->graph_steps == steps && e->graph_cur == e->cur) return PM_OK;
    if (e->graph) {
        (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    size_t max_lds = 0;
    for (const pm_member &s : e->spec) {
        const size_t edge = std::max(s.nx, std::max(s.ny, s.nz));
        if ((int)edge <= e->resident_edge && resident_bytes(s) <= e->lds_limit) max_lds = std::max(max_lds, resident_bytes(s));
    }
    if (max_lds > 0)
        PM_HIP(hipFuncSetAttribute((const void *)ens_resident, hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds));
    hipGraph_t g = nullptr;
    const int c0 = e->cur;
    PM_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    // Whatever the recording returns, the capture is ended (and a partial graph destroyed), so a
    // failure leaves the stream usable and reports its own error.
    const int rc = ens_record(e, steps, max_lds, c0);
    const hipError_t end = hipStreamEndCapture(e->stream, &g);
    if (rc != PM_OK || end != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc != PM_OK) return rc;
        set_error("hipStreamEndCapture failed: %s", hipGetErrorString(end));
        return PM_ERR_DEVICE;
    }
    hipError_t err = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess) {
        set_error("hipGraphInstantiate failed: %s", hipGetErrorString(err));
        return PM_ERR_DEVICE;
    }
    e->graph_steps = steps;
    e->graph_cur = c0;
    return PM_OK;
}

static int ens_run(pm_ensemble *e, uint64_t steps, double *elapsed_ms) {
    PM_CHECK(ens_capture(e, steps));
    PM_HIP(hipEventRecord(e->ev0, e->stream));
    PM_HIP(hipGraphLaunch(e->graph, e->stream));
    PM_HIP(hipEventRecord(e->ev1, e->stream));
    PM_HIP(hipStreamSynchronize(e->stream));
    if (elapsed_ms) {
        float ms = 0.0f;
        PM_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        *elapsed_ms = ms;
    }
    e->cur = (int)((e->cur + steps) & 1);
    return PM_OK;
}

}  // namespace pm

extern "C" {

int pm_ensemble_create(const pm_member *members, size_t count, pm_ensemble **out) {
    if (!out || (!members && count)) return PM_ERR_ARGUMENT;
    *out = nullptr;
    PM_CHECK(require_device());
    std::unique_ptr<pm_ensemble> e(new pm_ensemble());
    e->spec.assign(members, members + count);
    int64_t off = 0;
    for (size_t m = 0; m < count; m++) {
        pm_member &s = e->spec[m];
        s.nx = std::max<size_t>(s.nx, 1), s.ny = std::max<size_t>(s.ny, 1), s.nz = std::max<size_t>(s.nz, 1);
        if (s.nx * s.ny > (1u << 30) || s.nz > (1u << 30) || s.nx * s.ny * s.nz >= ((size_t)1 << 29)) {
            set_error("pm_ensemble_create: member %zu is too large (2^29 cells or more)", m);
            return PM_ERR_ARGUMENT;
        }
        Member d;
        d.off = off;
        d.nx = (int)s.nx, d.ny = (int)s.ny, d.nz = (int)s.nz, d.pad = 0;
        tile_shape(d.nx, d.ny, 0, &d.tx, &d.ty);
        if (!tile_fits(d)) {
            set_error("pm_ensemble_create: no streamed tile fits member %zu", m);
            return PM_ERR_ARGUMENT;
        }
        const double k = s.substance.conductivity;
        const double sum = k + k;
        d.kface = sum > 0.0 ? 2.0 * k * k / sum : 0.0;
        const double volume = s.dx * s.dx * s.dx;
        d.cap = (s.substance.density * volume) * s.substance.specific_heat;
        const double mob = s.dx / d.cap;
        d.f = s.dt * mob;
        d.tref = s.initial;
        e->members.push_back(d);
        // worst_rate over the 27 boundary classes of a uniform block (solid.rs:757-782).
        double worst = 0.0;
        bool nan = false;
        auto reps = [](int64_t n) {
            std::vector<int64_t> r{0};
            if (n > 2) r.push_back(1);
            if (n > 1) r.push_back(n - 1);
            return r;
        };
        for (int64_t kk : reps((int64_t)s.nz))
            for (int64_t jj : reps((int64_t)s.ny))
                for (int64_t ii : reps((int64_t)s.nx)) {
                    double f0 = ii > 0 ? d.kface : 0.0, f1 = ii + 1 < (int64_t)s.nx ? d.kface : 0.0;
                    double f2 = jj > 0 ? d.kface : 0.0, f3 = jj + 1 < (int64_t)s.ny ? d.kface : 0.0;
                    double f4 = kk > 0 ? d.kface : 0.0, f5 = kk + 1 < (int64_t)s.nz ? d.kface : 0.0;
                    double r = mob * (f0 + f1 + f2 + f3 + f4 + f5);
                    if (std::isnan(r)) nan = true; else worst = rmax(worst, r);
                }
        e->worst_rate.push_back(nan ? NAN : worst);
        off += (int64_t)(s.nx * s.ny * s.nz);
    }
    e->cells = off;
    int dev = 0;
    PM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    PM_HIP(hipGetDeviceProperties(&prop, dev));
    e->lds_limit = prop.sharedMemPerBlock;
    PM_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    PM_HIP(hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking));
    PM_HIP(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
    PM_HIP(hipEventCreateWithFlags(&e->join, hipEventDisableTiming));
    PM_HIP(hipEventCreate(&e->ev0));
    PM_HIP(hipEventCreate(&e->ev1));
    PM_CHECK(e->t[0].alloc(sizeof(double) * (size_t)std::max<int64_t>(off, 1)));
    PM_CHECK(e->t[1].alloc(sizeof(double) * (size_t)std::max<int64_t>(off, 1)));
    PM_CHECK(e->dmembers.alloc(sizeof(Member) * std::max<size_t>(count, 1)));
    PM_CHECK(e->dstored.alloc(sizeof(double) * std::max<size_t>(count, 1)));
    if (count) PM_HIP(hipMemcpy(e->dmembers.d(), e->members.data(), sizeof(Member) * count, hipMemcpyHostToDevice));
    // Every cell starts at its member's initial temperature (Solid3D::new).
    std::vector<double> init((size_t)std::max<int64_t>(off, 1), 0.0);
    for (size_t m = 0; m < count; m++)
        std::fill(init.begin() + e->members[m].off, init.begin() + e->members[m].off + (int64_t)(e->spec[m].nx * e->spec[m].ny * e->spec[m].nz), e->spec[m].initial);
    PM_HIP(hipMemcpy(e->t[0].d(), init.data(), sizeof(double) * (size_t)off, hipMemcpyHostToDevice));
    e->resident_edge = 20;
    PM_CHECK(ens_plan(e.get()));
    *out = e.release();
    return PM_OK;
}

void pm_ensemble_destroy(pm_ensemble *e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    for (hipEvent_t ev : {e->fork, e->join, e->ev0, e->ev1})
        if (ev) (void)hipEventDestroy(ev);
    hipStream_t s1 = e->stream, s2 = e->side;
    delete e;
    if (s1) (void)hipStreamDestroy(s1);
    if (s2) (void)hipStreamDestroy(s2);
}

int pm_ensemble_set_cells(pm_ensemble *e, size_t member, const double *cells) {
    if (!e || !cells || member >= e->spec.size()) return PM_ERR_ARGUMENT;
    const pm_member &s = e->spec[member];
    PM_HIP(hipMemcpy(e->t[e->cur].d() + e->members[member].off, cells, sizeof(double) * s.nx * s.ny * s.nz, hipMemcpyHostToDevice));
    return PM_OK;
}

int pm_ensemble_get_cells(pm_ensemble *e, size_t member, double *cells) {
    if (!e || !cells || member >= e->spec.size()) return PM_ERR_ARGUMENT;
    const pm_member &s = e->spec[member];
    PM_HIP(hipStreamSynchronize(e->stream));
    PM_HIP(hipMemcpy(cells, e->t[e->cur].d() + e->members[member].off, sizeof(double) * s.nx * s.ny * s.nz, hipMemcpyDeviceToHost));
    return PM_OK;
}

int pm_ensemble_run(pm_ensemble *e, uint64_t steps, double *elapsed_ms, pm_violation *v) {
    if (!e) return PM_ERR_ARGUMENT;
    for (size_t m = 0; m < e->spec.size(); m++) {
        const double ratio = e->spec[m].dt * e->worst_rate[m];
        if (std::isnan(ratio)) {
            violation_at(v, "ensemble member", "substance has no diffusivity", INFINITY, (int64_t)m);
            return PM_VIOLATION;
        }
        if (ratio > 1.0 + 6e-12) {
            const double F = 1.0 / 6.0;
            fill_violation(v, "Fourier number", "ensemble member (explicit 3D conduction)", F, ratio * F, F, 1e-12, (int64_t)m);
            return PM_VIOLATION;
        }
    }
    if (steps == 0 || e->spec.empty()) return PM_OK;
    return ens_run(e, steps, elapsed_ms);
}

int pm_ensemble_ledgers(pm_ensemble *e, double *stored) {
    if (!e || !stored) return PM_ERR_ARGUMENT;
    if (e->spec.empty()) return PM_OK;
    ens_stored<<<(unsigned)e->spec.size(), 256, 0, e->stream>>>(e->dmembers.as<Member>(), e->t[e->cur].d(), e->dstored.d());
    PM_HIP(hipGetLastError());
    PM_HIP(hipMemcpyAsync(stored, e->dstored.d(), sizeof(double) * e->spec.size(), hipMemcpyDeviceToHost, e->stream));
    PM_HIP(hipStreamSynchronize(e->stream));
    return PM_OK;
}

int pm_ensemble_set_tiles(pm_ensemble *e, int policy) {
    if (!e || policy < 0 || policy >= ENS_POLICIES) {
        set_error("pm_ensemble_set_tiles: policy must be 0..%d", ENS_POLICIES - 1);
        return PM_ERR_ARGUMENT;
    }
    e->tile_policy = policy;
    return ens_plan(e);
}

int pm_ensemble_set_schedule(pm_ensemble *e, int resident_edge, int z_chunk) {
    if (!e || z_chunk < 1) return PM_ERR_ARGUMENT;
    e->resident_edge = resident_edge;
    e->z_chunk = z_chunk;
    return ens_plan(e);
}

int pm_ensemble_set_steps_per_pass(pm_ensemble *e, int steps) {
    if (!e || (steps != 1 && steps != 2)) {
        set_error("pm_ensemble_set_steps_per_pass: steps must be 1 or 2");
        return PM_ERR_ARGUMENT;
    }
    e->steps_per_pass = steps;
    if (e->graph) {
        (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    return PM_OK;
}

int pm_ensemble_tune(pm_ensemble *e, char *report, size_t report_len) {
    if (!e) return PM_ERR_ARGUMENT;
    if (e->spec.empty()) return PM_OK;
    // Keep the state: tuning steps a copy and puts the original back.
    DevBuf keep;
    PM_CHECK(keep.alloc(sizeof(double) * (size_t)e->cells));
    PM_HIP(hipMemcpy(keep.d(), e->t[e->cur].d(), sizeof(double) * (size_t)e->cells, hipMemcpyDeviceToDevice));
    const int keep_cur = e->cur;
    struct Cand { int policy, edge, zc; };
    std::vector<Cand> cands;
    for (int policy = 0; policy < ENS_POLICIES; policy++)
        for (int edge : {0, 16, 20, 24})
            for (int zc : {16, 32, 64, 128}) cands.push_back({policy, edge, zc});
    const uint64_t trial = 20;
    double best = INFINITY;
    Cand pick = {e->tile_policy, e->resident_edge, e->z_chunk};
    std::string text;
    for (const Cand &c : cands) {
        e->tile_policy = c.policy, e->resident_edge = c.edge, e->z_chunk = c.zc;
        PM_CHECK(ens_plan(e));
        double ms = 0.0, ms2 = 0.0;
        PM_CHECK(ens_run(e, trial, &ms));   // warm (includes capture)
        PM_CHECK(ens_run(e, trial, &ms2));  // timed replay
        char line[160];
        std::snprintf(line, sizeof line,
                      "tiles=%d resident_edge=%d z_chunk=%d resident=%lld tiles=%lld: %.4f ms/step\n",
                      c.policy, c.edge, c.zc, (long long)e->resident_count,
                      (long long)e->tile_count, ms2 / trial);
        text += line;
        if (ms2 < best) best = ms2, pick = c;
    }
    e->tile_policy = pick.policy, e->resident_edge = pick.edge, e->z_chunk = pick.zc;
    PM_CHECK(ens_plan(e));
    e->cur = keep_cur;
    PM_HIP(hipMemcpy(e->t[e->cur].d(), keep.d(), sizeof(double) * (size_t)e->cells, hipMemcpyDeviceToDevice));
    char line[128];
    std::snprintf(line, sizeof line, "picked tiles=%d resident_edge=%d z_chunk=%d\n", pick.policy,
                  pick.edge, pick.zc);
    text += line;
    if (report && report_len) std::snprintf(report, report_len, "%s", text.c_str());
    return PM_OK;
}

}  // extern "C"
