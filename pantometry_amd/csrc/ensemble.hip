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
//   streamed   every other member is cut into work items (256 xy-columns by a z-chunk); one
//              launch per step covers every item of every streamed member, and the launches of
//              a run are captured once into a hipGraph and replayed.
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
    int nx, ny, nz, pad;
    double kface;         // series(k, k)
    double f;             // dt * dx / C
    double cap;           // C
    double tref;          // ledger reference
};

struct Item {
    int member;
    int col0;  // first xy column (i + nx*j)
    int z0, z1;
};

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

// Streamed schedule: one step of every work item.  A thread owns one xy column of its item and
// marches z with its own column in registers; x and y neighbours come through the L1/L2.
__global__ __launch_bounds__(256) void ens_step(const Member *__restrict__ members,
                                                const Item *__restrict__ items,
                                                const double *__restrict__ told,
                                                double *__restrict__ tnew) {
    const Item it = items[blockIdx.x];
    const Member m = members[it.member];
    const int64_t nx = m.nx, ny = m.ny, nz = m.nz, plane = nx * ny;
    const int64_t col = it.col0 + threadIdx.x;
    if (col >= plane) return;
    const int64_t i = col % nx, j = col / nx;
    const double *T = told + m.off;
    double *Tn = tnew + m.off;
    const double kxl = i > 0 ? m.kface : 0.0, kxh = i + 1 < nx ? m.kface : 0.0;
    const double kyl = j > 0 ? m.kface : 0.0, kyh = j + 1 < ny ? m.kface : 0.0;
    const int64_t dw = i > 0 ? -1 : 0, de = i + 1 < nx ? 1 : 0;
    const int64_t ds = j > 0 ? -nx : 0, dn = j + 1 < ny ? nx : 0;
    const int64_t z0 = it.z0, z1 = it.z1;
    double t_m = T[col + plane * (z0 > 0 ? z0 - 1 : 0)];
    double t_c = T[col + plane * z0];
    for (int64_t z = z0; z < z1; z++) {
        const int64_t c = col + plane * z;
        const double t_p = T[col + plane * (z + 1 < nz ? z + 1 : z)];
        const double kzl = z > 0 ? m.kface : 0.0, kzh = z + 1 < nz ? m.kface : 0.0;
        Tn[c] = uniform_cell(t_c, T[c + dw], T[c + de], T[c + ds], T[c + dn], t_m, t_p, kxl, kxh,
                             kyl, kyh, kzl, kzh, m.f);
        t_m = t_c, t_c = t_p;
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
    DevBuf dmembers, ditems, dresident, dstored;
    int resident_edge = 0;  // members with max edge <= this run resident
    int z_chunk = 16;
    int64_t item_count = 0, resident_count = 0;
    size_t lds_limit = 0;
    hipGraphExec_t graph = nullptr;
    uint64_t graph_steps = 0;
    int graph_cur = -1;
};

namespace pm {

static size_t resident_bytes(const pm_member &m) { return 2 * sizeof(double) * m.nx * m.ny * m.nz; }

static int ens_plan(pm_ensemble *e) {
    std::vector<Item> items;
    std::vector<int> resident;
    // Largest members first, so the long items start early.
    std::vector<size_t> order(e->spec.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return e->spec[a].nx * e->spec[a].ny * e->spec[a].nz > e->spec[b].nx * e->spec[b].ny * e->spec[b].nz;
    });
    for (size_t idx : order) {
        const pm_member &s = e->spec[idx];
        const size_t edge = std::max(s.nx, std::max(s.ny, s.nz));
        if ((int)edge <= e->resident_edge && resident_bytes(s) <= e->lds_limit) {
            resident.push_back((int)idx);
            continue;
        }
        const int plane = (int)(s.nx * s.ny);
        for (int z0 = 0; z0 < (int)s.nz; z0 += e->z_chunk)
            for (int c0 = 0; c0 < plane; c0 += 256)
                items.push_back({(int)idx, c0, z0, std::min<int>(z0 + e->z_chunk, (int)s.nz)});
    }
    e->item_count = (int64_t)items.size();
    e->resident_count = (int64_t)resident.size();
    if (!items.empty()) {
        PM_CHECK(e->ditems.alloc(sizeof(Item) * items.size()));
        PM_HIP(hipMemcpy(e->ditems.d(), items.data(), sizeof(Item) * items.size(), hipMemcpyHostToDevice));
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
    PM_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    PM_HIP(hipEventRecord(e->fork, e->stream));
    PM_HIP(hipStreamWaitEvent(e->side, e->fork, 0));
    const int c0 = e->cur;
    const int cend = (int)((c0 + steps) & 1);
    if (e->resident_count > 0)
        ens_resident<<<(unsigned)e->resident_count, 1024, max_lds, e->side>>>(
            e->dmembers.as<Member>(), e->dresident.as<int>(), e->t[c0].d(), e->t[cend].d(), (int)steps);
    int c = c0;
    for (uint64_t s = 0; s < steps && e->item_count > 0; s++) {
        ens_step<<<(unsigned)e->item_count, 256, 0, e->stream>>>(e->dmembers.as<Member>(), e->ditems.as<Item>(),
                                                                  e->t[c].d(), e->t[c ^ 1].d());
        c ^= 1;
    }
    PM_HIP(hipEventRecord(e->join, e->side));
    PM_HIP(hipStreamWaitEvent(e->stream, e->join, 0));
    PM_HIP(hipStreamEndCapture(e->stream, &g));
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
        if (s.nx * s.ny > (1u << 30) || s.nz > (1u << 30)) {
            set_error("pm_ensemble_create: member %zu is too large", m);
            return PM_ERR_ARGUMENT;
        }
        Member d;
        d.off = off;
        d.nx = (int)s.nx, d.ny = (int)s.ny, d.nz = (int)s.nz, d.pad = 0;
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

int pm_ensemble_set_schedule(pm_ensemble *e, int resident_edge, int z_chunk) {
    if (!e || z_chunk < 1) return PM_ERR_ARGUMENT;
    e->resident_edge = resident_edge;
    e->z_chunk = z_chunk;
    return ens_plan(e);
}

int pm_ensemble_tune(pm_ensemble *e, char *report, size_t report_len) {
    if (!e) return PM_ERR_ARGUMENT;
    if (e->spec.empty()) return PM_OK;
    // Keep the state: tuning steps a copy and puts the original back.
    DevBuf keep;
    PM_CHECK(keep.alloc(sizeof(double) * (size_t)e->cells));
    PM_HIP(hipMemcpy(keep.d(), e->t[e->cur].d(), sizeof(double) * (size_t)e->cells, hipMemcpyDeviceToDevice));
    const int keep_cur = e->cur;
    struct Cand { int edge, zc; };
    std::vector<Cand> cands;
    for (int edge : {0, 16, 20, 24})
        for (int zc : {8, 16, 32}) cands.push_back({edge, zc});
    const uint64_t trial = 20;
    double best = INFINITY;
    Cand pick = {e->resident_edge, e->z_chunk};
    std::string text;
    for (const Cand &c : cands) {
        e->resident_edge = c.edge, e->z_chunk = c.zc;
        PM_CHECK(ens_plan(e));
        double ms = 0.0, ms2 = 0.0;
        PM_CHECK(ens_run(e, trial, &ms));   // warm (includes capture)
        PM_CHECK(ens_run(e, trial, &ms2));  // timed replay
        char line[160];
        std::snprintf(line, sizeof line, "resident_edge=%d z_chunk=%d resident=%lld items=%lld: %.4f ms/step\n",
                      c.edge, c.zc, (long long)e->resident_count, (long long)e->item_count, ms2 / trial);
        text += line;
        if (ms2 < best) best = ms2, pick = c;
    }
    e->resident_edge = pick.edge, e->z_chunk = pick.zc;
    PM_CHECK(ens_plan(e));
    e->cur = keep_cur;
    PM_HIP(hipMemcpy(e->t[e->cur].d(), keep.d(), sizeof(double) * (size_t)e->cells, hipMemcpyDeviceToDevice));
    char line[128];
    std::snprintf(line, sizeof line, "picked resident_edge=%d z_chunk=%d\n", pick.edge, pick.zc);
    text += line;
    if (report && report_len) std::snprintf(report, report_len, "%s", text.c_str());
    return PM_OK;
}

}  // extern "C"
