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
// node's share of the invariant and the shares of the faces it owns.
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
    int64_t src;         // source node or -1
    double source;       // its increment this step
    double source_prev;  // its increment the step before
    int nx, ny, nz;
};

__device__ __forceinline__ double wall(int64_t k, int64_t n) { return (k == 0 || k + 1 == n) ? 2.0 : 1.0; }

template <int TY>
__global__ __launch_bounds__(64 * TY) void hall_step_kernel(HallArgs a) {
    constexpr int NT = 64 * TY;
    __shared__ double red[NT / 64];
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    const int64_t j = (int64_t)blockIdx.y * TY + threadIdx.y;
    const int64_t k = blockIdx.z;
    const int tid = threadIdx.x + 64 * threadIdx.y;
    const int64_t bid = blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
    double acc = 0.0;
    if (a.audit->halted) {
        if (i < nx && j < ny) {
            const int64_t n = i + nx * (j + ny * k);
            a.p_out[n] = a.p[n];
            if (i + 1 < nx) a.vx_out[i + (nx - 1) * (j + ny * k)] = a.vx[i + (nx - 1) * (j + ny * k)];
            if (j + 1 < ny) a.vy_out[i + nx * (j + (ny - 1) * k)] = a.vy[i + nx * (j + (ny - 1) * k)];
            if (k + 1 < nz) a.vz_out[n] = a.vz[n];
        }
        if (tid == 0) a.partials[bid] = 0.0;
        return;
    }
    if (i < nx && j < ny) {
        const int64_t n = i + nx * (j + ny * k);
        const double pc = a.p[n];
        const double g = a.gain;
        // New face velocities, hall.rs:433-457: v -= gain * (p[hi] - p[lo]).
        double hi_x = 0.0, lo_x = 0.0, hi_y = 0.0, lo_y = 0.0, hi_z = 0.0, lo_z = 0.0;
        const int64_t fxi = i + (nx - 1) * (j + ny * k);
        const int64_t fyi = i + nx * (j + (ny - 1) * k);
        if (i + 1 < nx) hi_x = a.vx[fxi] - g * (a.p[n + 1] - pc);
        if (i > 0) lo_x = a.vx[fxi - 1] - g * (pc - a.p[n - 1]);
        if (j + 1 < ny) hi_y = a.vy[fyi] - g * (a.p[n + nx] - pc);
        if (j > 0) lo_y = a.vy[fyi - nx] - g * (pc - a.p[n - nx]);
        if (k + 1 < nz) hi_z = a.vz[n] - g * (a.p[n + nx * ny] - pc);
        if (k > 0) lo_z = a.vz[n - nx * ny] - g * (pc - a.p[n - nx * ny]);
        const double wi = wall(i, nx), wj = wall(j, ny), wk = wall(k, nz);
        // Pressure, hall.rs:494-504: p -= scale * ((wi dx + wj dy) + wk dz).
        double pn = pc - a.scale * (wi * (hi_x - lo_x) + wj * (hi_y - lo_y) + wk * (hi_z - lo_z));
        if (n == a.src) {
            pn = pn + a.source;
            // An increment d at step n moves the invariant by d*p^(n-1) at once and by d*p^n at
            // step n+1 (whose energy identity uses p^n - p^(n-1)); both use this step's p_old.
            const double share = 1.0 / (wi * wj * wk);
            *a.supplied_step = share * (a.source + a.source_prev) * pc / a.two_rc2 * a.volume;
        }
        a.p_out[n] = pn;
        if (i + 1 < nx) a.vx_out[fxi] = hi_x;
        if (j + 1 < ny) a.vy_out[fyi] = hi_y;
        if (k + 1 < nz) a.vz_out[n] = hi_z;
        // The invariant's shares, hall.rs:346-381: the node's, then each owned face's.
        const double share = 1.0 / (wi * wj * wk);
        acc += share * pn * pc / a.two_rc2 * a.volume;
        if (i + 1 < nx) acc += 1.0 / (wj * wk) * a.half_rho * hi_x * hi_x * a.volume;
        if (j + 1 < ny) acc += 1.0 / (wi * wk) * a.half_rho * hi_y * hi_y * a.volume;
        if (k + 1 < nz) acc += 1.0 / (wi * wj) * a.half_rho * hi_z * hi_z * a.volume;
    }
    const double total = block_sum<NT>(acc, red);
    if (tid == 0) a.partials[bid] = total;
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

struct pm_hall {
    std::string name;
    int64_t nx = 2, ny = 1, nz = 1;
    double dx = 0.0, speed = 0.0, density = 0.0;
    bool staggered = false;
    double energy_offset = 0.0, energy_datum = 0.0;
    hipStream_t stream = nullptr;
    DevBuf p[2], vx[2], vy[2], vz[2];
    int cur = 0;
    DevBuf saved_p, saved_vx, saved_vy, saved_vz;
    bool saved_valid = false, saved_staggered = false;
    DevBuf partials, audit, scalars;  // scalars: [0] invariant, [1] supplied total, [2] supplied step
    int64_t partial_count = 0;
    DevBuf src_inc;
    int64_t src = -1;
    uint64_t src_count = 0, src_cursor = 0;
    std::vector<double> src_host;
    double supplied_host = 0.0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    int64_t nodes() const { return nx * ny * nz; }
    int64_t nfx() const { return (nx - 1) * ny * nz; }
    int64_t nfy() const { return nx * (ny - 1) * nz; }
    int64_t nfz() const { return nx * ny * (nz - 1); }
};

namespace pm {

static constexpr int HALL_TY = 4;

static dim3 hall_grid(const pm_hall *h) {
    return dim3((unsigned)((h->nx + 63) / 64), (unsigned)((h->ny + HALL_TY - 1) / HALL_TY), (unsigned)h->nz);
}

static int hall_invariant(pm_hall *h, const double *p, const double *pp, const double *vx,
                          const double *vy, const double *vz, double *out) {
    const double rc2 = h->density * h->speed * h->speed;
    if (rc2 <= 0.0) {
        *out = 0.0;
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
    return hall_invariant(h, h->p[h->cur].d(), h->p[h->cur ^ 1].d(), h->vx[h->cur].d(),
                          h->vy[h->cur].d(), h->vz[h->cur].d(), out);
}

static int hall_check(const pm_hall *h, double dt, pm_violation *v, int64_t step) {
    const double courant = h->speed * dt * std::sqrt(3.0) / h->dx;
    if (courant > 1.0 + 1e-12) {
        std::string site = h->name + " (three-dimensional wave equation)";
        fill_violation(v, "Courant number", site.c_str(), 1.0, courant, 1.0, 1e-12, step);
        return PM_VIOLATION;
    }
    return PM_OK;
}

// One fused step from buffer `cur` into `cur^1`, then the finalize (audit when step >= 0).
static int hall_sweep(pm_hall *h, double hstep, double vh, int64_t step, double *totals_dev, double offset) {
    const dim3 g = hall_grid(h);
    const int64_t nparts = (int64_t)g.x * g.y * g.z;
    if (h->partial_count < nparts) {
        PM_CHECK(h->partials.alloc(sizeof(double) * (size_t)nparts));
        h->partial_count = nparts;
    }
    const double rc2 = h->density * h->speed * h->speed;
    HallArgs a;
    const int c = h->cur, o = h->cur ^ 1;
    a.p = h->p[c].d(), a.vx = h->vx[c].d(), a.vy = h->vy[c].d(), a.vz = h->vz[c].d();
    a.p_out = h->p[o].d(), a.vx_out = h->vx[o].d(), a.vy_out = h->vy[o].d(), a.vz_out = h->vz[o].d();
    a.partials = h->partials.d();
    a.supplied_step = h->scalars.d() + 2;
    a.audit = h->audit.as<AuditState>();
    a.gain = vh / (h->density * h->dx);
    a.scale = hstep * rc2 / h->dx;
    a.two_rc2 = 2.0 * rc2;
    a.volume = h->dx * h->dx * h->dx;
    a.half_rho = 0.5 * h->density;
    a.src = -1;
    a.source = 0.0;
    a.source_prev = 0.0;
    bool sourced = false;
    if (h->src >= 0 && h->src_cursor <= h->src_count) {
        a.src = h->src;
        if (h->src_cursor < h->src_count) a.source = h->src_host[h->src_cursor];
        if (h->src_cursor > 0) a.source_prev = h->src_host[h->src_cursor - 1];
        sourced = true;
    }
    h->src_cursor++;
    a.nx = (int)h->nx, a.ny = (int)h->ny, a.nz = (int)h->nz;
    if (sourced) {
        hall_zero<<<1, 64, 0, h->stream>>>(h->scalars.d() + 2, 1);
        PM_HIP(hipGetLastError());
    }
    hall_step_kernel<HALL_TY><<<g, dim3(64, HALL_TY), 0, h->stream>>>(a);
    PM_HIP(hipGetLastError());
    hall_finalize<<<1, 1024, 0, h->stream>>>(h->partials.d(), nparts, h->audit.as<AuditState>(), offset,
                                             h->scalars.d() + 1, sourced ? h->scalars.d() + 2 : nullptr, step,
                                             totals_dev, h->scalars.d());
    PM_HIP(hipGetLastError());
    h->cur = o;
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
    h->cur = 0;
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
    return h->dx / (h->speed * std::sqrt(3.0));
}

double pm_hall_courant(const pm_hall *h, double dt) {
    if (!h) return NAN;
    return h->speed * dt * std::sqrt(3.0) / h->dx;
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

int pm_hall_energy(pm_hall *h, double *invariant, double *offset, double *supplied) {
    if (!h) return PM_ERR_ARGUMENT;
    double inv = 0.0;
    PM_CHECK(hall_invariant_now(h, &inv));
    if (invariant) *invariant = inv;
    if (offset) *offset = h->energy_offset;
    if (supplied) *supplied = h->supplied_host;
    return PM_OK;
}

int pm_hall_advance(pm_hall *h, double dt, uint64_t steps, double tolerance, double *totals,
                    double *elapsed_ms, pm_violation *v) {
    if (!h) return PM_ERR_ARGUMENT;
    PM_CHECK(hall_check(h, dt, v, 0));
    if (dt <= 0.0 || steps == 0) return PM_OK;
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
    for (uint64_t n = first; n < steps; n++) PM_CHECK(hall_sweep(h, dt, dt, (int64_t)n, dtot.d(), h->energy_offset));
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
        fill_violation(v, "energy", "simulation", st.before, st.after, st.scale, tolerance, st.breach_step);
        return PM_VIOLATION;
    }
    return PM_OK;
}

int pm_hall_get_fields(pm_hall *h, double *p, double *pp, double *vx, double *vy, double *vz) {
    if (!h) return PM_ERR_ARGUMENT;
    const int c = h->cur;
    if (p) PM_HIP(hipMemcpyAsync(p, h->p[c].d(), sizeof(double) * (size_t)h->nodes(), hipMemcpyDeviceToHost, h->stream));
    if (pp) PM_HIP(hipMemcpyAsync(pp, h->p[c ^ 1].d(), sizeof(double) * (size_t)h->nodes(), hipMemcpyDeviceToHost, h->stream));
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
    const int c = h->cur;
    PM_CHECK(hall_copy(h, h->saved_p, h->p[c], h->nodes()));
    PM_CHECK(hall_copy(h, h->saved_vx, h->vx[c], h->nfx()));
    PM_CHECK(hall_copy(h, h->saved_vy, h->vy[c], h->nfy()));
    PM_CHECK(hall_copy(h, h->saved_vz, h->vz[c], h->nfz()));
    PM_HIP(hipStreamSynchronize(h->stream));
    h->saved_valid = true;
    h->saved_staggered = h->staggered;
    return PM_OK;
}

int pm_hall_restore(pm_hall *h) {
    if (!h) return PM_ERR_ARGUMENT;
    if (!h->saved_valid) return PM_OK;  // `if let Some(..) = self.saved` (hall.rs:547)
    const int c = h->cur;
    PM_CHECK(hall_copy(h, h->p[c], h->saved_p, h->nodes()));
    PM_CHECK(hall_copy(h, h->p[c ^ 1], h->saved_p, h->nodes()));  // pressure_prev = p
    PM_CHECK(hall_copy(h, h->vx[c], h->saved_vx, h->nfx()));
    PM_CHECK(hall_copy(h, h->vy[c], h->saved_vy, h->nfy()));
    PM_CHECK(hall_copy(h, h->vz[c], h->saved_vz, h->nfz()));
    PM_HIP(hipStreamSynchronize(h->stream));
    h->staggered = h->saved_staggered;
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
