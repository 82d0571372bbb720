// Bar1D and Tube on gfx950: the one-dimensional domains of
// crates/pantometry-thermal/src/lib.rs:425-660 (explicit conduction, insulated ends) and
// crates/pantometry-acoustic/src/lib.rs:136-563 (the staggered acoustic leapfrog with closed, open
// or impedance-matched ends).
//
// A line is tens to thousands of cells and its steps are many, so a line lives in one workgroup:
// its arrays in LDS when they fit (LINE_LDS_CELLS cells), in global scratch otherwise, and a
// launch runs `steps` steps with workgroup barriers between them.  Each step's ledger is summed in
// the same launch (per-thread partials in index order, then the fixed-order block sum) and
// audited on the device; a breach stops the launch with the state the breaking step made.  The
// arithmetic of every cell is the reference's, in its order, so the fields are bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace pm {

static constexpr int LINE_THREADS = 1024;
static constexpr int LINE_LDS_CELLS = 2048;
static constexpr int LINE_MAX_STEPS = 1 << 16;  // steps per launch

__global__ void line_audit_reset(AuditState *st, double total, double scale, double tol) {
    st->prev_total = total, st->prev_scale = scale, st->tolerance = tol;
    st->halted = 0, st->breach_step = -1;
    st->before = st->after = st->scale = 0.0;
}

// ---- Bar1D ------------------------------------------------------------------------------------

struct BarArgs {
    double *cells;    // in and out
    double *scratch;  // n doubles when the line does not fit in LDS
    const double *joules;  // per-cell joules of the launch's first step, or null
    double gained;         // bus heat into cell 0 on the launch's first step
    double f, capacity, reference;
    AuditState *st;
    double *totals;  // ledger total after each step, indexed from step0
    int64_t step0;
    int n, steps;
};

// Bar1D::step after its guard (lib.rs:584-626) and Bar1D::stored_heat (lib.rs:535-538).
template <bool LDS>
__global__ __launch_bounds__(LINE_THREADS) void bar_run(BarArgs a) {
    __shared__ double lds[LDS ? 2 * LINE_LDS_CELLS : 1];
    __shared__ double red[LINE_THREADS / 64];
    __shared__ int halt;
    const int n = a.n, last = n - 1, t = threadIdx.x;
    if (a.st && a.st->halted) return;  // a breach earlier in the run froze the state
    double *x = LDS ? lds : a.cells;
    double *y = LDS ? lds + LINE_LDS_CELLS : a.scratch;
    for (int i = t; i < n; i += LINE_THREADS) {
        double v = a.cells[i];
        // The bus share lands in cell 0, then every face's joules in its cell (lib.rs:589-608).
        if (i == 0) v += a.gained / a.capacity;
        if (a.joules) v += a.joules[i] / a.capacity;
        x[i] = v;
    }
    __syncthreads();
    for (int s = 0; s < a.steps; s++) {
        for (int i = t; i < n; i += LINE_THREADS) {
            const double left = x[i == 0 ? 0 : i - 1];
            const double right = x[i < last ? i + 1 : last];
            y[i] = x[i] + a.f * (left - 2.0 * x[i] + right);
        }
        __syncthreads();
        double *const w = x;
        x = y, y = w;
        double acc = 0.0;
        for (int i = t; i < n; i += LINE_THREADS) acc += x[i] - a.reference;
        const double sum = block_sum<LINE_THREADS>(acc, red);
        if (t == 0) {
            const double stored = a.capacity * sum;
            if (a.st) audit_step(a.st, stored, std::fabs(stored), a.step0 + s, a.totals);
            halt = a.st ? a.st->halted : 0;
        }
        __syncthreads();
        if (halt) break;
    }
    if (x != a.cells)
        for (int i = t; i < n; i += LINE_THREADS) a.cells[i] = x[i];
}

// ---- Tube -------------------------------------------------------------------------------------

struct TubeArgs {
    double *p, *pp, *u;  // in and out: pressure, pressure_prev, face velocity
    double cv_first, cv;  // vh/(rho dx): the launch's first step, the rest
    double cp;            // dt rc2/dx
    double zl, zr;        // each end's impedance for the wall velocity
    int closed_l, closed_r;
    double area, dt, two_rc2, volume, density, offset;
    double *radiated;  // running total (device scalar)
    double *first;     // [0] joules absorbed, [1] invariant, after the launch's first step
    AuditState *st;
    double *totals;
    int64_t step0;
    int n, steps;
};

// Tube::wall_velocity (lib.rs:371-385) with the impedance resolved on the host: (outward p)/Z.
__device__ __forceinline__ double wall_velocity(int closed, double z, double p, double outward) {
    return closed ? 0.0 : outward * p / z;
}

// Tube::advance's sweeps and wall power (lib.rs:389-446) and Tube::invariant_si (lib.rs:300-327).
template <bool LDS>
__global__ __launch_bounds__(LINE_THREADS) void tube_run(TubeArgs a) {
    __shared__ double lds[LDS ? 3 * LINE_LDS_CELLS : 1];
    __shared__ double red[LINE_THREADS / 64];
    __shared__ int halt;
    const int n = a.n, last = n - 1, t = threadIdx.x;
    if (a.st && a.st->halted) return;
    double *p = LDS ? lds : a.p;
    double *q = LDS ? lds + LINE_LDS_CELLS : a.pp;
    double *u = LDS ? lds + 2 * LINE_LDS_CELLS : a.u;
    if (LDS) {
        for (int i = t; i < n; i += LINE_THREADS) p[i] = a.p[i];
        for (int i = t; i < last; i += LINE_THREADS) u[i] = a.u[i];
    }
    __syncthreads();
    for (int s = 0; s < a.steps; s++) {
        const double cv = s == 0 ? a.cv_first : a.cv;
        for (int i = t; i < last; i += LINE_THREADS) u[i] -= cv * (p[i + 1] - p[i]);
        __syncthreads();
        const double lw = wall_velocity(a.closed_l, a.zl, p[0], -1.0);
        const double rw = wall_velocity(a.closed_r, a.zr, p[last], 1.0);
        for (int i = t; i < n; i += LINE_THREADS) {
            const double inflow = i == 0 ? lw : u[i - 1];
            const double outflow = i == last ? rw : u[i];
            const double share = (i == 0 || i == last) ? 2.0 : 1.0;
            q[i] = p[i] - a.cp * share * (outflow - inflow);
        }
        __syncthreads();
        double *const w = p;  // q is now pressure_prev
        p = q, q = w;
        double pot = 0.0, kin = 0.0;
        for (int i = t; i < n; i += LINE_THREADS) {
            const double share = (i == 0 || i == last) ? 0.5 : 1.0;
            pot += share * p[i] * q[i] / a.two_rc2 * a.volume;
        }
        for (int i = t; i < last; i += LINE_THREADS) kin += a.density * u[i] * u[i] / 2.0 * a.volume;
        const double potential = block_sum<LINE_THREADS>(pot, red);
        __syncthreads();
        const double kinetic = block_sum<LINE_THREADS>(kin, red);
        if (t == 0) {
            double absorbed = 0.0;
            absorbed += std::fabs(p[0] * lw) * a.area * a.dt;
            absorbed += std::fabs(p[last] * rw) * a.area * a.dt;
            double radiated = *a.radiated;
            if (absorbed > 0.0) radiated += absorbed;
            *a.radiated = radiated;
            const double inv = a.two_rc2 > 0.0 ? potential + kinetic : 0.0;
            if (s == 0 && a.first) a.first[0] = absorbed, a.first[1] = inv;
            // The tube's books against what its ends put on the bus: energy, plus what has left.
            const double energy = inv + a.offset;
            double total = 0.0;
            total += energy;
            if (radiated != 0.0) total += radiated;
            if (a.st) audit_step(a.st, total, rmax(std::fabs(energy), std::fabs(radiated)), a.step0 + s, a.totals);
            halt = a.st ? a.st->halted : 0;
        }
        __syncthreads();
        if (halt) break;
    }
    if (LDS) {
        for (int i = t; i < n; i += LINE_THREADS) a.p[i] = p[i], a.pp[i] = q[i];
        for (int i = t; i < last; i += LINE_THREADS) a.u[i] = u[i];
    } else if (p != a.p) {
        for (int i = t; i < n; i += LINE_THREADS) {
            const double cur = p[i];
            a.pp[i] = q[i];
            a.p[i] = cur;
        }
    }
}

// Tube::invariant_si alone, for a released state.
__global__ __launch_bounds__(LINE_THREADS) void tube_invariant(const double *p, const double *pp,
                                                               const double *u, int n, double two_rc2,
                                                               double volume, double density,
                                                               double *out) {
    __shared__ double red[LINE_THREADS / 64];
    const int last = n - 1, t = threadIdx.x;
    double pot = 0.0, kin = 0.0;
    for (int i = t; i < n; i += LINE_THREADS) {
        const double share = (i == 0 || i == last) ? 0.5 : 1.0;
        pot += share * p[i] * pp[i] / two_rc2 * volume;
    }
    for (int i = t; i < last; i += LINE_THREADS) kin += density * u[i] * u[i] / 2.0 * volume;
    const double potential = block_sum<LINE_THREADS>(pot, red);
    __syncthreads();
    const double kinetic = block_sum<LINE_THREADS>(kin, red);
    if (t == 0) *out = two_rc2 > 0.0 ? potential + kinetic : 0.0;
}

// ScalarField::at for Bar1D, lib.rs:736-751: linear between cell centres, held past the ends
// (cells[0] for a NaN position).
__global__ __launch_bounds__(256) void bar_sample(const double *cells, int n, double dx,
                                                  const double *xs, int64_t count, double *out) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int last = n - 1;
        const double u = xs[k] / dx - 0.5;
        double v;
        if (isnan(u) || u <= 0.0) {
            v = cells[0];
        } else if (u >= (double)last) {
            v = cells[last];
        } else {
            const int i = (int)floor(u);
            const double f = u - (double)i;
            v = cells[i] * (1.0 - f) + cells[i + 1] * f;
        }
        out[k] = v;
    }
}

}  // namespace pm

using namespace pm;

struct pm_bar1d {
    std::string name;
    int64_t n = 0;
    double dx = 0.0, area = 0.0, alpha = 0.0, capacity = 0.0, reference = 0.0, absorbed = 0.0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf cells, scratch, saved, audit, totals;
    bool fits() const { return n <= LINE_LDS_CELLS; }
};

struct pm_tube {
    std::string name;
    int64_t n = 0;
    double dx = 0.0, area = 0.0, speed = 0.0, density = 0.0;
    int left = 0, right = 0;        // 0 closed, 1 open, 2 matched
    double z_left = 0.0, z_right = 0.0;
    bool staggered = false;
    double offset = 0.0, datum = 0.0, radiated_host = 0.0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf p, pp, u, saved_p, saved_u, audit, scalars;  // scalars: [0] radiated, [1..2] first step
    bool saved_valid = false, saved_staggered = false;
    bool fits() const { return n <= LINE_LDS_CELLS; }
    double impedance() const { return density * speed; }
    double rc2() const { return density * speed * speed; }
};

namespace pm {

// ---- Bar1D host side ---------------------------------------------------------------------------

static double bar_fourier(const pm_bar1d *b, double dt) {
    if (!std::isfinite(b->alpha)) return INFINITY;
    return b->alpha * dt / (b->dx * b->dx);
}

// The guard of Bar1D::step, lib.rs:571-587.
static int bar_guard(const pm_bar1d *b, double dt, pm_violation *v) {
    const double f = bar_fourier(b, dt);
    if (!std::isfinite(f)) {
        violation_at(v, b->name.c_str(), "substance has no diffusivity", f, -1);
        return PM_VIOLATION;
    }
    if (f > 0.5 + 1e-12) {
        std::string site = b->name + " (explicit conduction)";
        fill_violation(v, "Fourier number", site.c_str(), 0.5, f, 0.5, 1e-12, -1);
        return PM_VIOLATION;
    }
    return PM_OK;
}

static int bar_launch(pm_bar1d *b, double f, int steps, double gained, const double *joules,
                      AuditState *st, double *totals, int64_t step0) {
    BarArgs a;
    a.cells = b->cells.d(), a.scratch = b->scratch.d(), a.joules = joules, a.gained = gained;
    a.f = f, a.capacity = b->capacity, a.reference = b->reference;
    a.st = st, a.totals = totals, a.step0 = step0, a.n = (int)b->n, a.steps = steps;
    if (b->fits())
        bar_run<true><<<1, LINE_THREADS, 0, b->stream>>>(a);
    else
        bar_run<false><<<1, LINE_THREADS, 0, b->stream>>>(a);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

static int bar_stored(pm_bar1d *b, double *out) {
    std::vector<double> c((size_t)b->n);
    PM_HIP(hipMemcpyAsync(c.data(), b->cells.d(), sizeof(double) * c.size(), hipMemcpyDeviceToHost, b->stream));
    PM_HIP(hipStreamSynchronize(b->stream));
    double s = 0.0;
    for (double t : c) s += t - b->reference;
    *out = b->capacity * s;
    return PM_OK;
}

// ---- Tube host side ----------------------------------------------------------------------------

static double tube_z(const pm_tube *t, int end, double z) {
    if (end == 1 || z <= 0.0) return t->impedance();
    return z;
}

static int tube_invariant_now(pm_tube *t, double *out) {
    DevBuf d;
    PM_CHECK(d.alloc(sizeof(double)));
    tube_invariant<<<1, LINE_THREADS, 0, t->stream>>>(t->p.d(), t->pp.d(), t->u.d(), (int)t->n,
                                                      2.0 * t->rc2(), t->area * t->dx, t->density, d.d());
    PM_HIP(hipGetLastError());
    PM_HIP(hipMemcpyAsync(out, d.d(), sizeof(double), hipMemcpyDeviceToHost, t->stream));
    PM_HIP(hipStreamSynchronize(t->stream));
    return PM_OK;
}

static int tube_launch(pm_tube *t, double dt, int steps, bool first_is_half, AuditState *st,
                       double *totals, int64_t step0, double *first) {
    TubeArgs a;
    a.p = t->p.d(), a.pp = t->pp.d(), a.u = t->u.d();
    const double vh = first_is_half ? 0.5 * dt : dt;
    a.cv_first = vh / (t->density * t->dx);
    a.cv = dt / (t->density * t->dx);
    a.cp = dt * t->rc2() / t->dx;
    a.zl = tube_z(t, t->left, t->z_left), a.zr = tube_z(t, t->right, t->z_right);
    a.closed_l = t->left == 0, a.closed_r = t->right == 0;
    a.area = t->area, a.dt = dt, a.two_rc2 = 2.0 * t->rc2(), a.volume = t->area * t->dx;
    a.density = t->density, a.offset = t->offset;
    a.radiated = t->scalars.d(), a.first = first;
    a.st = st, a.totals = totals, a.step0 = step0, a.n = (int)t->n, a.steps = steps;
    if (t->fits())
        tube_run<true><<<1, LINE_THREADS, 0, t->stream>>>(a);
    else
        tube_run<false><<<1, LINE_THREADS, 0, t->stream>>>(a);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

// Domain::step for Tube, lib.rs:516-533 with Tube::advance's startup (lib.rs:449-463): one step,
// the joules its ends absorbed returned for the bus.
static int tube_one(pm_tube *t, double dt, double *published) {
    const bool starting = !t->staggered;
    t->staggered = true;
    double first[2] = {0.0, 0.0};
    PM_CHECK(tube_launch(t, dt, 1, starting, nullptr, nullptr, 0, t->scalars.d() + 1));
    PM_HIP(hipMemcpyAsync(first, t->scalars.d() + 1, sizeof first, hipMemcpyDeviceToHost, t->stream));
    double radiated = 0.0;
    PM_HIP(hipMemcpyAsync(&radiated, t->scalars.d(), sizeof(double), hipMemcpyDeviceToHost, t->stream));
    PM_HIP(hipStreamSynchronize(t->stream));
    t->radiated_host = radiated;
    if (published) *published = first[0] > 0.0 ? first[0] : 0.0;
    if (starting) {
        const double inv = first[1];
        const double off = t->datum - inv;
        const double scale = rmax(std::fabs(t->datum), std::fabs(inv));
        if (scale > 0.0 && std::fabs(off) / scale <= 0.25) t->offset = off;
    }
    return PM_OK;
}

}  // namespace pm

extern "C" {

// ---- Bar1D ------------------------------------------------------------------------------------

int pm_bar1d_create(const char *name, const pm_substance *s, size_t cells, double dx, double area,
                    double initial, pm_bar1d **out) {
    if (!out || !s) return PM_ERR_ARGUMENT;
    *out = nullptr;
    PM_CHECK(require_device());
    std::unique_ptr<pm_bar1d> b(new pm_bar1d());
    b->name = name ? name : "bar";
    b->n = (int64_t)std::max<size_t>(cells, 2);  // Bar1D::new, lib.rs:451
    if (b->n > (1 << 30)) {
        set_error("pm_bar1d_create: %zu cells is beyond the kernel's indexing", cells);
        return PM_ERR_ARGUMENT;
    }
    b->dx = dx, b->area = area, b->reference = initial;
    // Substance::diffusivity and heat_capacity(area dx) (substance.rs:211-222); a substance
    // without thermal data has no diffusivity and an infinite cell capacity (lib.rs:527-533).
    const bool thermal = std::isfinite(s->conductivity) && std::isfinite(s->specific_heat);
    b->alpha = thermal ? s->conductivity / (s->density * s->specific_heat) : NAN;
    b->capacity = std::isfinite(s->specific_heat) ? (s->density * (area * dx)) * s->specific_heat : INFINITY;
    PM_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    PM_HIP(hipEventCreate(&b->ev0));
    PM_HIP(hipEventCreate(&b->ev1));
    const size_t bytes = sizeof(double) * (size_t)b->n;
    PM_CHECK(b->cells.alloc(bytes));
    PM_CHECK(b->saved.alloc(bytes));
    if (!b->fits()) PM_CHECK(b->scratch.alloc(bytes));
    PM_CHECK(b->audit.alloc(sizeof(AuditState)));
    std::vector<double> init((size_t)b->n, initial);
    PM_HIP(hipMemcpyAsync(b->cells.d(), init.data(), bytes, hipMemcpyHostToDevice, b->stream));
    PM_HIP(hipMemcpyAsync(b->saved.d(), init.data(), bytes, hipMemcpyHostToDevice, b->stream));
    PM_HIP(hipStreamSynchronize(b->stream));
    *out = b.release();
    return PM_OK;
}

void pm_bar1d_destroy(pm_bar1d *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    hipStream_t st = b->stream;
    delete b;
    if (st) (void)hipStreamDestroy(st);
}

size_t pm_bar1d_cells(const pm_bar1d *b) { return b ? (size_t)b->n : 0; }

int pm_bar1d_set_cells(pm_bar1d *b, const double *cells) {
    if (!b || !cells) return PM_ERR_ARGUMENT;
    PM_HIP(hipMemcpyAsync(b->cells.d(), cells, sizeof(double) * (size_t)b->n, hipMemcpyHostToDevice, b->stream));
    PM_HIP(hipStreamSynchronize(b->stream));
    return PM_OK;
}

int pm_bar1d_get_cells(pm_bar1d *b, double *cells) {
    if (!b || !cells) return PM_ERR_ARGUMENT;
    PM_HIP(hipMemcpyAsync(cells, b->cells.d(), sizeof(double) * (size_t)b->n, hipMemcpyDeviceToHost, b->stream));
    PM_HIP(hipStreamSynchronize(b->stream));
    return PM_OK;
}

double pm_bar1d_fourier_number(const pm_bar1d *b, double dt) { return b ? bar_fourier(b, dt) : NAN; }

double pm_bar1d_max_stable_dt(const pm_bar1d *b) {
    if (!b) return NAN;
    if (!std::isfinite(b->alpha)) return INFINITY;
    return b->dx * b->dx / (2.0 * b->alpha);
}

int pm_bar1d_refuses(const pm_bar1d *b, double dt, pm_violation *v) {
    if (!b) return PM_ERR_ARGUMENT;
    return bar_guard(b, dt, v);
}

int pm_bar1d_step(pm_bar1d *b, double dt, double gained, const double *joules, pm_violation *v) {
    if (!b) return PM_ERR_ARGUMENT;
    PM_CHECK(bar_guard(b, dt, v));
    DevBuf dj;
    if (joules) {
        PM_CHECK(dj.alloc(sizeof(double) * (size_t)b->n));
        PM_HIP(hipMemcpyAsync(dj.d(), joules, sizeof(double) * (size_t)b->n, hipMemcpyHostToDevice, b->stream));
    }
    PM_CHECK(bar_launch(b, bar_fourier(b, dt), 1, gained, joules ? dj.d() : nullptr, nullptr, nullptr, 0));
    PM_HIP(hipStreamSynchronize(b->stream));
    b->absorbed += gained;
    if (joules) {
        double total = 0.0;  // Flux::total, crates/pantometry-core/src/scene.rs:221-223
        for (int64_t i = 0; i < b->n; i++) total += joules[i];
        b->absorbed += total;
    }
    return PM_OK;
}

int pm_bar1d_ledger(pm_bar1d *b, double *stored, double *absorbed) {
    if (!b) return PM_ERR_ARGUMENT;
    if (stored) PM_CHECK(bar_stored(b, stored));
    if (absorbed) *absorbed = b->absorbed;
    return PM_OK;
}

int pm_bar1d_advance(pm_bar1d *b, double dt, uint64_t steps, double tolerance, double *totals,
                     double *elapsed_ms, pm_violation *v) {
    if (!b) return PM_ERR_ARGUMENT;
    PM_CHECK(bar_guard(b, dt, v));
    if (steps == 0) return PM_OK;
    double stored0 = 0.0;
    PM_CHECK(bar_stored(b, &stored0));
    hipLaunchKernelGGL(line_audit_reset, dim3(1), dim3(1), 0, b->stream, b->audit.as<AuditState>(), stored0,
                       std::fabs(stored0), tolerance);
    PM_HIP(hipGetLastError());
    PM_CHECK(b->totals.alloc(sizeof(double) * (size_t)steps));
    PM_HIP(hipEventRecord(b->ev0, b->stream));
    const double f = bar_fourier(b, dt);
    for (uint64_t s = 0; s < steps; s += LINE_MAX_STEPS) {
        const int k = (int)std::min<uint64_t>(LINE_MAX_STEPS, steps - s);
        PM_CHECK(bar_launch(b, f, k, 0.0, nullptr, b->audit.as<AuditState>(), b->totals.d(), (int64_t)s));
    }
    PM_HIP(hipEventRecord(b->ev1, b->stream));
    AuditState st;
    PM_HIP(hipMemcpyAsync(&st, b->audit.d(), sizeof st, hipMemcpyDeviceToHost, b->stream));
    PM_HIP(hipStreamSynchronize(b->stream));
    const uint64_t done = st.halted ? (uint64_t)st.breach_step + 1 : steps;
    if (totals) {
        PM_HIP(hipMemcpy(totals, b->totals.d(), sizeof(double) * done, hipMemcpyDeviceToHost));
    }
    if (elapsed_ms) {
        float ms = 0.0f;
        PM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        *elapsed_ms = ms;
    }
    if (st.halted) {
        fill_violation(v, "energy", "simulation", st.before, st.after, st.scale, tolerance, st.breach_step);
        return PM_VIOLATION;
    }
    return PM_OK;
}

int pm_bar1d_sample(pm_bar1d *b, const double *x, size_t count, double *out) {
    if (!b || (count && (!x || !out))) return PM_ERR_ARGUMENT;
    if (count == 0) return PM_OK;
    DevBuf dx, dv;
    PM_CHECK(dx.alloc(sizeof(double) * count));
    PM_CHECK(dv.alloc(sizeof(double) * count));
    PM_HIP(hipMemcpyAsync(dx.d(), x, sizeof(double) * count, hipMemcpyHostToDevice, b->stream));
    bar_sample<<<(unsigned)std::min<size_t>((count + 255) / 256, 4096), 256, 0, b->stream>>>(
        b->cells.d(), (int)b->n, b->dx, dx.d(), (int64_t)count, dv.d());
    PM_HIP(hipGetLastError());
    PM_HIP(hipMemcpyAsync(out, dv.d(), sizeof(double) * count, hipMemcpyDeviceToHost, b->stream));
    PM_HIP(hipStreamSynchronize(b->stream));
    return PM_OK;
}

int pm_bar1d_readings(pm_bar1d *b, double out[3]) {
    if (!b || !out) return PM_ERR_ARGUMENT;
    std::vector<double> c((size_t)b->n);
    PM_HIP(hipMemcpyAsync(c.data(), b->cells.d(), sizeof(double) * c.size(), hipMemcpyDeviceToHost, b->stream));
    PM_HIP(hipStreamSynchronize(b->stream));
    double sum = 0.0, peak = -1.7976931348623157e308;  // f64::MIN
    for (double t : c) sum += t, peak = rmax(peak, t);
    out[0] = sum / (double)b->n;  // mean_temperature, lib.rs:504-506
    out[1] = peak;
    out[2] = b->absorbed;
    return PM_OK;
}

int pm_bar1d_checkpoint(pm_bar1d *b) {
    if (!b) return PM_ERR_ARGUMENT;
    PM_HIP(hipMemcpyAsync(b->saved.d(), b->cells.d(), sizeof(double) * (size_t)b->n, hipMemcpyDeviceToDevice, b->stream));
    PM_HIP(hipStreamSynchronize(b->stream));
    return PM_OK;
}

int pm_bar1d_restore(pm_bar1d *b) {
    if (!b) return PM_ERR_ARGUMENT;
    PM_HIP(hipMemcpyAsync(b->cells.d(), b->saved.d(), sizeof(double) * (size_t)b->n, hipMemcpyDeviceToDevice, b->stream));
    PM_HIP(hipStreamSynchronize(b->stream));
    return PM_OK;
}

// ---- Tube -------------------------------------------------------------------------------------

int pm_tube_create(const char *name, double length, size_t cells, double area, double density,
                   double speed, pm_tube **out) {
    if (!out) return PM_ERR_ARGUMENT;
    *out = nullptr;
    PM_CHECK(require_device());
    std::unique_ptr<pm_tube> t(new pm_tube());
    t->name = name ? name : "tube";
    t->n = (int64_t)std::max<size_t>(cells, 3);  // Tube::new, lib.rs:175
    if (t->n > (1 << 30)) {
        set_error("pm_tube_create: %zu cells is beyond the kernel's indexing", cells);
        return PM_ERR_ARGUMENT;
    }
    t->dx = length / (double)(t->n - 1);
    t->area = area, t->density = density, t->speed = speed;
    PM_HIP(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    PM_HIP(hipEventCreate(&t->ev0));
    PM_HIP(hipEventCreate(&t->ev1));
    const size_t bytes = sizeof(double) * (size_t)t->n;
    PM_CHECK(t->p.alloc(bytes));
    PM_CHECK(t->pp.alloc(bytes));
    PM_CHECK(t->u.alloc(bytes - sizeof(double)));
    PM_CHECK(t->audit.alloc(sizeof(AuditState)));
    PM_CHECK(t->scalars.alloc(sizeof(double) * 4));
    PM_HIP(hipMemsetAsync(t->p.d(), 0, bytes, t->stream));
    PM_HIP(hipMemsetAsync(t->pp.d(), 0, bytes, t->stream));
    PM_HIP(hipMemsetAsync(t->u.d(), 0, bytes - sizeof(double), t->stream));
    PM_HIP(hipMemsetAsync(t->scalars.d(), 0, sizeof(double) * 4, t->stream));
    PM_HIP(hipStreamSynchronize(t->stream));
    *out = t.release();
    return PM_OK;
}

void pm_tube_destroy(pm_tube *t) {
    if (!t) return;
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    if (t->ev0) (void)hipEventDestroy(t->ev0);
    if (t->ev1) (void)hipEventDestroy(t->ev1);
    hipStream_t st = t->stream;
    delete t;
    if (st) (void)hipStreamDestroy(st);
}

size_t pm_tube_cells(const pm_tube *t) { return t ? (size_t)t->n : 0; }
double pm_tube_spacing(const pm_tube *t) { return t ? t->dx : NAN; }

int pm_tube_set_ends(pm_tube *t, int left, double z_left, int right, double z_right) {
    if (!t || left < 0 || left > 2 || right < 0 || right > 2) return PM_ERR_ARGUMENT;
    t->left = left, t->z_left = z_left, t->right = right, t->z_right = z_right;
    return PM_OK;
}

int pm_tube_release(pm_tube *t, const double *pressure) {
    if (!t || !pressure) return PM_ERR_ARGUMENT;
    const size_t bytes = sizeof(double) * (size_t)t->n;
    PM_HIP(hipMemcpyAsync(t->p.d(), pressure, bytes, hipMemcpyHostToDevice, t->stream));
    PM_HIP(hipMemcpyAsync(t->pp.d(), pressure, bytes, hipMemcpyHostToDevice, t->stream));
    PM_HIP(hipMemsetAsync(t->u.d(), 0, bytes - sizeof(double), t->stream));
    t->staggered = false;
    t->offset = 0.0;
    PM_CHECK(tube_invariant_now(t, &t->datum));
    return PM_OK;
}

double pm_tube_max_stable_dt(const pm_tube *t) {
    if (!t) return NAN;
    if (t->speed <= 0.0) return INFINITY;
    // Domain::max_stable_dt for Tube, lib.rs:486-514: the wave's dx/c, and an absorbing end's
    // dx Z/(2 rho c^2).
    const double wave = t->dx / t->speed;
    const double rc2 = t->rc2();
    auto drain = [&](int end, double z) {
        if (end == 0) return (double)INFINITY;
        if (end == 1) return t->dx * t->impedance() / (2.0 * rc2);
        const double zz = z > 0.0 ? z : t->impedance();
        return t->dx * zz / (2.0 * rc2);
    };
    return std::fmin(std::fmin(wave, drain(t->left, t->z_left)), drain(t->right, t->z_right));
}

double pm_tube_courant(const pm_tube *t, double dt) { return t ? t->speed * dt / t->dx : NAN; }

static int tube_guard(const pm_tube *t, double dt, pm_violation *v, int64_t step) {
    const double courant = t->speed * dt / t->dx;
    if (courant > 1.0 + 1e-12) {
        std::string site = t->name + " (explicit wave equation)";
        fill_violation(v, "Courant number", site.c_str(), 1.0, courant, 1.0, 1e-12, step);
        return PM_VIOLATION;
    }
    return PM_OK;
}

int pm_tube_step(pm_tube *t, double dt, double *published, pm_violation *v) {
    if (!t) return PM_ERR_ARGUMENT;
    if (published) *published = 0.0;
    PM_CHECK(tube_guard(t, dt, v, -1));
    if (dt <= 0.0) return PM_OK;
    return tube_one(t, dt, published);
}

int pm_tube_energy(pm_tube *t, double *invariant, double *offset, double *radiated) {
    if (!t) return PM_ERR_ARGUMENT;
    if (invariant) PM_CHECK(tube_invariant_now(t, invariant));
    if (offset) *offset = t->offset;
    if (radiated) *radiated = t->radiated_host;
    return PM_OK;
}

int pm_tube_advance(pm_tube *t, double dt, uint64_t steps, double tolerance, double *totals,
                    double *elapsed_ms, pm_violation *v) {
    if (!t) return PM_ERR_ARGUMENT;
    PM_CHECK(tube_guard(t, dt, v, 0));
    if (dt <= 0.0 || steps == 0) return PM_OK;
    PM_HIP(hipEventRecord(t->ev0, t->stream));
    uint64_t first = 0;
    if (!t->staggered) {
        // The half step and the startup offset run alone; the audit of that step against the
        // released books holds by construction of the offset.
        PM_CHECK(tube_one(t, dt, nullptr));
        first = 1;
    }
    double inv = 0.0;
    PM_CHECK(tube_invariant_now(t, &inv));
    const double energy = inv + t->offset;
    double total0 = 0.0;
    total0 += energy;
    if (t->radiated_host != 0.0) total0 += t->radiated_host;
    if (first == 1 && totals) totals[0] = total0;
    hipLaunchKernelGGL(line_audit_reset, dim3(1), dim3(1), 0, t->stream, t->audit.as<AuditState>(), total0,
                       rmax(std::fabs(energy), std::fabs(t->radiated_host)), tolerance);
    PM_HIP(hipGetLastError());
    DevBuf dtot;
    PM_CHECK(dtot.alloc(sizeof(double) * (size_t)steps));
    for (uint64_t s = first; s < steps; s += LINE_MAX_STEPS) {
        const int k = (int)std::min<uint64_t>(LINE_MAX_STEPS, steps - s);
        PM_CHECK(tube_launch(t, dt, k, false, t->audit.as<AuditState>(), dtot.d(), (int64_t)s, nullptr));
    }
    PM_HIP(hipEventRecord(t->ev1, t->stream));
    AuditState st;
    PM_HIP(hipMemcpyAsync(&st, t->audit.d(), sizeof st, hipMemcpyDeviceToHost, t->stream));
    double radiated = 0.0;
    PM_HIP(hipMemcpyAsync(&radiated, t->scalars.d(), sizeof(double), hipMemcpyDeviceToHost, t->stream));
    PM_HIP(hipStreamSynchronize(t->stream));
    t->radiated_host = radiated;
    const uint64_t done = st.halted ? (uint64_t)st.breach_step + 1 : steps;
    if (totals && done > first)
        PM_HIP(hipMemcpy(totals + first, dtot.d() + first, sizeof(double) * (done - first), hipMemcpyDeviceToHost));
    if (elapsed_ms) {
        float ms = 0.0f;
        PM_HIP(hipEventElapsedTime(&ms, t->ev0, t->ev1));
        *elapsed_ms = ms;
    }
    if (st.halted) {
        fill_violation(v, "energy", "simulation", st.before, st.after, st.scale, tolerance, st.breach_step);
        return PM_VIOLATION;
    }
    return PM_OK;
}

int pm_tube_get_fields(pm_tube *t, double *p, double *pp, double *u) {
    if (!t) return PM_ERR_ARGUMENT;
    const size_t bytes = sizeof(double) * (size_t)t->n;
    if (p) PM_HIP(hipMemcpyAsync(p, t->p.d(), bytes, hipMemcpyDeviceToHost, t->stream));
    if (pp) PM_HIP(hipMemcpyAsync(pp, t->pp.d(), bytes, hipMemcpyDeviceToHost, t->stream));
    if (u) PM_HIP(hipMemcpyAsync(u, t->u.d(), bytes - sizeof(double), hipMemcpyDeviceToHost, t->stream));
    PM_HIP(hipStreamSynchronize(t->stream));
    return PM_OK;
}

int pm_tube_checkpoint(pm_tube *t) {
    if (!t) return PM_ERR_ARGUMENT;
    const size_t bytes = sizeof(double) * (size_t)t->n;
    if (!t->saved_p) {
        PM_CHECK(t->saved_p.alloc(bytes));
        PM_CHECK(t->saved_u.alloc(bytes - sizeof(double)));
    }
    PM_HIP(hipMemcpyAsync(t->saved_p.d(), t->p.d(), bytes, hipMemcpyDeviceToDevice, t->stream));
    PM_HIP(hipMemcpyAsync(t->saved_u.d(), t->u.d(), bytes - sizeof(double), hipMemcpyDeviceToDevice, t->stream));
    PM_HIP(hipStreamSynchronize(t->stream));
    t->saved_valid = true;
    t->saved_staggered = t->staggered;
    return PM_OK;
}

int pm_tube_restore(pm_tube *t) {
    if (!t) return PM_ERR_ARGUMENT;
    if (!t->saved_valid) return PM_OK;  // `if let Some(..) = self.saved` (lib.rs:546-553)
    const size_t bytes = sizeof(double) * (size_t)t->n;
    PM_HIP(hipMemcpyAsync(t->p.d(), t->saved_p.d(), bytes, hipMemcpyDeviceToDevice, t->stream));
    PM_HIP(hipMemcpyAsync(t->pp.d(), t->saved_p.d(), bytes, hipMemcpyDeviceToDevice, t->stream));
    PM_HIP(hipMemcpyAsync(t->u.d(), t->saved_u.d(), bytes - sizeof(double), hipMemcpyDeviceToDevice, t->stream));
    PM_HIP(hipStreamSynchronize(t->stream));
    t->staggered = t->saved_staggered;
    return PM_OK;
}

}  // extern "C"
