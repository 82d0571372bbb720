/*
 * CPU oracle for the explicit 3D grid-stencil path of pantometry.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C99 and IEEE binary64 with
 * round-to-nearest, the arithmetic of the reference's Rust implementation, in the reference's
 * own evaluation order: Rust never contracts `a*b+c` into an FMA, never reassociates, and
 * evaluates `a * b * c` as `(a*b)*c`.  It is compiled with `-O2 -ffp-contract=off` and without
 * `-ffast-math`.  Only tests, the smoke check and the benchmark's CPU-baseline leg load it; the
 * product library never links or calls anything here.
 *
 * Every function cites the reference file:line it follows.  Which test pins which function is
 * listed in oracle/README.md.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------ */
/* Rng: crates/pantometry-core/src/rng.rs:37-152                                               */
/* ------------------------------------------------------------------------------------------ */

/* SplitMix64 finaliser, rng.rs:37-42. */
uint64_t orc_mix64(uint64_t z) {
    z = z + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Rng::new, rng.rs:49-55: a zero seed is replaced. */
uint64_t orc_rng_new(uint64_t seed) { return seed == 0 ? 0x9E3779B97F4A7C15ULL : seed; }

/* Rng::for_index, rng.rs:66-68. */
uint64_t orc_rng_for_index(uint64_t seed, uint64_t index) {
    return orc_rng_new(orc_mix64(seed ^ orc_mix64(index)));
}

/* Rng::next_u64, rng.rs:81-88: xorshift64* . */
uint64_t orc_rng_next_u64(uint64_t *s) {
    uint64_t x = *s;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *s = x;
    return x * 0x2545F4914F6CDD1DULL;
}

/* Rng::unit, rng.rs:91-93. */
double orc_rng_unit(uint64_t *s) {
    return (double)(orc_rng_next_u64(s) >> 11) / (double)(1ULL << 53);
}

/* Rng::range, rng.rs:96-98. */
double orc_rng_range(uint64_t *s, double lo, double hi) { return lo + (hi - lo) * orc_rng_unit(s); }

/* Rng::gaussian, rng.rs:146-151 (Box-Muller through libm log, sqrt and cos). */
double orc_rng_gaussian(uint64_t *s) {
    double u1 = orc_rng_unit(s);
    if (!(u1 >= 2.2250738585072014e-308)) u1 = 2.2250738585072014e-308; /* f64::max(MIN_POSITIVE) */
    double u2 = orc_rng_unit(s);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Rng::split, rng.rs:75-78. */
uint64_t orc_rng_split(uint64_t *s) { return orc_rng_new(orc_mix64(orc_rng_next_u64(s))); }

/* The order-sensitive digest of rng.rs:295-305 (`the_stream_is_pinned`). */
uint64_t orc_rng_digest(uint64_t seed, int draws) {
    uint64_t s = orc_rng_new(seed);
    uint64_t hash = 0;
    for (int n = 0; n < draws; n++) {
        uint64_t rot = (hash << 7) | (hash >> 57);
        hash = (rot * 0x1000000001B3ULL) ^ (uint64_t)(orc_rng_unit(&s) * (double)(1ULL << 53));
    }
    return hash;
}

/* ------------------------------------------------------------------------------------------ */
/* Solid3D: crates/pantometry-thermal/src/solid.rs                                             */
/* ------------------------------------------------------------------------------------------ */

/* Substance::heat_capacity = mass_of(volume) * c_p = (rho*volume)*c_p,
 * crates/pantometry-core/src/substance.rs:219-239; volume = (dx*dx)*dx, solid.rs:591. */
double orc_heat_capacity(double density, double specific_heat, double dx) {
    double volume = dx * dx * dx;
    return (density * volume) * specific_heat;
}

/* The harmonic face mean `series`, solid.rs:706-713: `2.0 * a * b / sum` guarded at zero. */
double orc_series(double a, double b) {
    double sum = a + b;
    if (sum > 0.0) return 2.0 * a * b / sum;
    return 0.0;
}

/* Solid3D::resolve for the plain branch (no phase change, no void), solid.rs:704-782.
 *   kcell[N]   per-cell conductivity, the `mixed.0` of solid.rs:686-702
 *   cap[N]     per-cell heat capacity
 * Out: kx[(nx+1)ny nz], ky[nx(ny+1)nz], kz[nx ny(nz+1)], mobility[N], face_sum[N];
 * returns worst_rate (NaN when any rate is NaN, solid.rs:771-782). */
double orc_resolve_void(size_t nx, size_t ny, size_t nz, double dx, const double *kcell,
                        const double *cap, const unsigned char *voids, double *kx, double *ky,
                        double *kz, double *mobility, double *face_sum);

double orc_resolve(size_t nx, size_t ny, size_t nz, double dx, const double *kcell,
                   const double *cap, double *kx, double *ky, double *kz, double *mobility,
                   double *face_sum) {
    return orc_resolve_void(nx, ny, nz, dx, kcell, cap, NULL, kx, ky, kz, mobility, face_sum);
}

/* The same with void cells (solid.rs:686-702, 753-756): a void cell conducts and holds nothing
 * (its k and capacity are zero, so every face it touches is zero through `series`), and its
 * mobility is zero rather than dx/0.  `voids` may be NULL. */
double orc_resolve_void(size_t nx, size_t ny, size_t nz, double dx, const double *kcell,
                        const double *cap, const unsigned char *voids, double *kx, double *ky,
                        double *kz, double *mobility, double *face_sum) {
    memset(kx, 0, sizeof(double) * (nx + 1) * ny * nz);
    memset(ky, 0, sizeof(double) * nx * (ny + 1) * nz);
    memset(kz, 0, sizeof(double) * nx * ny * (nz + 1));
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                size_t c = i + nx * (j + ny * k);
                if (i > 0) kx[i + (nx + 1) * (j + ny * k)] = orc_series(kcell[c - 1], kcell[c]);
                if (j > 0) ky[i + nx * (j + (ny + 1) * k)] = orc_series(kcell[c - nx], kcell[c]);
                if (k > 0) kz[i + nx * (j + ny * k)] = orc_series(kcell[c - nx * ny], kcell[c]);
            }
    double worst = 0.0;
    int nowhere = 0;
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                size_t c = i + nx * (j + ny * k);
                mobility[c] = (voids && voids[c]) ? 0.0 : dx / cap[c];
                double sum = kx[i + (nx + 1) * (j + ny * k)] + kx[i + 1 + (nx + 1) * (j + ny * k)] +
                             ky[i + nx * (j + (ny + 1) * k)] + ky[i + nx * (j + 1 + (ny + 1) * k)] +
                             kz[i + nx * (j + ny * k)] + kz[i + nx * (j + ny * (k + 1))];
                face_sum[c] = sum;
                double rate = mobility[c] * sum;
                if (isnan(rate))
                    nowhere = 1;
                else
                    worst = fmax(worst, rate);
            }
    return nowhere ? NAN : worst;
}

/* One cell of the plain sweep, solid.rs:1776-1826 (Appendix A.1 of the contract). */
static inline double orc_cell(size_t nx, size_t ny, size_t nz, size_t i, size_t j, size_t k,
                              const double *kx, const double *ky, const double *kz,
                              const double *mobility, double dts, const double *old) {
    size_t c = i + nx * (j + ny * k);
    double t = old[c];
    size_t xr = (nx + 1) * (j + ny * k), yr = nx * (j + (ny + 1) * k);
    double flux = 0.0;
    if (i > 0) flux += kx[i + xr] * (old[c - 1] - t);
    if (i + 1 < nx) flux += kx[i + 1 + xr] * (old[c + 1] - t);
    if (j > 0) flux += ky[i + yr] * (old[c - nx] - t);
    if (j + 1 < ny) flux += ky[i + nx + yr] * (old[c + nx] - t);
    if (k > 0) flux += kz[c] * (old[c - nx * ny] - t);
    if (k + 1 < nz) flux += kz[c + nx * ny] * (old[c + nx * ny] - t);
    /* add_kelvin(c, dts * mobility[c] * (flux - shed)) with shed = 0 and latent = 0,
     * solid.rs:1818 and 1620-1635. */
    double rise = dts * mobility[c] * (flux - 0.0);
    return t + rise;
}

/* Solid3D::step, plain branch, single thread exactly as the reference (with the per-step clone
 * of solid.rs:1774 into `old`, which the caller provides as scratch of N doubles). */
void orc_solid_step(size_t nx, size_t ny, size_t nz, const double *kx, const double *ky,
                    const double *kz, const double *mobility, double dts, double *cells,
                    double *old) {
    memcpy(old, cells, sizeof(double) * nx * ny * nz);
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++)
                cells[i + nx * (j + ny * k)] =
                    orc_cell(nx, ny, nz, i, j, k, kx, ky, kz, mobility, dts, old);
}

/* The same sweep with OpenMP over z-planes.  Cells are independent, so the result is
 * bit-identical to orc_solid_step; used for full-size parity runs and the CPU baseline. */
void orc_solid_step_omp(size_t nx, size_t ny, size_t nz, const double *kx, const double *ky,
                        const double *kz, const double *mobility, double dts, double *cells,
                        double *old, int threads) {
    size_t n = nx * ny * nz;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (long long c = 0; c < (long long)n; c++) old[c] = cells[c];
#pragma omp parallel for num_threads(threads) schedule(static)
    for (long long k = 0; k < (long long)nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++)
                cells[i + nx * (j + ny * (size_t)k)] =
                    orc_cell(nx, ny, nz, i, j, (size_t)k, kx, ky, kz, mobility, dts, old);
}

/* ---- Exposed faces and volumetric sources ------------------------------------------------ */

/* STEFAN_BOLTZMANN, crates/pantometry-units/src/lib.rs:899. */
#define ORC_SIGMA 5.670374419e-8

/* x.powi(4) and x.powi(3) with a constant exponent, which Rust lowers to repeated squaring (the
 * same products compiler-rt's __powidf2 forms): (x*x)*(x*x) and x*(x*x). */
static inline double orc_pow4(double x) {
    double x2 = x * x;
    return x2 * x2;
}
static inline double orc_pow3(double x) { return x * (x * x); }

/* Environment::loss_from, crates/pantometry-thermal/src/lib.rs:104-110. */
double orc_loss_from(double t, double ambient, double h, double area, double emissivity) {
    double convective = h * area * (t - ambient);
    double radiative = emissivity * ORC_SIGMA * area * (orc_pow4(t) - orc_pow4(ambient));
    return convective + radiative;
}

/* series_with_half_cell, solid.rs:110-122. */
double orc_series_with_half_cell(double film, double conductivity, double area, double dx) {
    if (film <= 0.0) return 0.0;
    double half = 2.0 * conductivity * area / dx;
    if (!isfinite(half)) return film;
    if (half <= 0.0) return 0.0;
    return 1.0 / (1.0 / film + 1.0 / half);
}

/* Face::holds, solid.rs:158-169; faces in the enum's order XMin, XMax, YMin, YMax, ZMin, ZMax. */
static inline int orc_holds(int f, size_t i, size_t j, size_t k, size_t nx, size_t ny, size_t nz) {
    switch (f) {
        case 0: return i == 0;
        case 1: return i + 1 == nx;
        case 2: return j == 0;
        case 3: return j + 1 == ny;
        case 4: return k == 0;
        default: return k + 1 == nz;
    }
}

/* cells_on(face).max(1) as f64, solid.rs:1224-1231. */
static inline double orc_cells_on(int f, size_t nx, size_t ny, size_t nz) {
    size_t n = f < 2 ? ny * nz : f < 4 ? nx * nz : nx * ny;
    return (double)(n > 1 ? n : 1);
}

/* Solid3D::film_flux, solid.rs:1192-1221, for cell (i,j,k) at temperature t.  `on`, `ambient`,
 * `h`, `area` describe the six faces (the BTreeMap `exposed`, iterated in Face order);
 * `emissivity` and `conductivity` are the cell's substance's (0 and +inf when it does not say). */
double orc_film_flux(size_t nx, size_t ny, size_t nz, double dx, size_t i, size_t j, size_t k,
                     double t, const int *on, const double *ambient, const double *h,
                     const double *area, double emissivity, double conductivity) {
    double out = 0.0;
    for (int f = 0; f < 6; f++) {
        if (!on[f] || !orc_holds(f, i, j, k, nx, ny, nz)) continue;
        double share = area[f] / orc_cells_on(f, nx, ny, nz);
        double gap = t - ambient[f];
        if (gap == 0.0) continue;
        double bare = orc_loss_from(t, ambient[f], h[f], share, emissivity);
        double g = orc_series_with_half_cell(bare / gap, conductivity, share, dx);
        out += g * gap / dx;
    }
    return out;
}

/* Solid3D::loss_conductance_at, solid.rs:1262-1283: the film's tangent conductance at
 * max(at, ambient), in series with the half cell, summed over the exposed faces the cell is on. */
double orc_loss_conductance(size_t nx, size_t ny, size_t nz, double dx, size_t i, size_t j,
                            size_t k, double at, const int *on, const double *ambient,
                            const double *h, const double *area, double emissivity,
                            double conductivity) {
    double sum = 0.0;
    for (int f = 0; f < 6; f++) {
        if (!on[f] || !orc_holds(f, i, j, k, nx, ny, nz)) continue;
        double share = area[f] / orc_cells_on(f, nx, ny, nz);
        double hot = fmax(at, ambient[f]);
        double hh = h[f] + 4.0 * emissivity * ORC_SIGMA * orc_pow3(hot);
        sum += orc_series_with_half_cell(hh * share, conductivity, share, dx);
    }
    return sum;
}

/* Solid3D::worst_rate_now, solid.rs:1148-1185, for a block with exposed faces and no gaps. */
double orc_worst_rate_now(size_t nx, size_t ny, size_t nz, double dx, double worst_rate,
                          const double *mobility, const double *face_sum, const double *cells,
                          const int *on, const double *ambient, const double *h,
                          const double *area, const double *emissivity,
                          const double *conductivity) {
    int any = 0;
    for (int f = 0; f < 6; f++) any |= on[f];
    if (!any || isnan(worst_rate)) return worst_rate;
    double worst = worst_rate;
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                size_t c = i + nx * (j + ny * k);
                double loss = orc_loss_conductance(nx, ny, nz, dx, i, j, k, cells[c], on, ambient,
                                                   h, area, emissivity[c], conductivity[c]);
                if (loss == 0.0) continue;
                worst = fmax(worst, mobility[c] * (face_sum[c] + loss / dx));
            }
    return worst;
}

/* Solid3D::step's sweep with the film inside it and the sources after it, solid.rs:1767-1864
 * (plain branch: no void, no gaps, no phase change).  `source` (watts per cell) may be NULL.
 * Adds what the film removed to *lost and what the sources generated to *supplied, each a left
 * fold in cell order as the reference accumulates them. */
void orc_solid_step_ext(size_t nx, size_t ny, size_t nz, double dx, const double *kx,
                        const double *ky, const double *kz, const double *mobility,
                        const double *cap, double dts, double *cells, double *old, const int *on,
                        const double *ambient, const double *h, const double *area,
                        const double *emissivity, const double *conductivity,
                        const double *source, double *lost, double *supplied) {
    memcpy(old, cells, sizeof(double) * nx * ny * nz);
    int exposed = 0;
    for (int f = 0; f < 6; f++) exposed |= on[f];
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                size_t c = i + nx * (j + ny * k);
                double t = old[c];
                size_t xr = (nx + 1) * (j + ny * k), yr = nx * (j + (ny + 1) * k);
                double flux = 0.0;
                if (i > 0) flux += kx[i + xr] * (old[c - 1] - t);
                if (i + 1 < nx) flux += kx[i + 1 + xr] * (old[c + 1] - t);
                if (j > 0) flux += ky[i + yr] * (old[c - nx] - t);
                if (j + 1 < ny) flux += ky[i + nx + yr] * (old[c + nx] - t);
                if (k > 0) flux += kz[c] * (old[c - nx * ny] - t);
                if (k + 1 < nz) flux += kz[c + nx * ny] * (old[c + nx * ny] - t);
                double shed = exposed ? orc_film_flux(nx, ny, nz, dx, i, j, k, t, on, ambient, h,
                                                      area, emissivity[c], conductivity[c])
                                      : 0.0;
                cells[c] += dts * mobility[c] * (flux - shed);
                if (shed != 0.0) *lost += dts * mobility[c] * shed * cap[c];
            }
    if (!source) return;
    size_t n = nx * ny * nz;
    int supplies = 0;
    for (size_t c = 0; c < n && !supplies; c++) supplies = source[c] != 0.0;
    if (!supplies) return;
    for (size_t c = 0; c < n; c++) {
        if (source[c] == 0.0) continue;
        double joules = source[c] * dts;
        cells[c] += joules / cap[c];
        *supplied += joules;
    }
}

/* ScalarField::at for Solid3D, solid.rs:2045-2102 (no void cells): the temperature at point
 * (x, y, z) in metres, trilinear between cell centres, clamped at the outer half cells, the
 * weights renormalised over the corners that carry weight. */
double orc_sample(size_t nx, size_t ny, size_t nz, double dx, const double *cells, double x,
                  double y, double z) {
    double q[3] = {x / dx - 0.5, y / dx - 0.5, z / dx - 0.5};
    if (isnan(q[0]) || isnan(q[1]) || isnan(q[2])) return NAN;
    size_t n[3] = {nx, ny, nz}, at[3];
    double fr[3];
    for (int a = 0; a < 3; a++) {
        size_t last = n[a] > 0 ? n[a] - 1 : 0;
        if (q[a] <= 0.0) {
            at[a] = 0, fr[a] = 0.0;
        } else if (q[a] >= (double)last) {
            at[a] = last, fr[a] = 0.0;
        } else {
            double fl = floor(q[a]);
            at[a] = (size_t)fl, fr[a] = q[a] - fl;
        }
    }
    size_t i1 = at[0] + 1 < nx ? at[0] + 1 : nx - 1;
    size_t j1 = at[1] + 1 < ny ? at[1] + 1 : ny - 1;
    size_t k1 = at[2] + 1 < nz ? at[2] + 1 : nz - 1;
    size_t ia[2] = {at[0], i1}, jb[2] = {at[1], j1}, kc[2] = {at[2], k1};
    double wa[2] = {1.0 - fr[0], fr[0]}, wb[2] = {1.0 - fr[1], fr[1]}, wc[2] = {1.0 - fr[2], fr[2]};
    double sum = 0.0, weight = 0.0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                double w = wa[a] * wb[b] * wc[c];
                if (w > 0.0) {
                    sum += w * cells[ia[a] + nx * (jb[b] + ny * kc[c])];
                    weight += w;
                }
            }
    return weight > 0.0 ? sum / weight : NAN;
}

static inline double orc_wall(size_t k, size_t n);

/* Room::invariant_si, crates/pantometry-acoustic/src/room.rs:243-281: the scheme's conserved
 * energy of a two-dimensional room `depth` deep (nodes x-fastest, nx*ny; vx (nx-1) per row, vy
 * ny-1 rows), each group a left fold in storage order, then potential + kinetic_x + kinetic_y. */
double orc_room_invariant(size_t nx, size_t ny, double dx, double depth, double speed,
                          double density, const double *p, const double *pp, const double *vx,
                          const double *vy) {
    double volume = dx * dx * depth;
    double rc2 = density * speed * speed;
    if (rc2 <= 0.0) return 0.0;
    double potential = 0.0, kx = 0.0, ky = 0.0;
    for (size_t k = 0; k < nx * ny; k++) {
        size_t i = k % nx, j = k / nx;
        double share = 1.0 / (orc_wall(i, nx) * orc_wall(j, ny));
        potential += share * p[k] * pp[k] / (2.0 * rc2) * volume;
    }
    for (size_t k = 0; k < (nx - 1) * ny; k++)
        kx += density * vx[k] * vx[k] / 2.0 * volume / orc_wall(k / (nx - 1), ny);
    for (size_t k = 0; k < nx * (ny - 1); k++)
        ky += density * vy[k] * vy[k] / 2.0 * volume / orc_wall(k % nx, nx);
    return potential + kx + ky;
}

/* clamp_index, crates/pantometry-acoustic/src/hall.rs:684-694. */
static void orc_clamp_index(double q, size_t n, size_t *i, double *f) {
    if (isnan(q) || q <= 0.0) {
        *i = 0, *f = 0.0;
        return;
    }
    double last = (double)(n - 1);
    if (q >= last) {
        *i = n - 1, *f = 0.0;
        return;
    }
    double fl = floor(q);
    *i = (size_t)fl, *f = q - fl;
}

static inline double orc_lerp(double lo, double hi, double t) { return lo * (1.0 - t) + hi * t; }

/* ScalarField::at for Hall, hall.rs:590-617: the pressure at (x, y, z) metres from the corner
 * node, trilinear between nodes (x, then y, then z), clamped at the walls; p[0] for a NaN point. */
double orc_hall_sample(size_t nx, size_t ny, size_t nz, double dx, const double *p, double x,
                       double y, double z) {
    double qx = x / dx, qy = y / dx, qz = z / dx;
    if (isnan(qx) || isnan(qy) || isnan(qz)) return p[0];
    size_t i, j, k;
    double fx, fy, fz;
    orc_clamp_index(qx, nx, &i, &fx);
    orc_clamp_index(qy, ny, &j, &fy);
    orc_clamp_index(qz, nz, &k, &fz);
    size_t i1 = i + 1 < nx ? i + 1 : nx - 1, j1 = j + 1 < ny ? j + 1 : ny - 1;
    size_t k1 = k + 1 < nz ? k + 1 : nz - 1;
#define G(a, b, c) p[(a) + nx * ((b) + ny * (c))]
    double z0 = orc_lerp(orc_lerp(G(i, j, k), G(i1, j, k), fx), orc_lerp(G(i, j1, k), G(i1, j1, k), fx), fy);
    double z1 = orc_lerp(orc_lerp(G(i, j, k1), G(i1, j, k1), fx), orc_lerp(G(i, j1, k1), G(i1, j1, k1), fx), fy);
#undef G
    return orc_lerp(z0, z1, fz);
}

/* Placeless bus heat, solid.rs:1752-1765: a uniform rise gained / sum(capacity), the sum a left
 * fold in index order (`self.capacity.iter().sum::<f64>()`). */
void orc_solid_bus_heat(size_t n, const double *cap, double gained, double *cells) {
    double total = 0.0;
    for (size_t c = 0; c < n; c++) total += cap[c];
    double rise = gained / total;
    for (size_t c = 0; c < n; c++) cells[c] += rise;
}

/* The same with void cells: void takes no share of the rise (solid.rs:1756-1764). */
void orc_solid_bus_heat_void(size_t n, const double *cap, const unsigned char *voids,
                             double gained, double *cells) {
    double total = 0.0;
    for (size_t c = 0; c < n; c++) total += cap[c];
    double rise = gained / total;
    for (size_t c = 0; c < n; c++)
        if (!voids[c]) cells[c] += rise;
}

/* Solid3D::stored_heat, solid.rs:1557-1561 with enthalpy_joules' latent <= 0 branch
 * (solid.rs:1583-1589): a left fold of capacity*(t - reference) - 0. */
double orc_stored_heat(size_t n, const double *cap, const double *cells, double reference) {
    double acc = 0.0;
    for (size_t c = 0; c < n; c++) acc += cap[c] * (cells[c] - reference) - 0.0;
    return acc;
}

/* Neumaier-compensated form of the same sum: the diagnostic the audit is judged beside. */
double orc_stored_heat_compensated(size_t n, const double *cap, const double *cells,
                                   double reference) {
    double sum = 0.0, comp = 0.0;
    for (size_t c = 0; c < n; c++) {
        double x = cap[c] * (cells[c] - reference) - 0.0;
        double t = sum + x;
        if (fabs(sum) >= fabs(x))
            comp += (sum - t) + x;
        else
            comp += (x - t) + sum;
        sum = t;
    }
    return sum + comp;
}

/* The compensated sum over contiguous chunks on `threads` threads (Neumaier within each chunk),
 * the chunks' sums and compensations then combined in chunk order: within ~1e-16 relative of the
 * exact sum for the stored heat of the large blocks, where a single-thread pass is too slow to run
 * after every step.  Deterministic for a given thread count. */
double orc_stored_heat_compensated_omp(size_t n, const double *cap, const double *cells,
                                       double reference, int threads) {
    if (threads < 1) threads = 1;
    double sums[256], comps[256];
    if (threads > 256) threads = 256;
    size_t per = (n + (size_t)threads - 1) / (size_t)threads;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int w = 0; w < threads; w++) {
        size_t lo = (size_t)w * per, hi = lo + per < n ? lo + per : n;
        double sum = 0.0, comp = 0.0;
        for (size_t c = lo; c < hi; c++) {
            double x = cap[c] * (cells[c] - reference) - 0.0;
            double t = sum + x;
            if (fabs(sum) >= fabs(x))
                comp += (sum - t) + x;
            else
                comp += (x - t) + sum;
            sum = t;
        }
        sums[w] = sum;
        comps[w] = comp;
    }
    double sum = 0.0, comp = 0.0;
    for (int w = 0; w < threads; w++) {
        double x = sums[w];
        double t = sum + x;
        if (fabs(sum) >= fabs(x))
            comp += (sum - t) + x;
        else
            comp += (x - t) + sum;
        sum = t;
        comp += comps[w];
    }
    return sum + comp;
}

/* Solid3D::release_mode, solid.rs:1478-1500. */
void orc_release_mode(size_t nx, size_t ny, size_t nz, int a, int b, int cm, double mean,
                      double amplitude, double *cells) {
    const double PI = 3.141592653589793;
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                double px = nx <= 1 ? 1.0 : cos((double)a * PI * ((double)i + 0.5) / (double)nx);
                double py = ny <= 1 ? 1.0 : cos((double)b * PI * ((double)j + 0.5) / (double)ny);
                double pz = nz <= 1 ? 1.0 : cos((double)cm * PI * ((double)k + 0.5) / (double)nz);
                cells[i + nx * (j + ny * k)] = mean + amplitude * px * py * pz;
            }
}

/* Solid3D::mode_amplitude, solid.rs:1507-1535. */
double orc_mode_amplitude(size_t nx, size_t ny, size_t nz, int a, int b, int cm,
                          const double *cells) {
    const double PI = 3.141592653589793;
    double n0 = (nx <= 1 || a == 0) ? 1.0 : 0.5;
    double n1 = (ny <= 1 || b == 0) ? 1.0 : 0.5;
    double n2 = (nz <= 1 || cm == 0) ? 1.0 : 0.5;
    double norm = 1.0 * n0 * n1 * n2;
    double sum = 0.0;
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                double px = nx <= 1 ? 1.0 : cos((double)a * PI * ((double)i + 0.5) / (double)nx);
                double py = ny <= 1 ? 1.0 : cos((double)b * PI * ((double)j + 0.5) / (double)ny);
                double pz = nz <= 1 ? 1.0 : cos((double)cm * PI * ((double)k + 0.5) / (double)nz);
                sum += cells[i + nx * (j + ny * k)] * px * py * pz;
            }
    return sum / ((double)(nx * ny * nz) * norm);
}

/* mode_eigenvalue_dx2 and mode_amplification, solid.rs:1455-1474; fourier_number with
 * alpha = k/(rho*c_p), solid.rs:1423-1428 and substance.rs:211-216. */
double orc_mode_amplification(size_t nx, size_t ny, size_t nz, int a, int b, int cm,
                              double conductivity, double density, double specific_heat,
                              double dx, double dt) {
    const double PI = 3.141592653589793;
    double alpha = conductivity / (density * specific_heat);
    double f = alpha * dt / (dx * dx);
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (nx > 1) { double s = sin((double)a * PI / (2.0 * (double)nx)); t0 = -4.0 * s * s; }
    if (ny > 1) { double s = sin((double)b * PI / (2.0 * (double)ny)); t1 = -4.0 * s * s; }
    if (nz > 1) { double s = sin((double)cm * PI / (2.0 * (double)nz)); t2 = -4.0 * s * s; }
    return 1.0 + f * (t0 + t1 + t2);
}

/* ------------------------------------------------------------------------------------------ */
/* Hall: crates/pantometry-acoustic/src/hall.rs                                                */
/* ------------------------------------------------------------------------------------------ */

/* wall_weight, hall.rs:675-681. */
static inline double orc_wall(size_t k, size_t n) { return (k == 0 || k + 1 == n) ? 2.0 : 1.0; }

/* Hall::invariant_si, hall.rs:334-384: four sequential folds into one total. */
double orc_hall_invariant(size_t nx, size_t ny, size_t nz, double dx, double speed,
                          double density, const double *p, const double *pp, const double *vx,
                          const double *vy, const double *vz) {
    double volume = dx * dx * dx;
    double rc2 = density * speed * speed;
    if (rc2 <= 0.0) return 0.0;
    double total = 0.0;
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                double share = 1.0 / (orc_wall(i, nx) * orc_wall(j, ny) * orc_wall(k, nz));
                size_t idx = i + nx * (j + ny * k);
                total += share * p[idx] * pp[idx] / (2.0 * rc2) * volume;
            }
    double half_rho = 0.5 * density;
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i + 1 < nx; i++) {
                double v = vx[i + (nx - 1) * (j + ny * k)];
                double share = 1.0 / (orc_wall(j, ny) * orc_wall(k, nz));
                total += share * half_rho * v * v * volume;
            }
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j + 1 < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                double v = vy[i + nx * (j + (ny - 1) * k)];
                double share = 1.0 / (orc_wall(i, nx) * orc_wall(k, nz));
                total += share * half_rho * v * v * volume;
            }
    for (size_t k = 0; k + 1 < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                double v = vz[i + nx * (j + ny * k)];
                double share = 1.0 / (orc_wall(i, nx) * orc_wall(j, ny));
                total += share * half_rho * v * v * volume;
            }
    return total;
}

/* The same terms as orc_hall_invariant, each evaluated as there, summed with Neumaier's
 * compensation instead of a left fold: per (group, plane) on `threads` threads, the partial sums
 * and compensations then combined in group and plane order.  Within ~1e-16 relative of the exact
 * sum of the terms, where the reference's left fold over tens of millions of nodes drifts by
 * ~1e-10.  Deterministic for any thread count. */
static void orc_neumaier(double *sum, double *comp, double x) {
    double t = *sum + x;
    if (fabs(*sum) >= fabs(x))
        *comp += (*sum - t) + x;
    else
        *comp += (x - t) + *sum;
    *sum = t;
}

double orc_hall_invariant_compensated_omp(size_t nx, size_t ny, size_t nz, double dx, double speed,
                                          double density, const double *p, const double *pp,
                                          const double *vx, const double *vy, const double *vz,
                                          int threads) {
    double volume = dx * dx * dx;
    double rc2 = density * speed * speed;
    if (rc2 <= 0.0) return 0.0;
    if (threads < 1) threads = 1;
    double half_rho = 0.5 * density;
    double *part = (double *)calloc(8 * nz, sizeof(double));  /* [group][plane][sum, comp] */
    if (!part) return NAN;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (long long kk = 0; kk < (long long)nz; kk++) {
        size_t k = (size_t)kk;
        double s, c;
        s = 0.0, c = 0.0;
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                double share = 1.0 / (orc_wall(i, nx) * orc_wall(j, ny) * orc_wall(k, nz));
                size_t idx = i + nx * (j + ny * k);
                orc_neumaier(&s, &c, share * p[idx] * pp[idx] / (2.0 * rc2) * volume);
            }
        part[(0 * nz + k) * 2] = s, part[(0 * nz + k) * 2 + 1] = c;
        s = 0.0, c = 0.0;
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i + 1 < nx; i++) {
                double v = vx[i + (nx - 1) * (j + ny * k)];
                double share = 1.0 / (orc_wall(j, ny) * orc_wall(k, nz));
                orc_neumaier(&s, &c, share * half_rho * v * v * volume);
            }
        part[(1 * nz + k) * 2] = s, part[(1 * nz + k) * 2 + 1] = c;
        s = 0.0, c = 0.0;
        for (size_t j = 0; j + 1 < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                double v = vy[i + nx * (j + (ny - 1) * k)];
                double share = 1.0 / (orc_wall(i, nx) * orc_wall(k, nz));
                orc_neumaier(&s, &c, share * half_rho * v * v * volume);
            }
        part[(2 * nz + k) * 2] = s, part[(2 * nz + k) * 2 + 1] = c;
        s = 0.0, c = 0.0;
        if (k + 1 < nz)
            for (size_t j = 0; j < ny; j++)
                for (size_t i = 0; i < nx; i++) {
                    double v = vz[i + nx * (j + ny * k)];
                    double share = 1.0 / (orc_wall(i, nx) * orc_wall(j, ny));
                    orc_neumaier(&s, &c, share * half_rho * v * v * volume);
                }
        part[(3 * nz + k) * 2] = s, part[(3 * nz + k) * 2 + 1] = c;
    }
    double sum = 0.0, comp = 0.0;
    for (size_t q = 0; q < 4 * nz; q++) {
        orc_neumaier(&sum, &comp, part[2 * q]);
        comp += part[2 * q + 1];
    }
    free(part);
    return sum + comp;
}

/* The sweeps of Hall::step, hall.rs:430-505, after the Courant guard and the stagger choice
 * (the caller passes vh = h/2 on the first step, h afterwards).  pressure_prev receives the
 * pre-update pressure (hall.rs:463).  A point source adds `source` Pa at node `src` after the
 * pressure sweep (an extension: the reference hall has no source term); src = SIZE_MAX for
 * none.  Returns the energy the source moved into the invariant this step:
 * ((share*(source + source_prev))*pp)/(2 rc2)*volume, where source_prev is the previous step's
 * increment.  An increment d added to p^n changes the invariant once at once (d*p^(n-1)) and
 * once at the next step (d*p^n), because the leapfrog's energy identity for step n+1 uses
 * p^n - p^(n-1), which the increment has moved. */
double orc_hall_sweep(size_t nx, size_t ny, size_t nz, double dx, double speed, double density,
                      double h, double vh, double *p, double *pp, double *vx, double *vy,
                      double *vz, size_t src, double source, double source_prev,
                      int threads) {
    double rc2 = density * speed * speed;
    double gain = vh / (density * dx);
    (void)threads;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (long long k = 0; k < (long long)nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i + 1 < nx; i++) {
                size_t a = i + nx * (j + ny * (size_t)k);
                double d = p[a + 1] - p[a];
                vx[i + (nx - 1) * (j + ny * (size_t)k)] -= gain * d;
            }
#pragma omp parallel for num_threads(threads) schedule(static)
    for (long long k = 0; k < (long long)nz; k++)
        for (size_t j = 0; j + 1 < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                size_t a = i + nx * (j + ny * (size_t)k);
                double d = p[a + nx] - p[a];
                vy[i + nx * (j + (ny - 1) * (size_t)k)] -= gain * d;
            }
#pragma omp parallel for num_threads(threads) schedule(static)
    for (long long k = 0; k < (long long)nz - 1; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                size_t a = i + nx * (j + ny * (size_t)k);
                double d = p[a + nx * ny] - p[a];
                vz[a] -= gain * d;
            }
    memcpy(pp, p, sizeof(double) * nx * ny * nz);
    double scale = h * rc2 / dx;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (long long kk = 0; kk < (long long)nz; kk++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                size_t k = (size_t)kk;
                double lo_x = i > 0 ? vx[i - 1 + (nx - 1) * (j + ny * k)] : 0.0;
                double hi_x = i + 1 < nx ? vx[i + (nx - 1) * (j + ny * k)] : 0.0;
                double lo_y = j > 0 ? vy[i + nx * (j - 1 + (ny - 1) * k)] : 0.0;
                double hi_y = j + 1 < ny ? vy[i + nx * (j + (ny - 1) * k)] : 0.0;
                double lo_z = k > 0 ? vz[i + nx * (j + ny * (k - 1))] : 0.0;
                double hi_z = k + 1 < nz ? vz[i + nx * (j + ny * k)] : 0.0;
                size_t idx = i + nx * (j + ny * k);
                p[idx] -= scale * (orc_wall(i, nx) * (hi_x - lo_x) + orc_wall(j, ny) * (hi_y - lo_y) +
                                   orc_wall(k, nz) * (hi_z - lo_z));
            }
    if (src == (size_t)-1) return 0.0;
    p[src] = p[src] + source;
    size_t si = src % nx, sj = (src / nx) % ny, sk = src / (nx * ny);
    double share = 1.0 / (orc_wall(si, nx) * orc_wall(sj, ny) * orc_wall(sk, nz));
    double volume = dx * dx * dx;
    return share * (source + source_prev) * pp[src] / (2.0 * rc2) * volume;
}

/* Hall::released_in_mode's profile, hall.rs:190-210, evaluated at node positions
 * DVec3(i, j, k) * dx (hall.rs:170-180). */
void orc_hall_mode(size_t nx, size_t ny, size_t nz, double dx, int a, int b, int cm,
                   double amplitude, double *p) {
    const double PI = 3.141592653589793;
    double lx = (double)(nx - 1) * dx, ly = (double)(ny - 1) * dx, lz = (double)(nz - 1) * dx;
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                double x = (double)i * dx, y = (double)j * dx, z = (double)k * dx;
                double px = lx > 0.0 ? cos((double)a * PI * x / lx) : 1.0;
                double py = ly > 0.0 ? cos((double)b * PI * y / ly) : 1.0;
                double pz = lz > 0.0 ? cos((double)cm * PI * z / lz) : 1.0;
                p[i + nx * (j + ny * k)] = amplitude * px * py * pz;
            }
}

/* ------------------------------------------------------------------------------------------ */
/* Seeded workloads (the benchmark configurations), restated independently of the library's  */
/* generator so the two can be compared bit for bit.                                           */
/* ------------------------------------------------------------------------------------------ */

/* Gaussian hot spot plus uniform noise: T = 300 + amp*exp(-r2/(2 s^2)) + range(-0.1, 0.1),
 * r measured in cells from the block centre, noise from Rng::for_index(noise_seed, cell). */
void orc_workload_hot_spot(size_t nx, size_t ny, size_t nz, double base, double amp,
                           double sigma, double noise, uint64_t noise_seed, double *cells,
                           int threads) {
    (void)threads;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (long long kk = 0; kk < (long long)nz; kk++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                size_t k = (size_t)kk;
                size_t c = i + nx * (j + ny * k);
                double ddx = ((double)i + 0.5) - 0.5 * (double)nx;
                double ddy = ((double)j + 0.5) - 0.5 * (double)ny;
                double ddz = ((double)k + 0.5) - 0.5 * (double)nz;
                double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                uint64_t s = orc_rng_for_index(noise_seed, c);
                double t = base + amp * exp(-r2 / (2.0 * sigma * sigma));
                cells[c] = t + orc_rng_range(&s, -noise, noise);
            }
}

/* Heterogeneous block: k_c = exp(0 + 0.5 * gaussian) from Rng::for_index(k_seed, c); T = 300 +
 * sum over `spots` Gaussian hot spots whose amplitude, width and centre are drawn from
 * Rng::for_index(spot_seed, m). */
void orc_workload_heterogeneous(size_t nx, size_t ny, size_t nz, uint64_t k_seed,
                                uint64_t spot_seed, int spots, double *kcell, double *cells,
                                int threads) {
    double amp[64], sig[64], cx[64], cy[64], cz[64];
    if (spots > 64) spots = 64;
    for (int m = 0; m < spots; m++) {
        uint64_t s = orc_rng_for_index(spot_seed, (uint64_t)m);
        amp[m] = orc_rng_range(&s, 50.0, 200.0);
        sig[m] = orc_rng_range(&s, 4.0, 16.0);
        cx[m] = orc_rng_range(&s, 0.1 * (double)nx, 0.9 * (double)nx);
        cy[m] = orc_rng_range(&s, 0.1 * (double)ny, 0.9 * (double)ny);
        cz[m] = orc_rng_range(&s, 0.1 * (double)nz, 0.9 * (double)nz);
    }
    (void)threads;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (long long kk = 0; kk < (long long)nz; kk++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                size_t k = (size_t)kk;
                size_t c = i + nx * (j + ny * k);
                uint64_t s = orc_rng_for_index(k_seed, c);
                kcell[c] = exp(0.0 + 0.5 * orc_rng_gaussian(&s));
                double t = 300.0;
                for (int m = 0; m < spots; m++) {
                    double ddx = ((double)i + 0.5) - cx[m];
                    double ddy = ((double)j + 0.5) - cy[m];
                    double ddz = ((double)k + 0.5) - cz[m];
                    double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    t = t + amp[m] * exp(-r2 / (2.0 * sig[m] * sig[m]));
                }
                cells[c] = t;
            }
}

/* One ensemble member's shape: edge = round(exp(range(ln 16, ln 96))) and
 * alpha = exp(range(ln 1e-5, ln 1e-3)) from Rng::for_index(seed, i). */
void orc_workload_member(uint64_t seed, uint64_t i, double lo_edge, double hi_edge,
                         double lo_alpha, double hi_alpha, long long *edge, double *alpha) {
    uint64_t s = orc_rng_for_index(seed, i);
    *edge = (long long)round(exp(orc_rng_range(&s, log(lo_edge), log(hi_edge))));
    *alpha = exp(orc_rng_range(&s, log(lo_alpha), log(hi_alpha)));
}

/* ------------------------------------------------------------------------------------------ */
/* Bar1D, crates/pantometry-thermal/src/lib.rs:425-660                                       */
/* ------------------------------------------------------------------------------------------ */

/* Bar1D::step after its guard, lib.rs:584-626: the bus share into cell 0, per-face joules into
 * their cells, then the insulated three-point sweep on a copy of the previous cells. */
void orc_bar_step(size_t n, double f, double capacity, double gained, const double *joules,
                  double *cells, double *prev) {
    cells[0] += gained / capacity;
    if (joules)
        for (size_t i = 0; i < n; i++) cells[i] += joules[i] / capacity;
    memcpy(prev, cells, n * sizeof(double));
    size_t last = n - 1;
    for (size_t i = 0; i <= last; i++) {
        double left = prev[i == 0 ? 0 : i - 1];
        double right = prev[i + 1 < last ? i + 1 : last];
        cells[i] = prev[i] + f * (left - 2.0 * prev[i] + right);
    }
}

/* Bar1D::stored_heat, lib.rs:535-538: capacity times the left fold of (t - reference). */
double orc_bar_stored(size_t n, double capacity, const double *cells, double reference) {
    double s = 0.0;
    for (size_t i = 0; i < n; i++) s += cells[i] - reference;
    return capacity * s;
}

/* ScalarField::at for Bar1D, lib.rs:736-751: linear between cell centres, held past the ends. */
double orc_bar_sample(size_t n, double dx, const double *cells, double x) {
    size_t last = n - 1;
    double u = x / dx - 0.5;
    if (isnan(u) || u <= 0.0) return cells[0];
    if (u >= (double)last) return cells[last];
    size_t i = (size_t)floor(u);
    double f = u - (double)i;
    return cells[i] * (1.0 - f) + cells[i + 1] * f;
}

/* ------------------------------------------------------------------------------------------ */
/* Tube, crates/pantometry-acoustic/src/lib.rs:136-563                                        */
/* ------------------------------------------------------------------------------------------ */

/* Tube::wall_velocity, lib.rs:371-385: end 0 closed, 1 open, 2 matched to `z`. */
static double orc_wall_velocity(int end, double z, double impedance, double p, double outward) {
    if (end == 0) return 0.0;
    if (end == 1 || z <= 0.0) return outward * p / impedance;
    return outward * p / z;
}

/* Tube::invariant_si, lib.rs:300-327. */
double orc_tube_invariant(size_t n, double dx, double area, double speed, double density,
                          const double *p, const double *pp, const double *u) {
    double volume = area * dx;
    double rc2 = density * speed * speed;
    if (rc2 <= 0.0) return 0.0;
    size_t last = n - 1;
    double potential = 0.0, kinetic = 0.0;
    for (size_t i = 0; i < n; i++) {
        double share = (i == 0 || i == last) ? 0.5 : 1.0;
        potential += share * p[i] * pp[i] / (2.0 * rc2) * volume;
    }
    for (size_t i = 0; i + 1 < n; i++) kinetic += density * u[i] * u[i] / 2.0 * volume;
    return potential + kinetic;
}

/* Tube::advance's sweeps, lib.rs:389-446: faces, wall velocities from the pressures before the
 * update, cells with the end cells' half width, then the power through each wall.  Returns the
 * joules the walls absorbed this step (lib.rs:436-446). */
double orc_tube_sweep(size_t n, double dx, double area, double speed, double density, double dt,
                      double vh, int left, double zl, int right, double zr, double *p, double *pp,
                      double *u) {
    double rc2 = density * speed * speed;
    double impedance = density * speed;
    for (size_t i = 0; i + 1 < n; i++) u[i] -= vh / (density * dx) * (p[i + 1] - p[i]);
    double lw = orc_wall_velocity(left, zl, impedance, p[0], -1.0);
    double rw = orc_wall_velocity(right, zr, impedance, p[n - 1], 1.0);
    memcpy(pp, p, n * sizeof(double));
    for (size_t i = 0; i < n; i++) {
        double inflow = i == 0 ? lw : u[i - 1];
        double outflow = i == n - 1 ? rw : u[i];
        double share = (i == 0 || i == n - 1) ? 2.0 : 1.0;
        p[i] -= dt * rc2 / dx * share * (outflow - inflow);
    }
    double absorbed = 0.0;
    absorbed += fabs(p[0] * lw) * area * dt;
    absorbed += fabs(p[n - 1] * rw) * area * dt;
    return absorbed;
}

/* ------------------------------------------------------------------------------------------ */
/* Void cells and the gaps between them, crates/pantometry-thermal/src/solid.rs               */
/* ------------------------------------------------------------------------------------------ */

/* One grid line of Solid3D::find_gaps, solid.rs:803-831: a run of void with solid at both ends
 * pairs the two solid cells, coupled by sigma*dx^2/(1/ea + 1/eb - 1) when both radiate. */
static size_t orc_walk(const size_t *line, size_t len, const unsigned char *voids,
                       const double *emissivity, double area, size_t count, size_t *a,
                       size_t *b, double *coef) {
    long last = -1;
    int between = 0;
    for (size_t q = 0; q < len; q++) {
        size_t c = line[q];
        if (voids[c]) {
            between = last >= 0;
            continue;
        }
        if (last >= 0 && between) {
            double ea = emissivity[last], eb = emissivity[c];
            if (ea > 0.0 && eb > 0.0) {
                double resistance = 1.0 / ea + 1.0 / eb - 1.0;
                if (resistance > 0.0) {
                    a[count] = (size_t)last, b[count] = c;
                    coef[count] = ORC_SIGMA * area / resistance;
                    count++;
                }
            }
        }
        last = (long)c;
        between = 0;
    }
    return count;
}

/* Solid3D::find_gaps, solid.rs:791-847: x lines, then y lines, then z lines.  `emissivity` is
 * each cell's substance's (0 for one that does not say).  The outputs hold at most
 * 3*nx*ny*nz pairs; returns the count. */
size_t orc_find_gaps(size_t nx, size_t ny, size_t nz, double dx, const unsigned char *voids,
                     const double *emissivity, size_t *a, size_t *b, double *coef) {
    size_t n = nx * ny * nz, count = 0;
    int any = 0;
    for (size_t c = 0; c < n && !any; c++) any = voids[c];
    if (!any) return 0;
    double area = dx * dx;
    size_t longest = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
    size_t *line = (size_t *)malloc(sizeof(size_t) * longest);
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++) {
            for (size_t i = 0; i < nx; i++) line[i] = i + nx * (j + ny * k);
            count = orc_walk(line, nx, voids, emissivity, area, count, a, b, coef);
        }
    for (size_t k = 0; k < nz; k++)
        for (size_t i = 0; i < nx; i++) {
            for (size_t j = 0; j < ny; j++) line[j] = i + nx * (j + ny * k);
            count = orc_walk(line, ny, voids, emissivity, area, count, a, b, coef);
        }
    for (size_t j = 0; j < ny; j++)
        for (size_t i = 0; i < nx; i++) {
            for (size_t k = 0; k < nz; k++) line[k] = i + nx * (j + ny * k);
            count = orc_walk(line, nz, voids, emissivity, area, count, a, b, coef);
        }
    free(line);
    return count;
}

/* The gaps' exchange in Solid3D::step, solid.rs:1866-1878: in pair order, read off the step's
 * input `old`, antisymmetric. */
void orc_gap_exchange(size_t count, const size_t *a, const size_t *b, const double *coef,
                      const double *old, const double *cap, double dts, double *cells) {
    for (size_t g = 0; g < count; g++) {
        double ta = old[a[g]], tb = old[b[g]];
        double watts = coef[g] * (orc_pow4(ta) - orc_pow4(tb));
        if (watts == 0.0) continue;
        double joules = watts * dts;
        cells[a[g]] += -joules / cap[a[g]];
        cells[b[g]] += joules / cap[b[g]];
    }
}

/* The gaps' part of Solid3D::worst_rate_now, solid.rs:1159-1171: each pair's radiative
 * conductance linearised at the hotter cell, over each cell's capacity. */
double orc_gap_rates(size_t count, const size_t *a, const size_t *b, const double *coef,
                     const double *cells, const double *cap, double worst) {
    for (size_t g = 0; g < count; g++) {
        double hotter = fmax(cells[a[g]], cells[b[g]]);
        double gc = 4.0 * coef[g] * orc_pow3(hotter);
        if (cap[a[g]] > 0.0) worst = fmax(worst, gc / cap[a[g]]);
        if (cap[b[g]] > 0.0) worst = fmax(worst, gc / cap[b[g]]);
    }
    return worst;
}

/* ScalarField::at for Solid3D with void, solid.rs:2045-2102: NaN when the cell the point is in
 * (rounded, half away from zero) is void; otherwise the trilinear weights over the solid corners
 * that carry weight, renormalised. */
double orc_sample_void(size_t nx, size_t ny, size_t nz, double dx, const double *cells,
                       const unsigned char *voids, double x, double y, double z) {
    double q[3] = {x / dx - 0.5, y / dx - 0.5, z / dx - 0.5};
    if (isnan(q[0]) || isnan(q[1]) || isnan(q[2])) return NAN;
    size_t n[3] = {nx, ny, nz}, at[3], near[3];
    double fr[3];
    for (int a = 0; a < 3; a++) {
        size_t last = n[a] > 0 ? n[a] - 1 : 0;
        if (q[a] <= 0.0) {
            at[a] = 0, fr[a] = 0.0;
        } else if (q[a] >= (double)last) {
            at[a] = last, fr[a] = 0.0;
        } else {
            double fl = floor(q[a]);
            at[a] = (size_t)fl, fr[a] = q[a] - fl;
        }
        double r = round(q[a]);
        if (!(r > 0.0)) r = 0.0;
        near[a] = (size_t)r;
        if (near[a] > last) near[a] = last;
    }
    if (voids[near[0] + nx * (near[1] + ny * near[2])]) return NAN;
    size_t i1 = at[0] + 1 < nx ? at[0] + 1 : nx - 1;
    size_t j1 = at[1] + 1 < ny ? at[1] + 1 : ny - 1;
    size_t k1 = at[2] + 1 < nz ? at[2] + 1 : nz - 1;
    size_t ia[2] = {at[0], i1}, jb[2] = {at[1], j1}, kc[2] = {at[2], k1};
    double wa[2] = {1.0 - fr[0], fr[0]}, wb[2] = {1.0 - fr[1], fr[1]}, wc[2] = {1.0 - fr[2], fr[2]};
    double sum = 0.0, weight = 0.0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                double w = wa[a] * wb[b] * wc[c];
                size_t idx = ia[a] + nx * (jb[b] + ny * kc[c]);
                if (w > 0.0 && !voids[idx]) {
                    sum += w * cells[idx];
                    weight += w;
                }
            }
    return weight > 0.0 ? sum / weight : NAN;
}

/* ------------------------------------------------------------------------------------------ */
/* Phase change: the enthalpy method of Solid3D, solid.rs:1563-1648                           */
/* ------------------------------------------------------------------------------------------ */

/* Per-cell phase state: the melted fraction (written), the mixed capacity the step converts a
 * rise through, the melting point, the latent joules of the cell, and the solid and liquid
 * capacities. */
typedef struct {
    double *melted;
    const double *capacity;
    const double *point;
    const double *latent;
    const double *cap_solid;
    const double *cap_liquid;
} OrcPhase;

/* Solid3D::enthalpy_joules, solid.rs:1578-1593 (the latent > 0 branches). */
static double orc_enthalpy(const OrcPhase *ph, size_t c, double t) {
    double point = ph->point[c], latent = ph->latent[c];
    if (ph->melted[c] >= 1.0) return latent + ph->cap_liquid[c] * (t - point);
    if (ph->melted[c] <= 0.0) return ph->cap_solid[c] * (t - point);
    (void)point;
    return ph->melted[c] * latent;
}

/* Solid3D::add_kelvin and state_from_enthalpy, solid.rs:1620-1648. */
static void orc_add_kelvin(const OrcPhase *ph, size_t c, double rise, double *cells) {
    if (ph && ph->latent[c] > 0.0) {
        double joules = rise * ph->capacity[c];
        double e = orc_enthalpy(ph, c, cells[c]) + joules;
        double point = ph->point[c], latent = ph->latent[c];
        if (e <= 0.0) {
            cells[c] = point + e / ph->cap_solid[c];
            ph->melted[c] = 0.0;
        } else if (e >= latent) {
            cells[c] = point + (e - latent) / ph->cap_liquid[c];
            ph->melted[c] = 1.0;
        } else {
            cells[c] = point;
            ph->melted[c] = e / latent;
        }
    } else {
        cells[c] += rise;
    }
}

/* Solid3D::stored_heat with phases, solid.rs:1557-1614: a left fold of
 * enthalpy_joules - reference_enthalpy_joules in index order. */
double orc_stored_heat_phase(size_t n, const double *cap, const double *cells, double reference,
                             const double *melted, const double *point, const double *latent,
                             const double *cap_solid, const double *cap_liquid) {
    OrcPhase ph = {(double *)melted, cap, point, latent, cap_solid, cap_liquid};
    double acc = 0.0;
    for (size_t c = 0; c < n; c++) {
        double e, r;
        if (latent[c] <= 0.0) {
            e = cap[c] * (cells[c] - reference);
            r = 0.0;
        } else {
            e = orc_enthalpy(&ph, c, cells[c]);
            r = reference > point[c] ? latent[c] + cap_liquid[c] * (reference - point[c])
                                     : cap_solid[c] * (reference - point[c]);
        }
        acc += e - r;
    }
    return acc;
}

/* The bus share with phases and void: a uniform rise through add_kelvin (solid.rs:1752-1765). */
void orc_bus_heat_phase(size_t n, const double *cap, const unsigned char *voids, double gained,
                        double *cells, double *melted, const double *point, const double *latent,
                        const double *cap_solid, const double *cap_liquid) {
    OrcPhase ph = {melted, cap, point, latent, cap_solid, cap_liquid};
    double total = 0.0;
    for (size_t c = 0; c < n; c++) total += cap[c];
    double rise = gained / total;
    for (size_t c = 0; c < n; c++)
        if (!voids[c]) orc_add_kelvin(&ph, c, rise, cells);
}

/* Solid3D::step's sweep, film, sources and gaps with phases and void, solid.rs:1767-1878: every
 * change to a cell goes through add_kelvin; sources skip void cells; the gaps read `old`. */
void orc_solid_step_phase(size_t nx, size_t ny, size_t nz, double dx, const double *kx,
                          const double *ky, const double *kz, const double *mobility,
                          const double *cap, double dts, double *cells, double *old,
                          const int *on, const double *ambient, const double *h,
                          const double *area, const double *emissivity,
                          const double *conductivity, const double *source,
                          const unsigned char *voids, size_t gaps, const size_t *ga,
                          const size_t *gb, const double *gcoef, double *melted,
                          const double *point, const double *latent, const double *cap_solid,
                          const double *cap_liquid, double *lost, double *supplied) {
    OrcPhase ph = {melted, cap, point, latent, cap_solid, cap_liquid};
    size_t n = nx * ny * nz;
    memcpy(old, cells, sizeof(double) * n);
    int exposed = 0;
    for (int f = 0; f < 6; f++) exposed |= on[f];
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < nx; i++) {
                size_t c = i + nx * (j + ny * k);
                double t = old[c];
                size_t xr = (nx + 1) * (j + ny * k), yr = nx * (j + (ny + 1) * k);
                double flux = 0.0;
                if (i > 0) flux += kx[i + xr] * (old[c - 1] - t);
                if (i + 1 < nx) flux += kx[i + 1 + xr] * (old[c + 1] - t);
                if (j > 0) flux += ky[i + yr] * (old[c - nx] - t);
                if (j + 1 < ny) flux += ky[i + nx + yr] * (old[c + nx] - t);
                if (k > 0) flux += kz[c] * (old[c - nx * ny] - t);
                if (k + 1 < nz) flux += kz[c + nx * ny] * (old[c + nx * ny] - t);
                double shed = exposed ? orc_film_flux(nx, ny, nz, dx, i, j, k, t, on, ambient, h,
                                                      area, emissivity[c], conductivity[c])
                                      : 0.0;
                orc_add_kelvin(&ph, c, dts * mobility[c] * (flux - shed), cells);
                if (shed != 0.0) *lost += dts * mobility[c] * shed * cap[c];
            }
    int supplies = 0;
    for (size_t c = 0; c < n && source && !supplies; c++) supplies = source[c] != 0.0;
    if (supplies)
        for (size_t c = 0; c < n; c++) {
            if (source[c] == 0.0 || (voids && voids[c])) continue;
            double joules = source[c] * dts;
            orc_add_kelvin(&ph, c, joules / cap[c], cells);
            *supplied += joules;
        }
    for (size_t g = 0; g < gaps; g++) {
        double ta = old[ga[g]], tb = old[gb[g]];
        double watts = gcoef[g] * (orc_pow4(ta) - orc_pow4(tb));
        if (watts == 0.0) continue;
        double joules = watts * dts;
        orc_add_kelvin(&ph, ga[g], -joules / cap[ga[g]], cells);
        orc_add_kelvin(&ph, gb[g], joules / cap[gb[g]], cells);
    }
}

/* Solid3D::deposit with phases, solid.rs:1364-1372. */
void orc_deposit_phase(size_t c, double joules, double *cells, double *melted, const double *cap,
                       const double *point, const double *latent, const double *cap_solid,
                       const double *cap_liquid) {
    OrcPhase ph = {melted, cap, point, latent, cap_solid, cap_liquid};
    orc_add_kelvin(&ph, c, joules / cap[c], cells);
}
