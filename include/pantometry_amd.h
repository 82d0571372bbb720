/*
 * pantometry_amd — the C ABI of the MI355X (gfx950) backend for pantometry's explicit 3D
 * grid-stencil path: Solid3D's conduction sweep, Hall's staggered pressure/velocity leapfrog,
 * the conservation audit that totals heat or acoustic energy after each step, and a batched
 * ensemble of uneven heat grids.
 *
 * Every entry point takes plain pointers and sizes.  Each declaration cites the reference
 * interface it replaces (paths relative to the pantometry repository).  The Domain trait those
 * interfaces implement is crates/pantometry-core/src/sim.rs:77-252; a Rust binding implements
 * Domain by calling these entry points across extern "C".
 *
 * Conventions
 *   - Status: PM_OK (0); PM_VIOLATION (1) when a step is refused or an audit breaks, with the
 *     pm_violation out-parameter filled (Result<(), Violation>, conserved.rs:128-143); a
 *     PM_ERR_* code for device, argument or state failures, with pm_last_error() describing it.
 *   - A refused step mutates nothing (solid.rs:1718-1738, hall.rs:409-419).
 *   - Handles are single-threaded and not re-entrant; every handle owns its device buffers and
 *     one HIP stream.  Host arrays are caller-owned and copied in and out.
 *   - There is no CPU fallback: without a usable gfx950 device, create calls fail with
 *     PM_ERR_DEVICE.
 */
#ifndef PANTOMETRY_AMD_H
#define PANTOMETRY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PM_OK = 0,
    PM_VIOLATION = 1,
    PM_ERR_DEVICE = 2,
    PM_ERR_ARGUMENT = 3,
    PM_ERR_STATE = 4,
};

/* Violation, crates/pantometry-core/src/conserved.rs:128-143.  `step` names which step of a
 * batched run broke (-1 for a single call). */
typedef struct pm_violation {
    char quantity[64];
    char site[160];
    double before;
    double after;
    double scale;
    double tolerance;
    int64_t step;
} pm_violation;

/* The thermal part of Substance, crates/pantometry-core/src/substance.rs:43-239: density in
 * kg/m^3, conductivity in W/m/K, specific heat in J/kg/K, and the surface's emissivity (0..1,
 * ThermalProps::emissivity) that an exposed face radiates with.  NaN conductivity or specific
 * heat is a substance that does not say (Substance::thermal = None). */
typedef struct pm_substance {
    double density;
    double conductivity;
    double specific_heat;
    double emissivity;
    /* FusionProps, crates/pantometry-core/src/substance.rs:119-193: the melting point (K) and
     * the latent heat (J/kg), both NaN for a substance that does not melt; the liquid's
     * conductivity and specific heat, both NaN for the one-phase model (liquid = solid). */
    double melting_point;
    double latent_heat;
    double liquid_conductivity;
    double liquid_specific_heat;
} pm_substance;

/* The three entries of Solid3D::ledger, crates/pantometry-thermal/src/solid.rs:1904-1927:
 * Ledger::new().with(ENERGY, stored).with(ENERGY, lost).with(ENERGY, -supplied). */
typedef struct pm_books {
    double stored;
    double lost;
    double supplied;
} pm_books;

/* Message for the last PM_ERR_* status on this thread. */
const char *pm_last_error(void);

/* Name, compute-unit count and HBM size of the current device; PM_ERR_DEVICE without one. */
int pm_device_info(char *name, size_t name_len, int *compute_units, size_t *hbm_bytes);

/* A streaming kernel of known traffic, c[i] = a[i] + b[i] over n doubles (16 B read, 8 B
 * written per element), run `reps` times; the mean device time per launch.  Calibrates the HBM
 * counters and measures the bandwidth a read-read-write stream reaches.  No reference
 * counterpart (measurement only). */
int pm_stream_copy(size_t n, int reps, double *ms_per_launch);

/* ---------------------------------------------------------------------------------------- */
/* Rng, crates/pantometry-core/src/rng.rs:37-152 (host-side: seeded inputs are generated once */
/* on the host and uploaded, so libm's exp/ln/cos are never recomputed on the device).       */
/* ---------------------------------------------------------------------------------------- */

/* Rng::new (rng.rs:49-55) and Rng::for_index (rng.rs:66-68): both return the stream state. */
uint64_t pm_rng_new(uint64_t seed);
uint64_t pm_rng_for_index(uint64_t seed, uint64_t index);
/* Rng::unit (rng.rs:91-93), Rng::range (96-98), Rng::gaussian (146-151), Rng::split (75-78). */
double pm_rng_unit(uint64_t *state);
double pm_rng_range(uint64_t *state, double lo, double hi);
double pm_rng_gaussian(uint64_t *state);
uint64_t pm_rng_split(uint64_t *state);

/* Seeded workloads.  Gaussian hot spot plus uniform noise (the uniform heat block):
 * T = base + amp*exp(-r^2/(2 sigma^2)) + range(-noise, noise), r in cells from the centre,
 * noise from Rng::for_index(noise_seed, cell).  No reference counterpart: it writes the
 * initial condition a caller would otherwise set with Solid3D::set_temperature (solid.rs:1346). */
int pm_workload_hot_spot(size_t nx, size_t ny, size_t nz, double base, double amp, double sigma,
                         double noise, uint64_t noise_seed, double *cells);
/* Heterogeneous block: k_c = exp(0.5*gaussian) from Rng::for_index(k_seed, c); T = 300 + sum of
 * `spots` Gaussian hot spots drawn from Rng::for_index(spot_seed, m). */
int pm_workload_heterogeneous(size_t nx, size_t ny, size_t nz, uint64_t k_seed,
                              uint64_t spot_seed, int spots, double *conductivity,
                              double *cells);

/* ---------------------------------------------------------------------------------------- */
/* Solid3D, crates/pantometry-thermal/src/solid.rs                                          */
/* ---------------------------------------------------------------------------------------- */

typedef struct pm_solid3d pm_solid3d;

/* Solid3D::new, solid.rs:377-421: counts forced to at least one, every cell at `initial`, the
 * ledger measured from `initial` (the enthalpy reference).  PM_ERR_ARGUMENT for a count above
 * 2^30 or a plane (nx * ny) of 2^29 cells or more: the sweeps address a plane with 32-bit byte
 * offsets. */
int pm_solid3d_create(const char *name, const pm_substance *substance, size_t nx, size_t ny,
                      size_t nz, double dx, double initial, pm_solid3d **out);
void pm_solid3d_destroy(pm_solid3d *solid);
int pm_solid3d_counts(const pm_solid3d *solid, size_t counts[3]);

/* Solid3D::fill, solid.rs:519-543: give the cells where mask[c] != 0 another substance, then
 * resolve (solid.rs:588-784).  The caller evaluates the predicate into the mask. */
int pm_solid3d_fill(pm_solid3d *solid, const pm_substance *substance, const uint8_t *mask);
/* Per-cell conductivity with the capacity left as it is: arithmetically what one fill per
 * distinct conductivity produces (faces are the harmonic mean, solid.rs:706-734), without the
 * O(N^2) material search of fill.  No reference counterpart. */
int pm_solid3d_set_conductivity(pm_solid3d *solid, const double *conductivity);

/* Temperatures in kelvin: Solid3D::set_temperature over every cell (solid.rs:1346-1361) and
 * Solid3D::temperature_at over every cell (solid.rs:1323). */
int pm_solid3d_set_cells(pm_solid3d *solid, const double *cells);
int pm_solid3d_get_cells(pm_solid3d *solid, double *cells);
/* dst's temperatures := src's, on the device (same extents).  No reference counterpart: Solid3D
 * is Clone (solid.rs:172-173); this copies the field without the operator. */
int pm_solid3d_copy_cells(pm_solid3d *dst, pm_solid3d *src);
int pm_solid3d_set_temperature(pm_solid3d *solid, size_t i, size_t j, size_t k, double kelvin);
/* Solid3D::deposit, solid.rs:1368-1375: joules into one cell, counted as absorbed. */
int pm_solid3d_deposit(pm_solid3d *solid, size_t i, size_t j, size_t k, double joules);

/* Domain::max_stable_dt (solid.rs:1709-1715) and Solid3D::stability_ratio (solid.rs:1435). */
double pm_solid3d_max_stable_dt(const pm_solid3d *solid);
double pm_solid3d_stability_ratio(const pm_solid3d *solid, double dt);
/* The guard at the top of Domain::step (solid.rs:1718-1738), without stepping: PM_VIOLATION
 * with `v` filled if `dt` would be refused.  A binding calls this before taking its share of
 * the bus, so a refused step leaves the bus untouched as the reference does. */
int pm_solid3d_refuses(const pm_solid3d *solid, double dt, pm_violation *v);
/* Domain::step, solid.rs:1717-1901 (plain branch): `gained` is what the binding took from the
 * bus with Exchange::take_share(HEAT, dt) (solid.rs:1752); nonzero gains spread as a uniform
 * rise gained/sum(C) before the sweep. */
int pm_solid3d_step(pm_solid3d *solid, double dt, double gained, pm_violation *v);
/* Domain::ledger, solid.rs:1904-1927.  `stored` is summed on the device in an order fixed by the
 * grid (per thread, wave butterfly, workgroup, one block), not as the reference's left fold over
 * the cells: the same from run to run, within ~1e-12 relative of the exact sum, and up to ~2e-11
 * relative from the CPU left fold at 512^3 (that fold's own rounding). */
int pm_solid3d_ledger(pm_solid3d *solid, pm_books *books);
/* Device-resident loop of Simulation::advance (sim.rs:947-965) for this one domain with an
 * empty bus: `steps` steps of `dt`, each followed by the audit of audit_with (conserved.rs:
 * 323-363) against the previous audited ledger with relative tolerance `tolerance`.  The
 * ledger total after every step is written to totals[0..steps] when totals is not NULL.  On a
 * breach the run stops with the state as the breaking step left it, and PM_VIOLATION names the
 * step.  `elapsed_ms` (optional) receives the device time of the loop, from HIP events on the
 * handle's stream.  With per-cell conductivity and uniform capacity the loop runs two steps per
 * pass over the grid (see pm_solid3d_set_steps_per_pass) with the same arithmetic, results and
 * per-step audits. */
int pm_solid3d_advance(pm_solid3d *solid, double dt, uint64_t steps, double tolerance,
                       double *totals, double *elapsed_ms, pm_violation *v);
/* No counterpart in the reference: pm_solid3d_advance that allocates nothing, for callers that
 * cannot allocate device memory (tight serving loops, code under HIP graph capture, harnesses
 * that forbid hipMalloc).  Same kernels, arithmetic, results, per-step totals and breach
 * behaviour as pm_solid3d_advance; the audit state, the per-workgroup partials (at most 65536 per
 * launch) and the per-step totals (staged 8192 steps at a time) live in one static device array
 * per device (590,848 bytes, held until the process ends).  Plain blocks only: a block with a
 * film, sources, gaps or a phase change, or one whose launches need more partials, is refused
 * with PM_ERR_ARGUMENT before any device work.  Calls that overlap on one device share that array
 * and race: run them one at a time per device. */
int pm_solid3d_advance_static(pm_solid3d *solid, double dt, uint64_t steps, double tolerance,
                              double *totals, double *elapsed_ms, pm_violation *v);
/* Domain::checkpoint / Domain::restore, solid.rs:1935-1950. */
/* An environment a face loses heat to (Environment, crates/pantometry-thermal/src/lib.rs:38-49):
 * ambient temperature in kelvin, convective coefficient in W/m^2/K, and the face's whole area in
 * m^2 (each boundary cell is charged area / cells on the face). */
typedef struct pm_environment {
    double ambient;
    double convection_w_per_m2_k;
    double area;
} pm_environment;

/* Solid3D::losing_from, crates/pantometry-thermal/src/solid.rs:1038-1042: face 0..5 (XMin, XMax,
 * YMin, YMax, ZMin, ZMax) loses heat to `env` by convection and by radiation at each boundary
 * cell's own temperature and its substance's emissivity, inside the sweep (film_flux,
 * solid.rs:1192-1221); a later call for the same face replaces its environment.  The stability
 * limit then depends on the state (worst_rate_now, solid.rs:1148-1185) and is read on the device
 * before every step; what the film removes is kept in the `lost` entry of the books. */
int pm_solid3d_losing_from(pm_solid3d *solid, int face, const pm_environment *env);
/* Solid3D::dissipating, solid.rs:1058-1083: `watts` in total spread evenly over the cells `mask`
 * selects (x-fastest bytes, nonzero = selected), replacing what they generated before; added
 * after each sweep as (W*dt)/C and kept in the `supplied` entry of the books. */
int pm_solid3d_dissipating(pm_solid3d *solid, double watts, const uint8_t *mask);
/* Phase change (solid.rs:423-478, 1563-1648): a substance with a melting point and a latent heat
 * carries a melted fraction per cell and every heat that reaches the cell goes through its
 * enthalpy; a substance with a distinct liquid mixes its conductivity and capacity by that
 * fraction and the block's operator is rebuilt after every step. */
/* Solid3D::set_temperature (solid.rs:1346-1361) on every cell under `mask`: the region
 * initial temperature of a scene (crates/pantometry-world/src/lib.rs:2290-2305). */
int pm_solid3d_set_region(pm_solid3d *solid, const uint8_t *mask, double kelvin);
/* Solid3D::set_melted_fraction, solid.rs:448-466 (ignored for a cell that does not melt). */
int pm_solid3d_set_melted_fraction(pm_solid3d *solid, size_t i, size_t j, size_t k, double fraction);
/* melted_fraction_at (solid.rs:427-430) for every cell into out[n] (zeros for a block that
 * cannot melt), and melted_volume (solid.rs:476-478) in m^3. */
int pm_solid3d_melted(pm_solid3d *solid, double *out, double *volume);
/* How many cells can melt (latent heat above zero): the condition of the "melted" reading. */
size_t pm_solid3d_melting_cells(const pm_solid3d *solid);
/* Solid3D::empty, solid.rs:992-1005: the cells under `mask` hold nothing from now on (a fill
 * does not bring them back).  A void cell conducts and stores nothing; solid cells facing each
 * other across a run of void along a grid line exchange radiation sigma A (Ta^4 - Tb^4)/
 * (1/ea + 1/eb - 1) inside the step (solid.rs:791-847, 1866-1878), which also bounds the step
 * (solid.rs:1159-1171).  Void takes no share of the bus heat, generates nothing, has no
 * temperature in the readings or a sample. */
int pm_solid3d_empty(pm_solid3d *solid, const uint8_t *mask);
/* Solid3D::void_cells and is_void, solid.rs:1007-1015. */
size_t pm_solid3d_void_cells(const pm_solid3d *solid);
int pm_solid3d_is_void(const pm_solid3d *solid, size_t i, size_t j, size_t k);
/* The gap pairs in the reference's order (cell a, cell b, coefficient in W/K^4); at most
 * `capacity` written, the number of pairs returned. */
size_t pm_solid3d_gap_pairs(const pm_solid3d *solid, int64_t *a, int64_t *b, double *coef,
                            size_t capacity);
/* ScalarField::at for Solid3D, crates/pantometry-thermal/src/solid.rs:2045-2102: the temperature
 * (kelvin) at `count` points given as x, y, z triples in metres from the block's corner,
 * trilinear between cell centres and clamped at the outer half cells; NaN for a NaN point.  What
 * scene::capture reads a block through. */
int pm_solid3d_sample(pm_solid3d *solid, const double *xyz, size_t count, double *out);
/* Solid3D::generated_power, solid.rs:1119-1121: the watts all cells generate. */
int pm_solid3d_generated_power(const pm_solid3d *solid, double *watts);
int pm_solid3d_checkpoint(pm_solid3d *solid);
int pm_solid3d_restore(pm_solid3d *solid);
/* Domain::readings, solid.rs:1962-2012: peak, mean and coldest in degrees Celsius, then the
 * absorbed joules. */
int pm_solid3d_readings(pm_solid3d *solid, double out[4]);
/* peak_temperature, mean_temperature and coldest_temperature (solid.rs:1377-1405) in kelvin, over
 * the cells that are not void (the mean NaN when every cell is). */
int pm_solid3d_statistics(pm_solid3d *solid, double out[3]);
/* Solid3D::heat_capacity, solid.rs:1441-1443: the left-fold sum of the cell capacities. */
double pm_solid3d_heat_capacity(pm_solid3d *solid);
/* Shape of the sweep: rows per 64-wide tile (4, 8 or 16), z-planes per workgroup, and how many
 * planes ahead each thread prefetches (1-3); 0 keeps the current value.  Until set, the z-planes
 * per workgroup follow the extents: at most 128, at least 4, and about 1024 workgroups per
 * launch.  No reference counterpart. */
int pm_solid3d_set_tile(pm_solid3d *solid, int tile_y, int z_chunk, int prefetch);
/* How many steps pm_solid3d_advance takes per pass over the grid: 1, or 2 (the default), which
 * applies to per-cell conductivity with uniform capacity and runs consecutive pairs of steps in
 * one kernel (temporal blocking: each cell's T and k are read once per two steps; bit-identical
 * fields and the same audit after every step).  pm_solid3d_steps_per_pass reports what advance
 * will use for this block.  No reference counterpart. */
int pm_solid3d_set_steps_per_pass(pm_solid3d *solid, int steps);
int pm_solid3d_steps_per_pass(const pm_solid3d *solid);
/* When `on`, pm_solid3d_advance brackets every sweep launch and every audit launch with HIP
 * events on the handle's stream; pm_solid3d_kernel_times then gives the mean device duration of
 * each per launch over the last advance, and the steps one such launch advanced (a trailing
 * single step of a paired run is left out).  No reference counterpart (measurement only). */
int pm_solid3d_set_profiling(pm_solid3d *solid, int on);
int pm_solid3d_kernel_times(pm_solid3d *solid, double *sweep_ms, double *audit_ms,
                            int *steps_per_launch);

/* ---------------------------------------------------------------------------------------- */
/* Hall, crates/pantometry-acoustic/src/hall.rs                                             */
/* ---------------------------------------------------------------------------------------- */

typedef struct pm_hall pm_hall;

/* Hall::new, hall.rs:99-134: nx = max(across, 2), dx = width/(nx-1), height and depth
 * quantised to whole cells. */
int pm_hall_create(const char *name, double width, double height, double depth, size_t across,
                   double speed, double density, pm_hall **out);
void pm_hall_destroy(pm_hall *hall);
int pm_hall_nodes(const pm_hall *hall, size_t nodes[3]);
double pm_hall_spacing(const pm_hall *hall);
/* Hall::released_from, hall.rs:166-184, with the profile already evaluated at every node:
 * pressure set, pressure_prev = pressure, velocities zero, unstaggered, datum = invariant. */
int pm_hall_release(pm_hall *hall, const double *pressure);
/* A point source at one node: after the pressure sweep of step n the node gains
 * increments[n] Pa, and the energy that puts into the invariant is booked as supplied.  No
 * reference counterpart (the reference hall has no source term). */
int pm_hall_set_source(pm_hall *hall, size_t node, const double *increments, uint64_t count);
/* Domain::max_stable_dt (hall.rs:399-405) and Hall::courant (hall.rs:320-322). */
double pm_hall_max_stable_dt(const pm_hall *hall);
double pm_hall_courant(const pm_hall *hall, double dt);
/* Domain::step, hall.rs:408-530: Courant guard, half step the first time, velocity then
 * pressure sweeps, and the startup check of the first step. */
int pm_hall_step(pm_hall *hall, double dt, pm_violation *v);
/* ScalarField::at for Hall, crates/pantometry-acoustic/src/hall.rs:590-617: the pressure (Pa) at
 * `count` points given as x, y, z triples in metres from the corner node, trilinear between nodes
 * and clamped at the walls (p[0] for a NaN point). */
int pm_hall_sample(pm_hall *hall, const double *xyz, size_t count, double *out);
/* Hall::energy (hall.rs:249-251) split into the invariant (hall.rs:334-384) and the startup
 * offset, plus the energy a source has supplied. */
int pm_hall_energy(pm_hall *hall, double *invariant, double *offset, double *supplied);
/* Device-resident loop: `steps` steps, each audited (energy - supplied against the previous
 * step's, relative tolerance `tolerance`, after the first step's startup).  The ledger total
 * after every step goes to totals[0..steps] when not NULL. */
int pm_hall_advance(pm_hall *hall, double dt, uint64_t steps, double tolerance, double *totals,
                    double *elapsed_ms, pm_violation *v);
/* All five arrays: p and pressure_prev [nx ny nz], vx [(nx-1) ny nz], vy [nx (ny-1) nz],
 * vz [nx ny (nz-1)].  Any pointer may be NULL. */
int pm_hall_get_fields(pm_hall *hall, double *p, double *pp, double *vx, double *vy, double *vz);
/* Domain::checkpoint / restore, hall.rs:536-555.  With a source set, its booked energy and its
 * position in the increments go back with the fields. */
int pm_hall_checkpoint(pm_hall *hall);
int pm_hall_restore(pm_hall *hall);
/* Hall::peak_pressure, hall.rs:242-244. */
double pm_hall_peak_pressure(pm_hall *hall);
/* Shape of the step, single and paired: rows per 64-wide column tile (8, the default, or 16; 0
 * keeps the current) and z-planes per workgroup (0: chosen from the extents).  Results are the
 * same for every shape.  No reference counterpart. */
int pm_hall_set_tile(pm_hall *hall, int tile_y, int z_chunk);
/* Width of the paired step's column tiles: 64, 56 or 48 nodes, or 0 (the default: 64).  Results
 * are the same for every width.  No reference counterpart. */
int pm_hall_set_tile_width(pm_hall *hall, int tile_x);
/* How many leapfrog steps pm_hall_advance takes per pass over the grid: 1, or 2 (the default):
 * pairs of steps in one kernel (temporal blocking: p, vx, vy, vz read once per two steps), the
 * same arithmetic, fields and per-step audits.  The half step that starts the leapfrog always
 * runs alone.  No reference counterpart. */
int pm_hall_set_steps_per_pass(pm_hall *hall, int steps);

/* ---------------------------------------------------------------------------------------- */
/* Room, crates/pantometry-acoustic/src/room.rs: a hall one node deep                       */
/* ---------------------------------------------------------------------------------------- */

/* Room::new, room.rs:90-120: nx = max(cells_across, 3), dx = width/(nx-1),
 * ny = max(round(height/dx) + 1, 3), `depth` metres deep.  The handle is a pm_hall with nz = 1
 * and every pm_hall_* call applies, under the room's rules: the Courant number
 * c dt sqrt(2)/dx and its limit dx/(c sqrt 2) (room.rs:336-338, 361-366, 369-379), the room's
 * invariant for energy and startup (room.rs:243-281, 460-475), sampling in x and y only
 * (room.rs:587-600), no source.  pm_hall_advance audits each step's energy on the device. */
int pm_room_create(const char *name, double width, double height, size_t cells_across,
                   double depth, double density, double speed, pm_hall **out);
/* 2 for a room, 3 for a hall. */
int pm_hall_dimensions(const pm_hall *hall);

/* ---------------------------------------------------------------------------------------- */
/* Bar1D, crates/pantometry-thermal/src/lib.rs:425-660                                      */
/* ---------------------------------------------------------------------------------------- */

typedef struct pm_bar1d pm_bar1d;

/* Bar1D::new, lib.rs:443-467: max(cells, 2) cells of `dx` metres and `area` m^2, all at
 * `initial` kelvin, which is also the datum of the stored heat (lib.rs:535-538).  A substance
 * whose conductivity or specific heat is NaN has no diffusivity (its steps are refused). */
int pm_bar1d_create(const char *name, const pm_substance *substance, size_t cells, double dx,
                    double area, double initial, pm_bar1d **out);
void pm_bar1d_destroy(pm_bar1d *bar);
size_t pm_bar1d_cells(const pm_bar1d *bar);
/* Cell temperatures in and out (kelvin, cell 0 first). */
int pm_bar1d_set_cells(pm_bar1d *bar, const double *cells);
int pm_bar1d_get_cells(pm_bar1d *bar, double *cells);
/* Bar1D::fourier_number (lib.rs:546-551) and Domain::max_stable_dt (lib.rs:563-568). */
double pm_bar1d_fourier_number(const pm_bar1d *bar, double dt);
double pm_bar1d_max_stable_dt(const pm_bar1d *bar);
/* The guard of Bar1D::step (lib.rs:571-587) alone, so a binding can refuse before it takes its
 * bus share: PM_VIOLATION with "Fourier number" past alpha dt/dx^2 = 1/2. */
int pm_bar1d_refuses(const pm_bar1d *bar, double dt, pm_violation *v);
/* Domain::step for Bar1D, lib.rs:571-627: `gained` joules (the bus share) into cell 0, then
 * joules[i] into cell i when `joules` is not NULL (the resolved flux, one value per cell), then
 * the insulated three-point sweep.  Both count as absorbed. */
int pm_bar1d_step(pm_bar1d *bar, double dt, double gained, const double *joules, pm_violation *v);
/* Bar1D::ledger (lib.rs:629-631): the stored heat, and the absorbed total (lib.rs:541-543). */
int pm_bar1d_ledger(pm_bar1d *bar, double *stored, double *absorbed);
/* Device-resident loop: `steps` steps with no bus input in one launch per 65536 steps, the
 * stored heat audited after every step against the step before (relative tolerance), totals
 * after each step to totals[0..steps] when not NULL.  A breach stops with the breaking step's
 * state. */
int pm_bar1d_advance(pm_bar1d *bar, double dt, uint64_t steps, double tolerance, double *totals,
                     double *elapsed_ms, pm_violation *v);
/* ScalarField::at for Bar1D, lib.rs:736-751: the temperature (K) at `count` positions x in
 * metres along the bar, linear between cell centres and held past the ends. */
int pm_bar1d_sample(pm_bar1d *bar, const double *x, size_t count, double *out);
/* What Bar1D::readings reports (lib.rs:644-660), in kelvin and joules: the mean temperature
 * (left fold over the cells, over their count), the peak, and the absorbed total. */
int pm_bar1d_readings(pm_bar1d *bar, double out[3]);
/* Domain::checkpoint / restore, lib.rs:633-639 (the saved cells start as the initial ones). */
int pm_bar1d_checkpoint(pm_bar1d *bar);
int pm_bar1d_restore(pm_bar1d *bar);

/* ---------------------------------------------------------------------------------------- */
/* Tube, crates/pantometry-acoustic/src/lib.rs:136-563                                      */
/* ---------------------------------------------------------------------------------------- */

typedef struct pm_tube pm_tube;

/* Tube::new, lib.rs:167-193: max(cells, 3) pressure samples over `length` metres, closed ends,
 * at rest. */
int pm_tube_create(const char *name, double length, size_t cells, double area, double density,
                   double speed, pm_tube **out);
void pm_tube_destroy(pm_tube *tube);
size_t pm_tube_cells(const pm_tube *tube);
double pm_tube_spacing(const pm_tube *tube);
/* Tube::with_ends, lib.rs:211-215: each end 0 closed, 1 open, 2 matched to the given
 * impedance (Pa s/m; <= 0 means the tube's own, lib.rs:380-383). */
int pm_tube_set_ends(pm_tube *tube, int left, double z_left, int right, double z_right);
/* Tube::released_from, lib.rs:226-236, with the profile evaluated at every sample. */
int pm_tube_release(pm_tube *tube, const double *pressure);
/* Domain::max_stable_dt (lib.rs:486-514) and Tube::courant (lib.rs:358-360). */
double pm_tube_max_stable_dt(const pm_tube *tube);
double pm_tube_courant(const pm_tube *tube, double dt);
/* Domain::step, lib.rs:516-533: the Courant guard, then Tube::advance (lib.rs:389-464);
 * `published` gets the joules the ends absorbed, which the caller puts on the bus. */
int pm_tube_step(pm_tube *tube, double dt, double *published, pm_violation *v);
/* Tube::energy split into the invariant (lib.rs:300-327) and the startup offset, plus
 * Tube::radiated_energy (lib.rs:330-332). */
int pm_tube_energy(pm_tube *tube, double *invariant, double *offset, double *radiated);
/* Device-resident loop: `steps` steps, each audited (energy plus what the ends radiated,
 * against the step before), after the first step's startup.  Totals as pm_hall_advance. */
int pm_tube_advance(pm_tube *tube, double dt, uint64_t steps, double tolerance, double *totals,
                    double *elapsed_ms, pm_violation *v);
/* pressure [n], pressure_prev [n], velocity [n-1]; any pointer may be NULL. */
int pm_tube_get_fields(pm_tube *tube, double *p, double *pp, double *u);
/* Domain::checkpoint / restore, lib.rs:539-557. */
int pm_tube_checkpoint(pm_tube *tube);
int pm_tube_restore(pm_tube *tube);

/* ---------------------------------------------------------------------------------------- */
/* Ensemble of independent heat grids, crates/pantometry-core/src/ensemble.rs:102-147       */
/* ---------------------------------------------------------------------------------------- */

/* One member: a uniform Solid3D (counts, spacing, substance, initial = ledger reference) and
 * the step it takes. */
typedef struct pm_member {
    size_t nx, ny, nz;
    double dx;
    pm_substance substance;
    double initial;
    double dt;
} pm_member;

typedef struct pm_ensemble pm_ensemble;

/* Each member holds fewer than 2^29 cells (PM_ERR_ARGUMENT otherwise). */
int pm_ensemble_create(const pm_member *members, size_t count, pm_ensemble **out);
void pm_ensemble_destroy(pm_ensemble *ens);
/* Member cells in, out (kelvin, x-fastest). */
int pm_ensemble_set_cells(pm_ensemble *ens, size_t member, const double *cells);
int pm_ensemble_get_cells(pm_ensemble *ens, size_t member, double *cells);
/* `steps` steps of every member in one graph launch; each member's step is refused first if it
 * is past that member's limit (PM_VIOLATION names the member in `step`). */
int pm_ensemble_run(pm_ensemble *ens, uint64_t steps, double *elapsed_ms, pm_violation *v);
/* Per-member stored heat (Solid3D::ledger's stored entry), index order. */
int pm_ensemble_ledgers(pm_ensemble *ens, double *stored);
/* Fix the schedule: members whose largest edge is at most `resident_edge` (and whose two copies
 * fit in LDS) run resident, the rest streamed in z-chunks of `z_chunk` planes.  No reference
 * counterpart. */
int pm_ensemble_set_schedule(pm_ensemble *ens, int resident_edge, int z_chunk);
/* Streamed tile shapes: 0 (default) whole rows when a row is at most 128 cells, as many as fit
 * 512 cells, else 64 x 8; 1 half as many whole rows (else 64 x 4); 2 whole rows up to 64 cells,
 * wider rows in 64 x 8 tiles; 3 32-column tiles.  A member whose tile under the policy does not
 * fit (512 cells, 1024 with a two-cell halo) keeps policy 0.  Results are identical under every
 * policy.  No reference counterpart. */
int pm_ensemble_set_tiles(pm_ensemble *ens, int policy);
/* Steps per launch of the streamed members: 2 (default; temporal blocking, each launch advances
 * two steps with the same arithmetic) or 1.  Results are identical either way.  No reference
 * counterpart. */
int pm_ensemble_set_steps_per_pass(pm_ensemble *ens, int steps);
/* Time the candidate schedules on this ensemble (a few steps each, state restored afterwards):
 * every tile policy x resident edge (0, 16, 20, 24) x z-chunk (16, 32, 64, 128), and keep the
 * fastest.  `report` (optional, `report_len` bytes) receives one line per candidate. */
int pm_ensemble_tune(pm_ensemble *ens, char *report, size_t report_len);

/* ---- The learned heat surrogate ------------------------------------------------------------
 * No reference counterpart: pantometry has no learned component.  A four-level 3D U-Net
 * (channels 16-32-64-128, two 3x3x3 convolutions with tanh-GELU per level, 2x2x2 average pool,
 * nearest upsampling with skip concatenation, 1x1 head; 1,466,065 parameters) mapping a cubic
 * block's (T, log k) and the stencil's per-step coupling a = dt * dx / C to (T_16 - T_0) / 10 K,
 * the change over 16 stencil steps, trained with Adam on bf16 matrix-core kernels (fp32
 * accumulation, fp32 master weights).  Its sixteen input channels are (T - 300 K) / 100 K,
 * log k / 0.5, and per face (x-, x+, y-, y+, z-, z+) 10 a k_f and a k_f (T_nb - T) * 16 / 10 K
 * (k_f the harmonic face mean, zero on the insulated surface), their sum, and a constant one.
 * Edges must be powers of two from 8.  Parameters are one flat fp32 array, layer by layer:
 * weights [Cout][27][Cin] (tap t = 9 dz + 3 dy + dx) then biases [Cout]; the head's 16 weights
 * and 1 bias last. */
typedef struct pm_surrogate pm_surrogate;

int pm_surrogate_create(uint64_t seed, pm_surrogate **out);
void pm_surrogate_destroy(pm_surrogate *s);
int64_t pm_surrogate_parameter_count(const pm_surrogate *s);
int pm_surrogate_get_parameters(pm_surrogate *s, float *out);
int pm_surrogate_set_parameters(pm_surrogate *s, const float *in);
/* Reserves a device-resident training set of `samples` pairs of edge^3 cells (T_0 and k as fp32,
 * the normalised change as fp32, the coupling per pair). */
int pm_surrogate_reserve_set(pm_surrogate *s, int64_t samples, int64_t edge);
/* Pair `index`: T_0, conductivity and T_16 (kelvin, W/m/K; x fastest) and the coupling
 * a = dt * dx / C of the steps between them. */
int pm_surrogate_set_sample(pm_surrogate *s, int64_t index, const double *t0, const double *k,
                            const double *tk, double coupling);
/* Pair `index` taken from two Solid3D blocks on the device: the state before and after 16
 * stencil steps of `dt` (both cubic with the set's edge, the same per-cell conductivity and a
 * uniform capacity). */
int pm_surrogate_set_sample_from(pm_surrogate *s, int64_t index, pm_solid3d *before,
                                 pm_solid3d *after, double dt);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on the set's pairs `indices[0..batch)`; `loss`
 * receives 0.5 * mean squared error of the normalised change, before the step. */
int pm_surrogate_train_step(pm_surrogate *s, const int *indices, int batch, float lr,
                            double *loss);
/* The loss and its gradient with respect to every parameter (flat layout), no step taken. */
int pm_surrogate_gradients(pm_surrogate *s, const int *indices, int batch, float *grads,
                           double *loss);
/* The network's normalised output for one block of edge^3 cells and its coupling. */
int pm_surrogate_forward(pm_surrogate *s, int64_t edge, const double *t, const double *k,
                         double coupling, float *y);
/* One surrogate step, standing in for 16 stencil steps of `dt`, of a cubic block with per-cell
 * conductivity and uniform capacity, in place: T <- T + 10 K * net, shifted uniformly on the
 * device so the block's total heat is unchanged.  `drift` receives the relative change of
 * sum(T - reference) across the step, `elapsed_ms` the device time. */
int pm_surrogate_step_solid(pm_surrogate *s, pm_solid3d *block, double dt, double reference,
                            double *drift, double *elapsed_ms);
/* ||T_a - T_b||_2 / ||T_b - reference||_2 over two blocks of the same extents. */
int pm_solid3d_relative_l2(pm_solid3d *a, pm_solid3d *b, double reference, double *rel);
/* Its two norms: ||T_a - T_b||_2 into `diff` and ||T_b - reference||_2 into `norm`. */
int pm_solid3d_l2_parts(pm_solid3d *a, pm_solid3d *b, double reference, double *diff,
                        double *norm);
/* One bf16 convolution on host fp32 arrays (rounded to bf16 on the way in): x [B][D][H][W][Cin],
 * w [Cout][27][Cin], bias [Cout] -> y [B][D][H][W][Cout] before the activation.  `split` lets a
 * small grid use the split-reduction workspace (as the surrogate does); 0 runs it without. */
int pm_nn_conv3d(int B, int D, int H, int W, int Cin, int Cout, const float *x, const float *w,
                 const float *bias, float *y, int split);
/* Its gradients for an upstream gradient dy: dx (bf16-rounded), dw and db (fp32). */
int pm_nn_conv3d_backward(int B, int D, int H, int W, int Cin, int Cout, const float *x,
                          const float *w, const float *dy, float *dx, float *dw, float *db,
                          int split);

#ifdef __cplusplus
}
#endif

#endif
