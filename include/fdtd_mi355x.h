/*
 * C-ABI of the MI355X (gfx950) Yee-step library, libfdtd_mi355x.so.
 *
 * This is the drop-in boundary for the reference's FDTD class API
 * (include/FDTD/FDTD.h:8-41, include/FDTD_kokkos/FDTD_kokkos.h:7-36): plain
 * C types, opaque handles, int status codes.  No exception and no C++ or
 * torch type crosses it.  The C++ class FDTD_mi355x::FDTD (FDTD_mi355x.hpp)
 * is the header-only counterpart of the reference class built on these
 * entry points; Python binds them with ctypes (fdtd_method_amd/_lib.py).
 *
 * Layout: every field is a flat array of N = Ni*Nj*Nk cells, index
 * i + j*Ni + k*Ni*Nj (src/FDTD/FDTD.cpp:66-78).  Host buffers are always
 * FP64 (the reference's FP, include/FP.h:3); the device stores fields in the
 * precision chosen at creation.
 *
 * Threading: one HIP stream per grid; calls on one grid are serialised by
 * the caller.  Stepping is asynchronous; downloads and reductions
 * synchronise the grid's stream.
 */
#ifndef FDTD_MI355X_H
#define FDTD_MI355X_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The messages of the first two match the reference's
 * exceptions (src/FDTD/FDTD.cpp:6 and :149). */
#define FDTD_OK 0
#define FDTD_ERR_INVALID_PARAMETERS (-1) /* "ERROR: invalid parameters"       */
#define FDTD_ERR_INVALID_COMPONENT (-2)  /* "ERROR: Invalid field component"  */
#define FDTD_ERR_INVALID_ARGUMENT (-3)
#define FDTD_ERR_HIP (-4)
#define FDTD_ERR_UNSUPPORTED (-5)

/* Component ordinals of include/Enums.h:5. */
#define FDTD_EX 0
#define FDTD_EY 1
#define FDTD_EZ 2
#define FDTD_BX 3
#define FDTD_BY 4
#define FDTD_BZ 5
#define FDTD_JX 6
#define FDTD_JY 7
#define FDTD_JZ 8

/* Device storage precision.  FP64 is the parity path; FP32 stores and
 * computes in float; FP16 stores _Float16 with per-component power-of-two
 * scales and computes in float. */
#define FDTD_FP64 0
#define FDTD_FP32 1
#define FDTD_FP16 2

/* Stepping paths for fdtd_time_update_fields. */
#define FDTD_PATH_FUSED 0
#define FDTD_PATH_3SWEEP 1

/* include/Structures.h:23-38 (FDTD_struct::Parameters), same field order. */
typedef struct fdtd_parameters {
    int Ni, Nj, Nk;
    double ax, bx, ay, by, az, bz;
    double dx, dy, dz;
} fdtd_parameters;

/* A point current source on component comp (0=Jx, 1=Jy, 2=Jz) of cell
 * (i,j,k), evaluated in FP64 on the device at t = (n+1)*dt before step n
 * (the sampling of perf-tests/sample/sample.cpp:74):
 *   kind 0: amp
 *   kind 1: amp*sin(2*pi*(t-t0)/period + phase)*exp(-(t-t0)^2/(2*sigma^2))
 *           (no envelope when sigma <= 0); pi = 3.14159265358 as in
 *           include/Constants.h:11. */
typedef struct fdtd_point_source {
    int i, j, k, comp;
    int kind, reserved;
    double amp, period, t0, sigma, phase;
} fdtd_point_source;

typedef struct fdtd_grid fdtd_grid;   /* one Yee grid and its HIP stream */
typedef struct fdtd_batch fdtd_batch; /* many grids stepped by one launch */

/* Message of the last failing call on this host thread ("" if none). */
const char* fdtd_last_error(void);

/* FDTD(Parameters, FP dt) -- src/FDTD/FDTD.cpp:3-61,
 * src/FDTD_kokkos/FDTD_kokkos.cpp:5-67.  Fails with
 * FDTD_ERR_INVALID_PARAMETERS if any N <= 0 or dt <= 0.  All nine fields
 * start at zero. */
int fdtd_create(const fdtd_parameters* params, double dt, int precision, fdtd_grid** out);
int fdtd_destroy(fdtd_grid* grid);

/* FDTD_PML(Parameters, FP dt, FP pml_percent) -- src/FDTD/FDTD_PML.cpp:205-341,
 * src/FDTD_kokkos/FDTD_PML_kokkos.cpp:22-157: a grid whose outer
 * (int)(N*pml_percent) cells per side and axis form an absorbing layer
 * (split fields, graded conductivity sigma = SGm*(dist/width)^4).  Every
 * other call behaves as for fdtd_create's grids, with the class's step
 * (FDTD_PML.cpp:343-365): J acts on the main box only.  pml_percent must lie
 * in [0, 0.5] (the reference's layer slabs overlap beyond 0.5); when every
 * width is 0 the grid is a plain periodic one, as in the reference.  Split
 * fields start at zero.  Temporal blocking and batches take periodic grids
 * only (FDTD_ERR_UNSUPPORTED). */
int fdtd_create_pml(const fdtd_parameters* params, double dt, double pml_percent, int precision,
                    fdtd_grid** out);
/* How a PML grid's fused steps run (no counterpart in the reference, whose
 * step is FDTD_PML.cpp:343-365): path 0 (the default) as the single fused
 * PML sweep (shape: fdtd_set_tuning); path 1 as two in-place sweeps per
 * step, an E phase and a B phase, with V cells per lane, TY full rows per
 * workgroup and kchunk planes per z chunk (0: defaults; V*storage bytes at
 * most 16).  At 512^3 with a 0.2 layer the phases take 3.94 ms per FP32
 * step against the fused sweep's 3.56 (FP64: 7.51 both).  FP64
 * and FP32 results are bit-identical either way (and FP64 to the
 * reference); with FP16 storage the phase path's B update reads E as
 * stored where the fused sweep uses it before rounding, so the two agree
 * to FP16 rounding.  FP16 sweeps that rescale or track the max run the
 * fused sweep.  get: {path, V, TY, kchunk}. */
int fdtd_set_pml_path(fdtd_grid* grid, int path, int V, int TY, int kchunk);
int fdtd_get_pml_path(const fdtd_grid* grid, int out[4]);
/* Layer widths {p_i, p_j, p_k} (FDTD_PML.cpp:254-256) and pml_percent
 * (widths 0 and pml_percent 0 for fdtd_create's grids). */
int fdtd_get_pml(const fdtd_grid* grid, int widths[3], double* pml_percent);
/* FP64 copy of split field q of a PML grid, the reference's private members
 * Exy Exz Eyx Eyz Ezx Ezy (q = 0..5) and Bxy Bxz Byx Byz Bzx Bzy (6..11),
 * include/FDTD/FDTD_PML.h:12-13; zeros for a grid without a layer. */
int fdtd_pml_download_split(fdtd_grid* grid, int q, double* host, size_t count);
/* Host-only (no GPU needed): the layer profile along one axis of n cells
 * with spacing d -- conductivity sigma (FDTD_PML.cpp:3-61,316-338), decay
 * exp(-sigma*dt*C) (:63-65) and split coefficient (1-decay)/(sigma*d), or
 * C*dt/d where sigma = 0 (:98-111).  Any output may be NULL. */
int fdtd_pml_profile(int n, double pml_percent, double d, double dt, double* sigma,
                     double* decay, double* coef);

/* Write side of get_field(Component) (src/FDTD/FDTD.cpp:138-151): copy a
 * full FP64 host array into component `component` (0..8). count must be N. */
int fdtd_upload(fdtd_grid* grid, int component, const double* host, size_t count);
/* Read side of get_field(Component): FP64 copy of the current values. */
int fdtd_download(fdtd_grid* grid, int component, double* host, size_t count);

/* FP64 copy of plane k (Ni*Nj values, index i + j*Ni) of component 0..8:
 * a slice for output without moving the whole field (the CSV slices read
 * by python_script_legend/visualization.py:10-45 are planes). */
int fdtd_download_plane(fdtd_grid* grid, int component, int k, double* host, size_t count);

/* zeroed_currents() -- src/FDTD/FDTD.cpp:132-136: J := 0 (dense J and the
 * values set by fdtd_set_current_values; point sources are kept).  Once
 * J's device addresses are handed out (fdtd_device_field), the arrays are
 * cleared in place, queued on the grid's stream, as the Kokkos class's
 * deep_copy(J, 0) (src/FDTD_kokkos/FDTD_kokkos.cpp:69-73). */
int fdtd_zero_currents(fdtd_grid* grid);

/* Device side of get_field(Component) -- the Kokkos class returns device
 * Views (include/FDTD_kokkos/FDTD_kokkos.h:31, src/FDTD_kokkos/FDTD_kokkos.cpp:75-88)
 * that its driver fills in its own device loop
 * (perf-tests/kokkos_sample/kokkos_sample.cpp:82-108): the device address
 * of component 0..8's array of Ni*Nj*Nk elements, index i + j*Ni + k*Ni*Nj,
 * for the caller's own GPU kernels.
 *   J (6..8): elements of *elem_bytes = 8 (double) on FP64 grids, 4 (float)
 *     otherwise, physical units (*scale = 1).  J becomes dense (zeroed, the
 *     single-cell values set so far copied in) and stays dense at the same
 *     addresses for the grid's life: the steps read it, fdtd_zero_currents
 *     clears it in place.  Dense J turns temporal blocking off.
 *   E, B (0..5): elements of the storage type (8 double, 4 float, 2 _Float16
 *     holding scale * physical value).  The address holds the state between
 *     steps and is valid until the next stepping call (sweeps alternate
 *     between two buffers): fetch it again after stepping.  Each fetch drops
 *     the carried half step, since the caller may change E or B.
 * Ordering: the grid's work is queued on its own stream (fdtd_get_stream).
 * Launch on that stream, or synchronize: the grid's stream
 * (fdtd_synchronize) before touching the arrays, and your own work before
 * the next call on the grid. */
int fdtd_device_field(fdtd_grid* grid, int component, void** ptr, int* elem_bytes,
                      double* scale);
/* The grid's HIP stream (a hipStream_t) as void*: every call on the grid
 * queues its device work there, in call order. */
int fdtd_get_stream(fdtd_grid* grid, void** stream);

/* The reference's protected curl updates, for subclasses that compose their
 * own update_fields() (FDTD_PML::update_fields, src/FDTD/FDTD_PML.cpp:343-365
 * and src/FDTD_kokkos/FDTD_PML_kokkos.cpp:160-198):
 *   fdtd_update_b_box: update_B() -- src/FDTD/FDTD.cpp:99-130 (one B half
 *     step, B += c*dt/(2d) * (-curl E));
 *   fdtd_update_e_box: update_E() -- src/FDTD/FDTD.cpp:63-97 with the
 *     Kokkos J pairing Jx->Ex, Jy->Ey, Jz->Ez
 *     (include/FDTD_kokkos/kokkos_functors.h:64-90): one E step, with dense
 *     or single-cell J and the point sources of the current step;
 * over the cells [lo, hi) (the begin_main_* .. end_main_* bounds,
 * include/FDTD/FDTD.h:17-18), neighbours wrapped over the whole grid, in
 * place on E and B only (a PML grid's split fields are not touched).
 * Neither advances the step counter.  0 <= lo <= hi <= N per axis; an
 * empty box does nothing.  Queued on the grid's stream without a wait. */
int fdtd_update_b_box(fdtd_grid* grid, const int lo[3], const int hi[3]);
int fdtd_update_e_box(fdtd_grid* grid, const int lo[3], const int hi[3]);

/* Sparse form of get_field(JX|JY|JZ)[index] = value (the drivers'
 * per-step box source, perf-tests/sample/sample.cpp:67-81): sets J at n
 * cells; values persist until changed or zeroed like the reference's J.
 * Writes are applied in order (a cell written twice keeps the last value);
 * into dense J they go as one upload and one scatter kernel per call. */
int fdtd_set_current_values(fdtd_grid* grid, int n, const long long* index,
                            const int* component, const double* value);

/* nsteps steps, each after a batch of sparse current writes: before step s
 * the counts[s] next entries of (index, component, value) are set as by
 * fdtd_set_current_values (they persist until changed or zeroed).  Equal,
 * bit for bit, to that loop of fdtd_set_current_values and
 * fdtd_update_fields(grid, 1) -- the reference drivers' loop
 * (perf-tests/sample/sample.cpp:66-87) -- but the steps run as one call, so
 * temporally blocked passes apply: the cells the call writes (at most 64)
 * become sources whose per-step values come from a device table.  Dense J
 * or more cells fall back to the step-by-step loop. */
int fdtd_update_fields_with_currents(fdtd_grid* grid, int nsteps, const int* counts,
                                     const long long* index, const int* component,
                                     const double* value);

/* Launch counters since creation: {steps, temporally blocked passes,
 * single-step sweeps, B half-step prologues}. */
int fdtd_get_stats(const fdtd_grid* grid, long long out[4]);

/* Probes: the FP64 values (physical units, as stored) of n single field
 * values after every step of fdtd_update_fields, fdtd_update_fields_3sweep
 * and fdtd_update_fields_with_currents, one row of n values per step in a
 * device buffer of `capacity` rows, written by the stepping kernels
 * themselves (fused sweeps, strip passes) or gathered after each step
 * (three-sweep steps, PML grids), with no host wait.  Replaces reading
 * single values of get_field() (include/FDTD/FDTD.h:38) after each
 * update_fields() call of the reference's drivers.  Probe q is cell
 * index[q] (i + j*Ni + k*Ni*Nj) of component[q] (FDTD_EX..FDTD_BZ).  A
 * stepping call that would need more rows than remain fails with
 * FDTD_ERR_INVALID_ARGUMENT before any step runs.  n = 0 removes the
 * probes; a new list restarts at row 0.  Patch-blocked passes record no
 * probes: with probes set, runs take strip passes or single sweeps. */
int fdtd_set_probes(fdtd_grid* grid, int n, const long long* index, const int* component,
                    long long capacity);
/* Rows written so far (queued steps included). */
int fdtd_probe_rows(const fdtd_grid* grid, long long* rows);
/* Rows row0 .. row0+nrows-1 into host (nrows * n doubles, row-major);
 * waits for the grid's stream. */
int fdtd_read_probes(fdtd_grid* grid, long long row0, long long nrows, double* host);

/* Energy records: after every step whose number is a multiple of `every`,
 * the six sums of squares of fdtd_sum_squares (bitwise the same values),
 * reduced on the device into one row of a buffer of `capacity` rows, with
 * no host wait.  No counterpart in the reference, whose drivers would call
 * a full reduction between update_fields() calls.  every = 0 turns them
 * off; a new interval restarts at row 0.  A stepping call that would need
 * more rows than remain fails before any step runs.  Runs are split at the
 * recorded steps, bit-identically (FP16 grids with in-run rescaling take
 * their rescaling decisions within each part). */
int fdtd_set_energy_interval(fdtd_grid* grid, int every, long long capacity);
int fdtd_energy_rows(const fdtd_grid* grid, long long* rows);
/* Rows row0 .. row0+nrows-1: the step of each row into steps (may be
 * null), the six sums of each into sums (6 * nrows doubles); waits for the
 * grid's stream. */
int fdtd_read_energy(fdtd_grid* grid, long long row0, long long nrows, long long* steps,
                     double* sums);
/* Probe rows and energy records restart at row 0 (tables kept). */
int fdtd_rewind_records(fdtd_grid* grid);

/* Replace the list of waveform point sources (evaluated on the device). */
int fdtd_set_point_sources(fdtd_grid* grid, int n, const fdtd_point_source* sources);

/* nsteps calls of update_fields() -- src/FDTD/FDTD.cpp:153-157,
 * src/FDTD_kokkos/FDTD_kokkos.cpp:91-104 -- as one B half step plus
 * nsteps fused E-then-double-half-B sweeps.  Bit-identical to the
 * reference's B/2,E,B/2 sequence in FP64. */
int fdtd_update_fields(fdtd_grid* grid, int nsteps);
/* Same steps in the reference's structure: three in-place sweeps per step
 * (update_B, update_E, update_B -- src/FDTD/FDTD.cpp:63-130). */
int fdtd_update_fields_3sweep(fdtd_grid* grid, int nsteps);

/* Steps completed so far (sources are evaluated at (step+1)*dt). */
int fdtd_get_step(const fdtd_grid* grid, long long* step);
int fdtd_set_step(fdtd_grid* grid, long long step);
int fdtd_synchronize(fdtd_grid* grid);

/* Grid facts: dims[3], precision, cells. */
int fdtd_get_info(const fdtd_grid* grid, int dims[3], int* precision, long long* cells);

/* FP64 diagnostics over E and B (Ex,Ey,Ez,Bx,By,Bz), in physical units:
 * sum of squares (field energy is their total / (8 pi) * dx*dy*dz) and max |x|. */
int fdtd_sum_squares(fdtd_grid* grid, double out[6]);
int fdtd_max_abs(fdtd_grid* grid, double out[6]);
/* fdtd_sum_squares without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as tight
 * diagnostic loops and code that must not call hipMalloc.  Bit-identical to
 * fdtd_sum_squares.  Its partial sums live in one fixed 98,432-byte device
 * array per GPU (the library's, held until the process ends), which fits
 * every grid size (at most 1024 workgroups of partials per component; a
 * larger need would be rejected with FDTD_ERR_INVALID_ARGUMENT).  All calls
 * on one GPU share that array: calls that overlap on one GPU (from several
 * host threads) race.  Synchronises the grid's stream. */
int fdtd_sum_squares_noalloc(fdtd_grid* grid, double out[6]);
/* Non-finite check (no counterpart in the reference, which has no failure
 * detection): the number of NaN and +-Inf values of Ex..Bz as stored (an FP16
 * overflow shows here as Inf).  Synchronises the grid's stream. */
int fdtd_count_nonfinite(fdtd_grid* grid, long long out[6]);
/* Error of grid `test` (any precision) against FP64 grid `ref` of the same
 * shape, per component c: out[3c] = sum (t-r)^2, out[3c+1] = sum r^2,
 * out[3c+2] = max |t-r|. */
int fdtd_compare(fdtd_grid* test, fdtd_grid* ref, double out[18]);

/* Initial conditions: add a plane wave with integer wave numbers m (so it is
 * periodic on the grid; k = 2 pi (m0/Ni, m1/Nj, m2/Nk) per cell), E =
 * amplitude * e * cos(k.r + phase) with e the polarization made
 * perpendicular to k, and B = khat x E (Gaussian units), each component at
 * its Yee position.  Generalises the reference tests' axis-aligned wave
 * (src/FDTD/test_FDTD.cpp:5-51). */
int fdtd_add_plane_wave(fdtd_grid* grid, const int m[3], double amplitude, double phase,
                        const double polarization[3]);

/* Order-independent 64-bit checksum of the stored bit patterns of each of
 * Ex..Bz (sum mod 2^64 of a hash of (bits, index)): equal arrays give equal
 * checksums, so bit-exactness can be checked on the device at any size. */
int fdtd_checksum(fdtd_grid* grid, unsigned long long out[6]);

/* FP16 storage scales (stored = scale * physical), powers of two; the
 * current contents are rescaled exactly.  fdtd_autoscale picks for each
 * component the largest power of two that keeps max|stored| <= 2^target_log2
 * given an expected physical magnitude floor (used when a field is zero). */
int fdtd_set_scales(fdtd_grid* grid, const double scales[6]);
int fdtd_get_scales(const fdtd_grid* grid, double scales[6]);
int fdtd_autoscale(fdtd_grid* grid, int target_log2, double magnitude_floor);
/* FP16: during fdtd_update_fields, every `interval` sweeps rescale the
 * running state so each component's max|stored| lies in
 * [2^(target_log2-1), 2^target_log2) (interval 0 = off). */
int fdtd_set_autoscale_interval(fdtd_grid* grid, int interval, int target_log2);

/* Reset: E, B and J to zero, step counter to 0 (sources are kept). */
int fdtd_zero_fields(fdtd_grid* grid);

/* Carried half step (periodic grids, on by default): the final sweep of a
 * fdtd_update_fields call also stores the next call's first B half step,
 * (B_n + d) + d, in two extra B buffer sets allocated on first use, so
 * consecutive calls -- the reference drivers' one-step loops -- run no B/2
 * prologue.  Results are bit-identical either way; host writes to E or B
 * drop the carried step.  enable = 0 frees the buffers. */
int fdtd_set_step_carry(fdtd_grid* grid, int enable);

/* Fused-sweep launch shape: V cells per lane along x, TY rows per tile,
 * kchunk planes per z chunk (0 = keep current).  fdtd_autotune times the
 * candidate shapes on scratch buffers and keeps the fastest. */
int fdtd_set_tuning(fdtd_grid* grid, int V, int TY, int kchunk);
int fdtd_get_tuning(const fdtd_grid* grid, int out[3]);
int fdtd_autotune(fdtd_grid* grid, int reps, int out[3], float* best_ms);

/* Temporal blocking: fdtd_update_fields advances L steps per pass over HBM
 * (L = 2..4) with overlapped (ghost-zone) x/y tiles of lanes*V columns by
 * TYg rows (ghosts included) marching kchunk planes in z; L = 1 is the same
 * tiled kernel one step per pass (for rows too wide for the full-row
 * sweep); L = 0 returns to the full-row fused sweep.  0 picks a default for
 * V, lanes, TYg, kchunk.  Results are bit-identical to unblocked steps
 * (FP16 with in-run rescaling: to FP16 rounding, as it rescales between passes).
 * get: {L, V, lanes, TYg, kchunk}.  fdtd_autotune_blocking times the
 * full-row sweep and tiled/blocked candidates on scratch copies of the
 * state and keeps the fastest per step (best_ms: ms per step). */
int fdtd_set_time_blocking(fdtd_grid* grid, int L, int V, int lanes, int TYg, int kchunk);
int fdtd_get_time_blocking(const fdtd_grid* grid, int out[5]);
int fdtd_autotune_blocking(fdtd_grid* grid, int reps, int out[5], float* best_ms);
/* Autotuner cache (no counterpart in the reference): fdtd_autotune_blocking
 * keeps its picks in the text file `path`, one line per (shape, precision,
 * device architecture, CU count), and a later call on such a grid, in any
 * process, takes the recorded pick and its ms per step without timing.
 * NULL or "" turns the cache off; the default is $FDTD_MI355X_TUNE_CACHE
 * (unset: off).  Process-wide. */
int fdtd_set_tune_cache(const char* path);

/* Temporal blocking with the wavefront-strip kernel: tiles of 64*V columns
 * by W*R rows (one wavefront per R rows, R rows per lane in registers;
 * x neighbours by lane shifts, y neighbours across wavefronts through a
 * one-row LDS exchange), L = 1..4 steps per pass, kchunk planes per z
 * chunk (0 = default).  Shapes built: V*R*sizeof(compute type) <= 16 bytes
 * with V in {1,2}, R in {2,4} (FP32/FP16 compute in float, FP64 in double).
 * Planes (Ni*Nj stored values) must be below 2 GiB (32-bit plane offsets).
 * Results are bit-identical to unblocked steps (FP16 with in-run rescaling:
 * to FP16 rounding, as it rescales between passes).  get: {L, V, R, W, kchunk},
 * R = W = 0 when the ghost-zone patch kernel (or no blocking) is set. */
int fdtd_set_time_blocking_strips(fdtd_grid* grid, int L, int V, int R, int W, int kchunk);
int fdtd_get_time_blocking_strips(const fdtd_grid* grid, int out[5]);

/* Timed run on the grid's stream with HIP events: ms_total covers the whole
 * call (including the B half step of the fused path), ms_main only the
 * nsteps sweep launches (fused) or step triples (3-sweep). */
int fdtd_time_update_fields(fdtd_grid* grid, int nsteps, int path, float* ms_total, float* ms_main);
/* fdtd_update_fields(grid, nsteps) timed launch by launch (HIP events before
 * and after each stepping kernel on the grid's stream): ms[k] and count[k]
 * are the total duration and number of launches of kind k -- 0 the B half
 * step that starts a call without a carried step, 1 steady passes or sweeps,
 * 2 the call's last pass or sweep (which also stores the carried half step),
 * 3 the single sweeps that follow the passes of a blocked call.  No
 * counterpart in the reference. */
int fdtd_time_update_fields_kernels(fdtd_grid* grid, int nsteps, double ms[4], int count[4]);

/* Copy `bytes` (multiple of 16) device-to-device `iters` times with a
 * streaming kernel; ms = average per copy.  Calibrates byte counters. */
int fdtd_calibration_copy(size_t bytes, int iters, float* ms);

/* ---- Batched solves: many independent grids advanced by one persistent
 * launch per step (a work queue over every grid's tiles).  All grids must
 * share one precision and one step count; grids with waveform sources one
 * dt; currents must be point sources or sparse values (no dense J).  The
 * grids stay usable individually between batch calls. ---- */
int fdtd_batch_create(int n, fdtd_grid* const* grids, fdtd_batch** out);
int fdtd_batch_destroy(fdtd_batch* batch);
/* nsteps steps of every grid (same semantics as fdtd_update_fields). */
int fdtd_batch_update_fields(fdtd_batch* batch, int nsteps);
/* Capture nsteps steps into a HIP graph once and replay it `replays` times. */
int fdtd_batch_update_fields_graph(fdtd_batch* batch, int nsteps, int replays);
int fdtd_batch_synchronize(fdtd_batch* batch);
/* ms per step of the batched launch and of one launch per grid, timed with
 * HIP events over nsteps steps each (fields advance by 2*nsteps when
 * ms_sequential is requested, else by nsteps). */
int fdtd_batch_time(fdtd_batch* batch, int nsteps, float* ms_batched, float* ms_sequential);
/* Launch shape of the batched sweep: V cells per lane, threads per
 * workgroup, target cells per tile (sets each grid's z-chunk).
 * get: out = {V, threads, tile_cells, work items}.  The autotuner times the
 * candidate shapes on scratch copies of the state and keeps the fastest. */
int fdtd_batch_set_shape(fdtd_batch* batch, int V, int threads, long long tile_cells);
int fdtd_batch_get_shape(const fdtd_batch* batch, long long out[4]);
int fdtd_batch_autotune(fdtd_batch* batch, int reps, long long out[4], float* best_ms);
/* Temporal blocking for a batch (no counterpart in the reference): runs of
 * at least L steps advance in persistent strip passes of L steps over the
 * strip tiles of every grid (64*V columns by W*R rows, as
 * fdtd_set_time_blocking_strips; 0 picks V=1, R=2, W=16), the remaining
 * steps in batched single sweeps; bit-identical to single-grid runs.
 * L = 0 turns it off.  get: out = {L, V, R, W, work items}. */
int fdtd_batch_set_time_blocking_strips(fdtd_batch* batch, int L, int V, int R, int W);
int fdtd_batch_get_time_blocking_strips(const fdtd_batch* batch, int out[5]);

#ifdef __cplusplus
}
#endif

#endif /* FDTD_MI355X_H */
