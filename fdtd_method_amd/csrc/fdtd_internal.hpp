// Internal interface of libfdtd_mi355x.so shared by its translation units
// (fdtd_runtime.hip: single grids and the C-ABI; fdtd_batch.hip: batches).
// Not installed; callers use include/fdtd_mi355x.h.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <map>
#include <string>
#include <vector>

#include "fdtd_kernels.hpp"
#include "fdtd_mi355x.h"

namespace fdtd::rt {


constexpr double kC = 3e10;              // include/Constants.h:6
constexpr double kPiRef = 3.14159265358; // include/Constants.h:11
constexpr int kMaxSparseCurrents = 64;   // beyond this J is kept dense
constexpr size_t kLdsLimit = 160 * 1024; // gfx950 LDS per workgroup
constexpr int kReduceBlocks = 1024;
constexpr int kValTable = 256;            // steps of source values per table fill
constexpr double kPmlR = 1e-12;           // include/Constants.h:7  (layer reflection target)
constexpr double kPmlOrder = 4.0;         // include/Constants.h:10 (conductivity grading power)


// Records a message for fdtd_last_error() and returns code.
int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(FDTD_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// A roctx range over a scope (a run, a pass, a sweep): shown by
// rocprofv3 --marker-trace, free when no tool is attached.
struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

struct Tuning {
    int V = 1, TY = 1, kchunk = 1;
};

// Temporal blocking: L steps per pass (L < 2: off), V cells per lane,
// lanes x TYg patch per workgroup (ghosts included), kchunk planes per
// z chunk.
struct BlockTuning {
    int L = 0, V = 2, lanes = 64, TYg = 16, kchunk = 64;
    int R = 0;  // > 0: wavefront-strip kernel, R rows per lane, TYg / R wavefronts
};

inline size_t storage_size(int precision) {
    return precision == FDTD_FP64 ? 8 : (precision == FDTD_FP32 ? 4 : 2);
}

// x geometry of a strip tile: a window of 64*V columns with Gx ghost columns
// on each side (at least L, a multiple of V) and TX stored columns; tiles
// start at multiples of TX.  With 4-byte storage, TX is a multiple of 16
// columns, so every stored row is whole, aligned 64-byte granules: stores of
// 56-column (224-byte) rows that begin mid-granule ran at about half the
// bandwidth (measured at 512^3: the pass took 2.16 ms with 56 stored columns
// and 1.90 ms with 48, bit-identical, although the narrower tiles compute
// 10% more cells).  FP64 rows of 56 doubles are whole granules already; FP16
// keeps 56 columns (a granule-aligned FP16 tile would store 32).  Batches of
// small grids take the unaligned width (align false): most of their rows
// hold one to three tiles, whose first tile is aligned anyway, and the 11%
// more tiles of the aligned width cost more than the aligned stores save
// (256-grid batch, same box: 236.0 Gcell/s with 56 stored columns, 221.5
// with 48).
inline void strip_x_geometry(int L, int V, size_t esize, int* Gx, int* TX, bool align = true) {
    *Gx = (L + V - 1) / V * V;
    *TX = 64 * V - 2 * *Gx;
    if (align && esize == 4) {
        const int aligned = *TX / 16 * 16;
        if (aligned >= 16 && aligned % V == 0 && aligned < *TX) {
            *TX = aligned;
            *Gx = (64 * V - aligned) / 2;
        }
    }
}
inline size_t compute_size(int precision) { return precision == FDTD_FP64 ? 8 : 4; }
inline size_t current_size(int precision) { return precision == FDTD_FP64 ? 8 : 4; }


}  // namespace fdtd::rt

namespace fdtd::rt {
struct LaunchTrace;
}

struct fdtd_grid {
    fdtd_parameters params{};
    int Ni = 0, Nj = 0, Nk = 0;
    size_t N = 0;
    double dt = 0;
    int precision = FDTD_FP64;
    size_t esize = 8;
    void* E[2][3] = {};
    void* B[2][3] = {};
    int cur = 0;
    void* J[3] = {};
    bool j_dense_active = false;
    // J's device addresses were handed out (fdtd_device_field): J stays
    // dense and is cleared in place, since the caller may write it later.
    bool j_device = false;
    double* staging = nullptr;
    double scales[6] = {1, 1, 1, 1, 1, 1};
    hipStream_t stream = nullptr;
    int device = 0;  // the HIP device current when the grid was created
    long long step = 0;
    std::map<long long, double> sparse_j;  // key = index*3 + comp
    // Per-step current values of one fdtd_update_fields_with_currents call:
    // the sparse J cells of the call (keys as sparse_j) become kind-2
    // sources reading column c of row (step - jtab_step0) of d_jtab.
    std::vector<long long> jtab_keys;
    double* d_jtab = nullptr;
    size_t d_jtab_cap = 0;       // doubles
    // single-cell writes into dense J (fdtd_set_current_values)
    fdtd::CurrentWrite* d_jw = nullptr;
    size_t d_jw_cap = 0;         // writes
    long long jtab_step0 = 0;
    int jtab_rows = 0;
    // launches so far: {steps, blocked passes, single-step sweeps, B/2 prologues}
    long long stats[4] = {0, 0, 0, 0};
    std::vector<fdtd::PointSource> waveforms;
    fdtd::PointSource* d_src = nullptr;  // the list the next launches read (one of d_src_buf)
    double* d_vals = nullptr;  // source values, kValTable steps x nsrc
    int d_src_cap = 0;
    // Source lists are uploaded asynchronously from pinned staging, double
    // buffered so a new list never overwrites one that queued work reads.
    fdtd::PointSource* d_src_buf[2] = {};
    fdtd::PointSource* h_src[2] = {};
    int* d_kstart = nullptr;  // per-plane source ranges of d_src (Nk+1), one per buffer
    int* d_kst_buf[2] = {};
    int* h_kst[2] = {};
    hipEvent_t src_ev[2] = {};
    bool src_ev_used[2] = {false, false};
    int src_slot = 0;
    int nsrc = 0;
    bool src_dirty = true;
    fdtd::rt::Tuning tune;
    fdtd::rt::BlockTuning btune;
    bool fused_ok = false;
    int autoscale_interval = 0;  // FP16: rescale every this many sweeps (0 = off)
    int autoscale_log2 = 12;
    // FP16 strip passes: the next pass scales what it reads by these exact
    // powers of two (the in-place rescaling of the state, folded into the
    // pass's loads); set by run_blocked for one launch.
    bool pass_rescale = false;
    float pass_iscale[6] = {1, 1, 1, 1, 1, 1};
    double* d_part = nullptr;
    double* d_out = nullptr;
    unsigned* d_max = nullptr;  // per-component max |stored| tracked by a sweep (float bits)
    // Carried half step (periodic grids): BA[ba_cur] holds B_{n+1/2} =
    // (B_n + d) + d of the current state, written by the last run's final
    // sweep, so the next run starts without the B/2 prologue.  -1: none.
    void* BA[2][3] = {};
    int ba_cur = -1;
    bool carry = true;
    hipEvent_t ev[3] = {};
    // PML layer (include/FDTD/FDTD_PML.h): set when some layer width is > 0.
    bool pml = false;
    double pml_percent = 0;
    int pml_p[3] = {0, 0, 0};
    void* ES[2][6] = {};        // E splits Exy Exz Eyx Eyz Ezx Ezy, ping-pong with E
    void* BS[6] = {};           // B splits Bxy Bxz Byx Byz Bzx Bzy, in place
    void* pml_tab = nullptr;    // per-axis decay/coefficient tables (compute type)
    // PML steps: 0 = the fused PML sweep (shape tune); 1 = the two in-place
    // phase sweeps (pml_e_phase, pml_b_phase) with launch shape ptune.  FP16
    // sweeps that rescale or track the max take the fused sweep either way.
    int pml_path = 0;
    fdtd::rt::Tuning ptune;
    // set by fdtd_time_update_fields_kernels: the stepping launches are
    // bracketed by events on the grid's stream
    fdtd::rt::LaunchTrace* trace = nullptr;
    // Probes (fdtd_set_probes): FP64 values of single cells after every
    // step, one row of nprobe values per step in d_probe_out (probe_cap
    // rows).  Rows are written by the stepping kernels themselves (fused
    // sweeps, strip passes) or by a gather after each step (three-sweep
    // and PML steps); probe_rows counts the rows queued so far.
    fdtd::Probe* d_probe = nullptr;  // sorted by plane key (E at k: k, B at k: k+1 mod Nk)
    int* d_pstart = nullptr;         // Nk+1 ranges of d_probe
    int nprobe = 0;
    double* d_probe_out = nullptr;
    long long probe_cap = 0, probe_rows = 0;
    // Energy records (fdtd_set_energy_interval): after every step that is a
    // multiple of energy_every, the six sums of squares (physical units) in
    // d_energy, reduced on the device; energy_steps holds each row's step.
    int energy_every = 0;
    double* d_energy = nullptr;
    long long energy_cap = 0;
    std::vector<long long> energy_steps;
    // true while fdtd_update_fields runs its steps: only those record
    bool recording = false;
};

namespace fdtd::rt {

// Coefficients, src/FDTD/FDTD.cpp:43-53 (FP64, reference order), then the
// storage-scale ratios (exact powers of two) and a final rounding to C.
template <typename C>
fdtd::Coefs<C> make_coefs(const fdtd_grid* g) {
    const double cdt = kC * g->dt;
    const double cEx = cdt / g->params.dx, cEy = cdt / g->params.dy, cEz = cdt / g->params.dz;
    const double cBx = cdt / (2.0 * g->params.dx), cBy = cdt / (2.0 * g->params.dy),
                 cBz = cdt / (2.0 * g->params.dz);
    const double cur = -4.0 * kPiRef * g->dt;
    const double* s = g->scales;  // Ex Ey Ez Bx By Bz
    fdtd::Coefs<C> c;
    c.ex_bz = C(cEy * (s[0] / s[5]));
    c.ex_by = C(cEz * (s[0] / s[4]));
    c.ey_bx = C(cEz * (s[1] / s[3]));
    c.ey_bz = C(cEx * (s[1] / s[5]));
    c.ez_by = C(cEx * (s[2] / s[4]));
    c.ez_bx = C(cEy * (s[2] / s[3]));
    c.jx = C(cur * s[0]);
    c.jy = C(cur * s[1]);
    c.jz = C(cur * s[2]);
    c.bx_ey = C(cBz * (s[3] / s[1]));
    c.bx_ez = C(cBy * (s[3] / s[2]));
    c.by_ez = C(cBx * (s[4] / s[2]));
    c.by_ex = C(cBz * (s[4] / s[0]));
    c.bz_ex = C(cBy * (s[5] / s[0]));
    c.bz_ey = C(cBx * (s[5] / s[1]));
    return c;
}

// Layer widths, split buffers and coefficient tables of a PML grid; `in`
// and `out` select the E-split buffers a sweep reads and writes.
template <typename S, typename C>
fdtd::Pml<S, C> make_pml(const fdtd_grid* g, int in, int out) {
    fdtd::Pml<S, C> P{};
    P.pi = g->pml_p[0];
    P.pj = g->pml_p[1];
    P.pk = g->pml_p[2];
    for (int q = 0; q < 6; ++q) {
        P.es_in[q] = static_cast<const S*>(g->ES[in][q]);
        P.es_out[q] = static_cast<S*>(g->ES[out][q]);
        P.bs[q] = static_cast<S*>(g->BS[q]);
    }
    const C* t = static_cast<const C*>(g->pml_tab);
    P.tx = t;
    P.ty = t ? t + 5 * g->Ni : nullptr;
    P.tz = t ? t + 5 * (g->Ni + g->Nj) : nullptr;
    return P;
}

template <typename S>
fdtd::Fields<S> fields_of(fdtd_grid* g, int buf) {
    fdtd::Fields<S> f;
    for (int d = 0; d < 3; ++d) {
        f.e[d] = static_cast<S*>(g->E[buf][d]);
        f.b[d] = static_cast<S*>(g->B[buf][d]);
    }
    return f;
}

// Allows `kernel` to be launched with `bytes` of dynamic LDS (needed above
// 64 KiB); the largest value set so far is remembered per kernel.
int allow_dynamic_lds(const void* kernel, size_t bytes);

int grid_blocks(size_t n);
int max_threads(const fdtd_grid* g, int V);
int default_kchunk(int Nj, int Nk, int TY);
// Sparse J values (as kind-0 sources) followed by the waveform sources.
std::vector<fdtd::PointSource> host_sources(const fdtd_grid* g);
// Uploads the grid's source list to its device copy when it changed.
int flush_sources(fdtd_grid* g);
// B half step on the current buffers; on a PML grid `layer` adds the
// layer's full-step split update (the first B update of a step).
int plain_B(fdtd_grid* g, bool layer = true);
int zero_fill(void* p, size_t bytes, hipStream_t s);

}  // namespace fdtd::rt
