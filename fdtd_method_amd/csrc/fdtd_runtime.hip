// Host runtime of libfdtd_mi355x.so: grid objects in HBM, HIP streams,
// launches, diagnostics and the C-ABI declared in include/fdtd_mi355x.h.
//
// Reference counterparts (file:line):
//   fdtd_create          FDTD::FDTD            src/FDTD/FDTD.cpp:3-61
//   fdtd_upload/download FDTD::get_field       src/FDTD/FDTD.cpp:138-151
//   fdtd_zero_currents   FDTD::zeroed_currents src/FDTD/FDTD.cpp:132-136
//   fdtd_update_fields*  FDTD::update_fields   src/FDTD/FDTD.cpp:153-157
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "fdtd_internal.hpp"

namespace fdtd::rt {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

}  // namespace fdtd::rt

namespace fdtd::rt {


int allow_dynamic_lds(const void* kernel, size_t bytes) {
    static std::mutex mu;
    static std::map<const void*, size_t> allowed;
    if (bytes <= 64 * 1024) return FDTD_OK;
    std::lock_guard<std::mutex> lock(mu);
    size_t& cur = allowed[kernel];
    if (bytes <= cur) return FDTD_OK;
    const hipError_t e =
        hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(FDTD_ERR_HIP, "cannot allow %zu bytes of LDS: %s", bytes, hipGetErrorString(e));
    }
    cur = bytes;
    return FDTD_OK;
}

int grid_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    return static_cast<int>(std::min<size_t>(b, 8192));
}

// -------------------------------------------------------------------------
// Fused-sweep launch shapes.

int v_candidates(int precision, int* out) {
    if (precision == FDTD_FP64) { out[0] = 2; out[1] = 1; out[2] = 4; return 3; }
    if (precision == FDTD_FP32) { out[0] = 4; out[1] = 2; out[2] = 1; out[3] = 8; return 4; }
    out[0] = 4; out[1] = 2; out[2] = 8; out[3] = 1;
    return 4;
}

size_t sweep_lds_bytes(const fdtd_grid* g, int TY) {
    const int cs = (int)compute_size(g->precision);
    return g->pml ? fdtd::fused_pml_lds_bytes(TY, g->Ni, cs) : fdtd::fused_lds_bytes(TY, g->Ni, cs);
}

int max_threads(const fdtd_grid* g, int V) {
    const int cs = (int)compute_size(g->precision);
    return g->pml ? fdtd::fused_pml_max_threads(V, cs) : fdtd::fused_max_threads(V, cs);
}

bool shape_valid(const fdtd_grid* g, int V, int TY, int kchunk) {
    if (V <= 0 || TY <= 0 || kchunk <= 0) return false;
    if (V != 1 && V != 2 && V != 4 && V != 8) return false;
    if (g->precision == FDTD_FP64 && V > 4) return false;
    // the PML sweep holds the splits of its V cells too: wider lanes spill
    if (g->pml && V * (int)compute_size(g->precision) > 16) return false;
    if (g->Ni % V != 0) return false;
    const int lanes = g->Ni / V;
    if ((TY + 1) * lanes > max_threads(g, V)) return false;
    if (sweep_lds_bytes(g, TY) > kLdsLimit) return false;
    return true;
}

// z-chunk length for a tile height: balanced chunks (lengths differ by at
// most one plane), at least two tiles per CU where the grid allows it, the
// per-XCD tile count filling its 32 CUs in whole rounds, and otherwise
// longer chunks (fewer recomputed boundary planes).
int default_kchunk(int Nj, int Nk, int TY) {
    const int ntile_y = (Nj + TY - 1) / TY;
    int best_k = Nk;
    double best_score = -1.0;
    for (int nchunk = 1; nchunk <= Nk; ++nchunk) {
        const int k = (Nk + nchunk - 1) / nchunk;
        if ((Nk + k - 1) / k != nchunk) continue;  // not a distinct balanced split
        const int ntiles = ntile_y * nchunk;
        const int per_xcd = (ntiles + 7) / 8;
        const int rounds = (per_xcd + 31) / 32;
        const double fill = double(ntiles) / (rounds * 256.0);
        const double spread = std::min(1.0, ntiles / 512.0);
        const double score = fill * spread * (1.0 - 1.0 / (k + 1.0));
        if (score > best_score + 1e-9) {
            best_score = score;
            best_k = k;
        }
    }
    return best_k;
}

// Widest V giving 16-byte lanes that divides Ni, then the largest TY with at
// most 640 (FP32/FP16) or 768 (FP64) threads and 80 KiB of LDS per
// workgroup; the autotuner refines this.
bool default_tuning(const fdtd_grid* g, Tuning* t) {
    int vc[4];
    const int nv = v_candidates(g->precision, vc);
    const int cs = (int)compute_size(g->precision);
    for (int pass = 0; pass < 2; ++pass) {
        for (int q = 0; q < nv; ++q) {
            const int V = vc[q];
            if (g->Ni % V) continue;
            if (g->pml && V * cs > 16) continue;
            const int lanes = g->Ni / V;
            const int cap = cs == 8 ? 768 : 640;
            const int mt = pass == 0 ? std::min(cap, max_threads(g, V)) : max_threads(g, V);
            const size_t lds_cap = pass == 0 ? kLdsLimit / 2 : kLdsLimit;
            int best = 0;
            for (int ty = std::min(g->Nj, mt / lanes - 1); ty >= 1; --ty)
                if (sweep_lds_bytes(g, ty) <= lds_cap) { best = ty; break; }
            if (!best) continue;
            t->V = V;
            t->TY = best;
            t->kchunk = default_kchunk(g->Nj, g->Nk, best);
            return true;
        }
    }
    return false;
}

fdtd::SweepGeom make_geom(const fdtd_grid* g, const Tuning& t) {
    fdtd::SweepGeom s;
    s.Ni = g->Ni;
    s.Nj = g->Nj;
    s.Nk = g->Nk;
    s.TY = t.TY;
    s.kchunk = t.kchunk;
    s.ntile_y = (g->Nj + t.TY - 1) / t.TY;
    const int nchunk = (g->Nk + t.kchunk - 1) / t.kchunk;
    s.ntiles = s.ntile_y * nchunk;
    s.per_xcd = (s.ntiles + 7) / 8;
    return s;
}


std::vector<fdtd::PointSource> host_sources(const fdtd_grid* g) {
    std::vector<fdtd::PointSource> all;
    auto cell_source = [g](long long key) {
        const long long idx = key / 3;
        fdtd::PointSource s{};
        s.comp = (int)(key % 3);
        s.i = (int)(idx % g->Ni);
        s.j = (int)((idx / g->Ni) % g->Nj);
        s.k = (int)(idx / ((long long)g->Ni * g->Nj));
        return s;
    };
    if (!g->jtab_keys.empty()) {
        // a call with per-step values: every sparse cell of the call, in key
        // order (the order sparse_j gives), reads its column of the table
        all.reserve(g->jtab_keys.size() + g->waveforms.size());
        for (size_t c = 0; c < g->jtab_keys.size(); ++c) {
            fdtd::PointSource s = cell_source(g->jtab_keys[c]);
            s.kind = 2;
            s.col = (int)c;
            all.push_back(s);
        }
        all.insert(all.end(), g->waveforms.begin(), g->waveforms.end());
        return all;
    }
    all.reserve(g->sparse_j.size() + g->waveforms.size());
    for (const auto& kv : g->sparse_j) {
        fdtd::PointSource s = cell_source(kv.first);
        s.kind = 0;
        s.amp = kv.second;
        all.push_back(s);
    }
    all.insert(all.end(), g->waveforms.begin(), g->waveforms.end());
    return all;
}

// Push the combined source list to the device when it changed: copied into
// one of two pinned staging buffers and uploaded on the grid's stream with
// no host wait (the drivers change J every step).  The other device buffer
// may still be read by queued sweeps; a staging buffer is reused only after
// its previous upload has completed.
// The list is sorted by plane (stably, so sources sharing a cell keep their
// order) with a table of per-plane ranges: a sweep looks at the sources of
// the plane it updates only.
int flush_sources(fdtd_grid* g) {
    if (!g->src_dirty) return FDTD_OK;
    std::vector<fdtd::PointSource> all = host_sources(g);
    std::stable_sort(all.begin(), all.end(),
                     [](const fdtd::PointSource& x, const fdtd::PointSource& y) { return x.k < y.k; });
    std::vector<int> kstart(g->Nk + 1, 0);
    for (const auto& q : all) kstart[q.k + 1] += 1;
    for (int k = 0; k < g->Nk; ++k) kstart[k + 1] += kstart[k];
    for (int b = 0; b < 2; ++b) {  // per-plane tables: allocated once (each one left missing)
        if (!g->d_kst_buf[b]) HIP_TRY(hipMalloc(&g->d_kst_buf[b], sizeof(int) * (g->Nk + 1)));
        if (!g->h_kst[b])
            HIP_TRY(hipHostMalloc(&g->h_kst[b], sizeof(int) * (g->Nk + 1), hipHostMallocDefault));
    }
    if ((int)all.size() > g->d_src_cap) {
        HIP_TRY(hipStreamSynchronize(g->stream));
        for (int b = 0; b < 2; ++b) {
            if (g->d_src_buf[b]) HIP_TRY(hipFree(g->d_src_buf[b]));
            if (g->h_src[b]) HIP_TRY(hipHostFree(g->h_src[b]));
            g->d_src_buf[b] = nullptr;
            g->h_src[b] = nullptr;
            g->src_ev_used[b] = false;
        }
        if (g->d_vals) HIP_TRY(hipFree(g->d_vals));
        g->d_vals = nullptr;
        g->d_src = nullptr;
        g->d_src_cap = 0;  // until the new buffers exist
        const int cap = std::max<int>(16, 2 * (int)all.size());
        for (int b = 0; b < 2; ++b) {
            HIP_TRY(hipMalloc(&g->d_src_buf[b], sizeof(fdtd::PointSource) * cap));
            HIP_TRY(hipHostMalloc(&g->h_src[b], sizeof(fdtd::PointSource) * cap, hipHostMallocDefault));
        }
        HIP_TRY(hipMalloc(&g->d_vals, sizeof(double) * cap * kValTable));
        g->d_src_cap = cap;
    }
    if (!all.empty()) {
        const int slot = g->src_slot = 1 - g->src_slot;
        if (g->src_ev_used[slot]) HIP_TRY(hipEventSynchronize(g->src_ev[slot]));
        std::memcpy(g->h_src[slot], all.data(), sizeof(fdtd::PointSource) * all.size());
        std::memcpy(g->h_kst[slot], kstart.data(), sizeof(int) * kstart.size());
        HIP_TRY(hipMemcpyAsync(g->d_src_buf[slot], g->h_src[slot],
                               sizeof(fdtd::PointSource) * all.size(), hipMemcpyHostToDevice,
                               g->stream));
        HIP_TRY(hipMemcpyAsync(g->d_kst_buf[slot], g->h_kst[slot], sizeof(int) * kstart.size(),
                               hipMemcpyHostToDevice, g->stream));
        HIP_TRY(hipEventRecord(g->src_ev[slot], g->stream));
        g->src_ev_used[slot] = true;
        g->d_src = g->d_src_buf[slot];
        g->d_kstart = g->d_kst_buf[slot];
    }
    g->nsrc = (int)all.size();
    g->src_dirty = false;
    return FDTD_OK;
}

// J values of all point sources for steps step .. step+nsteps-1
// (t = (n+1)*dt) into the value table, nsteps <= kValTable.
int eval_sources(fdtd_grid* g, long long step, int nsteps) {
    if (g->nsrc == 0) return FDTD_OK;
    const int n = g->nsrc * nsteps;
    const double* tab = nullptr;
    if (g->jtab_rows > 0) {
        const long long row = step - g->jtab_step0;
        if (row < 0 || row + nsteps > g->jtab_rows)
            return fail(FDTD_ERR_INVALID_ARGUMENT, "steps outside the per-step current table");
        tab = g->d_jtab + row * (long long)g->jtab_keys.size();
    }
    hipLaunchKernelGGL(fdtd::eval_sources, dim3((n + 255) / 256), dim3(256), 0, g->stream,
                       g->d_src, g->nsrc, nsteps, (const long long*)nullptr, step, g->dt,
                       g->d_vals, tab, (int)g->jtab_keys.size());
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

// ---- plain kernels ----------------------------------------------------------

fdtd::BlockedGeom make_blocked_geom(const fdtd_grid* g, const BlockTuning& t) {
    fdtd::BlockedGeom G;
    G.Ni = g->Ni;
    G.Nj = g->Nj;
    G.Nk = g->Nk;
    G.TXg = t.lanes * t.V;
    G.TYg = t.TYg;
    G.TX = G.TXg - 2 * t.L;
    G.TY = G.TYg - 2 * t.L;
    G.kchunk = t.kchunk;
    G.ntx = (g->Ni + G.TX - 1) / G.TX;
    G.nty = (g->Nj + G.TY - 1) / G.TY;
    const int nchunk = (g->Nk + t.kchunk - 1) / t.kchunk;
    G.ntiles = G.ntx * G.nty * nchunk;
    G.per_xcd = (G.ntiles + 7) / 8;
    return G;
}

bool blocked_valid(const fdtd_grid* g, const BlockTuning& t) {
    if (t.L < 1 || t.L > 4 || (t.V != 1 && t.V != 2)) return false;
    if (g->Ni % t.V || (2 * t.L) % t.V) return false;
    if (t.lanes <= 0 || t.TYg <= 0 || t.kchunk <= 0) return false;
    if (t.lanes * t.V - 2 * t.L < 1 || t.TYg - 2 * t.L < 1) return false;
    if (t.lanes * t.TYg > fdtd::blocked_max_threads(t.V, t.L, (int)compute_size(g->precision)))
        return false;
    if (fdtd::blocked_lds_bytes(t.L, t.lanes * t.V, t.TYg, (int)compute_size(g->precision)) > kLdsLimit)
        return false;
    return true;
}

fdtd::StripGeom make_strip_geom(const fdtd_grid* g, const BlockTuning& t) {
    fdtd::StripGeom G;
    G.Ni = g->Ni;
    G.Nj = g->Nj;
    G.Nk = g->Nk;
    G.W = t.TYg / t.R;
    strip_x_geometry(t.L, t.V, g->esize, &G.Gx, &G.TX);
    G.TY = t.TYg - 2 * t.L;
    G.kchunk = t.kchunk;
    G.ntx = (g->Ni + G.TX - 1) / G.TX;
    G.nty = (g->Nj + G.TY - 1) / G.TY;
    const int nchunk = (g->Nk + t.kchunk - 1) / t.kchunk;
    G.ntiles = G.ntx * G.nty * nchunk;
    G.per_xcd = (G.ntiles + 7) / 8;
    return G;
}

// Strip variants built: cells per lane V*R with V*R*sizeof(C) <= 16 bytes
// (FP32/FP16: V=1 R=2, V=1 R=4, V=2 R=2; FP64: V=1 R=2).
constexpr bool strip_built(int csize, int V, int R) {
    return (V == 1 || V == 2) && (R == 2 || R == 4) && V * R * csize <= 16;
}

bool strip_valid(const fdtd_grid* g, const BlockTuning& t) {
    const int cs = (int)compute_size(g->precision);
    if (t.L < 1 || t.L > 4 || !strip_built(cs, t.V, t.R)) return false;
    if (g->Ni % t.V || t.TYg <= 0 || t.TYg % t.R || t.kchunk <= 0) return false;
    if ((unsigned long long)g->Ni * g->Nj * g->esize >= (1ull << 31)) return false;  // 32-bit plane offsets, buffer range checks
    const int W = t.TYg / t.R;
    if (W < 1 || W > 16 || 64 * W > fdtd::strip_max_threads(t.V, t.R, t.L, cs)) return false;
    int Gx = 0, TX = 0;
    strip_x_geometry(t.L, t.V, g->esize, &Gx, &TX);
    if (TX < 1 || t.TYg - 2 * t.L < 1) return false;
    return fdtd::strip_lds_bytes(t.L, W, t.V, cs) <= kLdsLimit;
}

template <typename S>
int launch_plain_B(fdtd_grid* g, bool layer) {
    using C = typename fdtd::ComputeOf<S>::type;
    const dim3 block(64, 4, 1);
    const dim3 grid((g->Ni + 63) / 64, (g->Nj + 3) / 4, g->Nk);
    if (g->pml) {
        const auto P = make_pml<S, C>(g, g->cur, g->cur);
        if (layer)
            hipLaunchKernelGGL((fdtd::pml_plain_update_B<S, C, true>), grid, block, 0, g->stream,
                               fields_of<S>(g, g->cur), P, make_coefs<C>(g), g->Ni, g->Nj, g->Nk);
        else
            hipLaunchKernelGGL((fdtd::pml_plain_update_B<S, C, false>), grid, block, 0, g->stream,
                               fields_of<S>(g, g->cur), P, make_coefs<C>(g), g->Ni, g->Nj, g->Nk);
    } else {
        hipLaunchKernelGGL((fdtd::plain_update_B<S, C>), grid, block, 0, g->stream,
                           fields_of<S>(g, g->cur), make_coefs<C>(g), g->Ni, g->Nj, g->Nk);
    }
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

template <typename S>
int launch_plain_E(fdtd_grid* g, long long step) {
    using C = typename fdtd::ComputeOf<S>::type;
    using JT = C;
    const dim3 block(64, 4, 1);
    const dim3 grid((g->Ni + 63) / 64, (g->Nj + 3) / 4, g->Nk);
    const JT* j[3] = {nullptr, nullptr, nullptr};
    if (g->j_dense_active)
        for (int d = 0; d < 3; ++d) j[d] = static_cast<const JT*>(g->J[d]);
    if (int rc = eval_sources(g, step, 1)) return rc;
    if (g->pml)
        hipLaunchKernelGGL((fdtd::pml_plain_update_E<S, C, JT>), grid, block, 0, g->stream,
                           fields_of<S>(g, g->cur), make_pml<S, C>(g, g->cur, g->cur), j[0], j[1],
                           j[2], g->d_src, g->d_vals, g->nsrc, make_coefs<C>(g), g->Ni, g->Nj,
                           g->Nk);
    else
        hipLaunchKernelGGL((fdtd::plain_update_E<S, C, JT>), grid, block, 0, g->stream,
                           fields_of<S>(g, g->cur), j[0], j[1], j[2], g->d_src, g->d_vals, g->nsrc,
                           make_coefs<C>(g), g->Ni, g->Nj, g->Nk);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

int plain_B(fdtd_grid* g, bool layer) {
    switch (g->precision) {
        case FDTD_FP64: return launch_plain_B<double>(g, layer);
        case FDTD_FP32: return launch_plain_B<float>(g, layer);
        default: return launch_plain_B<_Float16>(g, layer);
    }
}

int plain_E(fdtd_grid* g, long long step) {
    switch (g->precision) {
        case FDTD_FP64: return launch_plain_E<double>(g, step);
        case FDTD_FP32: return launch_plain_E<float>(g, step);
        default: return launch_plain_E<_Float16>(g, step);
    }
}

// One B half step (e_step false) or E step (true) of the plain update over
// the box; the E step adds dense J and the point sources of step g->step.
// The caller has checked 0 <= lo < hi <= N on every axis.
template <typename S>
int launch_box(fdtd_grid* g, const fdtd::Box& box, bool e_step) {
    using C = typename fdtd::ComputeOf<S>::type;
    using JT = C;
    const int nx = box.hi[0] - box.lo[0], ny = box.hi[1] - box.lo[1], nz = box.hi[2] - box.lo[2];
    const dim3 block(64, 4, 1);
    const dim3 grid((nx + 63) / 64, (ny + 3) / 4, nz);
    if (!e_step) {
        hipLaunchKernelGGL((fdtd::box_update_B<S, C>), grid, block, 0, g->stream,
                           fields_of<S>(g, g->cur), make_coefs<C>(g), g->Ni, g->Nj, g->Nk, box);
    } else {
        const JT* j[3] = {nullptr, nullptr, nullptr};
        if (g->j_dense_active)
            for (int d = 0; d < 3; ++d) j[d] = static_cast<const JT*>(g->J[d]);
        if (int rc = eval_sources(g, g->step, 1)) return rc;
        hipLaunchKernelGGL((fdtd::box_update_E<S, C, JT>), grid, block, 0, g->stream,
                           fields_of<S>(g, g->cur), j[0], j[1], j[2], g->d_src, g->d_vals,
                           g->nsrc, make_coefs<C>(g), g->Ni, g->Nj, g->Nk, box);
    }
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

int box_update(fdtd_grid* g, const fdtd::Box& box, bool e_step) {
    switch (g->precision) {
        case FDTD_FP64: return launch_box<double>(g, box, e_step);
        case FDTD_FP32: return launch_box<float>(g, box, e_step);
        default: return launch_box<_Float16>(g, box, e_step);
    }
}

// ---- PML layer ----------------------------------------------------------------

// Layer width of an axis of n cells, src/FDTD/FDTD_PML.cpp:254-256.
int pml_width(int n, double pml_percent) {
    return static_cast<int>(static_cast<double>(n) * pml_percent);
}

// Conductivity along one axis of n cells with layer width p and spacing d,
// src/FDTD/FDTD_PML.cpp:3-61 as applied by :316-338: SGm * (dist/p)^4 with
// dist = p - i in the lower band and i + 1 + p - n in the upper band (set
// second), SGm = -(4 + 1)/2 * ln(1e-12) / (p*d); zero between the bands.
std::vector<double> pml_sigma(int n, int p, double d) {
    std::vector<double> s(n, 0.0);
    if (p <= 0) return s;
    const double sgm = -(kPmlOrder + 1.0) / 2.0 * std::log(kPmlR) / (static_cast<double>(p) * d);
    for (int i = 0; i < p; ++i)
        s[i] = sgm * std::pow(static_cast<double>(p - i) / static_cast<double>(p), kPmlOrder);
    for (int i = n - p; i < n; ++i)
        s[i] = sgm * std::pow(static_cast<double>(i + 1 + p - n) / static_cast<double>(p), kPmlOrder);
    return s;
}

// Decay exp(-sigma*dt*C) and split coefficient (1 - decay)/(sigma*d), or
// C*dt/d where sigma == 0 (src/FDTD/FDTD_PML.cpp:63-65,98-111).
void pml_coef(double sigma, double dt, double d, double* decay, double* coef) {
    *decay = std::exp(-sigma * dt * kC);
    *coef = sigma != 0.0 ? (1.0 - std::exp(-sigma * dt * kC)) / (sigma * d) : kC * dt / d;
}

// Per-axis tables (see fdtd::Pml) in the compute type, with the storage
// scale ratio of each split's target and source folded in.
int build_pml_tables(fdtd_grid* g) {
    const int n[3] = {g->Ni, g->Nj, g->Nk};
    const double d[3] = {g->params.dx, g->params.dy, g->params.dz};
    const double* s = g->scales;  // Ex Ey Ez Bx By Bz
    // x: Eyx<-Bz, Ezx<-By, Byx<-Ez, Bzx<-Ey; y: Exy<-Bz, Ezy<-Bx, Bxy<-Ez,
    // Bzy<-Ex; z: Exz<-By, Eyz<-Bx, Bxz<-Ey, Byz<-Ex
    const double ratio[3][4] = {{s[1] / s[5], s[2] / s[4], s[4] / s[2], s[5] / s[1]},
                                {s[0] / s[5], s[2] / s[3], s[3] / s[2], s[5] / s[0]},
                                {s[0] / s[4], s[1] / s[3], s[3] / s[1], s[4] / s[0]}};
    std::vector<double> tab;
    tab.reserve(5 * (n[0] + n[1] + n[2]));
    for (int a = 0; a < 3; ++a) {
        const std::vector<double> sigma = pml_sigma(n[a], g->pml_p[a], d[a]);
        std::vector<double> dec(n[a]), coef(n[a]);
        for (int i = 0; i < n[a]; ++i) pml_coef(sigma[i], g->dt, d[a], &dec[i], &coef[i]);
        tab.insert(tab.end(), dec.begin(), dec.end());
        for (int r = 0; r < 4; ++r)
            for (int i = 0; i < n[a]; ++i) tab.push_back(coef[i] * ratio[a][r]);
    }
    // sweeps already queued may still read the old tables
    HIP_TRY(hipStreamSynchronize(g->stream));
    if (compute_size(g->precision) == 8) {
        HIP_TRY(hipMemcpy(g->pml_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(tab.begin(), tab.end());
        HIP_TRY(hipMemcpy(g->pml_tab, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return FDTD_OK;
}

// Storage of a PML grid's layer: split fields (zeroed) and tables.
int alloc_pml(fdtd_grid* g) {
    const size_t bytes = g->N * g->esize + 16;
    for (int q = 0; q < 6; ++q) {
        for (int b = 0; b < 2; ++b) {
            HIP_TRY(hipMalloc(&g->ES[b][q], bytes));
            if (int rc = zero_fill(g->ES[b][q], bytes, g->stream)) return rc;
        }
        HIP_TRY(hipMalloc(&g->BS[q], bytes));
        if (int rc = zero_fill(g->BS[q], bytes, g->stream)) return rc;
    }
    HIP_TRY(hipMalloc(&g->pml_tab, 5 * (size_t)(g->Ni + g->Nj + g->Nk) * compute_size(g->precision)));
    return build_pml_tables(g);
}

// Storage component (0..5) whose scale split q (0..5 E, 6..11 B) shares.
int split_component(int q) { return q < 6 ? q / 2 : 3 + (q - 6) / 2; }

void* split_ptr(fdtd_grid* g, int q) { return q < 6 ? g->ES[g->cur][q] : g->BS[q - 6]; }

void* field_ptr(fdtd_grid* g, int c);

// Scratch copies of a grid's current state (E, B and a PML grid's split
// fields) taken before tuning runs that sweep the grid's own buffers; the
// state is put back by restore() or, on any early return, by the destructor.
class StateSnapshot {
public:
    explicit StateSnapshot(fdtd_grid* g) : g_(g) {}
    StateSnapshot(const StateSnapshot&) = delete;
    StateSnapshot& operator=(const StateSnapshot&) = delete;
    ~StateSnapshot() {
        if (!restored_) (void)copy_back();
        for (void* p : saved_) (void)hipFree(p);
    }
    int save() {
        const int n = g_->pml ? 18 : 6;
        for (int d = 0; d < n; ++d) {
            void* p = nullptr;
            HIP_TRY(hipMalloc(&p, bytes()));
            saved_.push_back(p);
            HIP_TRY(hipMemcpyAsync(p, state_ptr(d), bytes(), hipMemcpyDeviceToDevice, g_->stream));
        }
        return FDTD_OK;
    }
    int restore() {
        restored_ = true;
        return copy_back();
    }

private:
    size_t bytes() const { return g_->N * g_->esize; }
    void* state_ptr(int d) const { return d < 6 ? field_ptr(g_, d) : split_ptr(g_, d - 6); }
    int copy_back() {
        for (size_t d = 0; d < saved_.size(); ++d)
            HIP_TRY(hipMemcpyAsync(state_ptr((int)d), saved_[d], bytes(), hipMemcpyDeviceToDevice,
                                   g_->stream));
        HIP_TRY(hipStreamSynchronize(g_->stream));
        return FDTD_OK;
    }
    fdtd_grid* g_;
    std::vector<void*> saved_;
    bool restored_ = false;
};

// ---- fused sweep --------------------------------------------------------------

// One sweep's inputs and outputs beyond the buffer pair: FP16 rescaling
// (oscale, maxbits), an input B other than B[in] (in_b: the carried half
// step), and where a final sweep stores the carried half step (ahead).
struct SweepExtra {
    const double* oscale = nullptr;
    unsigned* maxbits = nullptr;
    void* const* in_b = nullptr;
    void* const* ahead = nullptr;
};

// Probes of the steps a run queues from here: the next row.
bool probing(const fdtd_grid* g) { return g->recording && g->nprobe > 0; }

fdtd::ProbeArgs probe_args(const fdtd_grid* g) {
    fdtd::ProbeArgs p{};
    if (!probing(g)) return p;
    p.p = g->d_probe;
    p.pstart = g->d_pstart;
    p.n = g->nprobe;
    p.stride = g->nprobe;
    p.out = g->d_probe_out + static_cast<size_t>(g->probe_rows) * g->nprobe;
    for (int c = 0; c < 6; ++c) p.inv[c] = 1.0 / g->scales[c];
    return p;
}

// Probe row of the step just completed, read from the fields in memory.
template <typename S>
int gather_probes_s(fdtd_grid* g) {
    const fdtd::ProbeArgs p = probe_args(g);
    hipLaunchKernelGGL(fdtd::gather_probes<S>, dim3((p.n + 255) / 256), dim3(256), 0, g->stream,
                       fields_of<S>(g, g->cur), p, g->Ni, g->Nj);
    HIP_TRY(hipGetLastError());
    g->probe_rows += 1;
    return FDTD_OK;
}

int gather_probes(fdtd_grid* g) {
    if (!probing(g)) return FDTD_OK;
    TraceRange tr("probe_gather");
    switch (g->precision) {
        case FDTD_FP64: return gather_probes_s<double>(g);
        case FDTD_FP32: return gather_probes_s<float>(g);
        default: return gather_probes_s<_Float16>(g);
    }
}

template <typename S, int V, bool STEADY, bool AUTO, bool PROBE>
int launch_fused_a(fdtd_grid* g, const Tuning& t, int slot, int in, int out,
                   const SweepExtra& x) {
    using C = typename fdtd::ComputeOf<S>::type;
    using JT = C;
    auto kern = fdtd::fused_sweep<S, C, JT, V, STEADY, AUTO, PROBE>;
    if constexpr (!PROBE)
        if (g->pml) kern = fdtd::fused_pml_sweep<S, C, JT, V, STEADY, AUTO>;
    fdtd::SweepArgs<S, C, JT> a;
    a.in = fields_of<S>(g, in);
    a.out = fields_of<S>(g, out);
    for (int d = 0; d < 3; ++d) {
        if (x.in_b) a.in.b[d] = static_cast<S*>(x.in_b[d]);
        a.ahead[d] = x.ahead ? static_cast<S*>(x.ahead[d]) : nullptr;
        a.j[d] = g->j_dense_active ? static_cast<const JT*>(g->J[d]) : nullptr;
    }
    a.src = g->d_src;
    a.vals = g->d_vals ? g->d_vals + (size_t)slot * g->nsrc : nullptr;
    a.nsrc = g->nsrc;
    a.pad = 0;
    a.kstart = g->nsrc > 0 ? g->d_kstart : nullptr;
    a.c = make_coefs<C>(g);
    a.g = make_geom(g, t);
    a.pml = make_pml<S, C>(g, in, out);
    for (int c = 0; c < 6; ++c) a.oscale[c] = C(x.oscale ? x.oscale[c] : 1.0);
    a.maxbits = x.maxbits;
    a.probe = PROBE ? probe_args(g) : fdtd::ProbeArgs{};
    const size_t lds = sweep_lds_bytes(g, t.TY);
    if (int rc = allow_dynamic_lds(reinterpret_cast<const void*>(kern), lds)) return rc;
    const int threads = (t.TY + 1) * (g->Ni / V);
    hipLaunchKernelGGL(kern, dim3(8 * a.g.per_xcd), dim3(threads), lds, g->stream, a);
    HIP_TRY(hipGetLastError());
    if (PROBE) g->probe_rows += 1;
    return FDTD_OK;
}

// FP16 sweeps that rescale or track the max take the AUTO variant; sweeps
// of a run with probes (not on PML grids, which gather them after whole
// steps), the PROBE variant.
template <typename S, int V, bool STEADY>
int launch_fused_v(fdtd_grid* g, const Tuning& t, int slot, int in, int out,
                   const SweepExtra& x) {
    const bool probe = probing(g) && !g->pml;
    if constexpr (sizeof(S) == 2)
        if (x.oscale || x.maxbits)
            return probe ? launch_fused_a<S, V, STEADY, true, true>(g, t, slot, in, out, x)
                         : launch_fused_a<S, V, STEADY, true, false>(g, t, slot, in, out, x);
    return probe ? launch_fused_a<S, V, STEADY, false, true>(g, t, slot, in, out, x)
                 : launch_fused_a<S, V, STEADY, false, false>(g, t, slot, in, out, x);
}

template <typename S, bool STEADY>
int launch_fused_s(fdtd_grid* g, const Tuning& t, int slot, int in, int out,
                   const SweepExtra& x) {
    switch (t.V) {
        case 1: return launch_fused_v<S, 1, STEADY>(g, t, slot, in, out, x);
        case 2: return launch_fused_v<S, 2, STEADY>(g, t, slot, in, out, x);
        case 4: return launch_fused_v<S, 4, STEADY>(g, t, slot, in, out, x);
        case 8:
            if constexpr (sizeof(S) <= 4) return launch_fused_v<S, 8, STEADY>(g, t, slot, in, out, x);
            [[fallthrough]];
        default: return fail(FDTD_ERR_UNSUPPORTED, "unsupported V=%d", t.V);
    }
}

// One fused sweep; its source values are row `slot` of the value table.
int launch_fused(fdtd_grid* g, const Tuning& t, bool steady, int slot, int in, int out,
                 const SweepExtra& x = SweepExtra()) {
    switch (g->precision) {
        case FDTD_FP64:
            return steady ? launch_fused_s<double, true>(g, t, slot, in, out, x)
                          : launch_fused_s<double, false>(g, t, slot, in, out, x);
        case FDTD_FP32:
            return steady ? launch_fused_s<float, true>(g, t, slot, in, out, x)
                          : launch_fused_s<float, false>(g, t, slot, in, out, x);
        default:
            return steady ? launch_fused_s<_Float16, true>(g, t, slot, in, out, x)
                          : launch_fused_s<_Float16, false>(g, t, slot, in, out, x);
    }
}

// ---- two-phase PML steps --------------------------------------------------------

// Default shape of the phase sweeps: 16-byte lanes (FP64 V=2, else V=4; a
// smaller V when it does not divide Ni), as many full rows as the
// workgroup bound allows, and z chunks giving about 1024 tiles.
Tuning pml_phase_default(const fdtd_grid* g) {
    const int cs = (int)compute_size(g->precision);
    Tuning t;
    t.V = cs == 8 ? 2 : 4;
    while (t.V > 1 && g->Ni % t.V != 0) t.V /= 2;
    const int lanes = g->Ni / t.V;
    t.TY = std::max(1, std::min(g->Nj, fdtd::pml_phase_max_threads(t.V, cs) / std::max(1, lanes)));
    const int ntile_y = (g->Nj + t.TY - 1) / t.TY;
    const int nchunk = std::max(1, std::min(g->Nk / 8, (1024 + ntile_y - 1) / ntile_y));
    t.kchunk = (g->Nk + nchunk - 1) / nchunk;
    return t;
}

bool pml_phase_valid(const fdtd_grid* g, const Tuning& t) {
    const int cs = (int)compute_size(g->precision);
    if (t.V != 1 && t.V != 2 && t.V != 4) return false;
    if (t.V * g->esize > 16) return false;  // packs of at most 16 bytes
    if (g->Ni % t.V != 0 || t.TY < 1 || t.kchunk < 1) return false;
    if ((long long)t.TY * (g->Ni / t.V) > fdtd::pml_phase_max_threads(t.V, cs)) return false;
    return fdtd::pml_phase_lds_bytes(t.TY, g->Ni, cs) <= kLdsLimit;
}

// One step of a PML grid as the E phase and the B phase, in place on the
// current buffers (steady: B takes both half steps and the layer its split
// update; else the run's final B half step).
template <typename S, int V>
int launch_pml_phases_v(fdtd_grid* g, int slot, bool steady) {
    using C = typename fdtd::ComputeOf<S>::type;
    using JT = C;
    const Tuning& t = g->ptune;
    fdtd::PmlPhaseArgs<S, C, JT> a;
    a.f = fields_of<S>(g, g->cur);
    for (int d = 0; d < 3; ++d)
        a.j[d] = g->j_dense_active ? static_cast<const JT*>(g->J[d]) : nullptr;
    a.src = g->d_src;
    a.vals = g->d_vals ? g->d_vals + (size_t)slot * g->nsrc : nullptr;
    a.kstart = g->d_kstart;
    a.nsrc = g->nsrc;
    a.pad = 0;
    a.c = make_coefs<C>(g);
    a.g = make_geom(g, t);
    a.pml = make_pml<S, C>(g, g->cur, g->cur);
    auto ke = fdtd::pml_e_phase<S, C, JT, V>;
    auto kb = steady ? fdtd::pml_b_phase<S, C, JT, V, true> : fdtd::pml_b_phase<S, C, JT, V, false>;
    const size_t lds = fdtd::pml_phase_lds_bytes(t.TY, g->Ni, sizeof(C));
    if (int rc = allow_dynamic_lds(reinterpret_cast<const void*>(ke), lds)) return rc;
    if (int rc = allow_dynamic_lds(reinterpret_cast<const void*>(kb), lds)) return rc;
    const int threads = t.TY * (g->Ni / V);
    hipLaunchKernelGGL(ke, dim3(8 * a.g.per_xcd), dim3(threads), lds, g->stream, a);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(kb, dim3(8 * a.g.per_xcd), dim3(threads), lds, g->stream, a);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

template <typename S>
int launch_pml_phases_s(fdtd_grid* g, int slot, bool steady) {
    switch (g->ptune.V) {
        case 1: return launch_pml_phases_v<S, 1>(g, slot, steady);
        case 2: return launch_pml_phases_v<S, 2>(g, slot, steady);
        case 4:
            if constexpr (sizeof(S) <= 4) return launch_pml_phases_v<S, 4>(g, slot, steady);
            [[fallthrough]];
        default: return fail(FDTD_ERR_UNSUPPORTED, "unsupported V=%d", g->ptune.V);
    }
}

int launch_pml_phases(fdtd_grid* g, int slot, bool steady) {
    switch (g->precision) {
        case FDTD_FP64: return launch_pml_phases_s<double>(g, slot, steady);
        case FDTD_FP32: return launch_pml_phases_s<float>(g, slot, steady);
        default: return launch_pml_phases_s<_Float16>(g, slot, steady);
    }
}

// The carried half step's buffers (allocated on first use; false when the
// device has no room, and the run simply keeps its B/2 prologue).
bool ensure_carry(fdtd_grid* g) {
    if (g->BA[0][0]) return true;
    const size_t bytes = g->N * g->esize + 16;
    for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 3; ++d)
            if (hipMalloc(&g->BA[b][d], bytes) != hipSuccess) {
                (void)hipGetLastError();
                for (int q = 0; q < 2; ++q)
                    for (int e = 0; e < 3; ++e)
                        if (g->BA[q][e]) {
                            (void)hipFree(g->BA[q][e]);
                            g->BA[q][e] = nullptr;
                        }
                return false;
            }
    return true;
}

int rescale_to_target(fdtd_grid* g, bool current_only);
int apply_scales(fdtd_grid* g, const double s[6], bool current_only);
void target_scales(const fdtd_grid* g, const double mx[6], int target_log2, double floor_mag,
                   double s[6]);

// ---- temporally blocked sweeps ----------------------------------------------------

template <typename S, int V, int L, bool FINAL>
int launch_blocked_v(fdtd_grid* g, int slot, int in, int out) {
    using C = typename fdtd::ComputeOf<S>::type;
    auto kern = fdtd::blocked_sweep<S, C, V, L, FINAL>;
    const BlockTuning& t = g->btune;
    fdtd::BlockedArgs<S, C> a;
    a.in = fields_of<S>(g, in);
    a.out = fields_of<S>(g, out);
    a.src = g->d_src;
    a.vals = g->d_vals ? g->d_vals + (size_t)slot * g->nsrc : nullptr;
    a.nsrc = g->nsrc;
    a.pad = 0;
    a.c = make_coefs<C>(g);
    a.g = make_blocked_geom(g, t);
    const size_t lds = fdtd::blocked_lds_bytes(L, a.g.TXg, a.g.TYg, sizeof(C));
    if (int rc = allow_dynamic_lds(reinterpret_cast<const void*>(kern), lds)) return rc;
    hipLaunchKernelGGL(kern, dim3(8 * a.g.per_xcd), dim3(t.lanes * t.TYg), lds, g->stream, a);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

template <typename S, int V, int R, int L, bool FINAL>
int launch_strip_v(fdtd_grid* g, int slot, int in, int out, void* const* in_b, void* const* ahead) {
    using C = typename fdtd::ComputeOf<S>::type;
    if constexpr (!strip_built(sizeof(C), V, R)) {
        return fail(FDTD_ERR_UNSUPPORTED, "strip variant V=%d R=%d not built", V, R);
    } else {
        // FP16: the pass after an in-run rescaling decision scales its
        // reads; passes of a run with probes record them
        const bool probe = probing(g);
        auto kern = probe ? fdtd::strip_sweep<S, C, V, R, L, FINAL, true>
                          : fdtd::strip_sweep<S, C, V, R, L, FINAL, false>;
        if constexpr (sizeof(S) == 2)
            if (g->pass_rescale)
                kern = probe ? fdtd::strip_rescale_sweep<S, C, V, R, L, FINAL, true>
                             : fdtd::strip_rescale_sweep<S, C, V, R, L, FINAL, false>;
        const BlockTuning& t = g->btune;
        fdtd::StripArgs<S, C> a;
        a.in = fields_of<S>(g, in);
        if (in_b)
            for (int d = 0; d < 3; ++d) a.in.b[d] = static_cast<S*>(in_b[d]);
        a.out = fields_of<S>(g, out);
        for (int d = 0; d < 3; ++d) a.ahead[d] = ahead ? static_cast<S*>(ahead[d]) : nullptr;
        a.src = g->d_src;
        a.kstart = g->d_kstart;
        a.vals = g->d_vals ? g->d_vals + (size_t)slot * g->nsrc : nullptr;
        a.nsrc = g->nsrc;
        a.vstride = g->nsrc;
        a.c = make_coefs<C>(g);
        a.g = make_strip_geom(g, t);
        for (int c = 0; c < 6; ++c) a.iscale[c] = g->pass_iscale[c];
        a.probe = probe_args(g);
        const size_t lds = fdtd::strip_lds_bytes(L, a.g.W, V, sizeof(C));
        if (int rc = allow_dynamic_lds(reinterpret_cast<const void*>(kern), lds)) return rc;
        hipLaunchKernelGGL(kern, dim3(8 * a.g.per_xcd), dim3(64 * a.g.W), lds, g->stream, a);
        HIP_TRY(hipGetLastError());
        if (a.probe.n > 0) g->probe_rows += L;
        return FDTD_OK;
    }
}

// Workgroups of a strip variant resident per CU (0 if it cannot launch).
template <typename S, int V, int R, int L>
int strip_occupancy_v(int W) {
    using C = typename fdtd::ComputeOf<S>::type;
    if constexpr (!strip_built(sizeof(C), V, R)) {
        return 0;
    } else {
        auto kern = fdtd::strip_sweep<S, C, V, R, L, false>;
        const size_t lds = fdtd::strip_lds_bytes(L, W, V, sizeof(C));
        if (allow_dynamic_lds(reinterpret_cast<const void*>(kern), lds) != FDTD_OK) return 0;
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, 64 * W, lds) != hipSuccess) {
            (void)hipGetLastError();
            return 0;
        }
        return n;
    }
}

template <typename S, int V, int R>
int strip_occupancy_l(int L, int W) {
    switch (L) {
        case 1: return strip_occupancy_v<S, V, R, 1>(W);
        case 2: return strip_occupancy_v<S, V, R, 2>(W);
        case 3: return strip_occupancy_v<S, V, R, 3>(W);
        default: return strip_occupancy_v<S, V, R, 4>(W);
    }
}

template <typename S>
int strip_occupancy_s(const BlockTuning& t) {
    const int W = t.TYg / t.R;
    if (t.V == 1 && t.R == 2) return strip_occupancy_l<S, 1, 2>(t.L, W);
    if (t.V == 1 && t.R == 4) return strip_occupancy_l<S, 1, 4>(t.L, W);
    if (t.V == 2 && t.R == 2) return strip_occupancy_l<S, 2, 2>(t.L, W);
    return 0;
}

int strip_occupancy(const fdtd_grid* g, const BlockTuning& t) {
    switch (g->precision) {
        case FDTD_FP64: return strip_occupancy_s<double>(t);
        case FDTD_FP32: return strip_occupancy_s<float>(t);
        default: return strip_occupancy_s<_Float16>(t);
    }
}

// ---- autotuner cache ----------------------------------------------------------
// fdtd_autotune_blocking's picks, kept across processes in a text file (the
// path of fdtd_set_tune_cache, else $FDTD_MI355X_TUNE_CACHE; none: off).
// One line per pick, "<key> | L V lanes TYg kchunk R ms", the key naming
// the cache format, shape, precision, device architecture and CU count;
// the last line of a key wins, and a pick that no longer validates is
// timed again.
constexpr int kTuneCacheVersion = 1;
std::mutex g_tune_mu;
std::string g_tune_path;
bool g_tune_path_set = false;

std::string tune_cache_path() {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    if (!g_tune_path_set) {
        const char* e = std::getenv("FDTD_MI355X_TUNE_CACHE");
        g_tune_path = e ? e : "";
        g_tune_path_set = true;
    }
    return g_tune_path;
}

int device_cus();

std::string tune_key(const fdtd_grid* g) {
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    if (hipGetDeviceProperties(&prop, g->device) == hipSuccess) arch = prop.gcnArchName;
    for (char& c : arch)
        if (c == ' ' || c == '|') c = '_';
    char buf[256];
    std::snprintf(buf, sizeof buf, "v%d %d %d %d p%d %s cu%d", kTuneCacheVersion, g->Ni, g->Nj,
                  g->Nk, g->precision, arch.c_str(), device_cus());
    return buf;
}

bool tune_cache_lookup(const std::string& path, const std::string& key, BlockTuning* t,
                       float* ms) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    bool found = false;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        char* bar = std::strchr(line, '|');
        if (!bar) continue;
        std::string k(line, bar - line);
        while (!k.empty() && k.back() == ' ') k.pop_back();
        if (k != key) continue;
        BlockTuning c;
        float m = 0;
        if (std::sscanf(bar + 1, "%d %d %d %d %d %d %f", &c.L, &c.V, &c.lanes, &c.TYg, &c.kchunk,
                        &c.R, &m) == 7) {
            *t = c;
            *ms = m;
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

void tune_cache_store(const std::string& path, const std::string& key, const BlockTuning& t,
                      float ms) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    FILE* f = std::fopen(path.c_str(), "a");
    if (!f) return;  // a cache that cannot be written only costs the next run a retune
    std::fprintf(f, "%s | %d %d %d %d %d %d %.9g\n", key.c_str(), t.L, t.V, t.lanes, t.TYg,
                 t.kchunk, t.R, ms);
    std::fclose(f);
}

int device_cus() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            cus = n;
        else
            cus = 256;
    }
    return cus;
}

// z chunk length of a strip pass: every tile costs kchunk + 2(L-1) planes
// (the pipeline fills and drains over L-1 planes at each end), and the
// tiles run in rounds of (resident workgroups per CU x CUs); the chunking
// with the fewest plane-rounds wins (ties: fewer chunks).
int strip_default_kchunk(const fdtd_grid* g, const BlockTuning& t) {
    BlockTuning u = t;
    u.kchunk = g->Nk;
    const fdtd::StripGeom G = make_strip_geom(g, u);
    const long long nxy = (long long)G.ntx * G.nty;
    const long long slots = (long long)std::max(1, strip_occupancy(g, t)) * device_cus();
    const int maxchunks = std::max(1, g->Nk / 16);
    long long best_cost = -1;
    int best = g->Nk;
    for (int nchunk = 1; nchunk <= maxchunks; ++nchunk) {
        const int kc = (g->Nk + nchunk - 1) / nchunk;
        const int nc = (g->Nk + kc - 1) / kc;
        const long long rounds = (nxy * nc + slots - 1) / slots;
        const long long cost = rounds * (kc + 2 * (t.L - 1));
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = kc;
        }
    }
    return best;
}

template <typename S, int V, int R, bool FINAL>
int launch_strip_l(fdtd_grid* g, int slot, int in, int out, void* const* in_b, void* const* ahead) {
    switch (g->btune.L) {
        case 1: return launch_strip_v<S, V, R, 1, FINAL>(g, slot, in, out, in_b, ahead);
        case 2: return launch_strip_v<S, V, R, 2, FINAL>(g, slot, in, out, in_b, ahead);
        case 3: return launch_strip_v<S, V, R, 3, FINAL>(g, slot, in, out, in_b, ahead);
        case 4: return launch_strip_v<S, V, R, 4, FINAL>(g, slot, in, out, in_b, ahead);
        default: return fail(FDTD_ERR_UNSUPPORTED, "unsupported blocking depth %d", g->btune.L);
    }
}

template <typename S, bool FINAL>
int launch_strip_s(fdtd_grid* g, int slot, int in, int out, void* const* in_b, void* const* ahead) {
    const int V = g->btune.V, R = g->btune.R;
    if (V == 1 && R == 2) return launch_strip_l<S, 1, 2, FINAL>(g, slot, in, out, in_b, ahead);
    if (V == 1 && R == 4) return launch_strip_l<S, 1, 4, FINAL>(g, slot, in, out, in_b, ahead);
    if (V == 2 && R == 2) return launch_strip_l<S, 2, 2, FINAL>(g, slot, in, out, in_b, ahead);
    return fail(FDTD_ERR_UNSUPPORTED, "unsupported strip shape V=%d R=%d", V, R);
}

template <typename S, int V, bool FINAL>
int launch_blocked_l(fdtd_grid* g, int slot, int in, int out) {
    switch (g->btune.L) {
        case 1: return launch_blocked_v<S, V, 1, FINAL>(g, slot, in, out);
        case 2: return launch_blocked_v<S, V, 2, FINAL>(g, slot, in, out);
        case 3: return launch_blocked_v<S, V, 3, FINAL>(g, slot, in, out);
        case 4: return launch_blocked_v<S, V, 4, FINAL>(g, slot, in, out);
        default: return fail(FDTD_ERR_UNSUPPORTED, "unsupported blocking depth %d", g->btune.L);
    }
}

template <typename S, bool FINAL>
int launch_blocked_s(fdtd_grid* g, int slot, int in, int out, void* const* in_b,
                     void* const* ahead) {
    if (g->btune.R > 0) return launch_strip_s<S, FINAL>(g, slot, in, out, in_b, ahead);
    switch (g->btune.V) {
        case 1: return launch_blocked_l<S, 1, FINAL>(g, slot, in, out);
        case 2: return launch_blocked_l<S, 2, FINAL>(g, slot, in, out);
        default: return fail(FDTD_ERR_UNSUPPORTED, "unsupported blocked V=%d", g->btune.V);
    }
}

// One pass of btune.L fused steps; its source values are rows slot ..
// slot+L-1 of the value table.  Strip passes may read their B input from
// in_b (a carried half step) and, when final, also store the next half step
// to ahead; the patch kernel takes neither.
int launch_blocked(fdtd_grid* g, bool final, int slot, int in, int out,
                   void* const* in_b = nullptr, void* const* ahead = nullptr) {
    switch (g->precision) {
        case FDTD_FP64:
            return final ? launch_blocked_s<double, true>(g, slot, in, out, in_b, ahead)
                         : launch_blocked_s<double, false>(g, slot, in, out, in_b, ahead);
        case FDTD_FP32:
            return final ? launch_blocked_s<float, true>(g, slot, in, out, in_b, ahead)
                         : launch_blocked_s<float, false>(g, slot, in, out, in_b, ahead);
        default:
            return final ? launch_blocked_s<_Float16, true>(g, slot, in, out, in_b, ahead)
                         : launch_blocked_s<_Float16, false>(g, slot, in, out, in_b, ahead);
    }
}

// Per-launch timing (fdtd_time_update_fields_kernels): events before and
// after each stepping launch, with its kind (0 B/2 prologue, 1 steady pass
// or sweep, 2 a call's last pass or sweep, 3 the single sweeps that follow
// the passes of a blocked call).
struct LaunchTrace {
    struct Rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> pool;
    std::vector<Rec> recs;
    size_t next = 0;
    hipEvent_t take() {
        if (next == pool.size()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[next++];
    }
    ~LaunchTrace() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

hipEvent_t trace_begin(fdtd_grid* g) {
    if (!g->trace) return nullptr;
    hipEvent_t e = g->trace->take();
    if (e) (void)hipEventRecord(e, g->stream);
    return e;
}

// Launch counters (fdtd_get_stats) and the per-launch timing.
void trace_end(fdtd_grid* g, hipEvent_t a, int kind, bool pass) {
    g->stats[kind == 0 ? 3 : (pass ? 1 : 2)] += 1;
    if (!g->trace || !a) return;
    hipEvent_t e = g->trace->take();
    if (!e) return;
    (void)hipEventRecord(e, g->stream);
    g->trace->recs.push_back({kind, a, e});
}

// n steps with temporal blocking: B/2 (or, for strip passes, the carried
// half step), then n / L blocked passes, then the remaining n % L steps as
// single fused sweeps.  Source tables are filled in whole passes.  Strip
// runs carry the next half step out of their last pass or sweep, as the
// unblocked path does.
int run_blocked(fdtd_grid* g, int nsteps, hipEvent_t mid) {
    const bool strips = g->btune.R > 0;
    const int carried = strips ? g->ba_cur : -1;
    g->ba_cur = -1;
    int rc = FDTD_OK;
    if (carried < 0) {
        TraceRange tr("b_half_prologue");
        hipEvent_t t0 = trace_begin(g);
        rc = plain_B(g);
        trace_end(g, t0, 0, false);
    }
    if (rc) return rc;
    const bool carry_out = strips && g->carry && ensure_carry(g);
    const int ahead_buf = carried < 0 ? 0 : 1 - carried;
    if (mid) HIP_TRY(hipEventRecord(mid, g->stream));
    const int L = g->btune.L;
    const int npass = nsteps / L, rem = nsteps % L;
    const int chunk = (kValTable / L) * L;
    int s = 0;
    for (int p = 0; p < npass; ++p, s += L) {
        if (g->autoscale_interval > 0 && s > 0 && (s / L) % std::max(1, g->autoscale_interval / L) == 0) {
            if (strips) {
                // strip passes scale what they read (bit-identical to
                // rescaling the current state in place first)
                double mx[6], snew[6];
                if ((rc = fdtd_max_abs(g, mx))) return rc;
                target_scales(g, mx, g->autoscale_log2, 0.0, snew);
                for (int c = 0; c < 6; ++c) {
                    const double f = snew[c] / g->scales[c];
                    g->pass_iscale[c] = static_cast<float>(f);
                    g->pass_rescale = g->pass_rescale || f != 1.0;
                    g->scales[c] = snew[c];
                }
            } else if ((rc = rescale_to_target(g, true))) {
                return rc;
            }
        }
        if (s % chunk == 0)
            if ((rc = eval_sources(g, g->step + s, std::min(chunk, npass * L - s)))) return rc;
        const int in = g->cur, out = 1 - g->cur;
        const bool last = rem == 0 && p + 1 == npass;
        void* const* in_b = p == 0 && carried >= 0 ? g->BA[carried] : nullptr;
        void* const* ahead = last && carry_out ? g->BA[ahead_buf] : nullptr;
        TraceRange tr(strips ? "strip_pass" : "patch_pass");
        hipEvent_t t0 = trace_begin(g);
        rc = launch_blocked(g, last, s % chunk, in, out, in_b, ahead);
        trace_end(g, t0, last ? 2 : 1, true);
        g->pass_rescale = false;
        for (int c = 0; c < 6; ++c) g->pass_iscale[c] = 1.0f;
        if (rc) return rc;
        g->cur = out;
    }
    for (int q = 0; q < rem; ++q, ++s) {
        if ((rc = eval_sources(g, g->step + s, 1))) return rc;
        const int in = g->cur, out = 1 - g->cur;
        SweepExtra x;
        if (q + 1 == rem && carry_out) x.ahead = g->BA[ahead_buf];
        TraceRange tr("fused_sweep");
        hipEvent_t t0 = trace_begin(g);
        rc = launch_fused(g, g->tune, q + 1 < rem, 0, in, out, x);
        trace_end(g, t0, q + 1 < rem ? 3 : 2, false);
        if (rc) return rc;
        g->cur = out;
    }
    if (carry_out) g->ba_cur = ahead_buf;
    g->step += nsteps;
    return FDTD_OK;
}

// n steps: B half step in place, then n fused sweeps between the buffers.
// With FP16 autoscaling, the state between two sweeps, (E_n, B_{n+1/2}),
// is rescaled by exact powers of two every autoscale_interval sweeps.
// FP16 in-run rescaling: the sweep before a decision point tracks the max
// of what it stores; at the decision point the host reads it (one small
// copy), picks new power-of-two scales, and the next sweep stores its
// output multiplied by new/old (computing with the old coefficients), so
// the state is rescaled without another pass over HBM.  The last sweep of
// a run never rescales (a PML grid's final sweep leaves its B splits
// untouched).
int run_fused(fdtd_grid* g, int nsteps, hipEvent_t mid = nullptr) {
    // Blocked passes take point sources and sparse J values; a dense J array
    // (uploaded J) runs single sweeps, which read it.
    // Probes are recorded by strip passes, not by patch passes.
    if (g->btune.L >= 1 && nsteps >= g->btune.L && !g->pml && !g->j_dense_active &&
        (g->btune.R > 0 || !probing(g)))
        return run_blocked(g, nsteps, mid);
    // A carried half step replaces the B/2 prologue (bit-identical: the
    // final sweep of the previous run computed (B_n + d) + d from the same
    // E_n); otherwise B/2 in place.
    const int carried = g->ba_cur;
    int rc = FDTD_OK;
    if (carried < 0) {
        TraceRange tr("b_half_prologue");
        hipEvent_t t0 = trace_begin(g);
        rc = plain_B(g);
        trace_end(g, t0, 0, false);
    }
    if (rc) return rc;
    const bool carry_out = g->carry && !g->pml && ensure_carry(g);
    const int ahead_buf = carried < 0 ? 0 : 1 - carried;
    g->ba_cur = -1;
    if (mid) HIP_TRY(hipEventRecord(mid, g->stream));
    const int iv = g->precision == FDTD_FP16 ? g->autoscale_interval : 0;
    for (int s = 0; s < nsteps; ++s) {
        const bool decide = iv > 0 && s > 0 && s % iv == 0 && s + 1 < nsteps;
        const bool track = iv > 0 && (s + 1) % iv == 0 && s + 2 < nsteps;
        double f[6] = {1, 1, 1, 1, 1, 1}, snew[6];
        if (decide) {
            unsigned bits[6];
            HIP_TRY(hipMemcpyAsync(bits, g->d_max, sizeof bits, hipMemcpyDeviceToHost, g->stream));
            HIP_TRY(hipStreamSynchronize(g->stream));
            double mx[6];
            for (int c = 0; c < 6; ++c) {
                float m;
                std::memcpy(&m, &bits[c], sizeof m);
                mx[c] = static_cast<double>(m) / g->scales[c];
            }
            target_scales(g, mx, g->autoscale_log2, 0.0, snew);
            for (int c = 0; c < 6; ++c) f[c] = snew[c] / g->scales[c];
        }
        if (track)
            if ((rc = zero_fill(g->d_max, 32, g->stream))) return rc;
        if (s % kValTable == 0)
            if ((rc = eval_sources(g, g->step + s, std::min(kValTable, nsteps - s)))) return rc;
        if (g->pml && g->pml_path == 1 && !decide && !track) {
            // the step as two in-place phase sweeps
            TraceRange tr("pml_phases");
            hipEvent_t t0 = trace_begin(g);
            rc = launch_pml_phases(g, s % kValTable, s + 1 < nsteps);
            trace_end(g, t0, s + 1 < nsteps ? 1 : 2, false);
            if (rc) return rc;
            continue;
        }
        const int in = g->cur, out = 1 - g->cur;
        SweepExtra x;
        x.oscale = decide ? f : nullptr;
        x.maxbits = track ? g->d_max : nullptr;
        if (s == 0 && carried >= 0) x.in_b = g->BA[carried];
        if (s + 1 == nsteps && carry_out) x.ahead = g->BA[ahead_buf];
        TraceRange tr("fused_sweep");
        hipEvent_t t0 = trace_begin(g);
        rc = launch_fused(g, g->tune, s + 1 < nsteps, s % kValTable, in, out, x);
        trace_end(g, t0, s + 1 < nsteps ? 1 : 2, false);
        if (rc) return rc;
        g->cur = out;
        if (s + 1 == nsteps && carry_out) g->ba_cur = ahead_buf;
        if (decide) {
            bool changed = false;
            for (int c = 0; c < 6; ++c) {
                changed = changed || snew[c] != g->scales[c];
                g->scales[c] = snew[c];
            }
            if (changed && g->pml)
                if ((rc = build_pml_tables(g))) return rc;
        }
    }
    g->step += nsteps;
    return FDTD_OK;
}

int run_3sweep(fdtd_grid* g, int nsteps) {
    g->ba_cur = -1;
    for (int s = 0; s < nsteps; ++s) {
        TraceRange tr("three_sweep_step");
        int rc = plain_B(g, true);
        if (!rc) rc = plain_E(g, g->step);
        if (!rc) rc = plain_B(g, false);
        if (rc) return rc;
        g->step += 1;
        if ((rc = gather_probes(g))) return rc;
    }
    return FDTD_OK;
}

// Energy record of the current state: the six sums of squares (physical
// units) reduced into the next row of d_energy, queued without a wait, in
// fdtd_sum_squares' summation order.
int record_energy(fdtd_grid* g) {
    const long long row = static_cast<long long>(g->energy_steps.size());
    if (row >= g->energy_cap) return fail(FDTD_ERR_INVALID_ARGUMENT, "energy records full");
    TraceRange tr("energy_record");
    const int nb = std::min<int>(kReduceBlocks, grid_blocks(g->N));
    for (int c = 0; c < 6; ++c) {
        const void* p = field_ptr(g, c);
        const double inv = 1.0 / g->scales[c];
        double* part = g->d_part + (size_t)2 * kReduceBlocks * c;
        switch (g->precision) {
            case FDTD_FP64:
                hipLaunchKernelGGL(fdtd::sumsq_maxabs_partial<double>, dim3(nb), dim3(256), 0,
                                   g->stream, static_cast<const double*>(p), g->N, inv, part);
                break;
            case FDTD_FP32:
                hipLaunchKernelGGL(fdtd::sumsq_maxabs_partial<float>, dim3(nb), dim3(256), 0,
                                   g->stream, static_cast<const float*>(p), g->N, inv, part);
                break;
            default:
                hipLaunchKernelGGL(fdtd::sumsq_maxabs_partial<_Float16>, dim3(nb), dim3(256), 0,
                                   g->stream, static_cast<const _Float16*>(p), g->N, inv, part);
        }
        hipLaunchKernelGGL(fdtd::sum_partials_col0, dim3(1), dim3(256), 0, g->stream, part, nb, 2,
                           g->d_energy + 6 * row + c);
    }
    HIP_TRY(hipGetLastError());
    g->energy_steps.push_back(g->step);
    return FDTD_OK;
}

// Energy records a call of nsteps steps adds: the steps s in
// (step, step + nsteps] with s a multiple of energy_every.
long long energy_records_in(const fdtd_grid* g, long long nsteps) {
    if (g->energy_every <= 0 || nsteps <= 0) return 0;
    const long long e = g->energy_every;
    auto floor_div = [e](long long a) { return a >= 0 ? a / e : -((-a + e - 1) / e); };
    return floor_div(g->step + nsteps) - floor_div(g->step);
}

int check_record_capacity(const fdtd_grid* g, long long nsteps) {
    if (g->nprobe > 0 && g->probe_rows + nsteps > g->probe_cap)
        return fail(FDTD_ERR_INVALID_ARGUMENT,
                    "probe rows full: %lld of %lld used, %lld more steps asked",
                    g->probe_rows, g->probe_cap, nsteps);
    if ((long long)g->energy_steps.size() + energy_records_in(g, nsteps) > g->energy_cap &&
        g->energy_every > 0)
        return fail(FDTD_ERR_INVALID_ARGUMENT, "energy records full: %zu of %lld used",
                    g->energy_steps.size(), g->energy_cap);
    return FDTD_OK;
}

// The steps of the stepping entry points, recording probes and energy.
// Energy records split the run at the recorded steps: every run leaves the
// state (E_n, B_n) in memory, and the carried half step makes the split
// bit-identical.  Probes of PML grids are gathered after single-step runs
// (their sweeps keep B half a step ahead in memory).
int run_steps(fdtd_grid* g, int nsteps, bool three_sweep) {
    auto segment = [&](int n) -> int {
        if (three_sweep || !g->fused_ok) return run_3sweep(g, n);
        if (probing(g) && g->pml) {
            for (int s = 0; s < n; ++s) {
                if (int rc = run_fused(g, 1)) return rc;
                if (int rc = gather_probes(g)) return rc;
            }
            return FDTD_OK;
        }
        return run_fused(g, n);
    };
    if (int rc = check_record_capacity(g, nsteps)) return rc;
    TraceRange tr(three_sweep ? "update_fields_3sweep" : "update_fields");
    g->recording = true;
    int rc = FDTD_OK;
    if (g->energy_every <= 0) {
        rc = segment(nsteps);
    } else {
        const long long e = g->energy_every;
        for (int done = 0; done < nsteps && rc == FDTD_OK;) {
            const long long ph = ((g->step % e) + e) % e;
            const int m = static_cast<int>(std::min<long long>(nsteps - done, e - ph));
            rc = segment(m);
            done += m;
            if (rc == FDTD_OK && ((g->step % e) + e) % e == 0) rc = record_energy(g);
        }
    }
    g->recording = false;
    return rc;
}

// ---- conversions and reductions ------------------------------------------------

int ensure_staging(fdtd_grid* g) {
    if (!g->staging) HIP_TRY(hipMalloc(&g->staging, g->N * sizeof(double)));
    return FDTD_OK;
}

template <typename S>
void launch_from_f64(const double* src, void* dst, size_t n, double scale, hipStream_t s) {
    hipLaunchKernelGGL(fdtd::convert_from_f64<S>, dim3(grid_blocks(n)), dim3(256), 0, s, src,
                       static_cast<S*>(dst), n, scale);
}

template <typename S>
void launch_to_f64(const void* src, double* dst, size_t n, double inv, hipStream_t s) {
    hipLaunchKernelGGL(fdtd::convert_to_f64<S>, dim3(grid_blocks(n)), dim3(256), 0, s,
                       static_cast<const S*>(src), dst, n, inv);
}

void from_f64(int prec, const double* src, void* dst, size_t n, double scale, hipStream_t s) {
    if (prec == FDTD_FP32) launch_from_f64<float>(src, dst, n, scale, s);
    else launch_from_f64<_Float16>(src, dst, n, scale, s);
}

void to_f64(int prec, const void* src, double* dst, size_t n, double inv, hipStream_t s) {
    if (prec == FDTD_FP32) launch_to_f64<float>(src, dst, n, inv, s);
    else launch_to_f64<_Float16>(src, dst, n, inv, s);
}

int zero_fill(void* p, size_t bytes, hipStream_t s) {
    const size_t n16 = bytes / 16;
    if (n16)
        hipLaunchKernelGGL(fdtd::zero_bytes, dim3(grid_blocks(n16)), dim3(256), 0, s,
                           static_cast<uint4*>(p), n16);
    const size_t tail = bytes - n16 * 16;
    if (tail) HIP_TRY(hipMemsetAsync(static_cast<char*>(p) + n16 * 16, 0, tail, s));
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

int ensure_dense_j(fdtd_grid* g) {
    const size_t bytes = g->N * current_size(g->precision);
    for (int d = 0; d < 3; ++d)
        if (!g->J[d]) HIP_TRY(hipMalloc(&g->J[d], bytes + 16));
    return FDTD_OK;
}

void* field_ptr(fdtd_grid* g, int c);

// sum x^2 and max|x| of the six E and B components (physical units): all
// partial and final passes queued back to back, one copy and one wait.
// parts: 12 * kReduceBlocks doubles of per-workgroup partials (each one
// written before it is read); res: the 12 final values.
int reduce_fields_into(fdtd_grid* g, double* parts, double* res, double sumsq[6],
                       double maxabs[6]) {
    const int nb = std::min<int>(kReduceBlocks, grid_blocks(g->N));
    for (int c = 0; c < 6; ++c) {
        const void* p = field_ptr(g, c);
        const double inv = 1.0 / g->scales[c];
        double* part = parts + (size_t)2 * kReduceBlocks * c;
        switch (g->precision) {
            case FDTD_FP64:
                hipLaunchKernelGGL(fdtd::sumsq_maxabs_partial<double>, dim3(nb), dim3(256), 0,
                                   g->stream, static_cast<const double*>(p), g->N, inv, part);
                break;
            case FDTD_FP32:
                hipLaunchKernelGGL(fdtd::sumsq_maxabs_partial<float>, dim3(nb), dim3(256), 0,
                                   g->stream, static_cast<const float*>(p), g->N, inv, part);
                break;
            default:
                hipLaunchKernelGGL(fdtd::sumsq_maxabs_partial<_Float16>, dim3(nb), dim3(256), 0,
                                   g->stream, static_cast<const _Float16*>(p), g->N, inv, part);
        }
        hipLaunchKernelGGL(fdtd::reduce_partials, dim3(1), dim3(256), 0, g->stream, part, nb, 2,
                           2u, res + 2 * c);
    }
    HIP_TRY(hipGetLastError());
    double r[12];
    HIP_TRY(hipMemcpyAsync(r, res, sizeof r, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    for (int c = 0; c < 6; ++c) {
        if (sumsq) sumsq[c] = r[2 * c];
        if (maxabs) maxabs[c] = r[2 * c + 1];
    }
    return FDTD_OK;
}

int reduce_fields(fdtd_grid* g, double sumsq[6], double maxabs[6]) {
    return reduce_fields_into(g, g->d_part, g->d_out, sumsq, maxabs);
}

int check_component(int c) {
    if (c < 0 || c > 8) return fail(FDTD_ERR_INVALID_COMPONENT, "ERROR: Invalid field component");
    return FDTD_OK;
}

void* field_ptr(fdtd_grid* g, int c) {
    return c < 3 ? g->E[g->cur][c] : g->B[g->cur][c - 3];
}

}  // namespace fdtd::rt

// ---- allocation-free diagnostics --------------------------------------------
// Scratch of fdtd_sum_squares_noalloc: per-workgroup partial sums and maxima
// of the six components (at most kReduceBlocks workgroups each, every one
// written before it is read) and the 12 final values: 98,432 bytes.  Each GPU
// holds its own copy until the process ends; every call on a GPU shares it,
// so calls that overlap on one GPU (two host threads, or grids on different
// streams) race.  Host code uses the per-device address from
// hipGetSymbolAddress, never the array's name.
namespace {

constexpr size_t kNoAllocScratchDoubles = 12 * fdtd::rt::kReduceBlocks + 16;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on device dev, which must be current; looked
// up on the first call on each device.
int scratch_on_device(int dev, double** out) {
    if (dev < 0 || dev >= kMaxDevices)
        return fdtd::rt::fail(FDTD_ERR_UNSUPPORTED, "device %d beyond the %d supported", dev,
                              kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fdtd::rt::fail(FDTD_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                                  hipGetErrorString(e));
        g_scratch_addr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_addr[dev];
    return FDTD_OK;
}

}  // namespace

// =============================================================================
// C-ABI
// =============================================================================

using namespace fdtd::rt;

extern "C" {

const char* fdtd_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

namespace fdtd::rt {

// Common constructor of fdtd_create and fdtd_create_pml (pml_percent < 0:
// no layer).
int create_grid(const fdtd_parameters* p, double dt, double pml_percent, int precision,
                fdtd_grid** out) {
    if (!p || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    // src/FDTD/FDTD.cpp:5-7
    if (p->Ni <= 0 || p->Nj <= 0 || p->Nk <= 0 || !(dt > 0))
        return fail(FDTD_ERR_INVALID_PARAMETERS, "ERROR: invalid parameters");
    if (precision < FDTD_FP64 || precision > FDTD_FP16)
        return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid precision %d", precision);
    const long long cells = (long long)p->Ni * p->Nj * p->Nk;
    if (cells >= (1LL << 31))
        return fail(FDTD_ERR_INVALID_PARAMETERS, "grid of %lld cells exceeds 2^31-1", cells);
    fdtd_grid* g = new fdtd_grid();
    g->params = *p;
    g->Ni = p->Ni;
    g->Nj = p->Nj;
    g->Nk = p->Nk;
    g->N = (size_t)cells;
    g->dt = dt;
    g->precision = precision;
    g->esize = storage_size(precision);
    if (pml_percent >= 0) {
        g->pml_percent = pml_percent;
        g->pml_p[0] = pml_width(p->Ni, pml_percent);
        g->pml_p[1] = pml_width(p->Nj, pml_percent);
        g->pml_p[2] = pml_width(p->Nk, pml_percent);
        // all widths zero: the class is the plain FDTD one
        g->pml = g->pml_p[0] > 0 || g->pml_p[1] > 0 || g->pml_p[2] > 0;
    }
    auto cleanup = [&](int rc) {
        fdtd_destroy(g);
        return rc;
    };
    if (hipGetDevice(&g->device) != hipSuccess)
        return cleanup(fail(FDTD_ERR_HIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(FDTD_ERR_HIP, "hipStreamCreate failed"));
    const size_t bytes = g->N * g->esize + 16;
    for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 3; ++d) {
            if (hipMalloc(&g->E[b][d], bytes) != hipSuccess ||
                hipMalloc(&g->B[b][d], bytes) != hipSuccess)
                return cleanup(fail(FDTD_ERR_HIP, "out of device memory for %lld cells", cells));
        }
    for (int d = 0; d < 3; ++d) {
        int rc = zero_fill(g->E[0][d], bytes, g->stream);
        if (!rc) rc = zero_fill(g->B[0][d], bytes, g->stream);
        if (rc) return cleanup(rc);
    }
    if (g->pml) {
        if (int rc = alloc_pml(g)) return cleanup(rc);
        g->ptune = pml_phase_default(g);
    }
    if (hipMalloc(&g->d_part, sizeof(double) * 12 * kReduceBlocks) != hipSuccess ||
        hipMalloc(&g->d_out, sizeof(double) * 16) != hipSuccess ||
        hipMalloc(&g->d_max, 32) != hipSuccess)
        return cleanup(fail(FDTD_ERR_HIP, "out of device memory"));
    for (auto& e : g->ev)
        if (hipEventCreate(&e) != hipSuccess) return cleanup(fail(FDTD_ERR_HIP, "hipEventCreate"));
    for (auto& e : g->src_ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(FDTD_ERR_HIP, "hipEventCreate"));
    g->fused_ok = default_tuning(g, &g->tune);
    // Periodic grids large enough to fill the device with strip tiles step
    // in temporally blocked passes of 4 steps by default (bit-identical to
    // single sweeps; FP16 with in-run rescaling to FP16 rounding), with the
    // shape the autotuner picks at 512^3; fdtd_set_time_blocking_strips(grid,
    // 0, ...) returns to single sweeps.
    if (!g->pml && g->fused_ok && g->Ni >= 128 && g->Nj >= 64 && g->Nk >= 32) {
        BlockTuning t;
        t.L = 4;
        t.V = 1;
        t.R = 2;
        t.lanes = 64;
        t.TYg = (precision == FDTD_FP64 ? 8 : 16) * t.R;
        t.kchunk = 1;
        if (strip_valid(g, t)) {
            t.kchunk = strip_default_kchunk(g, t);
            g->btune = t;
        }
    }
    if (hipStreamSynchronize(g->stream) != hipSuccess)
        return cleanup(fail(FDTD_ERR_HIP, "initialisation failed"));
    *out = g;
    return FDTD_OK;
}

}  // namespace fdtd::rt

extern "C" {

int fdtd_create(const fdtd_parameters* p, double dt, int precision, fdtd_grid** out) {
    return create_grid(p, dt, -1.0, precision, out);
}

int fdtd_create_pml(const fdtd_parameters* p, double dt, double pml_percent, int precision,
                    fdtd_grid** out) {
    if (out) *out = nullptr;
    // 2*width <= N keeps the six layer slabs of FDTD_PML.cpp:346-362
    // disjoint (the reference updates overlapping cells twice).
    if (!(pml_percent >= 0.0 && pml_percent <= 0.5))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "pml_percent %g outside [0, 0.5]", pml_percent);
    return create_grid(p, dt, pml_percent, precision, out);
}

int fdtd_destroy(fdtd_grid* g) {
    if (!g) return FDTD_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 3; ++d) {
            if (g->E[b][d]) (void)hipFree(g->E[b][d]);
            if (g->B[b][d]) (void)hipFree(g->B[b][d]);
        }
    for (int q = 0; q < 6; ++q) {
        for (int b = 0; b < 2; ++b)
            if (g->ES[b][q]) (void)hipFree(g->ES[b][q]);
        if (g->BS[q]) (void)hipFree(g->BS[q]);
    }
    if (g->pml_tab) (void)hipFree(g->pml_tab);
    for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 3; ++d)
            if (g->BA[b][d]) (void)hipFree(g->BA[b][d]);
    for (int d = 0; d < 3; ++d)
        if (g->J[d]) (void)hipFree(g->J[d]);
    if (g->staging) (void)hipFree(g->staging);
    if (g->d_jtab) (void)hipFree(g->d_jtab);
    if (g->d_jw) (void)hipFree(g->d_jw);
    for (int b = 0; b < 2; ++b) {
        if (g->d_src_buf[b]) (void)hipFree(g->d_src_buf[b]);
        if (g->h_src[b]) (void)hipHostFree(g->h_src[b]);
        if (g->d_kst_buf[b]) (void)hipFree(g->d_kst_buf[b]);
        if (g->h_kst[b]) (void)hipHostFree(g->h_kst[b]);
        if (g->src_ev[b]) (void)hipEventDestroy(g->src_ev[b]);
    }
    if (g->d_vals) (void)hipFree(g->d_vals);
    if (g->d_part) (void)hipFree(g->d_part);
    if (g->d_out) (void)hipFree(g->d_out);
    if (g->d_max) (void)hipFree(g->d_max);
    if (g->d_probe) (void)hipFree(g->d_probe);
    if (g->d_pstart) (void)hipFree(g->d_pstart);
    if (g->d_probe_out) (void)hipFree(g->d_probe_out);
    if (g->d_energy) (void)hipFree(g->d_energy);
    for (auto& e : g->ev)
        if (e) (void)hipEventDestroy(e);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return FDTD_OK;
}

int fdtd_get_pml(const fdtd_grid* g, int widths[3], double* pml_percent) {
    if (!g) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (widths)
        for (int a = 0; a < 3; ++a) widths[a] = g->pml_p[a];
    if (pml_percent) *pml_percent = g->pml_percent;
    return FDTD_OK;
}

int fdtd_pml_download_split(fdtd_grid* g, int q, double* host, size_t count) {
    if (!g || !host) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (q < 0 || q > 11) return fail(FDTD_ERR_INVALID_ARGUMENT, "split index %d outside 0..11", q);
    if (count != g->N)
        return fail(FDTD_ERR_INVALID_ARGUMENT, "count %zu != grid cells %zu", count, g->N);
    if (!g->pml) {
        std::memset(host, 0, g->N * sizeof(double));
        return FDTD_OK;
    }
    const void* src = split_ptr(g, q);
    if (g->precision == FDTD_FP64) {
        HIP_TRY(hipMemcpyAsync(host, src, g->N * 8, hipMemcpyDeviceToHost, g->stream));
    } else {
        if (int rc = ensure_staging(g)) return rc;
        to_f64(g->precision, src, g->staging, g->N, 1.0 / g->scales[split_component(q)], g->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(host, g->staging, g->N * 8, hipMemcpyDeviceToHost, g->stream));
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    return FDTD_OK;
}

int fdtd_pml_profile(int n, double pml_percent, double d, double dt, double* sigma,
                     double* decay, double* coef) {
    if (n <= 0 || !(d > 0) || !(dt > 0) || !(pml_percent >= 0.0 && pml_percent <= 0.5))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    const std::vector<double> s = pml_sigma(n, pml_width(n, pml_percent), d);
    for (int i = 0; i < n; ++i) {
        double a, c;
        pml_coef(s[i], dt, d, &a, &c);
        if (sigma) sigma[i] = s[i];
        if (decay) decay[i] = a;
        if (coef) coef[i] = c;
    }
    return FDTD_OK;
}

int fdtd_upload(fdtd_grid* g, int c, const double* host, size_t count) {
    if (!g || !host) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc = check_component(c)) return rc;
    if (count != g->N)
        return fail(FDTD_ERR_INVALID_ARGUMENT, "count %zu != grid cells %zu", count, g->N);
    if (c >= 6) {
        if (int rc = ensure_dense_j(g)) return rc;
        const int d = c - 6;
        if (g->precision == FDTD_FP64) {
            HIP_TRY(hipMemcpyAsync(g->J[d], host, g->N * 8, hipMemcpyHostToDevice, g->stream));
        } else {
            if (int rc = ensure_staging(g)) return rc;
            HIP_TRY(hipMemcpyAsync(g->staging, host, g->N * 8, hipMemcpyHostToDevice, g->stream));
            launch_from_f64<float>(g->staging, g->J[d], g->N, 1.0, g->stream);
        }
        if (!g->j_dense_active) {
            // the other two components start from zero
            for (int o = 0; o < 3; ++o)
                if (o != d)
                    if (int rc = zero_fill(g->J[o], g->N * current_size(g->precision), g->stream))
                        return rc;
            g->j_dense_active = true;
        }
        for (auto it = g->sparse_j.begin(); it != g->sparse_j.end();) {
            if (it->first % 3 == d) it = g->sparse_j.erase(it);
            else ++it;
        }
        g->src_dirty = true;
        HIP_TRY(hipStreamSynchronize(g->stream));
        return FDTD_OK;
    }
    g->ba_cur = -1;  // the carried half step no longer matches E, B
    void* dst = field_ptr(g, c);
    if (g->precision == FDTD_FP64) {
        HIP_TRY(hipMemcpyAsync(dst, host, g->N * 8, hipMemcpyHostToDevice, g->stream));
    } else {
        if (int rc = ensure_staging(g)) return rc;
        HIP_TRY(hipMemcpyAsync(g->staging, host, g->N * 8, hipMemcpyHostToDevice, g->stream));
        from_f64(g->precision, g->staging, dst, g->N, g->scales[c], g->stream);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    return FDTD_OK;
}

int fdtd_download(fdtd_grid* g, int c, double* host, size_t count) {
    if (!g || !host) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc = check_component(c)) return rc;
    if (count != g->N)
        return fail(FDTD_ERR_INVALID_ARGUMENT, "count %zu != grid cells %zu", count, g->N);
    if (c >= 6) {
        const int d = c - 6;
        if (g->j_dense_active) {
            if (g->precision == FDTD_FP64) {
                HIP_TRY(hipMemcpyAsync(host, g->J[d], g->N * 8, hipMemcpyDeviceToHost, g->stream));
            } else {
                if (int rc = ensure_staging(g)) return rc;
                launch_to_f64<float>(g->J[d], g->staging, g->N, 1.0, g->stream);
                HIP_TRY(hipMemcpyAsync(host, g->staging, g->N * 8, hipMemcpyDeviceToHost, g->stream));
            }
            HIP_TRY(hipStreamSynchronize(g->stream));
        } else {
            std::memset(host, 0, g->N * sizeof(double));
        }
        for (const auto& kv : g->sparse_j)
            if (kv.first % 3 == d) host[kv.first / 3] = kv.second;
        return FDTD_OK;
    }
    const void* src = field_ptr(g, c);
    if (g->precision == FDTD_FP64) {
        HIP_TRY(hipMemcpyAsync(host, src, g->N * 8, hipMemcpyDeviceToHost, g->stream));
    } else {
        if (int rc = ensure_staging(g)) return rc;
        to_f64(g->precision, src, g->staging, g->N, 1.0 / g->scales[c], g->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(host, g->staging, g->N * 8, hipMemcpyDeviceToHost, g->stream));
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    return FDTD_OK;
}

int fdtd_download_plane(fdtd_grid* g, int c, int k, double* host, size_t count) {
    if (!g || !host) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc = check_component(c)) return rc;
    const size_t plane = (size_t)g->Ni * g->Nj;
    if (k < 0 || k >= g->Nk) return fail(FDTD_ERR_INVALID_ARGUMENT, "plane %d outside 0..%d", k, g->Nk - 1);
    if (count != plane)
        return fail(FDTD_ERR_INVALID_ARGUMENT, "count %zu != plane cells %zu", count, plane);
    const size_t off = (size_t)k * plane;
    if (c >= 6) {  // J: dense values of the plane, then the sparse ones
        const int d = c - 6;
        std::memset(host, 0, plane * sizeof(double));
        if (g->j_dense_active) {
            if (g->precision == FDTD_FP64) {
                HIP_TRY(hipMemcpyAsync(host, static_cast<const double*>(g->J[d]) + off, plane * 8,
                                       hipMemcpyDeviceToHost, g->stream));
            } else {
                if (int rc = ensure_staging(g)) return rc;
                launch_to_f64<float>(static_cast<const float*>(g->J[d]) + off, g->staging, plane, 1.0,
                                     g->stream);
                HIP_TRY(hipMemcpyAsync(host, g->staging, plane * 8, hipMemcpyDeviceToHost, g->stream));
            }
            HIP_TRY(hipStreamSynchronize(g->stream));
        }
        for (const auto& kv : g->sparse_j) {
            const long long idx = kv.first / 3;
            if (kv.first % 3 == d && (size_t)idx >= off && (size_t)idx < off + plane)
                host[idx - off] = kv.second;
        }
        return FDTD_OK;
    }
    const char* src = static_cast<const char*>(field_ptr(g, c)) + off * g->esize;
    if (g->precision == FDTD_FP64) {
        HIP_TRY(hipMemcpyAsync(host, src, plane * 8, hipMemcpyDeviceToHost, g->stream));
    } else {
        if (int rc = ensure_staging(g)) return rc;
        to_f64(g->precision, src, g->staging, plane, 1.0 / g->scales[c], g->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(host, g->staging, plane * 8, hipMemcpyDeviceToHost, g->stream));
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    return FDTD_OK;
}

int fdtd_zero_currents(fdtd_grid* g) {
    if (!g) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (!g->sparse_j.empty()) g->src_dirty = true;
    g->sparse_j.clear();
    if (g->j_device) {
        // the caller holds J's addresses: clear the arrays in place, queued
        // after the steps already launched (Kokkos::deep_copy(J, 0),
        // src/FDTD_kokkos/FDTD_kokkos.cpp:69-73)
        for (int d = 0; d < 3; ++d)
            if (int rc = zero_fill(g->J[d], g->N * current_size(g->precision), g->stream)) return rc;
        return FDTD_OK;
    }
    g->j_dense_active = false;
    return FDTD_OK;
}

int fdtd_device_field(fdtd_grid* g, int c, void** ptr, int* elem_bytes, double* scale) {
    if (!g || !ptr) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc = check_component(c)) return rc;
    if (c >= FDTD_JX) {
        if (!g->j_device) {
            if (int rc = ensure_dense_j(g)) return rc;
            if (!g->j_dense_active) {
                for (int d = 0; d < 3; ++d)
                    if (int rc = zero_fill(g->J[d], g->N * current_size(g->precision), g->stream))
                        return rc;
                g->j_dense_active = true;
                if (!g->sparse_j.empty()) {
                    // single-cell currents set so far move into the arrays
                    std::vector<long long> idx;
                    std::vector<int> cc;
                    std::vector<double> vv;
                    for (const auto& kv : g->sparse_j) {
                        idx.push_back(kv.first / 3);
                        cc.push_back((int)(kv.first % 3) + FDTD_JX);
                        vv.push_back(kv.second);
                    }
                    g->sparse_j.clear();
                    g->src_dirty = true;
                    if (int rc = fdtd_set_current_values(g, (int)idx.size(), idx.data(), cc.data(),
                                                         vv.data()))
                        return rc;
                }
            }
            g->j_device = true;
        }
        *ptr = g->J[c - FDTD_JX];
        if (elem_bytes) *elem_bytes = (int)current_size(g->precision);
        if (scale) *scale = 1.0;
        return FDTD_OK;
    }
    g->ba_cur = -1;  // the caller may change E or B before the next step
    *ptr = field_ptr(g, c);
    if (elem_bytes) *elem_bytes = (int)g->esize;
    if (scale) *scale = g->scales[c];
    return FDTD_OK;
}

int fdtd_get_stream(fdtd_grid* g, void** stream) {
    if (!g || !stream) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    *stream = static_cast<void*>(g->stream);
    return FDTD_OK;
}

static int box_entry(fdtd_grid* g, const int lo[3], const int hi[3], bool e_step) {
    if (!g || !lo || !hi) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    const int n[3] = {g->Ni, g->Nj, g->Nk};
    fdtd::Box box;
    for (int a = 0; a < 3; ++a) {
        if (lo[a] < 0 || hi[a] > n[a] || lo[a] > hi[a])
            return fail(FDTD_ERR_INVALID_ARGUMENT, "box [%d, %d) outside axis %d of %d cells",
                        lo[a], hi[a], a, n[a]);
        box.lo[a] = lo[a];
        box.hi[a] = hi[a];
    }
    g->ba_cur = -1;  // E, B change outside a run: the carried half step is stale
    if (int rc = flush_sources(g)) return rc;
    if (lo[0] == hi[0] || lo[1] == hi[1] || lo[2] == hi[2]) return FDTD_OK;
    TraceRange tr(e_step ? "update_E_box" : "update_B_box");
    return box_update(g, box, e_step);
}

int fdtd_update_b_box(fdtd_grid* g, const int lo[3], const int hi[3]) {
    return box_entry(g, lo, hi, false);
}

int fdtd_update_e_box(fdtd_grid* g, const int lo[3], const int hi[3]) {
    return box_entry(g, lo, hi, true);
}

int fdtd_set_current_values(fdtd_grid* g, int n, const long long* index, const int* comp,
                            const double* value) {
    if (!g || n < 0 || (n > 0 && (!index || !comp || !value)))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    for (int q = 0; q < n; ++q) {
        if (comp[q] < FDTD_JX || comp[q] > FDTD_JZ)
            return fail(FDTD_ERR_INVALID_COMPONENT, "ERROR: Invalid field component");
        if (index[q] < 0 || (size_t)index[q] >= g->N)
            return fail(FDTD_ERR_INVALID_ARGUMENT, "index %lld out of range", index[q]);
    }
    if (g->j_dense_active) {
        // scatter into the dense arrays: one upload and one kernel, each
        // cell once with the last value written to it (the order of the
        // writes, as the reference's element assignments)
        if (n == 0) return FDTD_OK;
        std::map<long long, double> last;
        for (int q = 0; q < n; ++q) last[index[q] * 3 + (comp[q] - FDTD_JX)] = value[q];
        std::vector<fdtd::CurrentWrite> w;
        w.reserve(last.size());
        for (const auto& kv : last)
            w.push_back({kv.first / 3, static_cast<int>(kv.first % 3), 0, kv.second});
        if (w.size() > g->d_jw_cap) {
            HIP_TRY(hipStreamSynchronize(g->stream));
            if (g->d_jw) HIP_TRY(hipFree(g->d_jw));
            g->d_jw = nullptr;
            g->d_jw_cap = 0;
            const size_t cap = std::max<size_t>(64, 2 * w.size());
            HIP_TRY(hipMalloc(&g->d_jw, cap * sizeof(fdtd::CurrentWrite)));
            g->d_jw_cap = cap;
        }
        HIP_TRY(hipMemcpyAsync(g->d_jw, w.data(), w.size() * sizeof(fdtd::CurrentWrite),
                               hipMemcpyHostToDevice, g->stream));
        const int nw = static_cast<int>(w.size());
        if (g->precision == FDTD_FP64)
            hipLaunchKernelGGL(fdtd::scatter_currents<double>, dim3((nw + 255) / 256), dim3(256), 0,
                               g->stream, static_cast<double*>(g->J[0]),
                               static_cast<double*>(g->J[1]), static_cast<double*>(g->J[2]),
                               g->d_jw, nw);
        else
            hipLaunchKernelGGL(fdtd::scatter_currents<float>, dim3((nw + 255) / 256), dim3(256), 0,
                               g->stream, static_cast<float*>(g->J[0]),
                               static_cast<float*>(g->J[1]), static_cast<float*>(g->J[2]),
                               g->d_jw, nw);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(g->stream));  // w is released on return
        return FDTD_OK;
    }
    for (int q = 0; q < n; ++q)
        g->sparse_j[index[q] * 3 + (comp[q] - FDTD_JX)] = value[q];
    g->src_dirty = true;
    if ((int)g->sparse_j.size() > kMaxSparseCurrents) {
        // Too many cells for the per-plane source scan: materialise dense J.
        if (int rc = ensure_dense_j(g)) return rc;
        for (int d = 0; d < 3; ++d)
            if (int rc = zero_fill(g->J[d], g->N * current_size(g->precision), g->stream)) return rc;
        g->j_dense_active = true;
        std::vector<long long> idx;
        std::vector<int> cc;
        std::vector<double> vv;
        for (const auto& kv : g->sparse_j) {
            idx.push_back(kv.first / 3);
            cc.push_back((int)(kv.first % 3) + FDTD_JX);
            vv.push_back(kv.second);
        }
        g->sparse_j.clear();
        return fdtd_set_current_values(g, (int)idx.size(), idx.data(), cc.data(), vv.data());
    }
    return FDTD_OK;
}

int fdtd_set_point_sources(fdtd_grid* g, int n, const fdtd_point_source* src) {
    if (!g || n < 0 || (n > 0 && !src)) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    std::vector<fdtd::PointSource> w;
    for (int q = 0; q < n; ++q) {
        const fdtd_point_source& s = src[q];
        if (s.i < 0 || s.i >= g->Ni || s.j < 0 || s.j >= g->Nj || s.k < 0 || s.k >= g->Nk)
            return fail(FDTD_ERR_INVALID_ARGUMENT, "source %d outside the grid", q);
        if (s.comp < 0 || s.comp > 2)
            return fail(FDTD_ERR_INVALID_COMPONENT, "ERROR: Invalid field component");
        if (s.kind != 0 && s.kind != 1)
            return fail(FDTD_ERR_INVALID_ARGUMENT, "source %d: unknown kind %d", q, s.kind);
        if (s.kind == 1 && !(s.period != 0.0))
            return fail(FDTD_ERR_INVALID_ARGUMENT, "source %d: zero period", q);
        fdtd::PointSource p{};
        p.i = s.i; p.j = s.j; p.k = s.k; p.comp = s.comp; p.kind = s.kind;
        p.amp = s.amp; p.period = s.period; p.t0 = s.t0; p.sigma = s.sigma; p.phase = s.phase;
        w.push_back(p);
    }
    g->waveforms = w;
    g->src_dirty = true;
    return FDTD_OK;
}

int fdtd_update_fields(fdtd_grid* g, int nsteps) {
    if (!g || nsteps < 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (nsteps == 0) return FDTD_OK;
    if (int rc = flush_sources(g)) return rc;
    g->stats[0] += nsteps;
    return run_steps(g, nsteps, false);
}

int fdtd_update_fields_with_currents(fdtd_grid* g, int nsteps, const int* counts,
                                     const long long* index, const int* comp, const double* value) {
    if (!g || nsteps < 0 || (nsteps > 0 && !counts))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    long long total = 0;
    for (int s = 0; s < nsteps; ++s) {
        if (counts[s] < 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "negative count at step %d", s);
        total += counts[s];
    }
    if (total == 0) return fdtd_update_fields(g, nsteps);
    if (!index || !comp || !value) return fail(FDTD_ERR_INVALID_ARGUMENT, "null current values");
    if (int rc = check_record_capacity(g, nsteps)) return rc;
    for (long long q = 0; q < total; ++q) {
        if (comp[q] < FDTD_JX || comp[q] > FDTD_JZ)
            return fail(FDTD_ERR_INVALID_COMPONENT, "ERROR: Invalid field component");
        if (index[q] < 0 || (size_t)index[q] >= g->N)
            return fail(FDTD_ERR_INVALID_ARGUMENT, "index %lld out of range", index[q]);
    }
    // The cells the call touches (the current sparse cells and every cell
    // written), and each step's values: a write persists until the next one.
    std::map<long long, double> cur = g->sparse_j;
    for (long long q = 0; q < total; ++q) cur.emplace(index[q] * 3 + (comp[q] - FDTD_JX), 0.0);
    const bool tabulate = !g->j_dense_active && (int)cur.size() <= kMaxSparseCurrents;
    if (!tabulate) {
        // dense J (or too many cells for the per-plane source scan): the
        // steps one at a time, each after its writes
        long long q = 0;
        for (int s = 0; s < nsteps; ++s) {
            if (counts[s] > 0) {
                if (int rc = fdtd_set_current_values(g, counts[s], index + q, comp + q, value + q))
                    return rc;
                q += counts[s];
            }
            if (int rc = fdtd_update_fields(g, 1)) return rc;
        }
        return FDTD_OK;
    }
    std::vector<long long> keys;
    std::map<long long, int> col;
    for (const auto& kv : cur) {
        col[kv.first] = (int)keys.size();
        keys.push_back(kv.first);
    }
    std::map<long long, double> now = g->sparse_j;  // values in effect
    std::vector<double> tab((size_t)nsteps * keys.size(), 0.0);
    long long q = 0;
    for (int s = 0; s < nsteps; ++s) {
        for (int e = 0; e < counts[s]; ++e, ++q) now[index[q] * 3 + (comp[q] - FDTD_JX)] = value[q];
        double* row = tab.data() + (size_t)s * keys.size();
        for (const auto& kv : now) row[col[kv.first]] = kv.second;
    }
    if (tab.size() > g->d_jtab_cap) {
        HIP_TRY(hipStreamSynchronize(g->stream));
        if (g->d_jtab) HIP_TRY(hipFree(g->d_jtab));
        g->d_jtab = nullptr;
        g->d_jtab_cap = 0;
        HIP_TRY(hipMalloc(&g->d_jtab, sizeof(double) * tab.size()));
        g->d_jtab_cap = tab.size();
    }
    // the table is read by this call's source evaluations, queued on the
    // grid's stream: a synchronous copy keeps the host buffer's lifetime simple
    HIP_TRY(hipStreamSynchronize(g->stream));
    HIP_TRY(hipMemcpy(g->d_jtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    g->jtab_keys = keys;
    g->jtab_step0 = g->step;
    g->jtab_rows = nsteps;
    g->src_dirty = true;
    int rc = fdtd_update_fields(g, nsteps);
    // the values in effect after the call, as single sparse values
    g->jtab_keys.clear();
    g->jtab_rows = 0;
    g->sparse_j = now;
    g->src_dirty = true;
    return rc;
}

int fdtd_set_probes(fdtd_grid* g, int n, const long long* index, const int* comp,
                    long long capacity) {
    if (!g || n < 0 || (n > 0 && (!index || !comp || capacity <= 0)))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (n > 0 && capacity > (long long)(SIZE_MAX / sizeof(double)) / n)
        return fail(FDTD_ERR_INVALID_ARGUMENT, "probe capacity too large");
    std::vector<fdtd::Probe> list(n);
    const size_t plane = (size_t)g->Ni * g->Nj;
    for (int q = 0; q < n; ++q) {
        if (comp[q] < FDTD_EX || comp[q] > FDTD_BZ)
            return fail(FDTD_ERR_INVALID_COMPONENT, "ERROR: Invalid field component");
        if (index[q] < 0 || (size_t)index[q] >= g->N)
            return fail(FDTD_ERR_INVALID_ARGUMENT, "index %lld out of range", index[q]);
        fdtd::Probe& p = list[q];
        p.i = (int)(index[q] % g->Ni);
        p.j = (int)((index[q] / g->Ni) % g->Nj);
        p.k = (int)(index[q] / (long long)plane);
        p.comp = comp[q];
        p.col = q;
        p.pad = 0;
    }
    // plane key: E at plane k under k, B at plane k under k+1 (mod Nk)
    auto key = [g](const fdtd::Probe& p) { return p.comp < 3 ? p.k : (p.k + 1) % g->Nk; };
    std::stable_sort(list.begin(), list.end(),
                     [&](const fdtd::Probe& a, const fdtd::Probe& b) { return key(a) < key(b); });
    std::vector<int> start(g->Nk + 1, 0);
    for (const auto& p : list) start[key(p) + 1] += 1;
    for (int k = 0; k < g->Nk; ++k) start[k + 1] += start[k];
    // queued work may read the old tables
    HIP_TRY(hipStreamSynchronize(g->stream));
    if (g->d_probe) HIP_TRY(hipFree(g->d_probe));
    if (g->d_pstart) HIP_TRY(hipFree(g->d_pstart));
    if (g->d_probe_out) HIP_TRY(hipFree(g->d_probe_out));
    g->d_probe = nullptr;
    g->d_pstart = nullptr;
    g->d_probe_out = nullptr;
    g->nprobe = 0;
    g->probe_cap = 0;
    g->probe_rows = 0;
    if (n == 0) return FDTD_OK;
    HIP_TRY(hipMalloc(&g->d_probe, sizeof(fdtd::Probe) * n));
    HIP_TRY(hipMalloc(&g->d_pstart, sizeof(int) * (g->Nk + 1)));
    HIP_TRY(hipMalloc(&g->d_probe_out, sizeof(double) * n * capacity));
    HIP_TRY(hipMemcpy(g->d_probe, list.data(), sizeof(fdtd::Probe) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g->d_pstart, start.data(), sizeof(int) * (g->Nk + 1), hipMemcpyHostToDevice));
    // rows not yet written read as NaN (all bits set)
    HIP_TRY(hipMemsetAsync(g->d_probe_out, 0xFF, sizeof(double) * n * capacity, g->stream));
    g->nprobe = n;
    g->probe_cap = capacity;
    return FDTD_OK;
}

int fdtd_probe_rows(const fdtd_grid* g, long long* rows) {
    if (!g || !rows) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    *rows = g->probe_rows;
    return FDTD_OK;
}

int fdtd_read_probes(fdtd_grid* g, long long row0, long long nrows, double* host) {
    if (!g || row0 < 0 || nrows < 0 || row0 + nrows > g->probe_rows || (nrows > 0 && !host))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid probe rows");
    if (nrows == 0) return FDTD_OK;
    HIP_TRY(hipMemcpyAsync(host, g->d_probe_out + (size_t)row0 * g->nprobe,
                           sizeof(double) * (size_t)nrows * g->nprobe, hipMemcpyDeviceToHost,
                           g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return FDTD_OK;
}

int fdtd_set_energy_interval(fdtd_grid* g, int every, long long capacity) {
    if (!g || every < 0 || (every > 0 && capacity <= 0) ||
        capacity > (long long)(SIZE_MAX / (6 * sizeof(double))))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    HIP_TRY(hipStreamSynchronize(g->stream));
    if (g->d_energy) HIP_TRY(hipFree(g->d_energy));
    g->d_energy = nullptr;
    g->energy_every = 0;
    g->energy_cap = 0;
    g->energy_steps.clear();
    if (every == 0) return FDTD_OK;
    HIP_TRY(hipMalloc(&g->d_energy, sizeof(double) * 6 * capacity));
    g->energy_every = every;
    g->energy_cap = capacity;
    return FDTD_OK;
}

int fdtd_energy_rows(const fdtd_grid* g, long long* rows) {
    if (!g || !rows) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    *rows = static_cast<long long>(g->energy_steps.size());
    return FDTD_OK;
}

int fdtd_read_energy(fdtd_grid* g, long long row0, long long nrows, long long* steps,
                     double* sums) {
    const long long have = static_cast<long long>(g ? g->energy_steps.size() : 0);
    if (!g || row0 < 0 || nrows < 0 || row0 + nrows > have || (nrows > 0 && !sums))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid energy rows");
    if (nrows == 0) return FDTD_OK;
    HIP_TRY(hipMemcpyAsync(sums, g->d_energy + (size_t)6 * row0, sizeof(double) * 6 * nrows,
                           hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    if (steps)
        for (long long r = 0; r < nrows; ++r) steps[r] = g->energy_steps[row0 + r];
    return FDTD_OK;
}

int fdtd_rewind_records(fdtd_grid* g) {
    if (!g) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    g->probe_rows = 0;
    g->energy_steps.clear();
    return FDTD_OK;
}

int fdtd_get_stats(const fdtd_grid* g, long long out[4]) {
    if (!g || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    for (int k = 0; k < 4; ++k) out[k] = g->stats[k];
    return FDTD_OK;
}

int fdtd_update_fields_3sweep(fdtd_grid* g, int nsteps) {
    if (!g || nsteps < 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (nsteps == 0) return FDTD_OK;
    if (int rc = flush_sources(g)) return rc;
    return run_steps(g, nsteps, true);
}

int fdtd_get_step(const fdtd_grid* g, long long* step) {
    if (!g || !step) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    *step = g->step;
    return FDTD_OK;
}

int fdtd_set_step(fdtd_grid* g, long long step) {
    if (!g || step < 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    g->step = step;
    return FDTD_OK;
}

int fdtd_synchronize(fdtd_grid* g) {
    if (!g) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipStreamSynchronize(g->stream));
    return FDTD_OK;
}

int fdtd_get_info(const fdtd_grid* g, int dims[3], int* precision, long long* cells) {
    if (!g) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (dims) { dims[0] = g->Ni; dims[1] = g->Nj; dims[2] = g->Nk; }
    if (precision) *precision = g->precision;
    if (cells) *cells = (long long)g->N;
    return FDTD_OK;
}

int fdtd_sum_squares(fdtd_grid* g, double out[6]) {
    if (!g || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    return reduce_fields(g, out, nullptr);
}

int fdtd_max_abs(fdtd_grid* g, double out[6]) {
    if (!g || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    return reduce_fields(g, nullptr, out);
}

int fdtd_sum_squares_noalloc(fdtd_grid* g, double out[6]) {
    if (!g || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    const int nb = std::min<int>(kReduceBlocks, grid_blocks(g->N));
    if ((size_t)nb * 12 + 12 > kNoAllocScratchDoubles)
        return fail(FDTD_ERR_INVALID_ARGUMENT,
                    "%d partial sums exceed the allocation-free scratch of %zu doubles", nb * 12,
                    kNoAllocScratchDoubles);
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != g->device) HIP_TRY(hipSetDevice(g->device));
    double* s = nullptr;
    int rc = scratch_on_device(g->device, &s);
    if (!rc) rc = reduce_fields_into(g, s, s + 12 * kReduceBlocks, out, nullptr);
    if (cur != g->device) (void)hipSetDevice(cur);
    return rc;
}

int fdtd_compare(fdtd_grid* t, fdtd_grid* r, double out[18]) {
    if (!t || !r || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (r->precision != FDTD_FP64) return fail(FDTD_ERR_INVALID_ARGUMENT, "reference grid must be FP64");
    if (t->Ni != r->Ni || t->Nj != r->Nj || t->Nk != r->Nk)
        return fail(FDTD_ERR_INVALID_ARGUMENT, "grid shapes differ");
    HIP_TRY(hipStreamSynchronize(r->stream));
    const int nb = std::min<int>(kReduceBlocks, grid_blocks(t->N));
    for (int c = 0; c < 6; ++c) {
        const double* b = static_cast<const double*>(field_ptr(r, c));
        const double inv = 1.0 / t->scales[c];
        const void* a = field_ptr(t, c);
        switch (t->precision) {
            case FDTD_FP64:
                hipLaunchKernelGGL(fdtd::diff_partial<double>, dim3(nb), dim3(256), 0, t->stream,
                                   static_cast<const double*>(a), inv, b, t->N, t->d_part);
                break;
            case FDTD_FP32:
                hipLaunchKernelGGL(fdtd::diff_partial<float>, dim3(nb), dim3(256), 0, t->stream,
                                   static_cast<const float*>(a), inv, b, t->N, t->d_part);
                break;
            default:
                hipLaunchKernelGGL(fdtd::diff_partial<_Float16>, dim3(nb), dim3(256), 0, t->stream,
                                   static_cast<const _Float16*>(a), inv, b, t->N, t->d_part);
        }
        hipLaunchKernelGGL(fdtd::reduce_partials, dim3(1), dim3(256), 0, t->stream, t->d_part, nb,
                           3, 4u, t->d_out);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out + 3 * c, t->d_out, 3 * sizeof(double), hipMemcpyDeviceToHost,
                               t->stream));
    }
    HIP_TRY(hipStreamSynchronize(t->stream));
    return FDTD_OK;
}

int fdtd_checksum(fdtd_grid* g, unsigned long long out[6]) {
    if (!g || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    const int nb = std::min<int>(kReduceBlocks, grid_blocks(g->N));
    auto* part = reinterpret_cast<unsigned long long*>(g->d_part);
    auto* res = reinterpret_cast<unsigned long long*>(g->d_out);
    for (int c = 0; c < 6; ++c) {
        const void* p = field_ptr(g, c);
        switch (g->precision) {
            case FDTD_FP64:
                hipLaunchKernelGGL(fdtd::checksum_partial<double>, dim3(nb), dim3(256), 0, g->stream,
                                   static_cast<const double*>(p), g->N, part);
                break;
            case FDTD_FP32:
                hipLaunchKernelGGL(fdtd::checksum_partial<float>, dim3(nb), dim3(256), 0, g->stream,
                                   static_cast<const float*>(p), g->N, part);
                break;
            default:
                hipLaunchKernelGGL(fdtd::checksum_partial<_Float16>, dim3(nb), dim3(256), 0,
                                   g->stream, static_cast<const _Float16*>(p), g->N, part);
        }
        hipLaunchKernelGGL(fdtd::checksum_final, dim3(1), dim3(256), 0, g->stream, part, nb, res + c);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(out, res, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return FDTD_OK;
}

int fdtd_count_nonfinite(fdtd_grid* g, long long out[6]) {
    if (!g || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    const int nb = std::min<int>(kReduceBlocks, grid_blocks(g->N));
    auto* part = reinterpret_cast<unsigned long long*>(g->d_part);
    auto* res = reinterpret_cast<unsigned long long*>(g->d_out);
    for (int c = 0; c < 6; ++c) {
        const void* p = field_ptr(g, c);
        switch (g->precision) {
            case FDTD_FP64:
                hipLaunchKernelGGL(fdtd::nonfinite_partial<double>, dim3(nb), dim3(256), 0,
                                   g->stream, static_cast<const double*>(p), g->N, part);
                break;
            case FDTD_FP32:
                hipLaunchKernelGGL(fdtd::nonfinite_partial<float>, dim3(nb), dim3(256), 0,
                                   g->stream, static_cast<const float*>(p), g->N, part);
                break;
            default:
                hipLaunchKernelGGL(fdtd::nonfinite_partial<_Float16>, dim3(nb), dim3(256), 0,
                                   g->stream, static_cast<const _Float16*>(p), g->N, part);
        }
        hipLaunchKernelGGL(fdtd::checksum_final, dim3(1), dim3(256), 0, g->stream, part, nb, res + c);
        HIP_TRY(hipGetLastError());
    }
    unsigned long long r[6];
    HIP_TRY(hipMemcpyAsync(r, res, sizeof r, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    for (int c = 0; c < 6; ++c) out[c] = static_cast<long long>(r[c]);
    return FDTD_OK;
}

int fdtd_add_plane_wave(fdtd_grid* g, const int m[3], double amplitude, double phase,
                        const double polarization[3]) {
    if (!g || !m || !polarization) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (!m[0] && !m[1] && !m[2]) return fail(FDTD_ERR_INVALID_ARGUMENT, "zero wave vector");
    const double kx = 2.0 * M_PI * m[0] / g->Ni, ky = 2.0 * M_PI * m[1] / g->Nj,
                 kz = 2.0 * M_PI * m[2] / g->Nk;
    const double kn = std::sqrt(kx * kx + ky * ky + kz * kz);
    const double kh[3] = {kx / kn, ky / kn, kz / kn};
    double e[3] = {polarization[0], polarization[1], polarization[2]};
    const double dot = e[0] * kh[0] + e[1] * kh[1] + e[2] * kh[2];
    for (int d = 0; d < 3; ++d) e[d] -= dot * kh[d];
    const double en = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    if (!(en > 1e-12)) return fail(FDTD_ERR_INVALID_ARGUMENT, "polarization parallel to k");
    for (int d = 0; d < 3; ++d) e[d] = amplitude * e[d] / en;
    g->ba_cur = -1;
    const double b[3] = {kh[1] * e[2] - kh[2] * e[1], kh[2] * e[0] - kh[0] * e[2],
                         kh[0] * e[1] - kh[1] * e[0]};
    const double* sc = g->scales;
    const dim3 grid(grid_blocks(g->N)), block(256);
    switch (g->precision) {
        case FDTD_FP64:
            hipLaunchKernelGGL(fdtd::add_plane_wave<double>, grid, block, 0, g->stream,
                               fields_of<double>(g, g->cur), g->Ni, g->Nj, g->Nk, kx, ky, kz, phase,
                               e[0], e[1], e[2], b[0], b[1], b[2], sc[0], sc[1], sc[2], sc[3], sc[4], sc[5]);
            break;
        case FDTD_FP32:
            hipLaunchKernelGGL(fdtd::add_plane_wave<float>, grid, block, 0, g->stream,
                               fields_of<float>(g, g->cur), g->Ni, g->Nj, g->Nk, kx, ky, kz, phase,
                               e[0], e[1], e[2], b[0], b[1], b[2], sc[0], sc[1], sc[2], sc[3], sc[4], sc[5]);
            break;
        default:
            hipLaunchKernelGGL(fdtd::add_plane_wave<_Float16>, grid, block, 0, g->stream,
                               fields_of<_Float16>(g, g->cur), g->Ni, g->Nj, g->Nk, kx, ky, kz,
                               phase, e[0], e[1], e[2], b[0], b[1], b[2], sc[0], sc[1], sc[2],
                               sc[3], sc[4], sc[5]);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(g->stream));
    return FDTD_OK;
}

int fdtd_set_scales(fdtd_grid* g, const double scales[6]) {
    if (!g || !scales) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    for (int c = 0; c < 6; ++c) {
        int e;
        const double m = std::frexp(scales[c], &e);
        if (!(scales[c] > 0) || m != 0.5)
            return fail(FDTD_ERR_INVALID_ARGUMENT, "scale %d (%g) is not a power of two", c, scales[c]);
    }
    if (g->precision != FDTD_FP16) {
        for (int c = 0; c < 6; ++c)
            if (scales[c] != 1.0)
                return fail(FDTD_ERR_UNSUPPORTED, "storage scales apply to FP16 grids only");
        return FDTD_OK;
    }
    if (int rc = apply_scales(g, scales, false)) return rc;
    HIP_TRY(hipStreamSynchronize(g->stream));
    return FDTD_OK;
}

int fdtd_get_scales(const fdtd_grid* g, double scales[6]) {
    if (!g || !scales) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    for (int c = 0; c < 6; ++c) scales[c] = g->scales[c];
    return FDTD_OK;
}

}  // extern "C"

namespace fdtd::rt {

// Power-of-two scales putting max|stored| in [2^(target-1), 2^target).
void target_scales(const fdtd_grid* g, const double mx[6], int target_log2, double floor_mag,
                   double s[6]) {
    for (int c = 0; c < 6; ++c) {
        const double m = std::max(mx[c], floor_mag);
        if (!(m > 0)) { s[c] = g->scales[c]; continue; }
        int e;
        std::frexp(m, &e);  // 2^(e-1) <= m < 2^e
        const int k = std::max(-100, std::min(100, target_log2 - e));
        s[c] = std::ldexp(1.0, k);
    }
}

// Rescales the FP16 state to the scales s (exact powers of two): E and B
// (both buffers, or only the current ones), the PML split fields with their
// totals' scales, and the PML tables' scale ratios.
int apply_scales(fdtd_grid* g, const double s[6], bool current_only) {
    auto scale = [&](void* p, double f) {
        hipLaunchKernelGGL(fdtd::scale_array<_Float16>, dim3(grid_blocks(g->N)), dim3(256), 0,
                           g->stream, static_cast<_Float16*>(p), g->N, (float)f);
    };
    bool changed = false;
    for (int c = 0; c < 6; ++c) {
        const double f = s[c] / g->scales[c];
        if (f == 1.0) continue;
        changed = true;
        g->ba_cur = -1;
        for (int b = 0; b < 2; ++b) {
            if (current_only && b != g->cur) continue;
            scale(c < 3 ? g->E[b][c] : g->B[b][c - 3], f);
        }
        if (g->pml)
            for (int q = 0; q < 12; ++q) {
                if (split_component(q) != c) continue;
                if (q < 6) {
                    for (int b = 0; b < 2; ++b)
                        if (!current_only || b == g->cur) scale(g->ES[b][q], f);
                } else {
                    scale(g->BS[q - 6], f);
                }
            }
        g->scales[c] = s[c];
    }
    HIP_TRY(hipGetLastError());
    if (changed && g->pml)
        if (int rc = build_pml_tables(g)) return rc;
    return FDTD_OK;
}

// Rescale the current FP16 state so every component's max sits just below
// 2^autoscale_log2 (components that are all zero keep their scale).
int rescale_to_target(fdtd_grid* g, bool current_only) {
    double mx[6], s[6];
    if (int rc = fdtd_max_abs(g, mx)) return rc;
    target_scales(g, mx, g->autoscale_log2, 0.0, s);
    return apply_scales(g, s, current_only);
}

// Sets a blocking pick read from the autotuner cache; false (and the
// grid's blocking unchanged) when it does not validate for this grid.
bool apply_blocking(fdtd_grid* g, const BlockTuning& t) {
    const BlockTuning keep = g->btune;
    int rc;
    if (t.L <= 0) {
        g->btune = t;  // the full-row sweep, as the timing left it
        g->btune.L = 0;
        return true;
    }
    if (t.R > 0 && t.TYg % t.R == 0)
        rc = fdtd_set_time_blocking_strips(g, t.L, t.V, t.R, t.TYg / t.R, t.kchunk);
    else if (t.R == 0)
        rc = fdtd_set_time_blocking(g, t.L, t.V, t.lanes, t.TYg, t.kchunk);
    else
        rc = FDTD_ERR_INVALID_ARGUMENT;
    if (rc != FDTD_OK) {
        g->btune = keep;
        g_last_error.clear();
        return false;
    }
    return true;
}

}  // namespace fdtd::rt

extern "C" {

int fdtd_autoscale(fdtd_grid* g, int target_log2, double floor_mag) {
    if (!g) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (g->precision != FDTD_FP16) return FDTD_OK;
    double mx[6];
    if (int rc = fdtd_max_abs(g, mx)) return rc;
    double s[6];
    target_scales(g, mx, target_log2, floor_mag, s);
    return fdtd_set_scales(g, s);
}

int fdtd_set_autoscale_interval(fdtd_grid* g, int interval, int target_log2) {
    if (!g || interval < 0 || target_log2 < 1 || target_log2 > 15)
        return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (g->precision != FDTD_FP16 && interval > 0)
        return fail(FDTD_ERR_UNSUPPORTED, "autoscaling applies to FP16 grids only");
    g->autoscale_interval = interval;
    g->autoscale_log2 = target_log2;
    return FDTD_OK;
}

int fdtd_zero_fields(fdtd_grid* g) {
    if (!g) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    g->ba_cur = -1;
    const size_t bytes = g->N * g->esize + 16;
    for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 3; ++d) {
            if (int rc = zero_fill(g->E[b][d], bytes, g->stream)) return rc;
            if (int rc = zero_fill(g->B[b][d], bytes, g->stream)) return rc;
        }
    if (g->pml)
        for (int q = 0; q < 6; ++q) {
            for (int b = 0; b < 2; ++b)
                if (int rc = zero_fill(g->ES[b][q], bytes, g->stream)) return rc;
            if (int rc = zero_fill(g->BS[q], bytes, g->stream)) return rc;
        }
    if (g->j_device) {
        for (int d = 0; d < 3; ++d)
            if (int rc = zero_fill(g->J[d], g->N * current_size(g->precision), g->stream)) return rc;
    } else {
        g->j_dense_active = false;
    }
    if (!g->sparse_j.empty()) g->src_dirty = true;
    g->sparse_j.clear();
    g->step = 0;
    HIP_TRY(hipStreamSynchronize(g->stream));
    return FDTD_OK;
}

int fdtd_set_step_carry(fdtd_grid* g, int enable) {
    if (!g) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    g->carry = enable != 0;
    if (!g->carry) {
        g->ba_cur = -1;
        HIP_TRY(hipStreamSynchronize(g->stream));
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 3; ++d)
                if (g->BA[b][d]) {
                    HIP_TRY(hipFree(g->BA[b][d]));
                    g->BA[b][d] = nullptr;
                }
    }
    return FDTD_OK;
}

int fdtd_set_tuning(fdtd_grid* g, int V, int TY, int kchunk) {
    if (!g) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    Tuning t = g->tune;
    if (V > 0) t.V = V;
    if (TY > 0) t.TY = TY;
    if (kchunk > 0) t.kchunk = kchunk;
    if (!shape_valid(g, t.V, t.TY, t.kchunk))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "launch shape V=%d TY=%d kchunk=%d invalid for Ni=%d",
                    t.V, t.TY, t.kchunk, g->Ni);
    g->tune = t;
    g->fused_ok = true;
    return FDTD_OK;
}

int fdtd_set_time_blocking(fdtd_grid* g, int L, int V, int lanes, int TYg, int kchunk) {
    if (!g) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (L <= 0) {
        g->btune.L = 0;
        return FDTD_OK;
    }
    if (g->pml) return fail(FDTD_ERR_UNSUPPORTED, "temporal blocking takes periodic grids (no PML layer)");
    BlockTuning t;
    t.L = L;
    t.V = V > 0 ? V : ((g->Ni % 2 == 0) ? 2 : 1);
    t.lanes = lanes > 0 ? lanes : 64;
    t.TYg = TYg > 0 ? TYg
                    : fdtd::blocked_max_threads(t.V, L, (int)compute_size(g->precision)) / t.lanes;
    t.kchunk = 1;
    if (!blocked_valid(g, t))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "blocking L=%d V=%d lanes=%d TYg=%d invalid for this grid",
                    t.L, t.V, t.lanes, t.TYg);
    if (kchunk > 0) {
        t.kchunk = kchunk;
    } else {
        // at least two tiles per CU when the grid allows, chunks of >= 16 planes
        const int TX = t.lanes * t.V - 2 * L, TY = t.TYg - 2 * L;
        const long long nxy = (long long)((g->Ni + TX - 1) / TX) * ((g->Nj + TY - 1) / TY);
        int nchunk = (int)std::max<long long>(1, (512 + nxy - 1) / nxy);
        nchunk = std::min(nchunk, std::max(1, g->Nk / 16));
        t.kchunk = (g->Nk + nchunk - 1) / nchunk;
    }
    if (!blocked_valid(g, t))
        return fail(FDTD_ERR_INVALID_ARGUMENT,
                    "blocking L=%d V=%d lanes=%d TYg=%d kchunk=%d invalid for this grid", t.L, t.V,
                    t.lanes, t.TYg, t.kchunk);
    g->btune = t;
    return FDTD_OK;
}

int fdtd_set_time_blocking_strips(fdtd_grid* g, int L, int V, int R, int W, int kchunk) {
    if (!g) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (L <= 0) {
        g->btune.L = 0;
        return FDTD_OK;
    }
    if (g->pml) return fail(FDTD_ERR_UNSUPPORTED, "temporal blocking takes periodic grids (no PML layer)");
    BlockTuning t;
    t.L = L;
    t.V = V > 0 ? V : 1;
    t.R = R > 0 ? R : 4;
    t.lanes = 64;
    const int cs = (int)compute_size(g->precision);
    const int wmax = std::min(16, fdtd::strip_max_threads(t.V, t.R, L, cs) / 64);
    t.TYg = (W > 0 ? W : std::min(8, wmax)) * t.R;
    t.kchunk = 1;
    if (!strip_valid(g, t))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "strip blocking L=%d V=%d R=%d W=%d invalid for this grid",
                    t.L, t.V, t.R, t.TYg / t.R);
    t.kchunk = kchunk > 0 ? kchunk : strip_default_kchunk(g, t);
    g->btune = t;
    return FDTD_OK;
}

int fdtd_set_pml_path(fdtd_grid* g, int path, int V, int TY, int kchunk) {
    if (!g || (path != 0 && path != 1)) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (!g->pml) return fail(FDTD_ERR_UNSUPPORTED, "the grid has no PML layer");
    if (path == 0) {
        g->pml_path = 0;
        return FDTD_OK;
    }
    Tuning t = pml_phase_default(g);
    if (V > 0) {
        t.V = V;
        if (TY <= 0)
            t.TY = std::max(1, std::min(g->Nj, fdtd::pml_phase_max_threads(V, (int)compute_size(g->precision)) /
                                                   std::max(1, g->Ni / V)));
    }
    if (TY > 0) t.TY = TY;
    if (kchunk > 0) t.kchunk = kchunk;
    if (!pml_phase_valid(g, t))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "PML phase shape V=%d TY=%d kchunk=%d invalid for this grid",
                    t.V, t.TY, t.kchunk);
    g->ptune = t;
    g->pml_path = 1;
    return FDTD_OK;
}

int fdtd_get_pml_path(const fdtd_grid* g, int out[4]) {
    if (!g || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    out[0] = g->pml ? g->pml_path : 0;
    out[1] = g->ptune.V;
    out[2] = g->ptune.TY;
    out[3] = g->ptune.kchunk;
    return FDTD_OK;
}

int fdtd_set_tune_cache(const char* path) {
    std::lock_guard<std::mutex> lock(fdtd::rt::g_tune_mu);
    fdtd::rt::g_tune_path = path ? path : "";
    fdtd::rt::g_tune_path_set = true;
    return FDTD_OK;
}

int fdtd_autotune_blocking(fdtd_grid* g, int reps, int out[5], float* best_ms) {
    if (!g || reps <= 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (g->pml) return fail(FDTD_ERR_UNSUPPORTED, "temporal blocking takes periodic grids (no PML layer)");
    const std::string cache = tune_cache_path();
    const std::string key = cache.empty() ? std::string() : tune_key(g);
    {
        BlockTuning t;
        float ms = 0;
        if (!cache.empty() && tune_cache_lookup(cache, key, &t, &ms) && apply_blocking(g, t)) {
            if (out) fdtd_get_time_blocking(g, out);
            if (best_ms) *best_ms = ms;
            return FDTD_OK;
        }
    }
    if (int rc = flush_sources(g)) return rc;
    // Candidate passes are timed on the grid's buffers and the state is
    // restored afterwards (on every exit path); the full-row kernel's time
    // per step is the reference.
    StateSnapshot snap(g);
    if (int rc = snap.save()) return rc;
    if (int rc = eval_sources(g, g->step, 4)) return rc;
    auto time_passes = [&](bool blocked, float* ms_step) -> int {
        const int per = blocked ? g->btune.L : 1;
        int rc = blocked ? launch_blocked(g, false, 0, g->cur, 1 - g->cur)
                         : launch_fused(g, g->tune, true, 0, g->cur, 1 - g->cur);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(g->ev[0], g->stream));
        for (int r = 0; r < reps; ++r) {
            const int in = (r & 1) ? 1 - g->cur : g->cur;
            rc = blocked ? launch_blocked(g, false, 0, in, 1 - in)
                         : launch_fused(g, g->tune, true, 0, in, 1 - in);
            if (rc) return rc;
        }
        HIP_TRY(hipEventRecord(g->ev[1], g->stream));
        HIP_TRY(hipEventSynchronize(g->ev[1]));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
        *ms_step = ms / (reps * per);
        return FDTD_OK;
    };
    const BlockTuning orig = g->btune;
    BlockTuning best;  // L = 0: full-row kernel
    float best_t = 1e30f;
    if (g->fused_ok) {
        float t;
        if (int rc = time_passes(false, &t)) return rc;
        best_t = t;
    }
    auto consider = [&]() -> int {
        float t;
        if (int rc = time_passes(true, &t)) {
            g->btune = orig;
            return rc;
        }
        if (t < best_t) {
            best_t = t;
            best = g->btune;
        }
        return FDTD_OK;
    };
    const int shapes[][3] = {{1, 32, 32}, {1, 64, 16}, {2, 64, 16}, {2, 32, 16}, {1, 64, 8},
                             {2, 64, 8}, {1, 32, 16}, {2, 32, 32}};
    for (int L = 1; L <= 4; ++L)
        for (const auto& sh : shapes) {
            if (fdtd_set_time_blocking(g, L, sh[0], sh[1], sh[2], 0) != FDTD_OK) continue;
            if (int rc = consider()) return rc;
        }
    const int strips[][3] = {{2, 2, 8}, {1, 2, 16}, {1, 2, 8}, {1, 4, 8}, {2, 2, 4}, {1, 4, 4}};
    for (int L = 1; L <= 4; ++L)
        for (const auto& sh : strips) {
            if (fdtd_set_time_blocking_strips(g, L, sh[0], sh[1], sh[2], 0) != FDTD_OK) continue;
            if (int rc = consider()) return rc;
        }
    g_last_error.clear();
    g->btune = best_t < 1e30f ? best : orig;
    if (int rc = snap.restore()) return rc;
    if (best_t >= 1e30f) return fail(FDTD_ERR_UNSUPPORTED, "no launch shape fits this grid");
    if (!cache.empty()) tune_cache_store(cache, key, g->btune, best_t);
    if (out) fdtd_get_time_blocking(g, out);
    if (best_ms) *best_ms = best_t;
    return FDTD_OK;
}

int fdtd_get_time_blocking(const fdtd_grid* g, int out[5]) {
    if (!g || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    out[0] = g->btune.L;
    out[1] = g->btune.V;
    out[2] = g->btune.lanes;
    out[3] = g->btune.TYg;
    out[4] = g->btune.kchunk;
    return FDTD_OK;
}

int fdtd_get_time_blocking_strips(const fdtd_grid* g, int out[5]) {
    if (!g || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    const bool strips = g->btune.L > 0 && g->btune.R > 0;
    out[0] = g->btune.L;
    out[1] = g->btune.V;
    out[2] = strips ? g->btune.R : 0;
    out[3] = strips ? g->btune.TYg / g->btune.R : 0;
    out[4] = g->btune.kchunk;
    return FDTD_OK;
}

int fdtd_get_tuning(const fdtd_grid* g, int out[3]) {
    if (!g || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    out[0] = g->tune.V;
    out[1] = g->tune.TY;
    out[2] = g->tune.kchunk;
    return g->fused_ok ? FDTD_OK : fail(FDTD_ERR_UNSUPPORTED, "no fused launch shape fits this grid");
}

int fdtd_time_update_fields(fdtd_grid* g, int nsteps, int path, float* ms_total, float* ms_main) {
    if (!g || nsteps <= 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (int rc = flush_sources(g)) return rc;
    HIP_TRY(hipEventRecord(g->ev[0], g->stream));
    int rc;
    if (path == FDTD_PATH_FUSED && g->fused_ok) {
        rc = run_fused(g, nsteps, g->ev[1]);
    } else {
        HIP_TRY(hipEventRecord(g->ev[1], g->stream));
        rc = run_3sweep(g, nsteps);
    }
    if (rc) return rc;
    HIP_TRY(hipEventRecord(g->ev[2], g->stream));
    HIP_TRY(hipEventSynchronize(g->ev[2]));
    float a = 0, b = 0;
    HIP_TRY(hipEventElapsedTime(&a, g->ev[0], g->ev[2]));
    HIP_TRY(hipEventElapsedTime(&b, g->ev[1], g->ev[2]));
    if (ms_total) *ms_total = a;
    if (ms_main) *ms_main = b;
    return FDTD_OK;
}

int fdtd_time_update_fields_kernels(fdtd_grid* g, int nsteps, double ms[4], int count[4]) {
    if (!g || nsteps <= 0 || !ms || !count) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    LaunchTrace trace;
    g->trace = &trace;
    int rc = fdtd_update_fields(g, nsteps);
    g->trace = nullptr;
    HIP_TRY(hipStreamSynchronize(g->stream));
    if (rc) return rc;
    for (int k = 0; k < 4; ++k) {
        ms[k] = 0.0;
        count[k] = 0;
    }
    for (const auto& r : trace.recs) {
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, r.a, r.b));
        ms[r.kind] += t;
        count[r.kind] += 1;
    }
    return FDTD_OK;
}

int fdtd_autotune(fdtd_grid* g, int reps, int out[3], float* best_ms) {
    if (!g || reps <= 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (int rc = flush_sources(g)) return rc;
    // Time sweeps between the two buffers of the grid; the fields (and PML
    // splits) are saved to and restored from scratch copies so tuning leaves
    // the state intact, also when a launch fails.
    StateSnapshot snap(g);
    if (int rc = snap.save()) return rc;
    int vc[4];
    const int nv = v_candidates(g->precision, vc);
    Tuning best = g->tune;
    float best_t = 1e30f;
    for (int q = 0; q < nv; ++q) {
        const int V = vc[q];
        if (g->Ni % V) continue;
        const int lanes = g->Ni / V;
        for (int TY = 1; TY <= std::min(g->Nj, 16); ++TY) {
            if ((TY + 1) * lanes > max_threads(g, V)) break;
            if (sweep_lds_bytes(g, TY) > kLdsLimit) break;
            const int kdef = default_kchunk(g->Nj, g->Nk, TY);
            const int kcs[3] = {kdef, std::max(1, kdef / 2), std::min(g->Nk, kdef * 2)};
            for (int kc : kcs) {
                Tuning t{V, TY, kc};
                if (!shape_valid(g, V, TY, kc)) continue;
                // warm-up then timed reps
                if (int rc = eval_sources(g, g->step, 1)) return rc;
                if (int rc = launch_fused(g, t, true, 0, g->cur, 1 - g->cur)) return rc;
                HIP_TRY(hipEventRecord(g->ev[0], g->stream));
                for (int r = 0; r < reps; ++r)
                    if (int rc = launch_fused(g, t, true, 0, (r & 1) ? 1 - g->cur : g->cur,
                                              (r & 1) ? g->cur : 1 - g->cur))
                        return rc;
                HIP_TRY(hipEventRecord(g->ev[1], g->stream));
                HIP_TRY(hipEventSynchronize(g->ev[1]));
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
                ms /= reps;
                if (ms < best_t) {
                    best_t = ms;
                    best = t;
                }
            }
        }
    }
    if (int rc = snap.restore()) return rc;
    if (best_t >= 1e30f) return fail(FDTD_ERR_UNSUPPORTED, "no fused launch shape fits this grid");
    g->tune = best;
    g->fused_ok = true;
    if (out) { out[0] = best.V; out[1] = best.TY; out[2] = best.kchunk; }
    if (best_ms) *best_ms = best_t;
    return FDTD_OK;
}

int fdtd_calibration_copy(size_t bytes, int iters, float* ms) {
    if (bytes == 0 || bytes % 16 || iters <= 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    // Everything created here is released on every return path.
    struct Resources {
        void* a = nullptr;
        void* b = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Resources() {
            if (s) (void)hipStreamSynchronize(s);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (s) (void)hipStreamDestroy(s);
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
        }
    } r;
    HIP_TRY(hipMalloc(&r.a, bytes));
    HIP_TRY(hipMalloc(&r.b, bytes));
    HIP_TRY(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&r.e0));
    HIP_TRY(hipEventCreate(&r.e1));
    const size_t n16 = bytes / 16;
    hipLaunchKernelGGL(fdtd::zero_bytes, dim3(grid_blocks(n16)), dim3(256), 0, r.s,
                       static_cast<uint4*>(r.a), n16);
    const int nb = 256 * 8 * 4;
    hipLaunchKernelGGL(fdtd::copy_calibration, dim3(nb), dim3(256), 0, r.s,
                       static_cast<const uint4*>(r.a), static_cast<uint4*>(r.b), n16);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(r.e0, r.s));
    for (int it = 0; it < iters; ++it)
        hipLaunchKernelGGL(fdtd::copy_calibration, dim3(nb), dim3(256), 0, r.s,
                           static_cast<const uint4*>(r.a), static_cast<uint4*>(r.b), n16);
    HIP_TRY(hipEventRecord(r.e1, r.s));
    HIP_TRY(hipEventSynchronize(r.e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, r.e0, r.e1));
    if (ms) *ms = t / iters;
    return FDTD_OK;
}

// Batched API: implemented in fdtd_batch.hip.

}  // extern "C"
