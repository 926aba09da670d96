// Batched solves: many independent Yee grids of uneven sizes advanced by one
// persistent launch per step (batch_fused_sweep), optionally as a captured
// HIP graph replayed with a device-side step counter.
//
// Semantics per grid are those of fdtd_update_fields (the reference's
// update_fields(), src/FDTD/FDTD.cpp:153-157, n times): one B half step,
// then n fused sweeps.  All grids of a batch share one precision and one
// step count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>
#include <queue>
#include <vector>

#include "fdtd_internal.hpp"

using namespace fdtd::rt;

namespace fdtd {
// The batch's device-side step counter (graph replays advance it).
static __global__ void advance_step(long long* step, long long by) { *step += by; }
}  // namespace fdtd

namespace {

constexpr int kBatchLdsCap = 80 * 1024;  // per workgroup: two per CU

struct BatchShape {
    int V = 4;                  // cells per lane along x
    int NT = 512;               // threads per workgroup
    long long target = 65536;   // cells per tile (sets each grid's z-chunk)
};

struct GridPlan {
    int TY = 1, kchunk = 1, ntile_y = 1, nchunk = 1;
};

}  // namespace

struct fdtd_batch {
    std::vector<fdtd_grid*> grids;
    int precision = FDTD_FP32;
    hipStream_t stream = nullptr;
    BatchShape shape;
    std::vector<GridPlan> plan;
    size_t lds_bytes = 0;
    int nwork = 0;
    int nblocks = 0;
    void* d_grids = nullptr;
    size_t grid_desc_bytes = 0;
    std::vector<unsigned char> h_grids;
    int4* d_work = nullptr;
    unsigned int* d_counters = nullptr;
    fdtd::PointSource* d_src = nullptr;
    double* d_vals = nullptr;
    int total_src = 0;
    int src_cap = 0;
    long long* d_step = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    int graph_swaps = 0;
    // What a captured graph bakes in besides its step count: source count
    // (value-table stride and eval_sources size), launch sizes and every
    // device buffer its launches read.  A replay re-captures when any of
    // them differs from the batch's current values.
    std::vector<long long> graph_key;
    hipEvent_t ev[4] = {};
    // temporal blocking: strip passes of sL steps (0: single sweeps only)
    int sL = 0, sV = 1, sR = 2, sW = 8;
    std::vector<fdtd::StripGeom> sgeom;
    int2* d_swork = nullptr;
    int snwork = 0;
    int snblocks = 0;
    size_t slds = 0;
    void* d_sgrids = nullptr;
    size_t sgrid_desc_bytes = 0;
    std::vector<unsigned char> h_sgrids;
    int* d_kstart = nullptr;
    size_t kstart_cap = 0;
};

namespace {

void drop_graph(fdtd_batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_steps = 0;
    b->graph_key.clear();
}

// The values a graph captured for nsteps steps depends on (see graph_key).
std::vector<long long> graph_key_of(const fdtd_batch* b, int nsteps) {
    auto addr = [](const void* p) { return static_cast<long long>(reinterpret_cast<uintptr_t>(p)); };
    return {nsteps, b->total_src, b->sL, b->sV, b->sR, b->sW, b->snwork, b->snblocks,
            b->nwork, b->nblocks, b->shape.V, b->shape.NT,
            addr(b->d_src), addr(b->d_vals), addr(b->d_kstart), addr(b->d_grids),
            addr(b->d_sgrids), addr(b->d_work), addr(b->d_swork), addr(b->d_counters),
            addr(b->d_step)};
}

int batch_v_candidates(int precision, int* out) {
    if (precision == FDTD_FP64) { out[0] = 2; out[1] = 1; return 2; }
    if (precision == FDTD_FP32) { out[0] = 4; out[1] = 2; out[2] = 1; return 3; }
    out[0] = 4; out[1] = 8; out[2] = 2; out[3] = 1;
    return 4;
}

// Per-grid tile height and z-chunk for a batch launch shape; false if a
// grid cannot be tiled with it.
bool make_plan(const fdtd_batch* b, const BatchShape& sh, std::vector<GridPlan>* plan,
               size_t* lds) {
    const int cs = (int)compute_size(b->precision);
    if (sh.NT > fdtd::fused_max_threads(sh.V, cs)) return false;
    plan->assign(b->grids.size(), GridPlan());
    *lds = 0;
    for (size_t q = 0; q < b->grids.size(); ++q) {
        const fdtd_grid* g = b->grids[q];
        if (g->Ni % sh.V) return false;
        const int lanes = g->Ni / sh.V;
        int ty = std::min(g->Nj, sh.NT / lanes - 1);
        while (ty >= 1 && fdtd::fused_lds_bytes(ty, g->Ni, cs) > (size_t)kBatchLdsCap) --ty;
        if (ty < 1) return false;
        GridPlan& p = (*plan)[q];
        p.TY = ty;
        p.ntile_y = (g->Nj + ty - 1) / ty;
        long long k = sh.target / ((long long)ty * g->Ni);
        k = std::max<long long>(4, std::min<long long>(k, g->Nk));
        p.nchunk = (int)((g->Nk + k - 1) / k);
        p.kchunk = (g->Nk + p.nchunk - 1) / p.nchunk;
        *lds = std::max(*lds, fdtd::fused_lds_bytes(ty, g->Ni, cs));
    }
    return true;
}

// Work list: every tile of every grid, most expensive first (cost ~ cells
// of the tile including its halo row and plane).
std::vector<int4> make_work(const fdtd_batch* b, const std::vector<GridPlan>& plan) {
    std::vector<int4> w;
    for (size_t q = 0; q < b->grids.size(); ++q) {
        const fdtd_grid* g = b->grids[q];
        const GridPlan& p = plan[q];
        for (int c = 0; c < p.nchunk; ++c)
            for (int t = 0; t < p.ntile_y; ++t) {
                const int ty = std::min(p.TY, g->Nj - t * p.TY);
                const int kc = std::min(p.kchunk, g->Nk - c * p.kchunk);
                w.push_back(make_int4((int)q, t, c, (ty + 1) * g->Ni * (kc + 1)));
            }
    }
    std::stable_sort(w.begin(), w.end(), [](const int4& a, const int4& c) { return a.w > c.w; });
    return w;
}

template <typename S>
void fill_descriptors(fdtd_batch* b, const std::vector<int>& src_off) {
    using C = typename fdtd::ComputeOf<S>::type;
    b->h_grids.resize(sizeof(fdtd::BatchGrid<S, C>) * b->grids.size());
    auto* d = reinterpret_cast<fdtd::BatchGrid<S, C>*>(b->h_grids.data());
    for (size_t q = 0; q < b->grids.size(); ++q) {
        fdtd_grid* g = b->grids[q];
        const GridPlan& p = b->plan[q];
        fdtd::BatchGrid<S, C> bg;
        bg.buf[0] = fields_of<S>(g, g->cur);
        bg.buf[1] = fields_of<S>(g, 1 - g->cur);
        bg.src = b->d_src ? b->d_src + src_off[q] : nullptr;
        bg.src_off = src_off[q];
        bg.nsrc = src_off[q + 1] - src_off[q];
        bg.c = make_coefs<C>(g);
        bg.g.Ni = g->Ni;
        bg.g.Nj = g->Nj;
        bg.g.Nk = g->Nk;
        bg.g.TY = p.TY;
        bg.g.kchunk = p.kchunk;
        bg.g.ntile_y = p.ntile_y;
        bg.g.ntiles = p.ntile_y * p.nchunk;
        bg.g.per_xcd = 0;
        d[q] = bg;
    }
}

template <typename S>
void fill_strip_descriptors(fdtd_batch* b, const std::vector<int>& src_off,
                            const std::vector<size_t>& kst_off) {
    using C = typename fdtd::ComputeOf<S>::type;
    b->h_sgrids.resize(sizeof(fdtd::BatchStripGrid<S, C>) * b->grids.size());
    auto* d = reinterpret_cast<fdtd::BatchStripGrid<S, C>*>(b->h_sgrids.data());
    for (size_t q = 0; q < b->grids.size(); ++q) {
        fdtd_grid* g = b->grids[q];
        fdtd::BatchStripGrid<S, C> bg;
        bg.buf[0] = fields_of<S>(g, g->cur);
        bg.buf[1] = fields_of<S>(g, 1 - g->cur);
        bg.src = b->d_src ? b->d_src + src_off[q] : nullptr;
        bg.kstart = b->d_kstart + kst_off[q];
        bg.src_off = src_off[q];
        bg.nsrc = src_off[q + 1] - src_off[q];
        bg.c = make_coefs<C>(g);
        bg.g = b->sgeom[q];
        d[q] = bg;
    }
}

template <typename S, int V, bool STEADY>
int occupancy_v(const fdtd_batch* b, int nt, size_t lds, int* blocks) {
    using C = typename fdtd::ComputeOf<S>::type;
    auto kern = fdtd::batch_fused_sweep<S, C, V, STEADY>;
    auto last = fdtd::batch_fused_sweep<S, C, V, !STEADY>;
    if (int rc = allow_dynamic_lds(reinterpret_cast<const void*>(kern), lds)) return rc;
    if (int rc = allow_dynamic_lds(reinterpret_cast<const void*>(last), lds)) return rc;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, kern, nt, lds));
    return FDTD_OK;
}

template <typename S, int V, bool STEADY>
int launch_batch_v(fdtd_batch* b, int parity, const double* vals_row) {
    using C = typename fdtd::ComputeOf<S>::type;
    fdtd::BatchArgs<S, C> a;
    a.grids = static_cast<const fdtd::BatchGrid<S, C>*>(b->d_grids);
    a.work = b->d_work;
    a.nwork = b->nwork;
    a.parity = parity;
    a.vals_row = vals_row;
    a.counters = b->d_counters;
    hipLaunchKernelGGL((fdtd::batch_fused_sweep<S, C, V, STEADY>), dim3(b->nblocks),
                       dim3(b->shape.NT), b->lds_bytes, b->stream, a);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

template <typename S, bool STEADY>
int launch_batch_s(fdtd_batch* b, int parity, const double* vals_row) {
    switch (b->shape.V) {
        case 1: return launch_batch_v<S, 1, STEADY>(b, parity, vals_row);
        case 2: return launch_batch_v<S, 2, STEADY>(b, parity, vals_row);
        case 4: return launch_batch_v<S, 4, STEADY>(b, parity, vals_row);
        case 8:
            if constexpr (sizeof(S) <= 4) return launch_batch_v<S, 8, STEADY>(b, parity, vals_row);
            [[fallthrough]];
        default: return fail(FDTD_ERR_UNSUPPORTED, "unsupported V=%d", b->shape.V);
    }
}

int launch_batch(fdtd_batch* b, bool steady, int parity, const double* vals_row) {
    switch (b->precision) {
        case FDTD_FP64:
            return steady ? launch_batch_s<double, true>(b, parity, vals_row)
                          : launch_batch_s<double, false>(b, parity, vals_row);
        case FDTD_FP32:
            return steady ? launch_batch_s<float, true>(b, parity, vals_row)
                          : launch_batch_s<float, false>(b, parity, vals_row);
        default:
            return steady ? launch_batch_s<_Float16, true>(b, parity, vals_row)
                          : launch_batch_s<_Float16, false>(b, parity, vals_row);
    }
}

// ---- strip passes --------------------------------------------------------

constexpr bool batch_strip_built(int csize, int V, int R) {
    return (V == 1 || V == 2) && (R == 2 || R == 4) && V * R * csize <= 16;
}

template <typename S, int V, int R, int L, bool FINAL>
const void* batch_strip_kernel() {
    using C = typename fdtd::ComputeOf<S>::type;
    if constexpr (!batch_strip_built(sizeof(C), V, R)) {
        return nullptr;
    } else {
        return reinterpret_cast<const void*>(fdtd::batch_strip_sweep<S, C, V, R, L, FINAL>);
    }
}

template <typename S, int V, int R, int L, bool FINAL>
int launch_batch_strip_v(fdtd_batch* b, int parity, const double* vals) {
    using C = typename fdtd::ComputeOf<S>::type;
    if constexpr (!batch_strip_built(sizeof(C), V, R)) {
        return fail(FDTD_ERR_UNSUPPORTED, "strip variant V=%d R=%d not built", V, R);
    } else {
        fdtd::BatchStripArgs<S, C> a;
        a.grids = static_cast<const fdtd::BatchStripGrid<S, C>*>(b->d_sgrids);
        a.work = b->d_swork;
        a.nwork = b->snwork;
        a.parity = parity;
        a.vals = vals;
        a.vstride = b->total_src;
        a.pad = 0;
        a.counters = b->d_counters;
        hipLaunchKernelGGL((fdtd::batch_strip_sweep<S, C, V, R, L, FINAL>), dim3(b->snblocks),
                           dim3(64 * b->sW), b->slds, b->stream, a);
        HIP_TRY(hipGetLastError());
        return FDTD_OK;
    }
}

template <typename S, int V, int R, bool FINAL>
int launch_batch_strip_l(fdtd_batch* b, int parity, const double* vals) {
    switch (b->sL) {
        case 1: return launch_batch_strip_v<S, V, R, 1, FINAL>(b, parity, vals);
        case 2: return launch_batch_strip_v<S, V, R, 2, FINAL>(b, parity, vals);
        case 3: return launch_batch_strip_v<S, V, R, 3, FINAL>(b, parity, vals);
        default: return launch_batch_strip_v<S, V, R, 4, FINAL>(b, parity, vals);
    }
}

template <typename S, bool FINAL>
int launch_batch_strip_s(fdtd_batch* b, int parity, const double* vals) {
    if (b->sV == 1 && b->sR == 2) return launch_batch_strip_l<S, 1, 2, FINAL>(b, parity, vals);
    if (b->sV == 1 && b->sR == 4) return launch_batch_strip_l<S, 1, 4, FINAL>(b, parity, vals);
    return launch_batch_strip_l<S, 2, 2, FINAL>(b, parity, vals);
}

// One strip pass of sL steps over every grid; source values are the rows
// starting at vals (null: no sources).
int launch_batch_strip(fdtd_batch* b, bool final, int parity, const double* vals) {
    switch (b->precision) {
        case FDTD_FP64:
            return final ? launch_batch_strip_s<double, true>(b, parity, vals)
                         : launch_batch_strip_s<double, false>(b, parity, vals);
        case FDTD_FP32:
            return final ? launch_batch_strip_s<float, true>(b, parity, vals)
                         : launch_batch_strip_s<float, false>(b, parity, vals);
        default:
            return final ? launch_batch_strip_s<_Float16, true>(b, parity, vals)
                         : launch_batch_strip_s<_Float16, false>(b, parity, vals);
    }
}

template <typename S, int V, int R>
const void* strip_kernel_l(int L, bool final) {
    switch (L) {
        case 1: return final ? batch_strip_kernel<S, V, R, 1, true>() : batch_strip_kernel<S, V, R, 1, false>();
        case 2: return final ? batch_strip_kernel<S, V, R, 2, true>() : batch_strip_kernel<S, V, R, 2, false>();
        case 3: return final ? batch_strip_kernel<S, V, R, 3, true>() : batch_strip_kernel<S, V, R, 3, false>();
        default: return final ? batch_strip_kernel<S, V, R, 4, true>() : batch_strip_kernel<S, V, R, 4, false>();
    }
}

template <typename S>
const void* strip_kernel_s(int V, int R, int L, bool final) {
    if (V == 1 && R == 2) return strip_kernel_l<S, 1, 2>(L, final);
    if (V == 1 && R == 4) return strip_kernel_l<S, 1, 4>(L, final);
    if (V == 2 && R == 2) return strip_kernel_l<S, 2, 2>(L, final);
    return nullptr;
}

const void* strip_kernel(int precision, int V, int R, int L, bool final) {
    switch (precision) {
        case FDTD_FP64: return strip_kernel_s<double>(V, R, L, final);
        case FDTD_FP32: return strip_kernel_s<float>(V, R, L, final);
        default: return strip_kernel_s<_Float16>(V, R, L, final);
    }
}

// Strip tiling of every grid (L, V, R rows per lane, W wavefronts), its
// work list (most expensive tiles first) and the persistent grid size.
// Every z chunk of a tile repeats the pipeline's 2(L-1) fill planes, while
// fewer, longer tiles fill the device's resident workgroups in coarser
// rounds.  For each candidate chunk length (whole grids down to about
// 128 Ki computed cells per tile) the queue's schedule is simulated, costliest
// tile first onto the workgroup that frees first; a shorter chunk length is
// taken only when it shortens the makespan by more than 2%, since each tile
// also has costs outside the model (descriptor loads, the queue, ramp-up).
int apply_strips(fdtd_batch* b, int L, int V, int R, int W) {
    const int cs = (int)compute_size(b->precision);
    if (L < 1 || L > 4 || !batch_strip_built(cs, V, R) || W < 1 || W > 16 ||
        64 * W > fdtd::strip_max_threads(V, R, L, cs))
        return fail(FDTD_ERR_INVALID_ARGUMENT, "strip blocking L=%d V=%d R=%d W=%d not available",
                    L, V, R, W);
    int Gx = 0, TX = 0;
    strip_x_geometry(L, V, storage_size(b->precision), &Gx, &TX, false);
    const int TY = W * R - 2 * L;
    if (TX < 1 || TY < 1)
        return fail(FDTD_ERR_INVALID_ARGUMENT, "strip tiles of W=%d R=%d leave no stored rows", W, R);
    for (size_t q = 0; q < b->grids.size(); ++q) {
        const fdtd_grid* g = b->grids[q];
        if (g->Ni % V || (unsigned long long)g->Ni * g->Nj * g->esize >= (1ull << 31))
            return fail(FDTD_ERR_INVALID_ARGUMENT, "grid %zu cannot take strips with V=%d", q, V);
    }
    const size_t lds = fdtd::strip_lds_bytes(L, W, V, cs);
    if (lds > kLdsLimit) return fail(FDTD_ERR_INVALID_ARGUMENT, "strip tiles need too much LDS");
    int per_cu = 0;
    for (bool final : {false, true}) {
        const void* kern = strip_kernel(b->precision, V, R, L, final);
        if (!kern) return fail(FDTD_ERR_UNSUPPORTED, "strip variant not built");
        if (int rc = allow_dynamic_lds(kern, lds)) return rc;
        if (!final) HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * W, lds));
    }
    if (per_cu < 1) return fail(FDTD_ERR_UNSUPPORTED, "strip tiles do not fit a CU");
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t slots = (size_t)per_cu * cus;
    const long long area = 64LL * V * W * R;
    const int kmin = (int)std::max<long long>(16, (128LL * 1024) / area);
    std::vector<fdtd::StripGeom> geom(b->grids.size()), try_geom(b->grids.size());
    std::vector<int2> work, try_work;
    std::vector<long long> cost, try_cost;
    long long best_span = -1;
    for (int kmax : {1 << 30, 512, 256, 192, 128, 96, kmin}) {
        if (kmax < kmin) continue;
        try_work.clear();
        try_cost.clear();
        for (size_t q = 0; q < b->grids.size(); ++q) {
            const fdtd_grid* g = b->grids[q];
            fdtd::StripGeom& G = try_geom[q];
            G.Ni = g->Ni;
            G.Nj = g->Nj;
            G.Nk = g->Nk;
            G.W = W;
            G.Gx = Gx;
            G.TX = TX;
            G.TY = TY;
            const int kc = std::min(kmax, g->Nk);
            const int nchunk = (g->Nk + kc - 1) / kc;
            G.kchunk = (g->Nk + nchunk - 1) / nchunk;
            G.ntx = (g->Ni + TX - 1) / TX;
            G.nty = (g->Nj + TY - 1) / TY;
            G.ntiles = G.ntx * G.nty * nchunk;
            G.per_xcd = (G.ntiles + 7) / 8;
            for (int t = 0; t < G.ntiles; ++t) {
                const int chunk = t / (G.ntx * G.nty);
                const int kcount = std::min(G.kchunk, g->Nk - chunk * G.kchunk) + 2 * (L - 1);
                try_work.push_back(make_int2((int)q, t));
                try_cost.push_back(area * kcount);
            }
        }
        std::vector<long long> sorted_cost = try_cost;
        std::sort(sorted_cost.begin(), sorted_cost.end(), std::greater<long long>());
        std::priority_queue<long long, std::vector<long long>, std::greater<long long>> free_at;
        for (size_t i = 0; i < std::min(slots, sorted_cost.size()); ++i) free_at.push(0);
        long long span = 0;
        for (long long c : sorted_cost) {
            const long long start = free_at.top();
            free_at.pop();
            free_at.push(start + c);
            span = std::max(span, start + c);
        }
        if (best_span < 0 || span * 50 < best_span * 49) {
            best_span = span;
            geom.swap(try_geom);
            work.swap(try_work);
            cost.swap(try_cost);
        }
    }
    std::vector<size_t> order(work.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return cost[x] > cost[y]; });
    std::vector<int2> sorted(work.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = work[order[i]];
    HIP_TRY(hipStreamSynchronize(b->stream));
    drop_graph(b);  // before freeing anything a captured graph reads
    if (b->d_swork) HIP_TRY(hipFree(b->d_swork));
    b->d_swork = nullptr;
    b->sL = 0;  // until the new work list exists
    HIP_TRY(hipMalloc(&b->d_swork, sizeof(int2) * sorted.size()));
    HIP_TRY(hipMemcpy(b->d_swork, sorted.data(), sizeof(int2) * sorted.size(),
                      hipMemcpyHostToDevice));
    b->sgeom = geom;
    b->snwork = (int)sorted.size();
    b->snblocks = std::max(1, std::min(b->snwork, per_cu * cus));
    b->slds = lds;
    b->sL = L;
    b->sV = V;
    b->sR = R;
    b->sW = W;
    return FDTD_OK;
}

template <typename S>
int occupancy(const fdtd_batch* b, int V, int nt, size_t lds, int* blocks) {
    switch (V) {
        case 1: return occupancy_v<S, 1, true>(b, nt, lds, blocks);
        case 2: return occupancy_v<S, 2, true>(b, nt, lds, blocks);
        case 4: return occupancy_v<S, 4, true>(b, nt, lds, blocks);
        case 8:
            if constexpr (sizeof(S) <= 4) return occupancy_v<S, 8, true>(b, nt, lds, blocks);
            [[fallthrough]];
        default: return fail(FDTD_ERR_UNSUPPORTED, "unsupported V=%d", V);
    }
}

// Apply a launch shape: per-grid plans, work list, persistent grid size.
int apply_shape(fdtd_batch* b, const BatchShape& sh) {
    std::vector<GridPlan> plan;
    size_t lds = 0;
    if (!make_plan(b, sh, &plan, &lds))
        return fail(FDTD_ERR_UNSUPPORTED, "batch shape V=%d NT=%d does not tile every grid",
                    sh.V, sh.NT);
    int per_cu = 0;
    int rc = b->precision == FDTD_FP64 ? occupancy<double>(b, sh.V, sh.NT, lds, &per_cu)
           : b->precision == FDTD_FP32 ? occupancy<float>(b, sh.V, sh.NT, lds, &per_cu)
                                       : occupancy<_Float16>(b, sh.V, sh.NT, lds, &per_cu);
    if (rc) return rc;
    if (per_cu < 1) return fail(FDTD_ERR_UNSUPPORTED, "batch shape does not fit a CU");
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const std::vector<int4> work = make_work(b, plan);
    HIP_TRY(hipStreamSynchronize(b->stream));
    drop_graph(b);  // before freeing anything a captured graph reads
    if (b->d_work) HIP_TRY(hipFree(b->d_work));
    b->d_work = nullptr;
    HIP_TRY(hipMalloc(&b->d_work, sizeof(int4) * work.size()));
    HIP_TRY(hipMemcpy(b->d_work, work.data(), sizeof(int4) * work.size(), hipMemcpyHostToDevice));
    b->shape = sh;
    b->plan = plan;
    b->lds_bytes = lds;
    b->nwork = (int)work.size();
    b->nblocks = std::max(1, std::min(b->nwork, per_cu * cus));
    return FDTD_OK;
}

// Sources of every grid into one device list (and value table), then the
// grid descriptors for the grids' current buffers.
int prepare(fdtd_batch* b) {
    std::vector<fdtd::PointSource> all;
    std::vector<int> off(b->grids.size() + 1, 0);
    std::vector<int> kst;  // per-grid plane ranges of the (plane-sorted) sources
    std::vector<size_t> kst_off(b->grids.size(), 0);
    for (size_t q = 0; q < b->grids.size(); ++q) {
        std::vector<fdtd::PointSource> s = host_sources(b->grids[q]);
        // sorted by plane, stably (sources on one cell keep their order)
        std::stable_sort(s.begin(), s.end(), [](const fdtd::PointSource& x,
                                                const fdtd::PointSource& y) { return x.k < y.k; });
        const int Nk = b->grids[q]->Nk;
        kst_off[q] = kst.size();
        kst.resize(kst.size() + Nk + 1, 0);
        int* ks = kst.data() + kst_off[q];
        for (const auto& x : s) ks[x.k + 1] += 1;
        for (int k = 0; k < Nk; ++k) ks[k + 1] += ks[k];
        off[q] = (int)all.size();
        all.insert(all.end(), s.begin(), s.end());
        if (b->grids[q]->j_dense_active)
            return fail(FDTD_ERR_UNSUPPORTED, "batched grids take point sources, not dense J");
        HIP_TRY(hipStreamSynchronize(b->grids[q]->stream));
    }
    off[b->grids.size()] = (int)all.size();
    HIP_TRY(hipStreamSynchronize(b->stream));
    if ((int)all.size() > b->src_cap) {
        drop_graph(b);
        if (b->d_src) HIP_TRY(hipFree(b->d_src));
        if (b->d_vals) HIP_TRY(hipFree(b->d_vals));
        b->d_src = nullptr;
        b->d_vals = nullptr;
        b->src_cap = 0;  // until the new buffers exist
        const int cap = std::max<int>(64, (int)all.size());
        HIP_TRY(hipMalloc(&b->d_src, sizeof(fdtd::PointSource) * cap));
        HIP_TRY(hipMalloc(&b->d_vals, sizeof(double) * cap * kValTable));
        b->src_cap = cap;
    }
    if (!all.empty())
        HIP_TRY(hipMemcpy(b->d_src, all.data(), sizeof(fdtd::PointSource) * all.size(),
                          hipMemcpyHostToDevice));
    b->total_src = (int)all.size();
    if (kst.size() > b->kstart_cap) {
        drop_graph(b);
        if (b->d_kstart) HIP_TRY(hipFree(b->d_kstart));
        b->d_kstart = nullptr;
        b->kstart_cap = 0;
        HIP_TRY(hipMalloc(&b->d_kstart, sizeof(int) * kst.size()));
        b->kstart_cap = kst.size();
    }
    HIP_TRY(hipMemcpy(b->d_kstart, kst.data(), sizeof(int) * kst.size(), hipMemcpyHostToDevice));
    switch (b->precision) {
        case FDTD_FP64: fill_descriptors<double>(b, off); break;
        case FDTD_FP32: fill_descriptors<float>(b, off); break;
        default: fill_descriptors<_Float16>(b, off);
    }
    if (b->sL > 0) {
        switch (b->precision) {
            case FDTD_FP64: fill_strip_descriptors<double>(b, off, kst_off); break;
            case FDTD_FP32: fill_strip_descriptors<float>(b, off, kst_off); break;
            default: fill_strip_descriptors<_Float16>(b, off, kst_off);
        }
        if (b->h_sgrids.size() != b->sgrid_desc_bytes) {
            drop_graph(b);
            if (b->d_sgrids) HIP_TRY(hipFree(b->d_sgrids));
            b->d_sgrids = nullptr;
            b->sgrid_desc_bytes = 0;
            HIP_TRY(hipMalloc(&b->d_sgrids, b->h_sgrids.size()));
            b->sgrid_desc_bytes = b->h_sgrids.size();
        }
        HIP_TRY(hipMemcpy(b->d_sgrids, b->h_sgrids.data(), b->h_sgrids.size(),
                          hipMemcpyHostToDevice));
    }
    if (b->h_grids.size() != b->grid_desc_bytes) {
        drop_graph(b);
        if (b->d_grids) HIP_TRY(hipFree(b->d_grids));
        b->d_grids = nullptr;
        b->grid_desc_bytes = 0;
        HIP_TRY(hipMalloc(&b->d_grids, b->h_grids.size()));
        b->grid_desc_bytes = b->h_grids.size();
    }
    HIP_TRY(hipMemcpy(b->d_grids, b->h_grids.data(), b->h_grids.size(), hipMemcpyHostToDevice));
    return FDTD_OK;
}

template <typename S>
int prologue_B_s(fdtd_batch* b) {
    using C = typename fdtd::ComputeOf<S>::type;
    const auto* desc = static_cast<const fdtd::BatchGrid<S, C>*>(b->d_grids);
    for (size_t q = 0; q < b->grids.size(); ++q) {
        const fdtd_grid* g = b->grids[q];
        const dim3 block(64, 4, 1);
        const dim3 grid((g->Ni + 63) / 64, (g->Nj + 3) / 4, g->Nk);
        hipLaunchKernelGGL((fdtd::batch_plain_update_B<S, C>), grid, block, 0, b->stream, desc,
                           (int)q);
    }
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

int prologue_B(fdtd_batch* b) {
    switch (b->precision) {
        case FDTD_FP64: return prologue_B_s<double>(b);
        case FDTD_FP32: return prologue_B_s<float>(b);
        default: return prologue_B_s<_Float16>(b);
    }
}

// Enqueue n steps on the batch stream: B/2 prologue of every grid, then n
// batched sweeps.  Source values come from step0 (host) or, when
// step_dev is non-null, from the device counter, which is advanced at the
// end (graph replays then continue in time).
int enqueue_steps(fdtd_batch* b, int nsteps, long long step0, bool device_step, int* swaps) {
    if (int rc = prologue_B(b)) return rc;
    // strip passes of sL steps first (when set and the run is long enough),
    // then the remaining steps as single sweeps; every launch swaps buffers
    const int L = b->sL;
    const int npass = L > 0 ? nsteps / L : 0;
    const int rem = nsteps - npass * L;
    int parity = 0;
    int s = 0;
    const int pchunk = L > 0 ? (kValTable / L) * L : kValTable;
    for (int p = 0; p < npass; ++p, s += L) {
        if (b->total_src > 0 && s % pchunk == 0) {
            const int chunk = std::min(pchunk, npass * L - s);
            const int n = b->total_src * chunk;
            hipLaunchKernelGGL(fdtd::eval_sources, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                               b->d_src, b->total_src, chunk,
                               device_step ? (const long long*)b->d_step : nullptr,
                               device_step ? 0LL : step0 + s, b->grids[0]->dt, b->d_vals);
            HIP_TRY(hipGetLastError());
            if (device_step) {
                hipLaunchKernelGGL(fdtd::advance_step, dim3(1), dim3(1), 0, b->stream, b->d_step,
                                   (long long)chunk);
                HIP_TRY(hipGetLastError());
            }
        }
        const double* vals = b->total_src > 0 ? b->d_vals + (size_t)(s % pchunk) * b->total_src
                                              : nullptr;
        if (int rc = launch_batch_strip(b, rem == 0 && p + 1 == npass, parity, vals)) return rc;
        parity ^= 1;
    }
    for (int q = 0; q < rem; ++q, ++s) {
        const int r = q;  // step index within the single sweeps
        if (b->total_src > 0 && r % kValTable == 0) {
            const int chunk = std::min(kValTable, rem - r);
            const int n = b->total_src * chunk;
            // values of steps step0+s ...; with a device counter the kernel
            // reads the run's first step and this launch adds s to it
            hipLaunchKernelGGL(fdtd::eval_sources, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                               b->d_src, b->total_src, chunk,
                               device_step ? (const long long*)b->d_step : nullptr,
                               device_step ? 0LL : step0 + s, b->grids[0]->dt, b->d_vals);
            HIP_TRY(hipGetLastError());
            if (device_step) {
                hipLaunchKernelGGL(fdtd::advance_step, dim3(1), dim3(1), 0, b->stream, b->d_step,
                                   (long long)chunk);
                HIP_TRY(hipGetLastError());
            }
        }
        const double* row = b->total_src > 0 ? b->d_vals + (size_t)(r % kValTable) * b->total_src
                                             : nullptr;
        if (int rc = launch_batch(b, q + 1 < rem, parity, row)) return rc;
        parity ^= 1;
    }
    if (swaps) *swaps = npass + rem;
    return FDTD_OK;
}

void finish_steps(fdtd_batch* b, int nsteps, int swaps) {
    for (fdtd_grid* g : b->grids) {
        g->cur = (g->cur + swaps) & 1;
        g->step += nsteps;
        g->ba_cur = -1;  // a batch run leaves no carried half step
    }
}

// The batch evaluates all waveform sources in one table with one dt, at
// the common step count of its grids.
int check_steps(fdtd_batch* b) {
    for (fdtd_grid* g : b->grids)
        if (g->step != b->grids[0]->step)
            return fail(FDTD_ERR_UNSUPPORTED, "grids of a batch must be at the same step");
    for (fdtd_grid* g : b->grids)
        if (!g->waveforms.empty() && g->dt != b->grids[0]->dt)
            return fail(FDTD_ERR_UNSUPPORTED, "grids with waveform sources must share one dt");
    for (fdtd_grid* g : b->grids)
        if (g->nprobe > 0 || g->energy_every > 0)
            return fail(FDTD_ERR_UNSUPPORTED,
                        "batches record no probes or energy: step such grids on their own");
    return FDTD_OK;
}

}  // namespace

extern "C" {

int fdtd_batch_create(int n, fdtd_grid* const* grids, fdtd_batch** out) {
    if (!out || n <= 0 || !grids) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    *out = nullptr;
    for (int q = 0; q < n; ++q) {
        if (!grids[q]) return fail(FDTD_ERR_INVALID_ARGUMENT, "null grid %d", q);
        if (grids[q]->precision != grids[0]->precision)
            return fail(FDTD_ERR_INVALID_ARGUMENT, "grids of a batch must share one precision");
        if (grids[q]->pml)
            return fail(FDTD_ERR_UNSUPPORTED, "grid %d has a PML layer; batches take periodic grids", q);
        for (int r = 0; r < q; ++r)
            if (grids[r] == grids[q]) return fail(FDTD_ERR_INVALID_ARGUMENT, "grid %d repeated", q);
    }
    fdtd_batch* b = new fdtd_batch();
    b->grids.assign(grids, grids + n);
    b->precision = grids[0]->precision;
    auto cleanup = [&](int rc) {
        fdtd_batch_destroy(b);
        return rc;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(FDTD_ERR_HIP, "hipStreamCreate failed"));
    for (auto& e : b->ev)
        if (hipEventCreate(&e) != hipSuccess) return cleanup(fail(FDTD_ERR_HIP, "hipEventCreate"));
    if (hipMalloc(&b->d_counters, 4 * sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&b->d_step, sizeof(long long)) != hipSuccess)
        return cleanup(fail(FDTD_ERR_HIP, "out of device memory"));
    if (hipMemset(b->d_counters, 0, 4 * sizeof(unsigned)) != hipSuccess)
        return cleanup(fail(FDTD_ERR_HIP, "hipMemset failed"));
    // default shape: widest valid V, 512 threads, 64 Ki cells per tile
    int vc[4];
    const int nv = batch_v_candidates(b->precision, vc);
    int rc = FDTD_ERR_UNSUPPORTED;
    for (int q = 0; q < nv && rc != FDTD_OK; ++q)
        for (int nt : {512, 1024, 256})
            if ((rc = apply_shape(b, BatchShape{vc[q], nt, 65536})) == FDTD_OK) break;
    if (rc) return cleanup(fail(FDTD_ERR_UNSUPPORTED, "no batch launch shape tiles every grid"));
    *out = b;
    return FDTD_OK;
}

int fdtd_batch_destroy(fdtd_batch* b) {
    if (!b) return FDTD_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->d_grids) (void)hipFree(b->d_grids);
    if (b->d_work) (void)hipFree(b->d_work);
    if (b->d_counters) (void)hipFree(b->d_counters);
    if (b->d_src) (void)hipFree(b->d_src);
    if (b->d_vals) (void)hipFree(b->d_vals);
    if (b->d_step) (void)hipFree(b->d_step);
    if (b->d_swork) (void)hipFree(b->d_swork);
    if (b->d_sgrids) (void)hipFree(b->d_sgrids);
    if (b->d_kstart) (void)hipFree(b->d_kstart);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return FDTD_OK;
}

int fdtd_batch_update_fields(fdtd_batch* b, int nsteps) {
    if (!b || nsteps < 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (nsteps == 0) return FDTD_OK;
    TraceRange tr("batch_update_fields");
    if (int rc = check_steps(b)) return rc;
    if (int rc = prepare(b)) return rc;
    int swaps = 0;
    if (int rc = enqueue_steps(b, nsteps, b->grids[0]->step, false, &swaps)) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    finish_steps(b, nsteps, swaps);
    return FDTD_OK;
}

int fdtd_batch_update_fields_graph(fdtd_batch* b, int nsteps, int replays) {
    if (!b || nsteps <= 0 || replays < 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (int rc = check_steps(b)) return rc;
    if (int rc = prepare(b)) return rc;
    TraceRange tr("batch_update_fields_graph");
    if (!b->graph || graph_key_of(b, nsteps) != b->graph_key) {
        drop_graph(b);
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int swaps = 0;
        const int rc = enqueue_steps(b, nsteps, 0, true, &swaps);
        hipError_t e = hipStreamEndCapture(b->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(FDTD_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(FDTD_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
        b->graph_steps = nsteps;
        b->graph_swaps = swaps;
        b->graph_key = graph_key_of(b, nsteps);
    }
    const long long step0 = b->grids[0]->step;
    HIP_TRY(hipMemcpy(b->d_step, &step0, sizeof step0, hipMemcpyHostToDevice));
    for (int r = 0; r < replays; ++r) {
        if (r > 0) {
            finish_steps(b, nsteps, b->graph_swaps);
            // an odd number of launches swaps every grid's buffers: re-point
            // the descriptors (prepare() waits for the previous replay first)
            if (b->graph_swaps & 1) {
                if (int rc = prepare(b)) return rc;
                if (graph_key_of(b, nsteps) != b->graph_key)
                    return fail(FDTD_ERR_HIP, "batch buffers changed between graph replays");
            }
        }
        HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (replays > 0) finish_steps(b, nsteps, b->graph_swaps);
    return FDTD_OK;
}

int fdtd_batch_synchronize(fdtd_batch* b) {
    if (!b) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return FDTD_OK;
}

int fdtd_batch_set_time_blocking_strips(fdtd_batch* b, int L, int V, int R, int W) {
    if (!b) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    if (L <= 0) {
        b->sL = 0;
        drop_graph(b);
        return FDTD_OK;
    }
    return apply_strips(b, L, V > 0 ? V : 1, R > 0 ? R : 2, W > 0 ? W : 16);
}

int fdtd_batch_get_time_blocking_strips(const fdtd_batch* b, int out[5]) {
    if (!b || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    out[0] = b->sL;
    out[1] = b->sL > 0 ? b->sV : 0;
    out[2] = b->sL > 0 ? b->sR : 0;
    out[3] = b->sL > 0 ? b->sW : 0;
    out[4] = b->sL > 0 ? b->snwork : 0;
    return FDTD_OK;
}

int fdtd_batch_set_shape(fdtd_batch* b, int V, int threads, long long tile_cells) {
    if (!b || V <= 0 || threads <= 0 || tile_cells <= 0)
        return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    return apply_shape(b, BatchShape{V, threads, tile_cells});
}

int fdtd_batch_get_shape(const fdtd_batch* b, long long out[4]) {
    if (!b || !out) return fail(FDTD_ERR_INVALID_ARGUMENT, "null argument");
    out[0] = b->shape.V;
    out[1] = b->shape.NT;
    out[2] = b->shape.target;
    out[3] = b->nwork;
    return FDTD_OK;
}

int fdtd_batch_time(fdtd_batch* b, int nsteps, float* ms_batched, float* ms_sequential) {
    if (!b || nsteps <= 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (int rc = check_steps(b)) return rc;
    if (int rc = prepare(b)) return rc;
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    int swaps = 0;
    if (int rc = enqueue_steps(b, nsteps, b->grids[0]->step, false, &swaps)) return rc;
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipEventSynchronize(b->ev[1]));
    finish_steps(b, nsteps, swaps);
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
    if (ms_batched) *ms_batched = t / nsteps;
    if (ms_sequential) {
        // the same steps as one fdtd_update_fields per grid, on one stream
        for (fdtd_grid* g : b->grids) {
            hipStream_t own = g->stream;
            g->stream = b->stream;
            const int rc = flush_sources(g);
            g->stream = own;
            if (rc) return rc;
        }
        HIP_TRY(hipEventRecord(b->ev[2], b->stream));
        for (fdtd_grid* g : b->grids) {
            hipStream_t own = g->stream;
            g->stream = b->stream;
            const int rc = fdtd_update_fields(g, nsteps);
            g->stream = own;
            if (rc) return rc;
        }
        HIP_TRY(hipEventRecord(b->ev[3], b->stream));
        HIP_TRY(hipEventSynchronize(b->ev[3]));
        HIP_TRY(hipEventElapsedTime(&t, b->ev[2], b->ev[3]));
        *ms_sequential = t / nsteps;
    }
    return FDTD_OK;
}

int fdtd_batch_autotune(fdtd_batch* b, int reps, long long out[4], float* best_ms) {
    if (!b || reps <= 0) return fail(FDTD_ERR_INVALID_ARGUMENT, "invalid argument");
    if (int rc = check_steps(b)) return rc;
    if (int rc = prepare(b)) return rc;
    // Steady sweeps are timed on the grids' own buffers; every grid's state
    // is copied aside first and put back on every return path, and a failed
    // run also restores the launch shape.
    struct Snapshot {
        fdtd_batch* b;
        BatchShape shape;
        std::vector<void*> saved;
        bool done = false;
        static void* field(fdtd_grid* g, int c) {
            return c < 3 ? g->E[g->cur][c] : g->B[g->cur][c - 3];
        }
        int save() {
            for (fdtd_grid* g : b->grids)
                for (int c = 0; c < 6; ++c) {
                    void* p = nullptr;
                    HIP_TRY(hipMalloc(&p, g->N * g->esize));
                    saved.push_back(p);
                    HIP_TRY(hipMemcpyAsync(p, field(g, c), g->N * g->esize,
                                           hipMemcpyDeviceToDevice, b->stream));
                }
            return FDTD_OK;
        }
        int restore() {
            done = true;
            size_t k = 0;
            for (fdtd_grid* g : b->grids)
                for (int c = 0; c < 6 && k < saved.size(); ++c, ++k)
                    HIP_TRY(hipMemcpyAsync(field(g, c), saved[k], g->N * g->esize,
                                           hipMemcpyDeviceToDevice, b->stream));
            HIP_TRY(hipStreamSynchronize(b->stream));
            return FDTD_OK;
        }
        ~Snapshot() {
            if (!done) {
                (void)restore();
                (void)apply_shape(b, shape);
            }
            for (void* p : saved) (void)hipFree(p);
        }
    } snap{b, b->shape};
    if (int rc = snap.save()) return rc;
    const BatchShape orig = b->shape;
    BatchShape best = orig;
    float best_t = 1e30f;
    int vc[4];
    const int nv = batch_v_candidates(b->precision, vc);
    for (int q = 0; q < nv; ++q)
        for (int nt : {256, 512, 768, 1024})
            for (long long tc : {16384LL, 32768LL, 65536LL, 131072LL}) {
                if (apply_shape(b, BatchShape{vc[q], nt, tc}) != FDTD_OK) continue;
                if (int rc = prepare(b)) return rc;
                if (b->total_src > 0) {
                    hipLaunchKernelGGL(fdtd::eval_sources, dim3((b->total_src + 255) / 256),
                                       dim3(256), 0, b->stream, b->d_src, b->total_src, 1,
                                       (const long long*)nullptr, b->grids[0]->step,
                                       b->grids[0]->dt, b->d_vals);
                }
                const double* row = b->total_src > 0 ? b->d_vals : nullptr;
                if (int rc = launch_batch(b, true, 0, row)) return rc;
                HIP_TRY(hipEventRecord(b->ev[0], b->stream));
                for (int r = 0; r < reps; ++r)
                    if (int rc = launch_batch(b, true, r & 1, row)) return rc;
                HIP_TRY(hipEventRecord(b->ev[1], b->stream));
                HIP_TRY(hipEventSynchronize(b->ev[1]));
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
                ms /= reps;
                if (ms < best_t) {
                    best_t = ms;
                    best = b->shape;
                }
            }
    if (int rc = snap.restore()) {
        (void)apply_shape(b, orig);
        return rc;
    }
    if (best_t >= 1e30f) {
        (void)apply_shape(b, orig);
        return fail(FDTD_ERR_UNSUPPORTED, "no batch launch shape fits");
    }
    if (int rc = apply_shape(b, best)) return rc;
    if (out) {
        out[0] = best.V;
        out[1] = best.NT;
        out[2] = best.target;
        out[3] = b->nwork;
    }
    if (best_ms) *best_ms = best_t;
    return FDTD_OK;
}

}  // extern "C"
