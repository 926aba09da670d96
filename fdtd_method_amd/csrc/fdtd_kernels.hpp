// gfx950 kernels for the Yee time step (B/2, E, B/2 per step) on a periodic
// grid stored as flat structure-of-arrays, index = i + j*Ni + k*Ni*Nj.
//
// Reference semantics (file:line):
//   B half step      include/FDTD_kokkos/kokkos_functors.h:128-151 (== src/FDTD/FDTD.cpp:99-130)
//   E full step      include/FDTD_kokkos/kokkos_functors.h:64-90   (Jx->Ex, Jy->Ey, Jz->Ez)
//   periodic wrap    include/FDTD/FDTD.h:24-31
//   step order       src/FDTD_kokkos/FDTD_kokkos.cpp:91-104
//
// Two kernel families:
//  * plain_update_B / plain_update_E: one thread per cell, in place, the
//    reference's three-sweep structure.  This is the FP64 parity path.
//  * fused_sweep: one sweep per time step.  Consecutive steps run
//    "... B/2 | B/2, E, B/2 | B/2 ..."; the two adjacent B half steps read
//    the same E, so their increment d is computed once and applied twice as
//    (B + d) + d -- bit-identical to the reference's two separate sweeps.
//    The steady-state sweep therefore maps (E_n, B_{n+1/2}) to
//    (E_{n+1}, B_{n+3/2}); the final sweep of a run applies d once, giving
//    (E_{n+1}, B_{n+1}).  A run of n steps is one plain B half step plus n
//    fused sweeps.
//
// All arithmetic keeps the reference's C++ evaluation order and the file is
// compiled with -ffp-contract=off, so FP64 results are bit-identical to the
// reference's CPU fields.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace fdtd {

// ---------------------------------------------------------------------------
// Types

template <typename S> struct ComputeOf { using type = S; };
template <> struct ComputeOf<_Float16> { using type = float; };

// Coefficients of one precision path.  With per-component storage scales
// s (stored = s * physical) each coefficient absorbs s_target / s_source;
// with unit scales they are exactly the reference's coefficients
// (src/FDTD/FDTD.cpp:43-53).
template <typename C> struct Coefs {
    // E update: Ex += jx*Jx + ex_bz*(Bz - Bz[j-1]) - ex_by*(By - By[k-1]), ...
    C ex_bz, ex_by, ey_bx, ey_bz, ez_by, ez_bx;
    C jx, jy, jz;
    // B update: Bx += bx_ey*(Ey[k+1] - Ey) - bx_ez*(Ez[j+1] - Ez), ...
    C bx_ey, bx_ez, by_ez, by_ex, bz_ex, bz_ey;
};

// a - k*x, the second curl term of every E and B update.  Float compute
// (FP32 and FP16 storage) forms it with one fused multiply-add (one
// rounding); double keeps the separate multiply and subtract of the
// reference's CPU code (this file is compiled with -ffp-contract=off), so
// FP64 stays bit-identical to it.  Every kernel forms the terms through
// this, so all float paths (three sweeps, fused and blocked sweeps,
// strips, batches, PML) stay bit-identical to each other.
template <typename C>
__device__ __forceinline__ C sub_mul(C a, C k, C x) {
    return a - k * x;
}
template <>
__device__ __forceinline__ float sub_mul<float>(float a, float k, float x) {
    return __builtin_fmaf(-k, x, a);
}

// A point current source.  Its J component is evaluated in FP64 at
// t = (n+1)*dt before step n (perf-tests/sample/sample.cpp:74):
//   kind 0: amp
//   kind 1: amp * sin(2*pi*(t - t0)/period + phase) * exp(-(t - t0)^2/(2 sigma^2))
//           (envelope omitted when sigma <= 0)
//   kind 2: column `col` of a table of values given per step by the host
//           (the values a caller set before each step of a call)
struct PointSource {
    int i, j, k, comp;
    int kind, col;
    double amp, period, t0, sigma, phase;
};

constexpr double kPi = 3.14159265358;  // include/Constants.h:11 (truncated)

__device__ __forceinline__ double source_value(const PointSource& s, double t) {
    if (s.kind == 0) return s.amp;
    const double tau = t - s.t0;
    double v = s.amp * sin(2.0 * kPi * tau / s.period + s.phase);
    if (s.sigma > 0.0) v = v * exp(-(tau * tau) / (2.0 * s.sigma * s.sigma));
    return v;
}

// Wrap for an index known to lie in [-1, N]: -1 -> N-1, N -> 0
// (same mapping as include/FDTD/FDTD.h:24-31).
__device__ __forceinline__ int wrap(int i, int n) {
    return i < 0 ? n - 1 : (i == n ? 0 : i);
}

template <typename S, int V> struct alignas(sizeof(S) * V) Pack { S v[V]; };

template <typename S, typename C, int V>
__device__ __forceinline__ void load_pack(const S* __restrict__ p, C (&out)[V]) {
    const Pack<S, V> pk = *reinterpret_cast<const Pack<S, V>*>(p);
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = static_cast<C>(pk.v[v]);
}

// A pack held as opaque machine words until it is widened where it is
// used: the compiler then cannot split FP16 halves (or convert) right after
// the load, which would make a prefetch wait for its own loads.
template <typename S, int V> struct RawPack {
    using W = typename std::conditional<(sizeof(S) * V >= 4), uint32_t, uint16_t>::type;
    static constexpr int kWords = static_cast<int>(sizeof(S) * V / sizeof(W));
    struct alignas(sizeof(S) * V) Words { W w[kWords]; };
    Words words;
};

template <typename S, int V>
__device__ __forceinline__ RawPack<S, V> load_raw(const S* __restrict__ p) {
    RawPack<S, V> r;
    r.words = *reinterpret_cast<const typename RawPack<S, V>::Words*>(p);
    return r;
}

template <typename S, typename C, int V>
__device__ __forceinline__ void widen(const RawPack<S, V>& raw, C (&out)[V]) {
    const Pack<S, V> pk = __builtin_bit_cast(Pack<S, V>, raw.words);
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = static_cast<C>(pk.v[v]);
}

// Sweep outputs are not read again within the sweep (and the state is far
// larger than the L2s), so they are stored non-temporally: the L2 keeps the
// input rows that neighbouring tiles re-read.  The per-element stores are
// merged into one V-wide store.
template <typename S, typename C, int V>
__device__ __forceinline__ void store_pack(S* __restrict__ p, const C (&in)[V]) {
#pragma unroll
    for (int v = 0; v < V; ++v) __builtin_nontemporal_store(static_cast<S>(in[v]), p + v);
}

// Raw buffer access to one plane: a resource over the plane's bytes and a
// 32-bit byte offset per lane.  The hardware range check makes an access at
// or beyond the resource's size read 0 and write nothing, so a lane outside
// the stored region takes kBufSkip as its offset and a plane that is not
// read or written gets a size-0 resource.  The memory instructions are then
// issued on every path through a loop, and the compiler can count them:
// waiting for a prefetch does not wait for stores issued after it.  Planes
// are at most 2 GiB, so kBufSkip lies beyond every resource.
constexpr int kBufDword3 = 0x00020000;   // raw (untyped) buffer, gfx9 family
constexpr unsigned kBufSkip = 0x80000000u;
constexpr int kBufNT = 2;                 // non-temporal store

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes),
                                             kBufDword3);
}

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt unconstrained): every vector
// memory access of the wavefront has completed.  Ends a prologue whose loads
// would otherwise still be outstanding where it joins a loop, making the
// loop's first uses of those registers wait on the loop's own stores.
__device__ __forceinline__ void wait_vmem_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }



template <typename S, int V>
__device__ __forceinline__ RawPack<S, V> buf_load_raw(__amdgpu_buffer_rsrc_t r, unsigned off) {
    using Words = typename RawPack<S, V>::Words;
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    constexpr int bytes = static_cast<int>(sizeof(S)) * V;
    RawPack<S, V> p;
    if constexpr (bytes == 2)
        p.words = __builtin_bit_cast(Words, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0));
    else if constexpr (bytes == 4)
        p.words = __builtin_bit_cast(Words, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    else if constexpr (bytes == 8)
        p.words = __builtin_bit_cast(Words, static_cast<u32x2>(__builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0)));
    else {
        static_assert(bytes == 16, "pack of 2, 4, 8 or 16 bytes");
        p.words = __builtin_bit_cast(Words, static_cast<u32x4>(__builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0)));
    }
    return p;
}

template <typename S>
__device__ __forceinline__ void buf_store_nt(S x, __amdgpu_buffer_rsrc_t r, unsigned off) {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    if constexpr (sizeof(S) == 2)
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, x), r, off, 0, kBufNT);
    else if constexpr (sizeof(S) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, off, 0, kBufNT);
    else {
        static_assert(sizeof(S) == 8, "2, 4 or 8 byte values");
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, off, 0, kBufNT);
    }
}

// V values stored as one pack through a buffer resource (non-temporal);
// an offset at or beyond the resource's size stores nothing.
template <typename S, typename C, int V>
__device__ __forceinline__ void buf_store_pack(const C (&x)[V], __amdgpu_buffer_rsrc_t r,
                                               unsigned off) {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    Pack<S, V> pk;
#pragma unroll
    for (int v = 0; v < V; ++v) pk.v[v] = static_cast<S>(x[v]);
    constexpr int bytes = static_cast<int>(sizeof(S)) * V;
    if constexpr (bytes == 2)
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, pk), r, off, 0, kBufNT);
    else if constexpr (bytes == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, pk), r, off, 0, kBufNT);
    else if constexpr (bytes == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, pk), r, off, 0, kBufNT);
    else {
        static_assert(bytes == 16, "pack of 2, 4, 8 or 16 bytes");
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, pk), r, off, 0, kBufNT);
    }
}

// A field value as the storage type holds it.  Blocked levels keep their
// intermediate fields in registers; with FP16 storage a level's output is
// rounded where the next unblocked sweep would read it back from memory
// (the next level's E and B inputs), while within a level the B update uses
// E unrounded, as the fused sweep does.
template <typename S, typename C>
__device__ __forceinline__ C as_stored(C x) {
    if constexpr (sizeof(S) < sizeof(C))
        return static_cast<C>(static_cast<S>(x));
    else
        return x;
}

// Field pointers of one grid (current buffers, or in/out buffer pairs).
template <typename S> struct Fields {
    S* e[3];
    S* b[3];
};

// ---------------------------------------------------------------------------
// Plain three-sweep kernels (one thread per cell, in place).

template <typename S, typename C>
__device__ __forceinline__ void plain_B_cell(const Fields<S>& f, const Coefs<C>& c, int Ni, int Nj,
                                             int Nk, int i, int j, int k) {
    const int plane = Ni * Nj;
    const int idx = i + j * Ni + k * plane;
    const int i_next = wrap(i + 1, Ni) + j * Ni + k * plane;
    const int j_next = i + wrap(j + 1, Nj) * Ni + k * plane;
    const int k_next = i + j * Ni + wrap(k + 1, Nk) * plane;
    const C ex = f.e[0][idx], ey = f.e[1][idx], ez = f.e[2][idx];
    const C dbx = sub_mul(c.bx_ey * (C(f.e[1][k_next]) - ey), c.bx_ez, C(f.e[2][j_next]) - ez);
    const C dby = sub_mul(c.by_ez * (C(f.e[2][i_next]) - ez), c.by_ex, C(f.e[0][k_next]) - ex);
    const C dbz = sub_mul(c.bz_ex * (C(f.e[0][j_next]) - ex), c.bz_ey, C(f.e[1][i_next]) - ey);
    f.b[0][idx] = static_cast<S>(C(f.b[0][idx]) + dbx);
    f.b[1][idx] = static_cast<S>(C(f.b[1][idx]) + dby);
    f.b[2][idx] = static_cast<S>(C(f.b[2][idx]) + dbz);
}

template <typename S, typename C>
__global__ void __launch_bounds__(256)
plain_update_B(Fields<S> f, Coefs<C> c, int Ni, int Nj, int Nk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= Ni || j >= Nj) return;
    plain_B_cell<S, C>(f, c, Ni, Nj, Nk, i, j, k);
}

// Source values of steps n0 .. n0+nsteps-1: vals[s*nsrc + q] = J of source q
// at t = (n0+s+1)*dt, in FP64 (n0 read from step_dev when it is non-null).
// The sweeps then only add precomputed values.
static __global__ void __launch_bounds__(256)
eval_sources(const PointSource* __restrict__ src, int nsrc, int nsteps,
             const long long* __restrict__ step_dev, long long step, double dt,
             double* __restrict__ vals, const double* __restrict__ tab = nullptr,
             int tab_stride = 0) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsrc * nsteps) return;
    const int s = t / nsrc, q = t - s * nsrc;
    if (src[q].kind == 2) {
        vals[t] = tab ? tab[static_cast<size_t>(s) * tab_stride + src[q].col] : 0.0;
        return;
    }
    const long long n = (step_dev ? *step_dev : step) + s;
    vals[t] = source_value(src[q], static_cast<double>(n + 1) * dt);
}

// Read-only kernel inputs through the constant address space: uniform
// reads become scalar loads instead of vector loads a wavefront waits on.
template <typename T> using cptr = const __attribute__((address_space(4))) T*;
template <typename T> __device__ __forceinline__ cptr<T> as_const(const T* p) {
    return (cptr<T>)(p);
}

// Sources q0..q1-1 of the list (those of plane k when the list is indexed
// by plane, else all of them): adds the J term of every source located in
// cells (ib..ib+V-1, j, k) to the E components' partial sums t[d][v] =
// first curl term, as the reference's (cur*J + curl term) --
// kokkos_functors.h:81-89.
template <typename C, int V>
__device__ __forceinline__ void add_point_sources(C (&t)[3][V], const Coefs<C>& c,
                                                  const PointSource* src_g,
                                                  const double* vals_g, int q0, int q1,
                                                  int ib, int j, int k) {
    const cptr<PointSource> src = as_const(src_g);
    const cptr<double> vals = as_const(vals_g);
    for (int q = q0; q < q1; ++q) {
        const int sk = src[q].k, sj = src[q].j, si = src[q].i;
        if (sk != k || sj != j || si < ib || si >= ib + V) continue;
        const int comp = src[q].comp;
        const double val = vals[q];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (si != ib + v) continue;
            if (comp == 0) t[0][v] = c.jx * C(val) + t[0][v];
            if (comp == 1) t[1][v] = c.jy * C(val) + t[1][v];
            if (comp == 2) t[2][v] = c.jz * C(val) + t[2][v];
        }
    }
}

// Probes: single field values recorded in FP64 after every step.  A probe
// of E at plane k is listed under plane key k, one of B at plane k under key
// k+1 (mod Nk): the sweep iteration (or strip level) that computes E at
// plane k also computes B at plane k-1.
struct Probe {
    int i, j, k, comp;  // comp 0..5: Ex..Bz
    int col, pad;       // output column
};

struct ProbeArgs {
    const Probe* p;      // sorted by plane key
    const int* pstart;   // probes of key k: [pstart[k], pstart[k+1])
    int n;               // 0: none
    int stride;          // output row length
    double* out;         // the launch's first step's row; step s at out + s*stride
    double inv[6];       // 1 / storage scale of Ex..Bz (values as computed in the launch)
};

// Records the probes of plane key `key` that lie in columns i0..i0+V-1 of
// row j: E components from e (when e_ok), B components from b (when b_ok),
// cell v only when vok[v]; values are as stored (rounded to S), in physical
// units.
template <typename S, typename C, int V>
__device__ __forceinline__ void record_probes(const ProbeArgs& pa, double* out, int key, int i0,
                                              int j, const C (&e)[3][V], const C (&b)[3][V],
                                              bool e_ok, bool b_ok, const bool (&vok)[V]) {
    const cptr<int> ps = as_const(pa.pstart);
    const int q0 = ps[key], q1 = ps[key + 1];
    if (q1 <= q0) return;
    const cptr<Probe> pr = as_const(pa.p);
    for (int q = q0; q < q1; ++q) {
        const int pi = pr[q].i;
        if (pr[q].j != j || pi < i0 || pi >= i0 + V) continue;
        const int comp = pr[q].comp;
        if (comp < 3 ? !e_ok : !b_ok) continue;
        C x = C(0);
        bool ok = false;
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (pi == i0 + v) {
                ok = vok[v];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (comp == d) x = e[d][v];
                    if (comp == 3 + d) x = b[d][v];
                }
            }
        // the scale by selects: an index into the argument array would
        // copy it to scratch
        double inv = pa.inv[0];
#pragma unroll
        for (int c = 1; c < 6; ++c)
            if (comp == c) inv = pa.inv[c];
        if (ok) out[pr[q].col] = static_cast<double>(static_cast<S>(x)) * inv;
    }
}

// One cell of the E update: dense J (type JT, may be null) plus point
// sources with values `vals`.
template <typename S, typename C, typename JT>
__device__ __forceinline__ void plain_E_cell(const Fields<S>& f, const JT* __restrict__ jx,
                                             const JT* __restrict__ jy, const JT* __restrict__ jz,
                                             const PointSource* __restrict__ src,
                                             const double* __restrict__ vals, int nsrc,
                                             const Coefs<C>& c, int Ni, int Nj, int Nk, int i,
                                             int j, int k) {
    const int plane = Ni * Nj;
    const int idx = i + j * Ni + k * plane;
    const int i_pred = wrap(i - 1, Ni) + j * Ni + k * plane;
    const int j_pred = i + wrap(j - 1, Nj) * Ni + k * plane;
    const int k_pred = i + j * Ni + wrap(k - 1, Nk) * plane;
    const C bx = f.b[0][idx], by = f.b[1][idx], bz = f.b[2][idx];
    C t[3][1];
    t[0][0] = c.ex_bz * (bz - C(f.b[2][j_pred]));
    t[1][0] = c.ey_bx * (bx - C(f.b[0][k_pred]));
    t[2][0] = c.ez_by * (by - C(f.b[1][i_pred]));
    if (jx != nullptr) {
        t[0][0] = c.jx * C(jx[idx]) + t[0][0];
        t[1][0] = c.jy * C(jy[idx]) + t[1][0];
        t[2][0] = c.jz * C(jz[idx]) + t[2][0];
    }
    if (nsrc > 0) add_point_sources<C, 1>(t, c, src, vals, 0, nsrc, i, j, k);
    const C ex = C(f.e[0][idx]) + sub_mul(t[0][0], c.ex_by, by - C(f.b[1][k_pred]));
    const C ey = C(f.e[1][idx]) + sub_mul(t[1][0], c.ey_bz, bz - C(f.b[2][i_pred]));
    const C ez = C(f.e[2][idx]) + sub_mul(t[2][0], c.ez_bx, bx - C(f.b[0][j_pred]));
    f.e[0][idx] = static_cast<S>(ex);
    f.e[1][idx] = static_cast<S>(ey);
    f.e[2][idx] = static_cast<S>(ez);
}

template <typename S, typename C, typename JT>
__global__ void __launch_bounds__(256)
plain_update_E(Fields<S> f, const JT* __restrict__ jx, const JT* __restrict__ jy,
               const JT* __restrict__ jz, const PointSource* __restrict__ src,
               const double* __restrict__ vals, int nsrc, Coefs<C> c, int Ni, int Nj, int Nk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= Ni || j >= Nj) return;
    plain_E_cell<S, C, JT>(f, jx, jy, jz, src, vals, nsrc, c, Ni, Nj, Nk, i, j, k);
}

// Single-cell current writes into dense J arrays: write q stores value q at
// cell index q of component comp (0..2).  The host passes each cell once.
struct CurrentWrite {
    long long index;
    int comp, pad;
    double value;
};

template <typename JT>
__global__ void __launch_bounds__(256)
scatter_currents(JT* __restrict__ jx, JT* __restrict__ jy, JT* __restrict__ jz,
                 const CurrentWrite* __restrict__ w, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const CurrentWrite x = w[q];
    JT* j = x.comp == 0 ? jx : (x.comp == 1 ? jy : jz);
    j[x.index] = static_cast<JT>(x.value);
}

// The plain updates over the box [lo, hi) of cells (the reference's
// begin_main_* .. end_main_* loop bounds, src/FDTD/FDTD.cpp:63-130,
// src/FDTD_kokkos/FDTD_PML_kokkos.cpp:160-198); neighbours still wrap over
// the whole grid.  Block (64, 4), grid over the box's extent, z per block.
struct Box {
    int lo[3], hi[3];
};

template <typename S, typename C>
__global__ void __launch_bounds__(256)
box_update_B(Fields<S> f, Coefs<C> c, int Ni, int Nj, int Nk, Box box) {
    const int i = box.lo[0] + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = box.lo[1] + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = box.lo[2] + blockIdx.z;
    if (i >= box.hi[0] || j >= box.hi[1]) return;
    plain_B_cell<S, C>(f, c, Ni, Nj, Nk, i, j, k);
}

template <typename S, typename C, typename JT>
__global__ void __launch_bounds__(256)
box_update_E(Fields<S> f, const JT* __restrict__ jx, const JT* __restrict__ jy,
             const JT* __restrict__ jz, const PointSource* __restrict__ src,
             const double* __restrict__ vals, int nsrc, Coefs<C> c, int Ni, int Nj, int Nk,
             Box box) {
    const int i = box.lo[0] + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = box.lo[1] + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = box.lo[2] + blockIdx.z;
    if (i >= box.hi[0] || j >= box.hi[1]) return;
    plain_E_cell<S, C, JT>(f, jx, jy, jz, src, vals, nsrc, c, Ni, Nj, Nk, i, j, k);
}

// ---------------------------------------------------------------------------
// PML absorbing layer (src/FDTD/FDTD_PML.cpp; Kokkos:
// include/FDTD_kokkos/kokkos_functors.h:189-415).
//
// Cells outside the main box [p_i, Ni-p_i) x [p_j, Nj-p_j) x [p_k, Nk-p_k)
// carry split fields (FDTD_PML.h:12-13).  Each split decays by
// a = exp(-sigma*dt*C) and takes one curl term with the coefficient
// (1 - a)/(sigma*d), or C*dt/d where sigma = 0 (FDTD_PML.cpp:63-65,98-111);
// a total is the sum of its two splits.  sigma_x depends on i only, sigma_y
// on j and sigma_z on k (the bands span the grid, FDTD_PML.cpp:326-338), so
// the coefficients are per-axis tables.  The layer takes one full B step
// where the main box takes the two half steps of a step, and its E update
// has no J term (FDTD_PML.cpp:343-365).

enum { kExy = 0, kExz, kEyx, kEyz, kEzx, kEzy };  // E split order, FDTD_PML.h:12
enum { kBxy = 0, kBxz, kByx, kByz, kBzx, kBzy };  // B split order, FDTD_PML.h:13

template <typename S, typename C> struct Pml {
    int pi, pj, pk, pad;
    const S* es_in[6];  // E splits read by this sweep
    S* es_out[6];       // E splits written (ping-pong partner of es_in; equal in place)
    S* bs[6];           // B splits, updated in place
    // Per-axis tables of 5 rows of N values: the decay a, then the
    // coefficients of the four splits that differentiate along that axis
    // (storage-scale ratios folded in):
    //   tx: a, Eyx<-Bz, Ezx<-By, Byx<-Ez, Bzx<-Ey
    //   ty: a, Exy<-Bz, Ezy<-Bx, Bxy<-Ez, Bzy<-Ex
    //   tz: a, Exz<-By, Eyz<-Bx, Bxz<-Ey, Byz<-Ex
    const C* tx;
    const C* ty;
    const C* tz;
};

__device__ __forceinline__ bool in_main(int i, int n, int p) { return i >= p && i < n - p; }

// Rounds a split to its storage type, as stored: a total is the sum of the
// stored splits (FDTD_PML.cpp:128-130,197-199).
template <typename S, typename C> __device__ __forceinline__ C stored(C x) {
    return static_cast<C>(static_cast<S>(x));
}

// Layer cell of the E update, FDTD_PML.cpp:113-130.
template <typename S, typename C>
__device__ __forceinline__ void pml_E_cell(const Fields<S>& f, const Pml<S, C>& P, int Ni, int Nj,
                                           int Nk, int i, int j, int k) {
    const int plane = Ni * Nj;
    const int idx = i + j * Ni + k * plane;
    const int i_pred = wrap(i - 1, Ni) + j * Ni + k * plane;
    const int j_pred = i + wrap(j - 1, Nj) * Ni + k * plane;
    const int k_pred = i + j * Ni + wrap(k - 1, Nk) * plane;
    const C bx = f.b[0][idx], by = f.b[1][idx], bz = f.b[2][idx];
    const C ax = P.tx[i], ay = P.ty[j], az = P.tz[k];
    const C eyx = stored<S, C>(C(P.es_in[kEyx][idx]) * ax - P.tx[Ni + i] * (bz - C(f.b[2][i_pred])));
    const C ezx = stored<S, C>(C(P.es_in[kEzx][idx]) * ax + P.tx[2 * Ni + i] * (by - C(f.b[1][i_pred])));
    const C exy = stored<S, C>(C(P.es_in[kExy][idx]) * ay + P.ty[Nj + j] * (bz - C(f.b[2][j_pred])));
    const C ezy = stored<S, C>(C(P.es_in[kEzy][idx]) * ay - P.ty[2 * Nj + j] * (bx - C(f.b[0][j_pred])));
    const C exz = stored<S, C>(C(P.es_in[kExz][idx]) * az - P.tz[Nk + k] * (by - C(f.b[1][k_pred])));
    const C eyz = stored<S, C>(C(P.es_in[kEyz][idx]) * az + P.tz[2 * Nk + k] * (bx - C(f.b[0][k_pred])));
    P.es_out[kEyx][idx] = static_cast<S>(eyx);
    P.es_out[kEzx][idx] = static_cast<S>(ezx);
    P.es_out[kExy][idx] = static_cast<S>(exy);
    P.es_out[kEzy][idx] = static_cast<S>(ezy);
    P.es_out[kExz][idx] = static_cast<S>(exz);
    P.es_out[kEyz][idx] = static_cast<S>(eyz);
    f.e[0][idx] = static_cast<S>(exz + exy);
    f.e[1][idx] = static_cast<S>(eyx + eyz);
    f.e[2][idx] = static_cast<S>(ezy + ezx);
}

// Layer cell of the (full-step) B update, FDTD_PML.cpp:182-199.
template <typename S, typename C>
__device__ __forceinline__ void pml_B_cell(const Fields<S>& f, const Pml<S, C>& P, int Ni, int Nj,
                                           int Nk, int i, int j, int k) {
    const int plane = Ni * Nj;
    const int idx = i + j * Ni + k * plane;
    const int i_next = wrap(i + 1, Ni) + j * Ni + k * plane;
    const int j_next = i + wrap(j + 1, Nj) * Ni + k * plane;
    const int k_next = i + j * Ni + wrap(k + 1, Nk) * plane;
    const C ex = f.e[0][idx], ey = f.e[1][idx], ez = f.e[2][idx];
    const C ax = P.tx[i], ay = P.ty[j], az = P.tz[k];
    const C byx = stored<S, C>(C(P.bs[kByx][idx]) * ax + P.tx[3 * Ni + i] * (C(f.e[2][i_next]) - ez));
    const C bzx = stored<S, C>(C(P.bs[kBzx][idx]) * ax - P.tx[4 * Ni + i] * (C(f.e[1][i_next]) - ey));
    const C bxy = stored<S, C>(C(P.bs[kBxy][idx]) * ay - P.ty[3 * Nj + j] * (C(f.e[2][j_next]) - ez));
    const C bzy = stored<S, C>(C(P.bs[kBzy][idx]) * ay + P.ty[4 * Nj + j] * (C(f.e[0][j_next]) - ex));
    const C bxz = stored<S, C>(C(P.bs[kBxz][idx]) * az + P.tz[3 * Nk + k] * (C(f.e[1][k_next]) - ey));
    const C byz = stored<S, C>(C(P.bs[kByz][idx]) * az - P.tz[4 * Nk + k] * (C(f.e[0][k_next]) - ex));
    P.bs[kByx][idx] = static_cast<S>(byx);
    P.bs[kBzx][idx] = static_cast<S>(bzx);
    P.bs[kBxy][idx] = static_cast<S>(bxy);
    P.bs[kBzy][idx] = static_cast<S>(bzy);
    P.bs[kBxz][idx] = static_cast<S>(bxz);
    P.bs[kByz][idx] = static_cast<S>(byz);
    f.b[0][idx] = static_cast<S>(bxy + bxz);
    f.b[1][idx] = static_cast<S>(byz + byx);
    f.b[2][idx] = static_cast<S>(bzx + bzy);
}

// B half step on the main box and, with LAYER, the full-step split update
// on the layer (one thread per cell, in place): LAYER is the first B
// update of a PML step (FDTD_PML.cpp:344-352), !LAYER the last (:364).
template <typename S, typename C, bool LAYER>
__global__ void __launch_bounds__(256)
pml_plain_update_B(Fields<S> f, Pml<S, C> P, Coefs<C> c, int Ni, int Nj, int Nk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= Ni || j >= Nj) return;
    if (in_main(i, Ni, P.pi) && in_main(j, Nj, P.pj) && in_main(k, Nk, P.pk))
        plain_B_cell<S, C>(f, c, Ni, Nj, Nk, i, j, k);
    else if (LAYER)
        pml_B_cell<S, C>(f, P, Ni, Nj, Nk, i, j, k);
}

// E on the main box (with J) and the split update on the layer,
// FDTD_PML.cpp:354-362.
template <typename S, typename C, typename JT>
__global__ void __launch_bounds__(256)
pml_plain_update_E(Fields<S> f, Pml<S, C> P, const JT* __restrict__ jx,
                   const JT* __restrict__ jy, const JT* __restrict__ jz,
                   const PointSource* __restrict__ src, const double* __restrict__ vals, int nsrc,
                   Coefs<C> c, int Ni, int Nj, int Nk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= Ni || j >= Nj) return;
    if (in_main(i, Ni, P.pi) && in_main(j, Nj, P.pj) && in_main(k, Nk, P.pk))
        plain_E_cell<S, C, JT>(f, jx, jy, jz, src, vals, nsrc, c, Ni, Nj, Nk, i, j, k);
    else
        pml_E_cell<S, C>(f, P, Ni, Nj, Nk, i, j, k);
}

// ---------------------------------------------------------------------------
// Fused sweep: LDS-staged 2.5D z-march over full-row tiles.
//
// A workgroup owns `ty` consecutive full x-rows [j0, j0+ty) of planes
// [k0, k1) and marches k upward.  Thread (r, x) holds V consecutive cells
// i = x*V .. x*V+V-1 of tile row r; row r == ty is the halo row j0+ty, whose
// E_{n+1} the B update of row ty-1 needs and which the next tile also owns.
// Per plane kk the workgroup
//   1. loads E_n(kk), B(kk) (coalesced, V cells per lane) and stages the B
//      components needed as neighbours in LDS,
//   2. computes E_{n+1}(kk) (x/y neighbours from LDS or registers, the k-1
//      neighbour from registers) and stores it if kk is an owned plane,
//   3. computes B(kk-1) from E_{n+1}(kk-1) (LDS, registers) and
//      E_{n+1}(kk) (registers) and stores it,
//   4. stages E_{n+1}(kk) in LDS for the next plane.
// Inputs and outputs are different buffers (ping-pong), so halo rows and
// the wrapped planes are recomputed from unmodified data.
//
// LDS rows (each Ni values of type C), only what neighbours need:
//   Bz, Bx: rows j0-1 .. j0+ty-1   (y neighbour j-1; Bz also x neighbour)
//   By    : rows j0   .. j0+ty     (x neighbour i-1)
//   Ex    : rows j0+1 .. j0+ty     (y neighbour j+1)
//   Ey    : rows j0   .. j0+ty-1   (x neighbour i+1)
//   Ez    : rows j0   .. j0+ty     (x and y neighbours)
// => (6*TY + 4) * Ni * sizeof(C) bytes for a tile height TY.

struct SweepGeom {
    int Ni, Nj, Nk;
    int TY;       // rows per tile
    int kchunk;   // planes per z chunk
    int ntile_y;  // ceil(Nj / TY)
    int ntiles;   // ntile_y * nchunk
    int per_xcd;  // ceil(ntiles / 8): tiles handled by each XCD
};

__host__ __device__ inline size_t fused_lds_bytes(int TY, int Ni, int csize) {
    return static_cast<size_t>(6 * TY + 4) * Ni * csize;
}

// The PML sweep also stages the five x-axis layer tables (5 rows of Ni).
__host__ __device__ inline size_t fused_pml_lds_bytes(int TY, int Ni, int csize) {
    return fused_lds_bytes(TY, Ni, csize) + static_cast<size_t>(5) * Ni * csize;
}

// Largest workgroup of the fused sweep for V cells of size csize per lane:
// wide lanes hold more registers, so their workgroups are capped lower.
__host__ __device__ constexpr int fused_max_threads(int V, int csize) {
    return V * csize >= 32 ? 512 : (csize == 8 && V == 2 ? 768 : 1024);
}

template <typename S, typename C, typename JT>
struct SweepArgs {
    Fields<S> in;
    Fields<S> out;
    const JT* j[3];                 // dense J or null
    const PointSource* src;         // point sources (device)
    const double* vals;             // their J values for this step
    int nsrc;
    int pad;
    // When set, the sources are sorted by plane and those of plane k are
    // src[kstart[k]] .. src[kstart[k+1]-1] (Nk+1 entries).
    const int* kstart;
    Coefs<C> c;
    SweepGeom g;
    // FP16 storage only: stored values are multiplied by oscale (per
    // component Ex..Bz; exact powers of two, 1 outside rescaling sweeps), and
    // when maxbits is set each wavefront folds max |stored value| per
    // component into it (float bits, atomic max).
    C oscale[6];
    unsigned* maxbits;
    // Final sweeps of periodic grids: when set, also store the next run's
    // first B half step, (B_n + d) + d, so that run needs no B/2 prologue.
    S* ahead[3];
    Pml<S, C> pml;  // used by PML sweeps only
    ProbeArgs probe;  // values after this sweep's step (n = 0: none)
};

// Lane -> pack (V cells) order within a row of a PML sweep: the packs that
// hold lower-layer cells (i < pi), then those holding upper-layer cells
// (i >= Ni - pi), then the rest, so that in main-box rows the layer cells
// share as few wavefronts as possible and the other wavefronts never take
// the layer branch.  A bijection of [0, npacks); the identity when the two
// layer groups cover the row.  Neighbours across packs come from LDS rows,
// so any lane order gives the same results.
__device__ __forceinline__ int pml_pack(int lane, int npacks, int pi, int V) {
    const int nlo = (pi + V - 1) / V;
    const int first_hi = (npacks * V - pi) / V;
    const int nhi = npacks - first_hi;
    if (nlo + nhi >= npacks) return lane;
    if (lane < nlo) return lane;
    if (lane < nlo + nhi) return first_hi + (lane - nlo);
    return lane - nhi;
}

template <typename S, typename C, typename JT, int V, bool STEADY, bool PML = false,
          bool AUTO = false, bool PROBE = false>
__device__ __forceinline__ void fused_tile(const SweepArgs<S, C, JT>& a, int tile_y, int chunk,
                                           C* __restrict__ lds) {
    static_assert(!(PML && PROBE), "PML sweeps record no probes");
    const int Ni = a.g.Ni, Nj = a.g.Nj, Nk = a.g.Nk, TY = a.g.TY;
    const int lanes = Ni / V;
    const int j0 = tile_y * TY;
    const int ty = min(TY, Nj - j0);
    const int k0 = chunk * a.g.kchunk;
    const int k1 = min(k0 + a.g.kchunk, Nk);
    const int tid = threadIdx.x;
    const int r = tid / lanes;
    const int ib = (PML ? pml_pack(tid - r * lanes, lanes, a.pml.pi, V) : tid - r * lanes) * V;
    const bool active = r <= ty;
    const bool own = r < ty;  // row r == ty is the halo row
    const int j = (j0 + r >= Nj) ? j0 + r - Nj : j0 + r;  // halo row wraps
    const int jm = (j0 == 0) ? Nj - 1 : j0 - 1;
    const int plane = Ni * Nj;
    const int im0 = (ib == 0) ? Ni - 1 : ib - 1;        // x neighbour of element 0
    const int ipl = (ib + V == Ni) ? 0 : ib + V;        // x neighbour of element V-1
    const int row_off = j * Ni + ib;
    const int rowm_off = jm * Ni + ib;

    C* const sBz = lds;
    C* const sBx = sBz + (TY + 1) * Ni;
    C* const sBy = sBx + (TY + 1) * Ni;
    C* const sEx = sBy + (TY + 1) * Ni;
    C* const sEy = sEx + TY * Ni;
    C* const sEz = sEy + TY * Ni;

    const Coefs<C> c = a.c;

    // PML sweeps: the x tables in LDS after the field rows, and this row's y
    // table values in registers (both constant over the z march).
    C* const sTx = sEz + (TY + 1) * Ni;
    C tyv[5];
    if constexpr (PML) {
        for (int q = tid; q < 5 * Ni; q += blockDim.x) sTx[q] = a.pml.tx[q];
#pragma unroll
        for (int q = 0; q < 5; ++q) tyv[q] = active ? a.pml.ty[q * Nj + j] : C(0);
    }

    // FP16 storage: rescaling factors and the running max of stored values.
    constexpr bool kAuto = AUTO && sizeof(S) < sizeof(C);
    C mx[6] = {C(0), C(0), C(0), C(0), C(0), C(0)};
    auto put = [&](S* p, const C (&x)[V], int comp) {
        if constexpr (kAuto) {
            C o[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                o[v] = x[v] * a.oscale[comp];
                mx[comp] = fmax(mx[comp], fabs(o[v]));
            }
            store_pack<S, C, V>(p, o);
        } else {
            store_pack<S, C, V>(p, x);
        }
    };

    C eP[3][V];  // E_{n+1} at plane kk-1 (own cells)
    C bP[3][V];  // B at plane kk-1 (own cells)
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < V; ++v) { eP[d][v] = C(0); bP[d][v] = C(0); }

    // Plane loads, software-pipelined one plane ahead and kept in the
    // storage type until the rotation at the end of the iteration.
    // Halo-row threads do not need Ey; its slot carries Bz of row j0-1
    // instead, and bm carries Bx of row j0-1.
    using PK = RawPack<S, V>;
    C en[3][V], bn[3][V], bm[V];
    // The Ey slot is one load for every thread with a selected address (a
    // branch whose two sides load into the same registers would make the
    // halo row wait for the whole prefetch).
    auto load_plane = [&](int kw, PK (&e)[3], PK (&b)[3], PK& m) {
        const int off = kw * plane + row_off;
        const int offm = kw * plane + rowm_off;
        b[0] = load_raw<S, V>(a.in.b[0] + off);
        b[1] = load_raw<S, V>(a.in.b[1] + off);
        b[2] = load_raw<S, V>(a.in.b[2] + off);
        e[0] = load_raw<S, V>(a.in.e[0] + off);
        e[2] = load_raw<S, V>(a.in.e[2] + off);
        e[1] = load_raw<S, V>(own ? a.in.e[1] + off : a.in.b[2] + offm);
        if (!own) m = load_raw<S, V>(a.in.b[0] + offm);
    };

    if (active) {
        const int kp = (k0 == 0) ? Nk - 1 : k0 - 1;
        const int off = kp * plane + row_off;
        load_pack<S, C, V>(a.in.b[0] + off, bP[0]);
        load_pack<S, C, V>(a.in.b[1] + off, bP[1]);
        PK e0[3], b0[3], m0;
        load_plane(k0, e0, b0, m0);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            widen<S, C, V>(e0[d], en[d]);
            widen<S, C, V>(b0[d], bn[d]);
        }
        widen<S, C, V>(m0, bm);
    }

    // PML sweeps: the split fields of iteration kn (E splits of its plane,
    // B splits of plane kn-1 for steady sweeps), for lanes holding layer
    // cells, loaded one iteration ahead as raw words so that their latency
    // overlaps a whole iteration.  A chunk's B splits are updated only by its
    // own workgroup, one plane behind these loads.
    PK hsr[6], hbr[6];
    auto load_splits = [&](int kn) {
        if constexpr (PML) {
            const Pml<S, C>& P = a.pml;
            const int kwn = (kn == Nk) ? 0 : kn;
            const int offn = kwn * plane + row_off;
            const bool rowm = in_main(kwn, Nk, P.pk) && in_main(j, Nj, P.pj);
            if (active && !(rowm && ib >= P.pi && ib + V <= Ni - P.pi)) {
                hsr[kExy] = load_raw<S, V>(P.es_in[kExy] + offn);
                hsr[kExz] = load_raw<S, V>(P.es_in[kExz] + offn);
                hsr[kEzx] = load_raw<S, V>(P.es_in[kEzx] + offn);
                hsr[kEzy] = load_raw<S, V>(P.es_in[kEzy] + offn);
                if (own) {
                    hsr[kEyx] = load_raw<S, V>(P.es_in[kEyx] + offn);
                    hsr[kEyz] = load_raw<S, V>(P.es_in[kEyz] + offn);
                }
            }
            if (STEADY && own && kn > k0) {
                const int kb = kn - 1;
                const bool rowb = in_main(kb, Nk, P.pk) && in_main(j, Nj, P.pj);
                if (!(rowb && ib >= P.pi && ib + V <= Ni - P.pi)) {
                    const int offb = kb * plane + row_off;
#pragma unroll
                    for (int q = 0; q < 6; ++q) hbr[q] = load_raw<S, V>(P.bs[q] + offb);
                }
            }
        }
    };
    if constexpr (PML) load_splits(k0);

    for (int kk = k0; kk <= k1; ++kk) {
        const int kw = (kk == Nk) ? 0 : kk;
        const int off = kw * plane + row_off;
        if (active) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                if (own) {
                    sBz[(r + 1) * Ni + ib + v] = bn[2][v];
                    sBx[(r + 1) * Ni + ib + v] = bn[0][v];
                } else {
                    sBz[ib + v] = en[1][v];
                    sBx[ib + v] = bm[v];
                }
                sBy[r * Ni + ib + v] = bn[1][v];
            }
        }
        __syncthreads();

        PK en2[3], bn2[3], bm2;
        if (active && kk < k1) load_plane((kk + 1 == Nk) ? 0 : kk + 1, en2, bn2, bm2);
        // PML sweeps: the split fields this iteration updates (E splits of
        // plane kw, B splits of plane kk-1), prefetched by the previous one.
        C hs[6][V];
        C hb[6][V];
        if constexpr (PML) {
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                widen<S, C, V>(hsr[q], hs[q]);
                widen<S, C, V>(hbr[q], hb[q]);
            }
            if (kk < k1) load_splits(kk + 1);
        }

        C e1[3][V];
        C b_after[3][V];  // PROBE: B of plane kk-1 after this sweep's step
        if (active) {
            // First curl term of each E component, then the J term in front
            // of it as in the reference: (cur*J + first) - second.
            C t[3][V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const C by_im = (v > 0) ? bn[1][v - 1] : sBy[r * Ni + im0];
                t[0][v] = c.ex_bz * (bn[2][v] - sBz[r * Ni + ib + v]);
                t[1][v] = c.ey_bx * (bn[0][v] - bP[0][v]);
                t[2][v] = c.ez_by * (bn[1][v] - by_im);
            }
            if (a.j[0] != nullptr) {
                C jv[V];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    load_pack<JT, C, V>(a.j[d] + off, jv);
                    const C cj = d == 0 ? c.jx : (d == 1 ? c.jy : c.jz);
#pragma unroll
                    for (int v = 0; v < V; ++v) t[d][v] = cj * jv[v] + t[d][v];
                }
            }
            if (a.nsrc > 0) {
                int q0 = 0, q1 = a.nsrc;
                if (a.kstart != nullptr) {
                    const cptr<int> ks = as_const(a.kstart);
                    q0 = ks[kw];
                    q1 = ks[kw + 1];
                }
                add_point_sources<C, V>(t, c, a.src, a.vals, q0, q1, ib, j, kw);
            }
#pragma unroll
            for (int v = 0; v < V; ++v) {
                e1[0][v] = en[0][v] + sub_mul(t[0][v], c.ex_by, bn[1][v] - bP[1][v]);
                e1[2][v] = en[2][v] + sub_mul(t[2][v], c.ez_bx, bn[0][v] - sBx[r * Ni + ib + v]);
                if (own) {
                    const C bz_im = (v > 0) ? bn[2][v - 1] : sBz[(r + 1) * Ni + im0];
                    e1[1][v] = en[1][v] + sub_mul(t[1][v], c.ey_bz, bn[2][v] - bz_im);
                } else {
                    e1[1][v] = C(0);
                }
            }
            if constexpr (PML) {
                // Layer cells replace e1 by the sums of their updated E
                // splits; lanes whose V cells are all in the main box skip.
                const Pml<S, C>& P = a.pml;
                const bool rowm = in_main(kw, Nk, P.pk) && in_main(j, Nj, P.pj);
                if (!(rowm && ib >= P.pi && ib + V <= Ni - P.pi)) {
                    C (&sxy)[V] = hs[kExy];
                    C (&sxz)[V] = hs[kExz];
                    C (&syx)[V] = hs[kEyx];
                    C (&syz)[V] = hs[kEyz];
                    C (&szx)[V] = hs[kEzx];
                    C (&szy)[V] = hs[kEzy];
                    const C ay = tyv[0], cy_xy = tyv[1], cy_zy = tyv[2];
                    const C az = P.tz[kw], cz_xz = P.tz[Nk + kw], cz_yz = P.tz[2 * Nk + kw];
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        const int i = ib + v;
                        if (rowm && in_main(i, Ni, P.pi)) continue;
                        const C ax = sTx[i], cx_yx = sTx[Ni + i], cx_zx = sTx[2 * Ni + i];
                        const C by_im = (v > 0) ? bn[1][v - 1] : sBy[r * Ni + im0];
                        sxy[v] = stored<S, C>(sxy[v] * ay + cy_xy * (bn[2][v] - sBz[r * Ni + ib + v]));
                        szy[v] = stored<S, C>(szy[v] * ay - cy_zy * (bn[0][v] - sBx[r * Ni + ib + v]));
                        sxz[v] = stored<S, C>(sxz[v] * az - cz_xz * (bn[1][v] - bP[1][v]));
                        szx[v] = stored<S, C>(szx[v] * ax + cx_zx * (bn[1][v] - by_im));
                        e1[0][v] = sxz[v] + sxy[v];
                        e1[2][v] = szy[v] + szx[v];
                        if (own) {
                            const C bz_im = (v > 0) ? bn[2][v - 1] : sBz[(r + 1) * Ni + im0];
                            syx[v] = stored<S, C>(syx[v] * ax - cx_yx * (bn[2][v] - bz_im));
                            syz[v] = stored<S, C>(syz[v] * az + cz_yz * (bn[0][v] - bP[0][v]));
                            e1[1][v] = syx[v] + syz[v];
                        }
                    }
                    if (own && kk < k1) {
                        put(P.es_out[kExy] + off, sxy, 0);
                        put(P.es_out[kExz] + off, sxz, 0);
                        put(P.es_out[kEyx] + off, syx, 1);
                        put(P.es_out[kEyz] + off, syz, 1);
                        put(P.es_out[kEzx] + off, szx, 2);
                        put(P.es_out[kEzy] + off, szy, 2);
                    }
                }
            }
            if (own && kk < k1) {
#pragma unroll
                for (int d = 0; d < 3; ++d) put(a.out.e[d] + off, e1[d], d);
            }
            if (own && kk > k0) {
                const int offb = (kk - 1) * plane + row_off;
                C nb[3][V];
                C na[3][V];  // the carried half step (final sweeps with a.ahead)
                C b_half[3][V];  // PROBE, steady sweeps: B after this sweep's step
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const C ey_ip = (v < V - 1) ? eP[1][v + 1] : sEy[r * Ni + ipl];
                    const C ez_ip = (v < V - 1) ? eP[2][v + 1] : sEz[r * Ni + ipl];
                    const C ex_jp = sEx[r * Ni + ib + v];
                    const C ez_jp = sEz[(r + 1) * Ni + ib + v];
                    const C dbx = sub_mul(c.bx_ey * (e1[1][v] - eP[1][v]), c.bx_ez, ez_jp - eP[2][v]);
                    const C dby = sub_mul(c.by_ez * (ez_ip - eP[2][v]), c.by_ex, e1[0][v] - eP[0][v]);
                    const C dbz = sub_mul(c.bz_ex * (ex_jp - eP[0][v]), c.bz_ey, ey_ip - eP[1][v]);
                    nb[0][v] = bP[0][v] + dbx;
                    nb[1][v] = bP[1][v] + dby;
                    nb[2][v] = bP[2][v] + dbz;
                    if constexpr (PROBE && STEADY) {
                        b_half[0][v] = nb[0][v];
                        b_half[1][v] = nb[1][v];
                        b_half[2][v] = nb[2][v];
                    }
                    if (STEADY) {
                        nb[0][v] = nb[0][v] + dbx;
                        nb[1][v] = nb[1][v] + dby;
                        nb[2][v] = nb[2][v] + dbz;
                    } else if (!PML && a.ahead[0] != nullptr) {
                        na[0][v] = nb[0][v] + dbx;
                        na[1][v] = nb[1][v] + dby;
                        na[2][v] = nb[2][v] + dbz;
                    }
                }
                if (!STEADY && !PML && a.ahead[0] != nullptr) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) put(a.ahead[d] + offb, na[d], 3 + d);
                }
                if constexpr (PROBE) {
#pragma unroll
                    for (int d = 0; d < 3; ++d)
#pragma unroll
                        for (int v = 0; v < V; ++v) b_after[d][v] = STEADY ? b_half[d][v] : nb[d][v];
                }
                if constexpr (PML) {
                    // Layer cells of plane kk-1: the next step's full-step B
                    // split update (steady sweeps), or no B update at all
                    // (the final sweep of a run).
                    const Pml<S, C>& P = a.pml;
                    const int kb = kk - 1;
                    const bool rowm = in_main(kb, Nk, P.pk) && in_main(j, Nj, P.pj);
                    if (!(rowm && ib >= P.pi && ib + V <= Ni - P.pi)) {
                        if (STEADY) {
                            C (&sxy)[V] = hb[kBxy];
                            C (&sxz)[V] = hb[kBxz];
                            C (&syx)[V] = hb[kByx];
                            C (&syz)[V] = hb[kByz];
                            C (&szx)[V] = hb[kBzx];
                            C (&szy)[V] = hb[kBzy];
                            const C ay = tyv[0], cy_xy = tyv[3], cy_zy = tyv[4];
                            const C az = P.tz[kb], cz_xz = P.tz[3 * Nk + kb], cz_yz = P.tz[4 * Nk + kb];
#pragma unroll
                            for (int v = 0; v < V; ++v) {
                                const int i = ib + v;
                                if (rowm && in_main(i, Ni, P.pi)) continue;
                                const C ax = sTx[i], cx_yx = sTx[3 * Ni + i], cx_zx = sTx[4 * Ni + i];
                                const C ey_ip = (v < V - 1) ? eP[1][v + 1] : sEy[r * Ni + ipl];
                                const C ez_ip = (v < V - 1) ? eP[2][v + 1] : sEz[r * Ni + ipl];
                                const C ex_jp = sEx[r * Ni + ib + v];
                                const C ez_jp = sEz[(r + 1) * Ni + ib + v];
                                syx[v] = stored<S, C>(syx[v] * ax + cx_yx * (ez_ip - eP[2][v]));
                                szx[v] = stored<S, C>(szx[v] * ax - cx_zx * (ey_ip - eP[1][v]));
                                sxy[v] = stored<S, C>(sxy[v] * ay - cy_xy * (ez_jp - eP[2][v]));
                                szy[v] = stored<S, C>(szy[v] * ay + cy_zy * (ex_jp - eP[0][v]));
                                sxz[v] = stored<S, C>(sxz[v] * az + cz_xz * (e1[1][v] - eP[1][v]));
                                syz[v] = stored<S, C>(syz[v] * az - cz_yz * (e1[0][v] - eP[0][v]));
                                nb[0][v] = sxy[v] + sxz[v];
                                nb[1][v] = syz[v] + syx[v];
                                nb[2][v] = szx[v] + szy[v];
                            }
                            put(P.bs[kBxy] + offb, sxy, 3);
                            put(P.bs[kBxz] + offb, sxz, 3);
                            put(P.bs[kByx] + offb, syx, 4);
                            put(P.bs[kByz] + offb, syz, 4);
                            put(P.bs[kBzx] + offb, szx, 5);
                            put(P.bs[kBzy] + offb, szy, 5);
                        } else {
#pragma unroll
                            for (int v = 0; v < V; ++v) {
                                if (rowm && in_main(ib + v, Ni, P.pi)) continue;
                                nb[0][v] = bP[0][v];
                                nb[1][v] = bP[1][v];
                                nb[2][v] = bP[2][v];
                            }
                        }
                    }
                }
#pragma unroll
                for (int d = 0; d < 3; ++d) put(a.out.b[d] + offb, nb[d], 3 + d);
            }
            if constexpr (PROBE) {
                // probes: E of plane kw (this chunk's planes), B of plane kk-1
                if (own) {
                    bool all[V];
#pragma unroll
                    for (int v = 0; v < V; ++v) all[v] = true;
                    record_probes<S, C, V>(a.probe, a.probe.out, kw, ib, j, e1, b_after, kk < k1,
                                           kk > k0, all);
                }
            }
        }
        __syncthreads();

        if (active && kk < k1) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                if (r >= 1) sEx[(r - 1) * Ni + ib + v] = e1[0][v];
                if (own) sEy[r * Ni + ib + v] = e1[1][v];
                sEz[r * Ni + ib + v] = e1[2][v];
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                eP[d][v] = e1[d][v];
                bP[d][v] = bn[d][v];
            }
            widen<S, C, V>(en2[d], en[d]);
            widen<S, C, V>(bn2[d], bn[d]);
        }
        widen<S, C, V>(bm2, bm);
    }
    if constexpr (kAuto) {
        if (a.maxbits != nullptr) {
#pragma unroll
            for (int c6 = 0; c6 < 6; ++c6) {
                float m = mx[c6];
                for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
                if ((threadIdx.x & 63) == 0) atomicMax(a.maxbits + c6, __float_as_uint(m));
            }
        }
    }
}

// Tile t of a sweep is handled by the workgroup whose XCD (blockIdx % 8,
// the hardware's round-robin dispatch) owns the contiguous tile range
// [xcd*per_xcd, (xcd+1)*per_xcd): tiles sharing halo rows and z-chunk
// boundary planes then share one L2.
// AUTO (FP16 storage only): the variant that applies SweepArgs::oscale to
// what it stores and tracks its max; run only around rescaling points.
// PROBE: the variant that records SweepArgs::probe (the step's B is held in
// registers until then, which the plain variant has no room for).
template <typename S, typename C, typename JT, int V, bool STEADY, bool AUTO = false,
          bool PROBE = false>
__global__ void __launch_bounds__(fused_max_threads(V, sizeof(C)))
fused_sweep(SweepArgs<S, C, JT> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int b = blockIdx.x;
    const int t = (b & 7) * a.g.per_xcd + (b >> 3);
    if (t >= a.g.ntiles) return;
    const int tile_y = t % a.g.ntile_y;
    const int chunk = t / a.g.ntile_y;
    fused_tile<S, C, JT, V, STEADY, false, AUTO, PROBE>(a, tile_y, chunk,
                                                       reinterpret_cast<C*>(lds_raw));
}

// The fused sweep of a grid with a PML layer (same tiles and LDS staging;
// layer cells also read and write their split fields).  Halo rows and
// chunk-boundary planes recompute layer E from the unmodified es_in, so the
// E splits are ping-ponged like E and B; B splits are only touched by
// their owning tile and stay in place.  Lanes of 8 bytes or less are
// bounded at 768 threads (three wavefronts per SIMD, up to 168 VGPRs): at
// 1024 the FP32 V=2 sweep spilled.
__host__ __device__ constexpr int fused_pml_max_threads(int V, int csize) {
    return V * csize >= 16 ? 512 : 768;
}

template <typename S, typename C, typename JT, int V, bool STEADY, bool AUTO = false>
__global__ void __launch_bounds__(fused_pml_max_threads(V, sizeof(C)))
fused_pml_sweep(SweepArgs<S, C, JT> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int b = blockIdx.x;
    const int t = (b & 7) * a.g.per_xcd + (b >> 3);
    if (t >= a.g.ntiles) return;
    const int tile_y = t % a.g.ntile_y;
    const int chunk = t / a.g.ntile_y;
    fused_tile<S, C, JT, V, STEADY, true, AUTO>(a, tile_y, chunk, reinterpret_cast<C*>(lds_raw));
}

// ---------------------------------------------------------------------------
// Two-phase PML step: a PML grid's step as two in-place z-marching sweeps
// over full-row tiles, the E phase (E and the E splits from B) and the B
// phase (B and the B splits from E).  A phase reads only the field it does
// not write, so it needs no halo rows, recomputes nothing and updates in
// place; each keeps the per-lane state of one field only, which the fused
// PML sweep (both fields, their splits and a halo row) has no room for.
// The state between steps is the fused sweep's: (E_n, B after the next
// step's first B update), so both forms give the same values (with FP16
// storage the B phase reads E_{n+1} as stored, rounded, where the fused
// sweep uses it before rounding).
//   E phase, plane k: main cells E += curl(B) + J (kokkos_functors.h:64-90);
//     layer cells E splits decay and take one curl term each, E = their sum
//     (FDTD_PML.cpp:113-130).  B neighbours at i-1, j-1 from LDS rows, at
//     k-1 from the previous plane's registers.
//   B phase, plane k (marching down): main cells (B + d) + d, d the half-step
//     increment from E_{n+1} (the last half step of step n and the first of
//     step n+1), or B + d in a run's final phase; layer cells the full-step
//     split update (FDTD_PML.cpp:182-199), untouched in the final phase.  E
//     neighbours at i+1, j+1 from LDS rows, at k+1 from the previous plane.
// Per cell: main 18 words (E phase 9, B phase 9), layer 36 (18 + 18): the
// fused sweep's 12 and 36, with 6 more for main cells.

template <typename S, typename C, typename JT>
struct PmlPhaseArgs {
    Fields<S> f;                 // updated in place
    const JT* j[3];              // dense J or null (E phase)
    const PointSource* src;      // point sources sorted by plane (E phase)
    const double* vals;
    const int* kstart;           // sources of plane k: [kstart[k], kstart[k+1])
    int nsrc;
    int pad;
    Coefs<C> c;
    SweepGeom g;                 // TY rows per tile (no halo row)
    Pml<S, C> pml;               // es_in == es_out
};

__host__ __device__ constexpr int pml_phase_max_threads(int V, int csize) {
    return V * csize >= 16 ? 512 : 1024;
}

// LDS of a phase: three neighbour components of TY + 1 rows and three
// x-table rows.
__host__ __device__ inline size_t pml_phase_lds_bytes(int TY, int Ni, int csize) {
    return static_cast<size_t>(3 * TY + 5) * Ni * csize;
}

template <typename S, typename C, typename JT, int V>
__global__ void __launch_bounds__(pml_phase_max_threads(V, sizeof(C)))
pml_e_phase(PmlPhaseArgs<S, C, JT> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int blk = blockIdx.x;
    const int t = (blk & 7) * a.g.per_xcd + (blk >> 3);
    if (t >= a.g.ntiles) return;  // whole workgroup
    const int Ni = a.g.Ni, Nj = a.g.Nj, Nk = a.g.Nk, TY = a.g.TY;
    const int tile_y = t % a.g.ntile_y;
    const int chunk = t / a.g.ntile_y;
    const int lanes = Ni / V;
    const int j0 = tile_y * TY;
    const int ty = min(TY, Nj - j0);
    const int k0 = chunk * a.g.kchunk;
    const int k1 = min(k0 + a.g.kchunk, Nk);
    const int tid = threadIdx.x;
    const int r = tid / lanes;
    const Pml<S, C>& P = a.pml;
    const int ib = pml_pack(tid - r * lanes, lanes, P.pi, V) * V;
    const bool own = r < ty;
    const bool top = r == 0;                       // also loads row j0-1
    const int j = own ? j0 + r : j0;
    const int jt = (j0 == 0) ? Nj - 1 : j0 - 1;
    const int plane = Ni * Nj;
    const int im0 = (ib == 0) ? Ni - 1 : ib - 1;
    const int row_off = j * Ni + ib;
    const int top_off = jt * Ni + ib;
    const bool jmain = in_main(j, Nj, P.pj);
    const bool xmain = ib >= P.pi && ib + V <= Ni - P.pi;   // every cell main in x
    const bool xlayer = ib + V <= P.pi || ib >= Ni - P.pi;  // every cell layer in x

    C* const lds = reinterpret_cast<C*>(lds_raw);
    C* const sBz = lds;                  // rows q = 0..TY: y = j0 - 1 + q
    C* const sBx = sBz + (TY + 1) * Ni;  // rows q = 0..TY: y = j0 - 1 + q
    C* const sBy = sBx + (TY + 1) * Ni;  // rows q = 0..TY-1: y = j0 + q
    C* const sTx = sBy + TY * Ni;        // x tables: a, Eyx<-Bz, Ezx<-By
    for (int q = tid; q < 3 * Ni; q += blockDim.x) sTx[q] = P.tx[q];
    const C ay = P.ty[j], cy_xy = P.ty[Nj + j], cy_zy = P.ty[2 * Nj + j];
    const Coefs<C> c = a.c;

    using PK = RawPack<S, V>;
    PK bq[3], tq[2], eq[3], sq[6];
    // Plane k's loads, issued on every lane through plane buffer resources
    // (a lane that needs a load has its byte offset, the others kBufSkip,
    // which reads 0 without a memory access): B (own rows), Bz and Bx of
    // row j0-1 (top row), E (lanes holding main cells), E splits (lanes
    // holding layer cells).  With no branch around them the compiler counts
    // them, so the next plane's loads stay in flight across the plane's
    // work (conditional loads made it wait for all of them).
    const unsigned pbytes = static_cast<unsigned>(plane * sizeof(S));
    const unsigned own_b = own ? static_cast<unsigned>(row_off * sizeof(S)) : kBufSkip;
    const unsigned top_b = (own && top) ? static_cast<unsigned>(top_off * sizeof(S)) : kBufSkip;
    auto rs = [&](const S* base, int k) { return plane_rsrc(base + static_cast<size_t>(k) * plane, pbytes); };
    auto load = [&](int k) {
        const bool kmain = jmain && in_main(k, Nk, P.pk);
        const unsigned e_b = (kmain && !xlayer) ? own_b : kBufSkip;
        const unsigned s_b = !(kmain && xmain) ? own_b : kBufSkip;
#pragma unroll
        for (int d = 0; d < 3; ++d) bq[d] = buf_load_raw<S, V>(rs(a.f.b[d], k), own_b);
        tq[0] = buf_load_raw<S, V>(rs(a.f.b[2], k), top_b);
        tq[1] = buf_load_raw<S, V>(rs(a.f.b[0], k), top_b);
#pragma unroll
        for (int d = 0; d < 3; ++d) eq[d] = buf_load_raw<S, V>(rs(a.f.e[d], k), e_b);
#pragma unroll
        for (int q = 0; q < 6; ++q) sq[q] = buf_load_raw<S, V>(rs(P.es_in[q], k), s_b);
    };

    C bP[2][V];  // Bx, By at plane k-1 (own cells)
    if (own) {
        const int off = ((k0 == 0) ? Nk - 1 : k0 - 1) * plane + row_off;
        load_pack<S, C, V>(a.f.b[0] + off, bP[0]);
        load_pack<S, C, V>(a.f.b[1] + off, bP[1]);
    }
    load(k0);
    wait_vmem_all();  // the loop's waits then count its own stores only
    // the z tables through the scalar cache: a vector load inside the layer
    // branch would make the branch wait for the next plane's loads as well
    const cptr<C> tz = as_const(P.tz);
    for (int kk = k0; kk < k1; ++kk) {
        const bool kmain = jmain && in_main(kk, Nk, P.pk);
        const bool many = kmain && !xlayer, mall = kmain && xmain;
        const int off = kk * plane + row_off;
        const C az = tz[kk], cz_xz = tz[Nk + kk], cz_yz = tz[2 * Nk + kk];
        C bn[3][V], en[3][V], hs[6][V];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            widen<S, C, V>(bq[d], bn[d]);
            widen<S, C, V>(eq[d], en[d]);
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) widen<S, C, V>(sq[q], hs[q]);
        if (own) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                sBz[(r + 1) * Ni + ib + v] = bn[2][v];
                sBx[(r + 1) * Ni + ib + v] = bn[0][v];
                sBy[r * Ni + ib + v] = bn[1][v];
            }
            if (top) {
                C bt[2][V];
                widen<S, C, V>(tq[0], bt[0]);
                widen<S, C, V>(tq[1], bt[1]);
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    sBz[ib + v] = bt[0][v];
                    sBx[ib + v] = bt[1][v];
                }
            }
        }
        if (kk + 1 < k1) load(kk + 1);
        __syncthreads();
        C e1[3][V];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < V; ++v) e1[d][v] = en[d][v];
        if (own) {
            if (many) {
                C t[3][V];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const C by_im = (v > 0) ? bn[1][v - 1] : sBy[r * Ni + im0];
                    t[0][v] = c.ex_bz * (bn[2][v] - sBz[r * Ni + ib + v]);
                    t[1][v] = c.ey_bx * (bn[0][v] - bP[0][v]);
                    t[2][v] = c.ez_by * (bn[1][v] - by_im);
                }
                if (a.j[0] != nullptr) {
                    C jv[V];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        load_pack<JT, C, V>(a.j[d] + off, jv);
                        const C cj = d == 0 ? c.jx : (d == 1 ? c.jy : c.jz);
#pragma unroll
                        for (int v = 0; v < V; ++v) t[d][v] = cj * jv[v] + t[d][v];
                    }
                }
                if (a.nsrc > 0) {
                    const cptr<int> ks = as_const(a.kstart);
                    add_point_sources<C, V>(t, c, a.src, a.vals, ks[kk], ks[kk + 1], ib, j, kk);
                }
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const C bz_im = (v > 0) ? bn[2][v - 1] : sBz[(r + 1) * Ni + im0];
                    e1[0][v] = en[0][v] + sub_mul(t[0][v], c.ex_by, bn[1][v] - bP[1][v]);
                    e1[1][v] = en[1][v] + sub_mul(t[1][v], c.ey_bz, bn[2][v] - bz_im);
                    e1[2][v] = en[2][v] + sub_mul(t[2][v], c.ez_bx, bn[0][v] - sBx[r * Ni + ib + v]);
                }
            }
            if (!mall) {
                C (&sxy)[V] = hs[kExy];
                C (&sxz)[V] = hs[kExz];
                C (&syx)[V] = hs[kEyx];
                C (&syz)[V] = hs[kEyz];
                C (&szx)[V] = hs[kEzx];
                C (&szy)[V] = hs[kEzy];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const int i = ib + v;
                    if (kmain && in_main(i, Ni, P.pi)) continue;
                    const C ax = sTx[i], cx_yx = sTx[Ni + i], cx_zx = sTx[2 * Ni + i];
                    const C by_im = (v > 0) ? bn[1][v - 1] : sBy[r * Ni + im0];
                    const C bz_im = (v > 0) ? bn[2][v - 1] : sBz[(r + 1) * Ni + im0];
                    sxy[v] = stored<S, C>(sxy[v] * ay + cy_xy * (bn[2][v] - sBz[r * Ni + ib + v]));
                    szy[v] = stored<S, C>(szy[v] * ay - cy_zy * (bn[0][v] - sBx[r * Ni + ib + v]));
                    sxz[v] = stored<S, C>(sxz[v] * az - cz_xz * (bn[1][v] - bP[1][v]));
                    szx[v] = stored<S, C>(szx[v] * ax + cx_zx * (bn[1][v] - by_im));
                    syx[v] = stored<S, C>(syx[v] * ax - cx_yx * (bn[2][v] - bz_im));
                    syz[v] = stored<S, C>(syz[v] * az + cz_yz * (bn[0][v] - bP[0][v]));
                    e1[0][v] = sxz[v] + sxy[v];
                    e1[1][v] = syx[v] + syz[v];
                    e1[2][v] = szy[v] + szx[v];
                }
            }
        }
        // Stores on every lane (skipped where the offset is kBufSkip), after
        // a wait for the next plane's loads (issued before the barrier, so
        // mostly complete): the wait at the top of the next iteration then
        // does not also wait for these stores.
        wait_vmem_all();
        {
            const unsigned s_b = !mall ? own_b : kBufSkip;
#pragma unroll
            for (int q = 0; q < 6; ++q) buf_store_pack<S, C, V>(hs[q], rs(P.es_out[q], kk), s_b);
#pragma unroll
            for (int d = 0; d < 3; ++d) buf_store_pack<S, C, V>(e1[d], rs(a.f.e[d], kk), own_b);
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            bP[0][v] = bn[0][v];
            bP[1][v] = bn[1][v];
        }
        __syncthreads();
    }
}

template <typename S, typename C, typename JT, int V, bool STEADY>
__global__ void __launch_bounds__(pml_phase_max_threads(V, sizeof(C)))
pml_b_phase(PmlPhaseArgs<S, C, JT> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int blk = blockIdx.x;
    const int t = (blk & 7) * a.g.per_xcd + (blk >> 3);
    if (t >= a.g.ntiles) return;  // whole workgroup
    const int Ni = a.g.Ni, Nj = a.g.Nj, Nk = a.g.Nk, TY = a.g.TY;
    const int tile_y = t % a.g.ntile_y;
    const int chunk = t / a.g.ntile_y;
    const int lanes = Ni / V;
    const int j0 = tile_y * TY;
    const int ty = min(TY, Nj - j0);
    const int k0 = chunk * a.g.kchunk;
    const int k1 = min(k0 + a.g.kchunk, Nk);
    const int tid = threadIdx.x;
    const int r = tid / lanes;
    const Pml<S, C>& P = a.pml;
    const int ib = pml_pack(tid - r * lanes, lanes, P.pi, V) * V;
    const bool own = r < ty;
    const bool bottom = r == ty - 1;               // also loads row j0+ty
    const int j = own ? j0 + r : j0;
    const int jb = (j0 + ty == Nj) ? 0 : j0 + ty;
    const int plane = Ni * Nj;
    const int ipl = (ib + V == Ni) ? 0 : ib + V;
    const int row_off = j * Ni + ib;
    const int bot_off = jb * Ni + ib;
    const bool jmain = in_main(j, Nj, P.pj);
    const bool xmain = ib >= P.pi && ib + V <= Ni - P.pi;
    const bool xlayer = ib + V <= P.pi || ib >= Ni - P.pi;

    C* const lds = reinterpret_cast<C*>(lds_raw);
    C* const sEz = lds;                  // rows q = 0..TY: y = j0 + q (row ty: j0+ty, wrapped)
    C* const sEx = sEz + (TY + 1) * Ni;  // rows q = 0..TY
    C* const sEy = sEx + (TY + 1) * Ni;  // rows q = 0..TY-1
    C* const sTx = sEy + TY * Ni;        // x tables: a, Byx<-Ez, Bzx<-Ey
    for (int q = tid; q < Ni; q += blockDim.x) {
        sTx[q] = P.tx[q];
        sTx[Ni + q] = P.tx[3 * Ni + q];
        sTx[2 * Ni + q] = P.tx[4 * Ni + q];
    }
    const C ay = P.ty[j], cy_xy = P.ty[3 * Nj + j], cy_zy = P.ty[4 * Nj + j];
    const Coefs<C> c = a.c;

    using PK = RawPack<S, V>;
    PK eq[3], uq[2], bq[3], sq[6];
    // Plane k's loads, on every lane through plane buffer resources as in
    // the E phase: E (own rows), Ex and Ez of row j0+ty (bottom row), B
    // (lanes holding main cells), B splits (steady phases, lanes holding
    // layer cells).
    const unsigned pbytes = static_cast<unsigned>(plane * sizeof(S));
    const unsigned own_b = own ? static_cast<unsigned>(row_off * sizeof(S)) : kBufSkip;
    const unsigned bot_b = (own && bottom) ? static_cast<unsigned>(bot_off * sizeof(S)) : kBufSkip;
    auto rs = [&](const S* base, int k) { return plane_rsrc(base + static_cast<size_t>(k) * plane, pbytes); };
    auto load = [&](int k) {
        const bool kmain = jmain && in_main(k, Nk, P.pk);
        const unsigned b_b = (kmain && !xlayer) ? own_b : kBufSkip;
        const unsigned s_b = (STEADY && !(kmain && xmain)) ? own_b : kBufSkip;
#pragma unroll
        for (int d = 0; d < 3; ++d) eq[d] = buf_load_raw<S, V>(rs(a.f.e[d], k), own_b);
        uq[0] = buf_load_raw<S, V>(rs(a.f.e[0], k), bot_b);
        uq[1] = buf_load_raw<S, V>(rs(a.f.e[2], k), bot_b);
#pragma unroll
        for (int d = 0; d < 3; ++d) bq[d] = buf_load_raw<S, V>(rs(a.f.b[d], k), b_b);
        if constexpr (STEADY) {
#pragma unroll
            for (int q = 0; q < 6; ++q) sq[q] = buf_load_raw<S, V>(rs(P.bs[q], k), s_b);
        }
    };

    C eN[2][V];  // Ex, Ey at plane k+1 (own cells)
    if (own) {
        const int off = ((k1 == Nk) ? 0 : k1) * plane + row_off;
        load_pack<S, C, V>(a.f.e[0] + off, eN[0]);
        load_pack<S, C, V>(a.f.e[1] + off, eN[1]);
    }
    load(k1 - 1);
    wait_vmem_all();  // the loop's waits then count its own stores only
    const cptr<C> tz = as_const(P.tz);  // scalar loads, as in the E phase
    for (int kk = k1 - 1; kk >= k0; --kk) {
        const bool kmain = jmain && in_main(kk, Nk, P.pk);
        const bool many = kmain && !xlayer, mall = kmain && xmain;
        const C az = tz[kk], cz_xz = tz[3 * Nk + kk], cz_yz = tz[4 * Nk + kk];
        C en[3][V], bn[3][V], hb[6][V];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            widen<S, C, V>(eq[d], en[d]);
            widen<S, C, V>(bq[d], bn[d]);
        }
        if constexpr (STEADY) {
#pragma unroll
            for (int q = 0; q < 6; ++q) widen<S, C, V>(sq[q], hb[q]);
        }
        if (own) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                sEz[r * Ni + ib + v] = en[2][v];
                sEx[r * Ni + ib + v] = en[0][v];
                sEy[r * Ni + ib + v] = en[1][v];
            }
            if (bottom) {
                C eu[2][V];
                widen<S, C, V>(uq[0], eu[0]);
                widen<S, C, V>(uq[1], eu[1]);
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    sEx[ty * Ni + ib + v] = eu[0][v];
                    sEz[ty * Ni + ib + v] = eu[1][v];
                }
            }
        }
        if (kk > k0) load(kk - 1);
        __syncthreads();
        // nb: the B to store (main cells: the half step(s); layer cells:
        // the split sums, or unchanged in a final phase)
        C nb[3][V];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < V; ++v) nb[d][v] = bn[d][v];
        if (own) {
            if (many) {
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    if (!in_main(ib + v, Ni, P.pi)) continue;
                    const C ey_ip = (v < V - 1) ? en[1][v + 1] : sEy[r * Ni + ipl];
                    const C ez_ip = (v < V - 1) ? en[2][v + 1] : sEz[r * Ni + ipl];
                    const C ex_jp = sEx[(r + 1) * Ni + ib + v];
                    const C ez_jp = sEz[(r + 1) * Ni + ib + v];
                    const C dbx = sub_mul(c.bx_ey * (eN[1][v] - en[1][v]), c.bx_ez, ez_jp - en[2][v]);
                    const C dby = sub_mul(c.by_ez * (ez_ip - en[2][v]), c.by_ex, eN[0][v] - en[0][v]);
                    const C dbz = sub_mul(c.bz_ex * (ex_jp - en[0][v]), c.bz_ey, ey_ip - en[1][v]);
                    nb[0][v] = bn[0][v] + dbx;
                    nb[1][v] = bn[1][v] + dby;
                    nb[2][v] = bn[2][v] + dbz;
                    if (STEADY) {
                        nb[0][v] = nb[0][v] + dbx;
                        nb[1][v] = nb[1][v] + dby;
                        nb[2][v] = nb[2][v] + dbz;
                    }
                }
            }
            if constexpr (STEADY) {
                if (!mall) {
                    C (&sxy)[V] = hb[kBxy];
                    C (&sxz)[V] = hb[kBxz];
                    C (&syx)[V] = hb[kByx];
                    C (&syz)[V] = hb[kByz];
                    C (&szx)[V] = hb[kBzx];
                    C (&szy)[V] = hb[kBzy];
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        const int i = ib + v;
                        if (kmain && in_main(i, Ni, P.pi)) continue;
                        const C ax = sTx[i], cx_yx = sTx[Ni + i], cx_zx = sTx[2 * Ni + i];
                        const C ey_ip = (v < V - 1) ? en[1][v + 1] : sEy[r * Ni + ipl];
                        const C ez_ip = (v < V - 1) ? en[2][v + 1] : sEz[r * Ni + ipl];
                        const C ex_jp = sEx[(r + 1) * Ni + ib + v];
                        const C ez_jp = sEz[(r + 1) * Ni + ib + v];
                        syx[v] = stored<S, C>(syx[v] * ax + cx_yx * (ez_ip - en[2][v]));
                        szx[v] = stored<S, C>(szx[v] * ax - cx_zx * (ey_ip - en[1][v]));
                        sxy[v] = stored<S, C>(sxy[v] * ay - cy_xy * (ez_jp - en[2][v]));
                        szy[v] = stored<S, C>(szy[v] * ay + cy_zy * (ex_jp - en[0][v]));
                        sxz[v] = stored<S, C>(sxz[v] * az + cz_xz * (eN[1][v] - en[1][v]));
                        syz[v] = stored<S, C>(syz[v] * az - cz_yz * (eN[0][v] - en[0][v]));
                        nb[0][v] = sxy[v] + sxz[v];
                        nb[1][v] = syz[v] + syx[v];
                        nb[2][v] = szx[v] + szy[v];
                    }
                }
            }
        }
        // Stores on every lane (skipped where the offset is kBufSkip): B of
        // lanes with main cells, and in steady phases of lanes with layer
        // cells too, with their splits; after a wait for the next plane's
        // loads, as in the E phase.
        wait_vmem_all();
        {
            const unsigned b_b = (many || (STEADY && !mall)) ? own_b : kBufSkip;
            if constexpr (STEADY) {
                const unsigned s_b = !mall ? own_b : kBufSkip;
#pragma unroll
                for (int q = 0; q < 6; ++q) buf_store_pack<S, C, V>(hb[q], rs(P.bs[q], kk), s_b);
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) buf_store_pack<S, C, V>(nb[d], rs(a.f.b[d], kk), b_b);
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            eN[0][v] = en[0][v];
            eN[1][v] = en[1][v];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Temporally blocked sweep: L fused steps per pass over HBM.
//
// A workgroup covers a TXg x TYg patch of every plane (lanes x V columns,
// TYg rows) whose outer L columns and rows on each side are ghost cells:
// all L levels are computed on the whole patch and only the inner
// TX x TY = (TXg-2L) x (TYg-2L) cells, valid after L levels, are stored
// (overlapped tiling).  Along z the levels form a pipeline: at iteration kk
// level l computes E_l at plane kk-l and B_l at plane kk-l-1, so every
// level's k+1 neighbour is ready one iteration before it is needed.
//
//   E_l(p) = E_{l-1}(p) + curl B_{l-1}(p)        (E_{-1}, B_{-1}: the input)
//   B_l(p) = B_{l-1}(p) + d + d, d = -c dt/2 curl E_l(p)   (d once for the
//            last level of the last pass of a run)
//
// with the same per-cell expressions as fused_tile, so results are
// bit-identical to L single fused sweeps.  Level l applies the currents of
// its own step (source-value row l).  Tile origins are multiples of TX, so
// a lane's V cells are contiguous in memory (V divides Ni and 2L).
//
// LDS: for every level, the B input planes (x/y neighbours of the E update)
// and the E output planes (x/y neighbours of the B update): 2L x 3 planes
// of TXg x TYg values of type C.

struct BlockedGeom {
    int Ni, Nj, Nk;
    int TXg, TYg;     // patch covered by a workgroup, ghosts included
    int TX, TY;       // stored cells per tile: TXg - 2L, TYg - 2L
    int kchunk;
    int ntx, nty;     // tiles along x and y
    int ntiles;       // ntx * nty * nchunk
    int per_xcd;
};

__host__ __device__ inline size_t blocked_lds_bytes(int L, int TXg, int TYg, int csize) {
    return static_cast<size_t>(2 * L * 3) * TXg * TYg * csize;
}

// Largest workgroup of the blocked sweep: variants holding more state per
// lane (V*L*csize >= 24 bytes) get 512 threads, hence more registers.
__host__ __device__ constexpr int blocked_max_threads(int V, int L, int csize) {
    return V * L * csize >= 24 ? 512 : 1024;
}

template <typename S, typename C>
struct BlockedArgs {
    Fields<S> in;
    Fields<S> out;
    const PointSource* src;
    const double* vals;  // source values of the pass's first step; step l at vals + l*nsrc
    int nsrc;
    int pad;
    Coefs<C> c;
    BlockedGeom g;
};

__device__ __forceinline__ int wrap_mod(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

template <typename S, typename C, int V, int L, bool FINAL>
__device__ __forceinline__ void blocked_tile(const BlockedArgs<S, C>& a, int tx, int ty, int chunk,
                                             C* __restrict__ lds) {
    const BlockedGeom& G = a.g;
    const int Ni = G.Ni, Nj = G.Nj, Nk = G.Nk, TXg = G.TXg, TYg = G.TYg;
    const int lanes = TXg / V;
    const int tid = threadIdx.x;
    const int r = tid / lanes;                 // patch row
    const int cl = (tid - r * lanes) * V;      // patch column of the lane's first cell
    const bool active = r < TYg;
    const int px0 = tx * G.TX;                 // patch column 0 <-> global column px0 mod Ni
    const int gi0 = (px0 + cl) % Ni;           // global column of the first cell
    const int qrow = ty * G.TY + r - L;        // output row position of this patch row
    const int gj = wrap_mod(qrow, Nj);
    const bool row_out = active && r >= L && r < L + G.TY && qrow < Nj;
    bool col_out[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int c = cl + v;
        col_out[v] = c >= L && c < L + G.TX && px0 + c - L < Ni;
    }
    const int k0 = chunk * G.kchunk;
    const int k1 = min(k0 + G.kchunk, Nk);
    const int plane = Ni * Nj;
    const int row_off = gj * Ni + gi0;
    const int P = TXg * TYg;
    // neighbours inside the patch (clamped at its border: only ghost cells
    // read clamped values, and no stored cell depends on them)
    const int rm = r > 0 ? r - 1 : 0, rp = r + 1 < TYg ? r + 1 : TYg - 1;
    const int cm = cl > 0 ? cl - 1 : 0, cp = cl + V < TXg ? cl + V : TXg - 1;
    auto sB = [&](int l, int d) { return lds + (l * 6 + d) * P; };
    auto sE = [&](int l, int d) { return lds + (l * 6 + 3 + d) * P; };
    const Coefs<C> c = a.c;

    C bP[L][3][V];  // B input of level l at its previous plane
    C eP[L][3][V];  // E output of level l at the previous iteration
#pragma unroll
    for (int l = 0; l < L; ++l)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < V; ++v) { bP[l][d][v] = C(0); eP[l][d][v] = C(0); }

    const int kfirst = k0 - L + 1, klast = k1 + L - 1;
    C en[3][V], bn[3][V];
    auto load_plane = [&](int kk, C (&e)[3][V], C (&b)[3][V]) {
        const int off = wrap_mod(kk, Nk) * plane + row_off;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            load_pack<S, C, V>(a.in.b[d] + off, b[d]);
            load_pack<S, C, V>(a.in.e[d] + off, e[d]);
        }
    };
    if (active) {
        C e0[3][V];
        load_plane(kfirst - 1, e0, bP[0]);
        load_plane(kfirst, en, bn);
        // the E planes are read (for ghost-plane results only) before their
        // first write: start them at zero
#pragma unroll
        for (int l = 0; l < L; ++l)
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int v = 0; v < V; ++v) sE(l, d)[r * TXg + cl + v] = C(0);
    }

    for (int kk = kfirst; kk <= klast; ++kk) {
        if (active) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int v = 0; v < V; ++v) sB(0, d)[r * TXg + cl + v] = bn[d][v];
        }
        __syncthreads();
        C en2[3][V], bn2[3][V];
        if (active && kk < klast) load_plane(kk + 1, en2, bn2);

        C bin[3][V];   // B input of the current level at its E plane
        C ein[3][V];   // E input of the current level
        C enew[L][3][V];
        C bnew[3][V];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < V; ++v) { bin[d][v] = bn[d][v]; ein[d][v] = en[d][v]; }
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const int pe = kk - l;  // this level's E plane; its B plane is pe - 1
            C* Bz_ = sB(l, 2);
            C* Bx_ = sB(l, 0);
            C* By_ = sB(l, 1);
            if (active) {
                C t[3][V];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const C bz_jm = Bz_[rm * TXg + cl + v];
                    const C by_im = (v > 0) ? bin[1][v - 1] : By_[r * TXg + cm];
                    t[0][v] = c.ex_bz * (bin[2][v] - bz_jm);
                    t[1][v] = c.ey_bx * (bin[0][v] - bP[l][0][v]);
                    t[2][v] = c.ez_by * (bin[1][v] - by_im);
                }
                if (a.nsrc > 0)
                    add_point_sources<C, V>(t, c, a.src, a.vals + l * a.nsrc, 0, a.nsrc, gi0, gj,
                                            wrap_mod(pe, Nk));
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const C bx_jm = Bx_[rm * TXg + cl + v];
                    const C bz_im = (v > 0) ? bin[2][v - 1] : Bz_[r * TXg + cm];
                    enew[l][0][v] = ein[0][v] + sub_mul(t[0][v], c.ex_by, bin[1][v] - bP[l][1][v]);
                    enew[l][1][v] = ein[1][v] + sub_mul(t[1][v], c.ey_bz, bin[2][v] - bz_im);
                    enew[l][2][v] = ein[2][v] + sub_mul(t[2][v], c.ez_bx, bin[0][v] - bx_jm);
                }
                // B_l at plane pe-1 from E_l(pe-1) (eP, LDS) and E_l(pe) (enew)
                C* Ex_ = sE(l, 0);
                C* Ey_ = sE(l, 1);
                C* Ez_ = sE(l, 2);
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const C ey_ip = (v < V - 1) ? eP[l][1][v + 1] : Ey_[r * TXg + cp];
                    const C ez_ip = (v < V - 1) ? eP[l][2][v + 1] : Ez_[r * TXg + cp];
                    const C ex_jp = Ex_[rp * TXg + cl + v];
                    const C ez_jp = Ez_[rp * TXg + cl + v];
                    const C dbx = sub_mul(c.bx_ey * (enew[l][1][v] - eP[l][1][v]), c.bx_ez, ez_jp - eP[l][2][v]);
                    const C dby = sub_mul(c.by_ez * (ez_ip - eP[l][2][v]), c.by_ex, enew[l][0][v] - eP[l][0][v]);
                    const C dbz = sub_mul(c.bz_ex * (ex_jp - eP[l][0][v]), c.bz_ey, ey_ip - eP[l][1][v]);
                    bnew[0][v] = bP[l][0][v] + dbx;
                    bnew[1][v] = bP[l][1][v] + dby;
                    bnew[2][v] = bP[l][2][v] + dbz;
                    if (!(FINAL && l == L - 1)) {
                        bnew[0][v] = bnew[0][v] + dbx;
                        bnew[1][v] = bnew[1][v] + dby;
                        bnew[2][v] = bnew[2][v] + dbz;
                    }
#pragma unroll
                    for (int d = 0; d < 3; ++d) bnew[d][v] = as_stored<S>(bnew[d][v]);
                }
            }
            if (l < L - 1) {
                // bnew = B_l(pe-1) is the next level's B input at its E plane
                if (active) {
#pragma unroll
                    for (int d = 0; d < 3; ++d)
#pragma unroll
                        for (int v = 0; v < V; ++v) sB(l + 1, d)[r * TXg + cl + v] = bnew[d][v];
                }
                __syncthreads();
#pragma unroll
                for (int d = 0; d < 3; ++d)
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        // level l+1 inputs: E_l(pe-1) = eP[l] (previous iteration) as
                        // it is stored, B_l(pe-1)
                        ein[d][v] = as_stored<S>(eP[l][d][v]);
                        bP[l][d][v] = bin[d][v];   // level l's B input at pe: next iteration's previous plane
                        bin[d][v] = bnew[d][v];
                    }
            } else {
#pragma unroll
                for (int d = 0; d < 3; ++d)
#pragma unroll
                    for (int v = 0; v < V; ++v) bP[l][d][v] = bin[d][v];
            }
        }
        // stores of the last level: E_{L-1} at plane kk-L+1, B_{L-1} at kk-L
        if (row_out) {
            const int pe = kk - L + 1;
            if (pe >= k0 && pe < k1) {
                const int off = pe * plane + row_off;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    S* dst = a.out.e[d] + off;
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        if (col_out[v])
                            __builtin_nontemporal_store(static_cast<S>(enew[L - 1][d][v]), dst + v);
                }
            }
            if (pe - 1 >= k0 && pe - 1 < k1) {
                const int off = (pe - 1) * plane + row_off;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    S* dst = a.out.b[d] + off;
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        if (col_out[v]) __builtin_nontemporal_store(static_cast<S>(bnew[d][v]), dst + v);
                }
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int l = 0; l < L; ++l)
#pragma unroll
                for (int d = 0; d < 3; ++d)
#pragma unroll
                    for (int v = 0; v < V; ++v) sE(l, d)[r * TXg + cl + v] = enew[l][d][v];
        }
#pragma unroll
        for (int l = 0; l < L; ++l)
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int v = 0; v < V; ++v) eP[l][d][v] = enew[l][d][v];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < V; ++v) {
                en[d][v] = en2[d][v];
                bn[d][v] = bn2[d][v];
            }
    }
}

template <typename S, typename C, int V, int L, bool FINAL>
__global__ void __launch_bounds__(blocked_max_threads(V, L, sizeof(C)))
blocked_sweep(BlockedArgs<S, C> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int b = blockIdx.x;
    const int t = (b & 7) * a.g.per_xcd + (b >> 3);
    if (t >= a.g.ntiles) return;
    const int txy = a.g.ntx * a.g.nty;
    const int chunk = t / txy;
    const int rem = t - chunk * txy;
    const int ty = rem / a.g.ntx;
    const int tx = rem - ty * a.g.ntx;
    blocked_tile<S, C, V, L, FINAL>(a, tx, ty, chunk, reinterpret_cast<C*>(lds_raw));
}

// ---------------------------------------------------------------------------
// Wavefront-strip temporally blocked sweep: L fused steps per pass over HBM.
//
// A workgroup of W wavefronts covers a tile of 64*V columns by W*R rows of
// every plane.  Lane l of wavefront w owns columns l*V .. l*V+V-1 of tile
// rows w*R .. w*R+R-1 and keeps them in registers, so of the x and y
// neighbours a cell update reads:
//   - x: the neighbouring lane's cell, moved by a DPP wavefront shift
//     (wave_shr:1 / wave_shl:1) -- no LDS;
//   - y: the lane's own next/previous row, except across a wavefront's
//     first and last row, whose two needed components (Bx, Bz below; Ex, Ez
//     above) go through a small LDS row exchange.
// The outer Gx columns (L rounded up to a multiple of V) and L rows on each
// side are ghost cells: computed on every level, never stored (overlapped
// tiling), since after L levels only cells at least L from the tile edge
// hold the values of L unblocked steps.  Along z the levels form the same
// pipeline as blocked_sweep: at iteration kk level l computes E_l at plane
// kk-l and B_l at plane kk-l-1, with the per-cell expressions of
// fused_tile, so results are bit-identical to L single fused sweeps.
//
// LDS (per level l): sB[l] holds each wavefront's last-row Bx, Bz input of
// level l (published before level l runs); sE[l] holds each wavefront's
// first-row Ex, Ez of E_l one plane below (published one phase after the
// reads of the previous value).  One barrier per level and plane (two for
// L = 1).

// p advanced by a 32-bit unsigned byte offset (zero-extended): lets the
// compiler address memory as a scalar base plus a 32-bit lane offset.
template <typename T>
__device__ __forceinline__ T* at(T* p, unsigned byte_off) {
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T*>(reinterpret_cast<B*>(p) + byte_off);
}

// Global-memory pointers built from wavefront-uniform addresses passed
// through readfirstlane: the compiler keeps each plane's base in scalar
// registers and addresses a row as that base plus a 32-bit lane offset,
// instead of turning per-plane addresses into 64-bit per-lane pointers
// stepped through the loop (two VGPRs per row and component).
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

template <typename T>
__device__ __forceinline__ gptr<T> uniform_gptr(T* p) {
    const unsigned long long u = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(u));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(u >> 32));
    return reinterpret_cast<gptr<T>>((static_cast<unsigned long long>(hi) << 32) | lo);
}

template <typename T>
__device__ __forceinline__ gptr<T> at(gptr<T> p, unsigned byte_off) {
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<gptr<T>>(reinterpret_cast<gptr<B>>(p) + byte_off);
}

// Back to a generic pointer for the access (an address space cast the
// compiler folds, keeping the global access).
template <typename T>
__device__ __forceinline__ T* generic(gptr<T> p) {
    return (T*)p;
}

struct StripGeom {
    int Ni, Nj, Nk;
    int W;            // wavefronts per tile (R rows each)
    int Gx;           // ghost columns per side: L rounded up to a multiple of V
    int TX, TY;       // stored columns and rows per tile: 64V - 2Gx, W*R - 2L
    int kchunk;
    int ntx, nty;
    int ntiles;
    int per_xcd;
};

__host__ __device__ inline size_t strip_lds_bytes(int L, int W, int V, int csize) {
    return static_cast<size_t>(4) * L * W * 64 * V * csize;
}

// Workgroup size bound of a strip variant: lanes holding more state get
// fewer threads per workgroup, hence more registers.
__host__ __device__ constexpr int strip_max_threads(int V, int R, int L, int csize) {
    return (12 * L + 24) * V * R * csize / 4 > 144 ? 512 : 1024;
}

template <typename S, typename C>
struct StripArgs {
    Fields<S> in;
    Fields<S> out;
    const PointSource* src;   // sorted by plane
    const int* kstart;        // sources of plane k: [kstart[k], kstart[k+1])
    const double* vals;       // source values of the pass's first step; step l at vals + l*vstride
    int nsrc;
    int vstride;              // value-table row length (nsrc for one grid)
    S* ahead[3];              // FINAL passes: also store the next half step (B + d) + d here
    Coefs<C> c;
    StripGeom g;
    // FP16 storage, RESCALE passes: every value read (E: iscale[0..2],
    // B: iscale[3..5], exact powers of two) is scaled and rounded to FP16
    // before use -- the same values as rescaling the state in place first
    float iscale[6];
    ProbeArgs probe;  // values after each of the pass's L steps (n = 0: none)
};

// Value of the same register in lane-1 (lane_prev) or lane+1 (lane_next) of
// the wavefront; the end lane (0, resp. 63) reads 0.
__device__ __forceinline__ int dpp_prev(int x) {
    return __builtin_amdgcn_mov_dpp(x, 0x138 /* wave_shr:1 */, 0xF, 0xF, true);
}
__device__ __forceinline__ int dpp_next(int x) {
    return __builtin_amdgcn_mov_dpp(x, 0x130 /* wave_shl:1 */, 0xF, 0xF, true);
}
__device__ __forceinline__ float lane_prev(float x) {
    return __builtin_bit_cast(float, dpp_prev(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ float lane_next(float x) {
    return __builtin_bit_cast(float, dpp_next(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double lane_prev(double x) {
    int2 h = __builtin_bit_cast(int2, x);
    h.x = dpp_prev(h.x);
    h.y = dpp_prev(h.y);
    return __builtin_bit_cast(double, h);
}
__device__ __forceinline__ double lane_next(double x) {
    int2 h = __builtin_bit_cast(int2, x);
    h.x = dpp_next(h.x);
    h.y = dpp_next(h.y);
    return __builtin_bit_cast(double, h);
}

// One tile (index < a.g.ntiles) of a strip pass; every thread of the
// workgroup calls it with the same tile.  RESCALE (FP16 storage only): the
// pass after an in-run rescaling decision, which scales what it reads.
template <typename S, typename C, int V, int R, int L, bool FINAL, bool RESCALE = false,
          bool PROBE = false>
__device__ __forceinline__ void strip_tile(const StripArgs<S, C>& a, const int tile,
                                           C* __restrict__ lds) {
    const StripGeom& G = a.g;
    const int txy = G.ntx * G.nty;
    const int chunk = tile / txy;
    const int rem = tile - chunk * txy;
    const int ty = rem / G.ntx;
    const int tx = rem - ty * G.ntx;

    constexpr int RW = 64 * V;
    const int Ni = G.Ni, Nj = G.Nj, Nk = G.Nk, W = G.W;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int cx0 = lane * V;                          // tile column of the lane's first cell
    const int px0 = tx * G.TX;                         // global column of the first stored column
    const int gi0 = wrap_mod(px0 - G.Gx + cx0, Ni);    // V | Gx, V | TX: packs never straddle the wrap
    const int py0 = ty * G.TY;
    // Row offsets within a plane are unsigned 32-bit byte offsets (a plane
    // is below 2 GiB), so every access is a uniform (scalar) plane base plus
    // one 32-bit lane offset per row.
    int gj[R];
    unsigned row_off[R];
    bool row_out[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int tr = w * R + r;
        gj[r] = wrap_mod(py0 - L + tr, Nj);
        row_off[r] = static_cast<unsigned>(gj[r] * Ni + gi0) * static_cast<unsigned>(sizeof(S));
        row_out[r] = tr >= L && tr < L + G.TY && py0 + tr - L < Nj;
    }
    bool col_out[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int cc = cx0 + v;
        col_out[v] = cc >= G.Gx && cc < G.Gx + G.TX && px0 + cc - G.Gx < Ni;
    }
    const int k0 = chunk * G.kchunk;
    const int k1 = min(k0 + G.kchunk, Nk);
    const int plane = Ni * Nj;
    // Memory path.  FP64 and FP16 storage prefetch and store through plane
    // buffer resources, issued on every path, so that waiting for the
    // prefetch does not wait for the stores issued after it (measured:
    // FP64 passes about 11%, FP16 passes 12-16% faster than with global
    // accesses).
    // FP32 keeps global accesses under branches, whose waits do include the
    // previous plane's stores: its buffer form measured 4-5% slower.
    constexpr bool kBuf = sizeof(S) != 4;
    const unsigned pbytes = static_cast<unsigned>(plane) * static_cast<unsigned>(sizeof(S));
    // store offsets: kBufSkip outside the stored rows and columns
    unsigned st_off[R][V];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int v = 0; v < V; ++v)
            st_off[r][v] = kBuf && row_out[r] && col_out[v]
                               ? row_off[r] + static_cast<unsigned>(v * sizeof(S)) : kBufSkip;
    // wavefront whose last row lies below / first row above this one's
    // (the end wavefronts read their own: those rows are ghost rows)
    const int wm = w > 0 ? w - 1 : w, wp = w + 1 < W ? w + 1 : w;
    // LDS rows grouped by wavefront, so that the level and component parts
    // of an address are compile-time offsets from one of three bases (this,
    // the previous and the next wavefront's block)
    using lds_c = __attribute__((address_space(3))) C*;
    auto sB = [&](int l, int c, int ww) {
        return (lds_c)lds + ww * (4 * L * RW) + (l * 2 + c) * RW + cx0;
    };
    auto sE = [&](int l, int c, int ww) {
        return (lds_c)lds + ww * (4 * L * RW) + ((L + l) * 2 + c) * RW + cx0;
    };
    const Coefs<C> c = a.c;
    const cptr<int> kst = as_const(a.kstart);

    C bP[L][3][R][V];  // level l: B input one plane below its E plane
    C eP[L][3][R][V];  // level l: E output of the previous iteration
#pragma unroll
    for (int l = 0; l < L; ++l)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    bP[l][d][r][v] = C(0);
                    eP[l][d][r][v] = C(0);
                }

    const int kfirst = k0 - L + 1, klast = k1 + L - 1;
    C en[3][R][V], bn[3][R][V];
    {
        const size_t offm = static_cast<size_t>(wrap_mod(kfirst - 1, Nk)) * plane;
        const size_t off = static_cast<size_t>(wrap_mod(kfirst, Nk)) * plane;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const gptr<const S> bm = uniform_gptr<const S>(a.in.b[d] + offm);
            const gptr<const S> b0 = uniform_gptr<const S>(a.in.b[d] + off);
            const gptr<const S> e0 = uniform_gptr<const S>(a.in.e[d] + off);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                load_pack<S, C, V>(generic(at(bm, row_off[r])), bP[0][d][r]);
                load_pack<S, C, V>(generic(at(b0, row_off[r])), bn[d][r]);
                load_pack<S, C, V>(generic(at(e0, row_off[r])), en[d][r]);
            }
        }
    }
    if constexpr (RESCALE) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    bP[0][d][r][v] = as_stored<S>(bP[0][d][r][v] * a.iscale[3 + d]);
                    bn[d][r][v] = as_stored<S>(bn[d][r][v] * a.iscale[3 + d]);
                    en[d][r][v] = as_stored<S>(en[d][r][v] * a.iscale[d]);
                }
    }
    // E rows are read (for ghost-plane results only) before their first write
#pragma unroll
    for (int l = 0; l < L; ++l)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int v = 0; v < V; ++v) sE(l, q, w)[v] = C(0);
    // level 0's B input row for the wavefront below: the buffer path
    // publishes it at the end of the previous iteration (after the barriers
    // that follow its last reads), once the prefetch has been waited for
    if constexpr (kBuf) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            sB(0, 0, w)[v] = bn[0][R - 1][v];
            sB(0, 1, w)[v] = bn[2][R - 1][v];
        }
    }
    // the first plane's loads have completed, so that the loop's waits are
    // for its own prefetches (and, in the global path, not for its stores)
    wait_vmem_all();

    // One plane iteration: level l reads its state of the previous
    // iteration from (ePi, bPi) and leaves the new one in (ePo, bPo) -- the
    // same arrays here: each element is read before it is overwritten.
    // (Alternating two state sets over an unrolled pair of iterations
    // removes the register copies of the pipeline rotation, but measured
    // slower: FP32 L=4 spills at 128 VGPRs, FP16 -5%, FP64 +3%.)
    auto plane_step = [&](const int kk, const C (&ePi)[L][3][R][V], const C (&bPi)[L][3][R][V],
                          C (&ePo)[L][3][R][V], C (&bPo)[L][3][R][V])
                          __attribute__((always_inline)) {
        // lane offsets of this iteration's global accesses, redefined each
        // iteration (an empty asm) so that the compiler forms every address
        // as a scalar plane base plus a 32-bit lane offset instead of
        // hoisting 64-bit per-lane sums out of the loop
        unsigned lo[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            lo[r] = row_off[r];
            asm volatile("" : "+v"(lo[r]));
        }
        if constexpr (!kBuf) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                sB(0, 0, w)[v] = bn[0][R - 1][v];
                sB(0, 1, w)[v] = bn[2][R - 1][v];
            }
        }
        // prefetch of plane kk+1 (buffer path: a size-0 resource after the
        // last plane)
        RawPack<S, V> rawE[3][R], rawB[3][R];
        const bool more = kk < klast;
        if constexpr (!kBuf) {
            if (more) {
                const size_t off = static_cast<size_t>(wrap_mod(kk + 1, Nk)) * plane;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const gptr<const S> bb = uniform_gptr<const S>(a.in.b[d] + off);
                    const gptr<const S> eb = uniform_gptr<const S>(a.in.e[d] + off);
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        rawB[d][r] = load_raw<S, V>(generic(at(bb, lo[r])));
                        rawE[d][r] = load_raw<S, V>(generic(at(eb, lo[r])));
                    }
                }
            }
        } else {
            const unsigned nb = more ? pbytes : 0u;
            const size_t off = static_cast<size_t>(wrap_mod(kk + 1, Nk)) * plane;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const __amdgpu_buffer_rsrc_t bb = plane_rsrc(a.in.b[d] + off, nb);
                const __amdgpu_buffer_rsrc_t eb = plane_rsrc(a.in.e[d] + off, nb);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    rawB[d][r] = buf_load_raw<S, V>(bb, row_off[r]);
                    rawE[d][r] = buf_load_raw<S, V>(eb, row_off[r]);
                }
            }
        }
        __syncthreads();

        C bin[3][R][V];  // B input of the current level at its E plane
        C ein[3][R][V];  // E input of the current level
        // FINAL passes with a carry buffer also store the next half step
        // (B + d) + d of the last level's B plane kk-L, as it is computed
        const bool carry_plane = FINAL && a.ahead[0] != nullptr && kk - L >= k0 && kk - L < k1;
        const size_t carry_off = static_cast<size_t>(carry_plane ? kk - L : k0) * plane;
        C cy[3][R][V];  // global path, FINAL: the carried half step of the last level
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    bin[d][r][v] = bn[d][r][v];
                    ein[d][r][v] = en[d][r][v];
                }
#pragma unroll
        for (int l = 0; l < L; ++l) {
            // this level's E plane is kk - l (pkl wrapped); its B plane is kk - l - 1
            if (L > 1) {
                // publish the first-row Ex, Ez of the level whose readers
                // have all passed the last barrier
                const int lp = l == 0 ? L - 1 : l - 1;
                const C (&prevE)[3][R][V] = l == 0 ? ePi[L - 1] : ePo[(l + L - 1) % L];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    sE(lp, 0, w)[v] = prevE[0][0][v];
                    sE(lp, 1, w)[v] = prevE[2][0][v];
                }
            }
            // point sources of plane pkl: [kstart[pkl], kstart[pkl + 1])
            const int pkl = wrap_mod(kk - l, Nk);
            int q0 = 0, q1 = 0;
            if (a.nsrc > 0) {
                q0 = kst[pkl];
                q1 = kst[pkl + 1];
            }
            C enew[3][R][V];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                C t[3][V];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const C bz_jm = r > 0 ? bin[2][r - 1][v] : sB(l, 1, wm)[v];
                    const C by_im = v > 0 ? bin[1][r][v - 1] : lane_prev(bin[1][r][V - 1]);
                    t[0][v] = c.ex_bz * (bin[2][r][v] - bz_jm);
                    t[1][v] = c.ey_bx * (bin[0][r][v] - bPi[l][0][r][v]);
                    t[2][v] = c.ez_by * (bin[1][r][v] - by_im);
                }
                if (q1 > q0)
                    add_point_sources<C, V>(t, c, a.src, a.vals + l * a.vstride, q0, q1, gi0, gj[r],
                                            pkl);
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const C bx_jm = r > 0 ? bin[0][r - 1][v] : sB(l, 0, wm)[v];
                    const C bz_im = v > 0 ? bin[2][r][v - 1] : lane_prev(bin[2][r][V - 1]);
                    enew[0][r][v] = ein[0][r][v] + sub_mul(t[0][v], c.ex_by, bin[1][r][v] - bPi[l][1][r][v]);
                    enew[1][r][v] = ein[1][r][v] + sub_mul(t[1][v], c.ey_bz, bin[2][r][v] - bz_im);
                    enew[2][r][v] = ein[2][r][v] + sub_mul(t[2][v], c.ez_bx, bin[0][r][v] - bx_jm);
                }
            }
            // B_l at plane pe-1 from E_l(pe-1) (eP) and E_l(pe) (enew)
            C bnew[3][R][V];
            C bh[3][R][V];  // B after step l of the pass (the first half step's sum)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const C ey_ip = v < V - 1 ? ePi[l][1][r][v + 1] : lane_next(ePi[l][1][r][0]);
                    const C ez_ip = v < V - 1 ? ePi[l][2][r][v + 1] : lane_next(ePi[l][2][r][0]);
                    const C ex_jp = r < R - 1 ? ePi[l][0][r + 1][v] : sE(l, 0, wp)[v];
                    const C ez_jp = r < R - 1 ? ePi[l][2][r + 1][v] : sE(l, 1, wp)[v];
                    const C dbx = sub_mul(c.bx_ey * (enew[1][r][v] - ePi[l][1][r][v]), c.bx_ez, ez_jp - ePi[l][2][r][v]);
                    const C dby = sub_mul(c.by_ez * (ez_ip - ePi[l][2][r][v]), c.by_ex, enew[0][r][v] - ePi[l][0][r][v]);
                    const C dbz = sub_mul(c.bz_ex * (ex_jp - ePi[l][0][r][v]), c.bz_ey, ey_ip - ePi[l][1][r][v]);
                    bnew[0][r][v] = bPi[l][0][r][v] + dbx;
                    bnew[1][r][v] = bPi[l][1][r][v] + dby;
                    bnew[2][r][v] = bPi[l][2][r][v] + dbz;
                    bh[0][r][v] = bnew[0][r][v];
                    bh[1][r][v] = bnew[1][r][v];
                    bh[2][r][v] = bnew[2][r][v];
                    if (!(FINAL && l == L - 1)) {
                        bnew[0][r][v] = bnew[0][r][v] + dbx;
                        bnew[1][r][v] = bnew[1][r][v] + dby;
                        bnew[2][r][v] = bnew[2][r][v] + dbz;
                    } else if (kBuf && a.ahead[0] != nullptr) {
                        const unsigned nb = carry_plane ? pbytes : 0u;
                        buf_store_nt(static_cast<S>(bnew[0][r][v] + dbx), plane_rsrc(a.ahead[0] + carry_off, nb), st_off[r][v]);
                        buf_store_nt(static_cast<S>(bnew[1][r][v] + dby), plane_rsrc(a.ahead[1] + carry_off, nb), st_off[r][v]);
                        buf_store_nt(static_cast<S>(bnew[2][r][v] + dbz), plane_rsrc(a.ahead[2] + carry_off, nb), st_off[r][v]);
                    } else if (!kBuf) {
                        // stored with the level's B plane below, after the
                        // wait for the prefetch
                        cy[0][r][v] = bnew[0][r][v] + dbx;
                        cy[1][r][v] = bnew[1][r][v] + dby;
                        cy[2][r][v] = bnew[2][r][v] + dbz;
                    }
#pragma unroll
                    for (int d = 0; d < 3; ++d) bnew[d][r][v] = as_stored<S>(bnew[d][r][v]);
                }
            if constexpr (PROBE) {
                // probes of step l: E of plane pkl, B of plane pkl-1, cells
                // this tile stores
                const int pe = kk - l;
                const bool e_ok = pe >= k0 && pe < k1, b_ok = pe - 1 >= k0 && pe - 1 < k1;
                double* const pout = a.probe.out + static_cast<size_t>(l) * a.probe.stride;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    C er[3][V], br[3][V];
                    bool vok[V];
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        vok[v] = row_out[r] && col_out[v];
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            er[d][v] = enew[d][r][v];
                            br[d][v] = bh[d][r][v];
                        }
                    }
                    record_probes<S, C, V>(a.probe, pout, pkl, gi0, gj[r], er, br, e_ok, b_ok, vok);
                }
            }
            // level l+1 inputs: E_l(pe-1) (previous iteration) and B_l(pe-1)
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        ein[d][r][v] = as_stored<S>(ePi[l][d][r][v]);
                        ePo[l][d][r][v] = enew[d][r][v];
                        bPo[l][d][r][v] = bin[d][r][v];
                        bin[d][r][v] = bnew[d][r][v];
                    }
            if (l < L - 1) {
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    sB(l + 1, 0, w)[v] = bin[0][R - 1][v];
                    sB(l + 1, 1, w)[v] = bin[2][R - 1][v];
                }
                __syncthreads();
            }
        }
        // stores of the last level: E_{L-1} at plane kk-L+1, B_{L-1} at kk-L
        // (buffer path: size-0 resources for planes outside the chunk)
        if constexpr (!kBuf) {
            // the prefetch of plane kk+1 has arrived before this plane's
            // stores are issued: the wait at the end of the iteration then
            // leaves the stores in flight (their count is not known to the
            // compiler, so its own wait would include them)
            wait_vmem_all();
            const int pe = kk - L + 1;
            if (pe >= k0 && pe < k1) {
                const size_t off = static_cast<size_t>(pe) * plane;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (row_out[r])
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            unsigned so = lo[r];
                            asm volatile("" : "+v"(so));
                            S* dst = generic(at(uniform_gptr(a.out.e[d] + off), so));
#pragma unroll
                            for (int v = 0; v < V; ++v)
                                if (col_out[v])
                                    __builtin_nontemporal_store(static_cast<S>(ePo[L - 1][d][r][v]), dst + v);
                        }
            }
            if (pe - 1 >= k0 && pe - 1 < k1) {
                const size_t off = static_cast<size_t>(pe - 1) * plane;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (row_out[r])
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            unsigned so = lo[r];
                            asm volatile("" : "+v"(so));
                            S* dst = generic(at(uniform_gptr(a.out.b[d] + off), so));
#pragma unroll
                            for (int v = 0; v < V; ++v)
                                if (col_out[v])
                                    __builtin_nontemporal_store(static_cast<S>(bin[d][r][v]), dst + v);
                            if (FINAL && a.ahead[0] != nullptr) {
                                S* cdst = generic(at(uniform_gptr(a.ahead[d] + off), so));
#pragma unroll
                                for (int v = 0; v < V; ++v)
                                    if (col_out[v])
                                        __builtin_nontemporal_store(static_cast<S>(cy[d][r][v]), cdst + v);
                            }
                        }
            }
        } else {
            const int pe = kk - L + 1;
            const bool se = pe >= k0 && pe < k1, sb = pe - 1 >= k0 && pe - 1 < k1;
            const size_t offe = static_cast<size_t>(se ? pe : k0) * plane;
            const size_t offb = static_cast<size_t>(sb ? pe - 1 : k0) * plane;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const __amdgpu_buffer_rsrc_t re = plane_rsrc(a.out.e[d] + offe, se ? pbytes : 0u);
                const __amdgpu_buffer_rsrc_t rb = plane_rsrc(a.out.b[d] + offb, sb ? pbytes : 0u);
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        buf_store_nt(static_cast<S>(ePo[L - 1][d][r][v]), re, st_off[r][v]);
                        buf_store_nt(static_cast<S>(bin[d][r][v]), rb, st_off[r][v]);
                    }
            }
        }
        if (L == 1) {
            __syncthreads();
#pragma unroll
            for (int v = 0; v < V; ++v) {
                sE(0, 0, w)[v] = ePo[0][0][0][v];
                sE(0, 1, w)[v] = ePo[0][2][0][v];
            }
        }
        if (kBuf || more) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    widen<S, C, V>(rawB[d][r], bn[d][r]);
                    widen<S, C, V>(rawE[d][r], en[d][r]);
                }
            if constexpr (RESCALE) {
#pragma unroll
                for (int d = 0; d < 3; ++d)
#pragma unroll
                    for (int r = 0; r < R; ++r)
#pragma unroll
                        for (int v = 0; v < V; ++v) {
                            bn[d][r][v] = as_stored<S>(bn[d][r][v] * a.iscale[3 + d]);
                            en[d][r][v] = as_stored<S>(en[d][r][v] * a.iscale[d]);
                        }
            }
        }
        if constexpr (kBuf) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                sB(0, 0, w)[v] = bn[0][R - 1][v];
                sB(0, 1, w)[v] = bn[2][R - 1][v];
            }
        }
    };
    for (int kk = kfirst; kk <= klast; ++kk) plane_step(kk, eP, bP, eP, bP);
}

// PROBE: the variant that records StripArgs::probe after each level (the
// level's B is held in registers until then: up to 35 more VGPRs, which the
// plain variant does without).
template <typename S, typename C, int V, int R, int L, bool FINAL, bool PROBE = false>
__global__ void __launch_bounds__(strip_max_threads(V, R, L, sizeof(C)))
strip_sweep(StripArgs<S, C> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int blk = blockIdx.x;
    const int tile = (blk & 7) * a.g.per_xcd + (blk >> 3);
    if (tile >= a.g.ntiles) return;  // whole workgroup
    strip_tile<S, C, V, R, L, FINAL, false, PROBE>(a, tile, reinterpret_cast<C*>(lds_raw));
}

// FP16 storage: the strip pass after an in-run rescaling decision, reading
// the state scaled by StripArgs::iscale (a kernel of its own, so that the
// other passes carry none of it).
template <typename S, typename C, int V, int R, int L, bool FINAL, bool PROBE = false>
__global__ void __launch_bounds__(strip_max_threads(V, R, L, sizeof(C)))
strip_rescale_sweep(StripArgs<S, C> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int blk = blockIdx.x;
    const int tile = (blk & 7) * a.g.per_xcd + (blk >> 3);
    if (tile >= a.g.ntiles) return;  // whole workgroup
    strip_tile<S, C, V, R, L, FINAL, true, PROBE>(a, tile, reinterpret_cast<C*>(lds_raw));
}

// ---------------------------------------------------------------------------
// Batched sweep: many independent grids, one persistent launch per step.
//
// The tiles of all grids form one work list (most expensive first); each
// workgroup takes the next tile from a global counter until the list is
// exhausted, then the last workgroup to finish resets the counters for the
// next launch.  Every tile runs the same fused_tile as a single grid.

template <typename S, typename C> struct BatchGrid {
    Fields<S> buf[2];        // buf[p]: the grid's buffer pair used at parity p
    const PointSource* src;  // this grid's sources
    int src_off;             // offset of its source values in a value-table row
    int nsrc;
    Coefs<C> c;
    SweepGeom g;             // per-grid tile height and z-chunk
};

template <typename S, typename C> struct BatchArgs {
    const BatchGrid<S, C>* grids;
    const int4* work;        // (grid, tile_y, chunk, cost), most expensive first
    int nwork;
    int parity;              // the sweep reads buf[parity], writes buf[1-parity]
    const double* vals_row;  // source values of this step (null: no sources)
    unsigned int* counters;  // [0] next work item, [1] finished workgroups
};

// Values that are the same for every lane, moved into scalar registers
// (readfirstlane).  Loads of the batch descriptors inside the work loop
// cannot be proven unclobbered, so the compiler keeps them as vector loads;
// these helpers make the results scalar again.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uni(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(b));
    const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                static_cast<unsigned int>(lo));
}
template <typename P> __device__ __forceinline__ P* uni(P* v) {
    const unsigned long long b = reinterpret_cast<unsigned long long>(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<int>(b));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32));
    return reinterpret_cast<P*>((static_cast<unsigned long long>(hi) << 32) | lo);
}

// Sweep arguments of one grid of a batch at a given parity, in scalar
// registers (the parity selects buffers by address, not by indexing a
// local array, which would go to scratch memory).
template <typename S, typename C>
__device__ __forceinline__ SweepArgs<S, C, C> load_batch_grid(const BatchGrid<S, C>* p, int parity,
                                                              const double* vals_row) {
    SweepArgs<S, C, C> a;
    const Fields<S>* in = p->buf + parity;
    const Fields<S>* out = p->buf + (1 - parity);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        a.in.e[d] = uni(in->e[d]);
        a.in.b[d] = uni(in->b[d]);
        a.out.e[d] = uni(out->e[d]);
        a.out.b[d] = uni(out->b[d]);
        a.j[d] = nullptr;
    }
    a.maxbits = nullptr;
    a.kstart = nullptr;
#pragma unroll
    for (int c6 = 0; c6 < 6; ++c6) a.oscale[c6] = C(1);
#pragma unroll
    for (int d = 0; d < 3; ++d) a.ahead[d] = nullptr;
    a.src = uni(p->src);
    a.vals = vals_row ? vals_row + uni(p->src_off) : nullptr;
    a.nsrc = uni(p->nsrc);
    a.pad = 0;
    const Coefs<C>& c = p->c;
    a.c.ex_bz = uni(c.ex_bz); a.c.ex_by = uni(c.ex_by); a.c.ey_bx = uni(c.ey_bx);
    a.c.ey_bz = uni(c.ey_bz); a.c.ez_by = uni(c.ez_by); a.c.ez_bx = uni(c.ez_bx);
    a.c.jx = uni(c.jx); a.c.jy = uni(c.jy); a.c.jz = uni(c.jz);
    a.c.bx_ey = uni(c.bx_ey); a.c.bx_ez = uni(c.bx_ez); a.c.by_ez = uni(c.by_ez);
    a.c.by_ex = uni(c.by_ex); a.c.bz_ex = uni(c.bz_ex); a.c.bz_ey = uni(c.bz_ey);
    const SweepGeom& q = p->g;
    a.g.Ni = uni(q.Ni); a.g.Nj = uni(q.Nj); a.g.Nk = uni(q.Nk); a.g.TY = uni(q.TY);
    a.g.kchunk = uni(q.kchunk); a.g.ntile_y = uni(q.ntile_y); a.g.ntiles = uni(q.ntiles);
    a.g.per_xcd = uni(q.per_xcd);
    a.probe = ProbeArgs{};  // batches record no probes
    return a;
}

template <typename S, typename C, int V, bool STEADY>
__global__ void __launch_bounds__(fused_max_threads(V, sizeof(C)))
batch_fused_sweep(BatchArgs<S, C> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ int item;
    for (;;) {
        if (threadIdx.x == 0) item = static_cast<int>(atomicAdd(&a.counters[0], 1u));
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(item);
        __syncthreads();
        if (t >= a.nwork) break;
        const int4 wv = a.work[t];
        const int4 w = make_int4(uni(wv.x), uni(wv.y), uni(wv.z), 0);
        const SweepArgs<S, C, C> sa = load_batch_grid(a.grids + w.x, a.parity, a.vals_row);
        fused_tile<S, C, C, V, STEADY>(sa, w.y, w.z, reinterpret_cast<C*>(lds_raw));
    }
    if (threadIdx.x == 0) {
        const unsigned done = atomicAdd(&a.counters[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&a.counters[0], 0u);
            atomicExch(&a.counters[1], 0u);
        }
    }
}

// Batched temporal blocking: one persistent launch per pass of L steps
// over the strip tiles of every grid (the same strip_tile as a single grid).
template <typename S, typename C> struct BatchStripGrid {
    Fields<S> buf[2];        // buf[p]: the grid's buffer pair used at parity p
    const PointSource* src;  // this grid's sources, sorted by plane
    const int* kstart;       // its per-plane ranges (Nk + 1 entries, local indices)
    int src_off;             // offset of its source values in a value-table row
    int nsrc;
    Coefs<C> c;
    StripGeom g;             // its strip tiling
};

template <typename S, typename C> struct BatchStripArgs {
    const BatchStripGrid<S, C>* grids;
    const int2* work;        // (grid, tile), most expensive first
    int nwork;
    int parity;              // the pass reads buf[parity], writes buf[1-parity]
    const double* vals;      // source values of the pass's first step (null: no sources)
    int vstride;             // value-table row length (all sources of the batch)
    int pad;
    unsigned int* counters;  // [0] next work item, [1] finished workgroups
};

template <typename S, typename C>
__device__ __forceinline__ StripArgs<S, C> load_batch_strip_grid(const BatchStripGrid<S, C>* p,
                                                                 int parity, const double* vals,
                                                                 int vstride) {
    StripArgs<S, C> a;
    const Fields<S>* in = p->buf + parity;
    const Fields<S>* out = p->buf + (1 - parity);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        a.in.e[d] = uni(in->e[d]);
        a.in.b[d] = uni(in->b[d]);
        a.out.e[d] = uni(out->e[d]);
        a.out.b[d] = uni(out->b[d]);
        a.ahead[d] = nullptr;
    }
    a.src = uni(p->src);
    a.kstart = uni(p->kstart);
    a.vals = vals ? vals + uni(p->src_off) : nullptr;
    a.nsrc = uni(p->nsrc);
    a.vstride = vstride;
#pragma unroll
    for (int c6 = 0; c6 < 6; ++c6) a.iscale[c6] = 1.0f;
    a.probe = ProbeArgs{};  // batches record no probes
    const Coefs<C>& c = p->c;
    a.c.ex_bz = uni(c.ex_bz); a.c.ex_by = uni(c.ex_by); a.c.ey_bx = uni(c.ey_bx);
    a.c.ey_bz = uni(c.ey_bz); a.c.ez_by = uni(c.ez_by); a.c.ez_bx = uni(c.ez_bx);
    a.c.jx = uni(c.jx); a.c.jy = uni(c.jy); a.c.jz = uni(c.jz);
    a.c.bx_ey = uni(c.bx_ey); a.c.bx_ez = uni(c.bx_ez); a.c.by_ez = uni(c.by_ez);
    a.c.by_ex = uni(c.by_ex); a.c.bz_ex = uni(c.bz_ex); a.c.bz_ey = uni(c.bz_ey);
    const StripGeom& q = p->g;
    a.g.Ni = uni(q.Ni); a.g.Nj = uni(q.Nj); a.g.Nk = uni(q.Nk); a.g.W = uni(q.W);
    a.g.Gx = uni(q.Gx); a.g.TX = uni(q.TX); a.g.TY = uni(q.TY); a.g.kchunk = uni(q.kchunk);
    a.g.ntx = uni(q.ntx); a.g.nty = uni(q.nty); a.g.ntiles = uni(q.ntiles);
    a.g.per_xcd = uni(q.per_xcd);
    return a;
}

template <typename S, typename C, int V, int R, int L, bool FINAL>
__global__ void __launch_bounds__(strip_max_threads(V, R, L, sizeof(C)))
batch_strip_sweep(BatchStripArgs<S, C> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ int item;
    for (;;) {
        if (threadIdx.x == 0) item = static_cast<int>(atomicAdd(&a.counters[0], 1u));
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(item);
        __syncthreads();
        if (t >= a.nwork) break;
        const int2 wv = a.work[t];
        const StripArgs<S, C> sa = load_batch_strip_grid(a.grids + uni(wv.x), a.parity, a.vals,
                                                         a.vstride);
        strip_tile<S, C, V, R, L, FINAL>(sa, uni(wv.y), reinterpret_cast<C*>(lds_raw));
    }
    if (threadIdx.x == 0) {
        const unsigned done = atomicAdd(&a.counters[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&a.counters[0], 0u);
            atomicExch(&a.counters[1], 0u);
        }
    }
}

// B half step of grid q of a batch, reading its current buffers (buf[0])
// and coefficients from the descriptor table, so that a captured launch
// follows the descriptors when they are re-pointed between graph replays.
template <typename S, typename C>
__global__ void __launch_bounds__(256)
batch_plain_update_B(const BatchGrid<S, C>* __restrict__ grids, int q) {
    const BatchGrid<S, C>* p = grids + q;
    const int Ni = p->g.Ni, Nj = p->g.Nj, Nk = p->g.Nk;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= Ni || j >= Nj || k >= Nk) return;
    const Fields<S> f = p->buf[0];
    const Coefs<C> c = p->c;
    plain_B_cell<S, C>(f, c, Ni, Nj, Nk, i, j, k);
}

// ---------------------------------------------------------------------------
// Diagnostics (FP64 accumulation, deterministic two-pass reductions).

constexpr int kReduceThreads = 256;

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = kReduceThreads / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] += sh[tid + s];
        __syncthreads();
    }
    const T r = sh[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = kReduceThreads / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] = fmax(sh[tid], sh[tid + s]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Per-block partials of sum(x^2) (scaled by inv_scale) and max|x| for one array.
template <typename S>
__global__ void __launch_bounds__(kReduceThreads)
sumsq_maxabs_partial(const S* __restrict__ x, size_t n, double inv_scale,
                     double* __restrict__ part) {
    // x is an allocation's start (16-byte aligned): 16-byte packs (8 FP16,
    // 4 FP32 or 2 FP64 values), four in flight per thread, then the tail
    // element-wise.  A fixed grid keeps the summation order fixed.
    __shared__ double sh[kReduceThreads];
    constexpr int P = 16 / static_cast<int>(sizeof(S));
    using Q = Pack<S, P>;
    const Q* __restrict__ xq = reinterpret_cast<const Q*>(x);
    const size_t npk = n / P;
    const size_t stride = static_cast<size_t>(gridDim.x) * kReduceThreads;
    double s = 0.0, m = 0.0;
    auto add = [&](S e) {
        const double v = static_cast<double>(e) * inv_scale;
        s += v * v;
        m = fmax(m, fabs(v));
    };
    size_t i = blockIdx.x * static_cast<size_t>(kReduceThreads) + threadIdx.x;
    for (; i + 3 * stride < npk; i += 4 * stride) {
        Q q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = xq[i + u * stride];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int e = 0; e < P; ++e) add(q[u].v[e]);
    }
    for (; i < npk; i += stride) {
        const Q q = xq[i];
#pragma unroll
        for (int e = 0; e < P; ++e) add(q.v[e]);
    }
    for (size_t t = npk * P + blockIdx.x * static_cast<size_t>(kReduceThreads) + threadIdx.x;
         t < n; t += stride)
        add(x[t]);
    s = block_sum(s, sh);
    m = block_max(m, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s;
        part[2 * blockIdx.x + 1] = m;
    }
}

// Per-block partials of sum((a - b)^2), sum(b^2), max|a - b| where a is any
// precision (divided by its storage scale) and b is FP64.
template <typename S>
__global__ void __launch_bounds__(kReduceThreads)
diff_partial(const S* __restrict__ a, double inv_scale, const double* __restrict__ b, size_t n,
             double* __restrict__ part) {
    __shared__ double sh[kReduceThreads];
    double sd = 0.0, sb = 0.0, m = 0.0;
    for (size_t i = blockIdx.x * static_cast<size_t>(kReduceThreads) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kReduceThreads) {
        const double bv = b[i];
        const double d = static_cast<double>(a[i]) * inv_scale - bv;
        sd += d * d;
        sb += bv * bv;
        m = fmax(m, fabs(d));
    }
    sd = block_sum(sd, sh);
    sb = block_sum(sb, sh);
    m = block_max(m, sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = sd;
        part[3 * blockIdx.x + 1] = sb;
        part[3 * blockIdx.x + 2] = m;
    }
}

// Final pass over `nblk` partial records of width w; column c is a sum
// unless (max_mask >> c) & 1.  Deterministic: one block, fixed order.
static __global__ void __launch_bounds__(kReduceThreads)
reduce_partials(const double* __restrict__ part, int nblk, int w, unsigned max_mask,
                double* __restrict__ out) {
    __shared__ double sh[kReduceThreads];
    for (int col = 0; col < w; ++col) {
        const bool is_max = (max_mask >> col) & 1u;
        double acc = 0.0;
        for (int b = threadIdx.x; b < nblk; b += kReduceThreads) {
            const double v = part[b * w + col];
            acc = is_max ? fmax(acc, v) : acc + v;
        }
        const double r = is_max ? block_max(acc, sh) : block_sum(acc, sh);
        if (threadIdx.x == 0) out[col] = r;
    }
}

// out[0] = the sum of column 0 of `nblk` partial records of width w, in
// reduce_partials' order (energy records equal fdtd_sum_squares bitwise).
static __global__ void __launch_bounds__(kReduceThreads)
sum_partials_col0(const double* __restrict__ part, int nblk, int w, double* __restrict__ out) {
    __shared__ double sh[kReduceThreads];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kReduceThreads) acc = acc + part[b * w];
    const double r = block_sum(acc, sh);
    if (threadIdx.x == 0) out[0] = r;
}

// Probe values read from the fields in memory (the state after a step of
// the three-sweep or PML path): probe q's value goes to pa.out[pa.p[q].col].
template <typename S>
__global__ void __launch_bounds__(256)
gather_probes(Fields<S> f, ProbeArgs pa, int Ni, int Nj) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= pa.n) return;
    const Probe r = pa.p[q];
    const size_t idx = static_cast<size_t>(r.i) + static_cast<size_t>(r.j) * Ni +
                       static_cast<size_t>(r.k) * Ni * Nj;
    const S* src = f.e[0];
    double inv = pa.inv[0];
#pragma unroll
    for (int c = 1; c < 6; ++c)
        if (r.comp == c) {
            src = c < 3 ? f.e[c] : f.b[c - 3];
            inv = pa.inv[c];
        }
    pa.out[r.col] = static_cast<double>(src[idx]) * inv;
}

// Element-wise conversions between the FP64 host layout and storage types.
template <typename S>
__global__ void __launch_bounds__(256)
convert_from_f64(const double* __restrict__ src, S* __restrict__ dst, size_t n, double scale) {
    for (size_t i = blockIdx.x * static_cast<size_t>(256) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * 256)
        dst[i] = static_cast<S>(src[i] * scale);
}

template <typename S>
__global__ void __launch_bounds__(256)
convert_to_f64(const S* __restrict__ src, double* __restrict__ dst, size_t n, double inv_scale) {
    for (size_t i = blockIdx.x * static_cast<size_t>(256) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * 256)
        dst[i] = static_cast<double>(src[i]) * inv_scale;
}

// Adds a plane wave amp*pol*cos(k.r + phase) to E and (khat x E) to B,
// each component sampled at its Yee position (Ex at (i+1/2,j,k), Ey at
// (i,j+1/2,k), Ez at (i,j,k+1/2), Bx at (i,j+1/2,k+1/2), By at
// (i+1/2,j,k+1/2), Bz at (i+1/2,j+1/2,k)); k in radians per cell, values in
// FP64, stored with the component's scale.
template <typename S>
__global__ void __launch_bounds__(256)
add_plane_wave(Fields<S> f, int Ni, int Nj, int Nk, double kx, double ky, double kz, double phase,
               double ex, double ey, double ez, double bx, double by, double bz,
               double s0, double s1, double s2, double s3, double s4, double s5) {
    const size_t n = static_cast<size_t>(Ni) * Nj * Nk;
    for (size_t idx = blockIdx.x * static_cast<size_t>(256) + threadIdx.x; idx < n;
         idx += static_cast<size_t>(gridDim.x) * 256) {
        const double i = static_cast<double>(idx % Ni);
        const double j = static_cast<double>((idx / Ni) % Nj);
        const double k = static_cast<double>(idx / (static_cast<size_t>(Ni) * Nj));
        const double amp[6] = {ex * s0, ey * s1, ez * s2, bx * s3, by * s4, bz * s5};
        const double ox[6] = {0.5, 0.0, 0.0, 0.0, 0.5, 0.5};
        const double oy[6] = {0.0, 0.5, 0.0, 0.5, 0.0, 0.5};
        const double oz[6] = {0.0, 0.0, 0.5, 0.5, 0.5, 0.0};
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            if (amp[c] == 0.0) continue;
            const double ph = kx * (i + ox[c]) + ky * (j + oy[c]) + kz * (k + oz[c]) + phase;
            S* dst = c < 3 ? f.e[c] : f.b[c - 3];
            dst[idx] = static_cast<S>(static_cast<double>(dst[idx]) + amp[c] * cos(ph));
        }
    }
}

// Multiply a storage array by an exact power of two (FP16 rescaling).
template <typename S>
__global__ void __launch_bounds__(256)
scale_array(S* __restrict__ x, size_t n, float factor) {
    // x is an allocation's start: 16-byte packs, then the tail
    constexpr int P = 16 / static_cast<int>(sizeof(S));
    using Q = Pack<S, P>;
    Q* __restrict__ xq = reinterpret_cast<Q*>(x);
    const size_t npk = n / P;
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i = blockIdx.x * static_cast<size_t>(256) + threadIdx.x; i < npk; i += stride) {
        Q q = xq[i];
#pragma unroll
        for (int e = 0; e < P; ++e) q.v[e] = static_cast<S>(static_cast<float>(q.v[e]) * factor);
        xq[i] = q;
    }
    for (size_t t = npk * P + blockIdx.x * static_cast<size_t>(256) + threadIdx.x; t < n;
         t += stride)
        x[t] = static_cast<S>(static_cast<float>(x[t]) * factor);
}

// Fill with zeros (used instead of hipMemsetAsync so that step sequences can
// be captured in HIP graphs).
static __global__ void __launch_bounds__(256)
zero_bytes(uint4* __restrict__ p, size_t n16) {
    for (size_t i = blockIdx.x * static_cast<size_t>(256) + threadIdx.x; i < n16;
         i += static_cast<size_t>(gridDim.x) * 256)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Order-independent checksum of the bit patterns of an array: the sum mod
// 2^64 of a 64-bit mix of (bits, index).  Integer addition makes the result
// independent of the reduction order, so equal arrays give equal sums.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename S>
__device__ __forceinline__ unsigned long long bits_of(S v) {
    if constexpr (sizeof(S) == 8) return __builtin_bit_cast(unsigned long long, v);
    else if constexpr (sizeof(S) == 4) return __builtin_bit_cast(unsigned int, v);
    else return __builtin_bit_cast(unsigned short, v);
}

template <typename S>
__global__ void __launch_bounds__(kReduceThreads)
checksum_partial(const S* __restrict__ x, size_t n, unsigned long long* __restrict__ part) {
    __shared__ unsigned long long sh[kReduceThreads];
    unsigned long long acc = 0;
    for (size_t i = blockIdx.x * static_cast<size_t>(kReduceThreads) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kReduceThreads)
        acc += mix64(bits_of(x[i]) ^ (static_cast<unsigned long long>(i) * 0xD6E8FEB86659FD93ull));
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Per-block counts of non-finite values (NaN, +-Inf) of one array.
template <typename S>
__global__ void __launch_bounds__(kReduceThreads)
nonfinite_partial(const S* __restrict__ x, size_t n, unsigned long long* __restrict__ part) {
    __shared__ unsigned long long sh[kReduceThreads];
    unsigned long long acc = 0;
    for (size_t i = blockIdx.x * static_cast<size_t>(kReduceThreads) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kReduceThreads)
    {
        bool finite;
        if constexpr (sizeof(S) == 2)
            finite = __builtin_isfinite(static_cast<float>(x[i]));  // exact widening
        else
            finite = __builtin_isfinite(x[i]);
        acc += finite ? 0u : 1u;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

static __global__ void __launch_bounds__(kReduceThreads)
checksum_final(const unsigned long long* __restrict__ part, int nblk,
               unsigned long long* __restrict__ out) {
    __shared__ unsigned long long sh[kReduceThreads];
    unsigned long long acc = 0;
    for (int b = threadIdx.x; b < nblk; b += kReduceThreads) acc += part[b];
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) *out = acc;
}

// Streaming copy of a known byte count, 16 B per lane: used to calibrate the
// HBM byte counters against an exactly known traffic.
static __global__ void __launch_bounds__(256)
copy_calibration(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
    for (size_t i = blockIdx.x * static_cast<size_t>(256) + threadIdx.x; i < n16;
         i += static_cast<size_t>(gridDim.x) * 256)
        dst[i] = src[i];
}

}  // namespace fdtd
