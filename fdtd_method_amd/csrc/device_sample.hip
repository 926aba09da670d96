// Counterpart of the reference's Kokkos perf-test driver
// (perf-tests/kokkos_sample/kokkos_sample.cpp:16-186) on the MI355X drop-in
// class: the same spherical wave as the OpenMP driver (the 8-cell cos^2 box
// current on Jx = Jy = Jz for the first period/dt = 40 steps), but the
// current is written into the class's device arrays by a GPU kernel of the
// driver's own before every step, as the Kokkos driver's parallel_for does
// (:88-128), and the printed slice is the Kokkos driver's (:135-150): Ex at
// x index Nk/2, one line per j in Nj/2-5 .. Nj/2+4, columns k in
// Ni/2-5 .. Ni/2+4.
//
//   device_sample                 N = 32, 100 iterations
//   device_sample N iters         as the reference
//   device_sample N iters [fp64|fp32|fp16] [--dump FILE]
//
// Prints "Execution time: X s" and the slice (setw(12), fixed,
// setprecision(5)).  --dump writes Ex (N^3 doubles, index i + j*N + k*N*N).
// The Kokkos driver's hwloc and configuration lines (:155-162) describe its
// Kokkos build and have no counterpart here.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "FDTD_mi355x.hpp"

using namespace FDTD_mi355x;

namespace {

struct Wave {
    double T, Tx, Ty, Tz, pi;  // period, spatial periods, FDTD_const::PI
    double dx, dy, dz;
    int Ni, Nj;
    int i0, j0, k0, ni, nj, nk;  // the box [i0, i0+ni) x [j0, j0+nj) x [k0, k0+nk)
};

// J at the box cells: the driver's current at t (kokkos_sample.cpp:91-110),
// or zero (:117-125).  One thread per box cell.
template <typename J>
__global__ void set_current(Wave w, double t, bool on, J* jx, J* jy, J* jz) {
    const int n = w.ni * w.nj * w.nk;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = w.i0 + q % w.ni;
    const int j = w.j0 + (q / w.ni) % w.nj;
    const int k = w.k0 + q / (w.ni * w.nj);
    const long long idx = i + (long long)j * w.Ni + (long long)k * w.Ni * w.Nj;
    double val = 0.0;
    if (on) {
        const double x = i * w.dx, y = j * w.dy, z = k * w.dz;
        val = sin(2.0 * w.pi * t / w.T) * pow(cos(2.0 * w.pi * x / w.Tx), 2.0) *
              pow(cos(2.0 * w.pi * y / w.Ty), 2.0) * pow(cos(2.0 * w.pi * z / w.Tz), 2.0);
    }
    jx[idx] = static_cast<J>(val);
    jy[idx] = static_cast<J>(val);
    jz[idx] = static_cast<J>(val);
}

bool hip_ok(hipError_t e) {
    if (e == hipSuccess) return true;
    std::cerr << "ERROR: " << hipGetErrorString(e) << std::endl;
    return false;
}

template <typename J>
bool run(FDTD& method, const Wave& w, int cur_time, int it, double dt) {
    J* jx = method.device_field_as<J>(Component::JX);
    J* jy = method.device_field_as<J>(Component::JY);
    J* jz = method.device_field_as<J>(Component::JZ);
    hipStream_t s = static_cast<hipStream_t>(method.stream());
    const int n = w.ni * w.nj * w.nk;
    if (n <= 0) {
        method.update_fields(it);
        return true;
    }
    const dim3 grid((n + 255) / 256), block(256);
    for (int t = 0; t < cur_time; ++t) {
        hipLaunchKernelGGL(set_current<J>, grid, block, 0, s, w, (t + 1) * dt, true, jx, jy, jz);
        if (!hip_ok(hipGetLastError())) return false;
        method.update_fields();
    }
    method.zeroed_currents();
    for (int t = cur_time; t < it; ++t) {
        hipLaunchKernelGGL(set_current<J>, grid, block, 0, s, w, 0.0, false, jx, jy, jz);
        if (!hip_ok(hipGetLastError())) return false;
        method.update_fields();
    }
    method.synchronize();
    return true;
}

int spherical_wave(int n, int it, Precision prec, const char* dump) {
    CurrentParameters cur_param{8, 4, 0.2};  // kokkos_sample.cpp:17-28
    cur_param.iterations = static_cast<int>(static_cast<double>(cur_param.period) / cur_param.dt);
    const double d = FDTD_const::C;  // :38-56
    const double boundary = static_cast<double>(n) / 2.0 * d;
    const Parameters params{n, n, n, -boundary, boundary, -boundary, boundary, -boundary,
                            boundary, d, d, d};
    FDTD method(params, cur_param.dt, prec);
    const int cur_time = std::min(cur_param.iterations, it);

    const double Tx = cur_param.period_x, Ty = cur_param.period_y, Tz = cur_param.period_z;
    const int start_i = static_cast<int>(std::floor((-Tx / 4.0 - params.ax) / params.dx));  // :60-66
    const int start_j = static_cast<int>(std::floor((-Ty / 4.0 - params.ay) / params.dy));
    const int start_k = static_cast<int>(std::floor((-Tz / 4.0 - params.az) / params.dz));
    const int max_i = static_cast<int>(std::floor((Tx / 4.0 - params.ax) / params.dx));
    const int max_j = static_cast<int>(std::floor((Ty / 4.0 - params.ay) / params.dy));
    const int max_k = static_cast<int>(std::floor((Tz / 4.0 - params.az) / params.dz));
    Wave w{static_cast<double>(cur_param.period), Tx, Ty, Tz, FDTD_const::PI,
           params.dx, params.dy, params.dz, params.Ni, params.Nj,
           start_i, start_j, start_k,
           std::max(0, max_i - start_i), std::max(0, max_j - start_j), std::max(0, max_k - start_k)};
    // the box must lie in the grid (the drivers index J without a check)
    if (w.ni > 0 && w.nj > 0 && w.nk > 0 &&
        (start_i < 0 || start_j < 0 || start_k < 0 || max_i > n || max_j > n || max_k > n)) {
        std::cerr << "ERROR: current box outside the grid" << std::endl;
        return 1;
    }

    auto start = std::chrono::high_resolution_clock::now();
    const bool ok = prec == Precision::FP64 ? run<double>(method, w, cur_time, it, cur_param.dt)
                                            : run<float>(method, w, cur_time, it, cur_param.dt);
    if (!ok) return 1;
    auto end = std::chrono::high_resolution_clock::now();
    std::chrono::duration<double> elapsed = end - start;
    std::cout << "Execution time: " << elapsed.count() << " s" << std::endl;

    Field& Ex = method.get_field(Component::EX);  // :135-150
    const int i = params.Nk / 2;
    for (int j = params.Nj / 2 - 5; j < params.Nj / 2 + 5; j++) {
        for (int k = params.Ni / 2 - 5; k < params.Ni / 2 + 5; k++) {
            const long long index = i + static_cast<long long>(j) * params.Ni +
                                    static_cast<long long>(k) * params.Ni * params.Nj;
            std::cout << std::setw(12) << std::fixed << std::setprecision(5)
                      << static_cast<double>(Ex[index]);
        }
        std::cout << std::endl;
    }
    std::cout << std::endl;

    if (dump) {
        std::ofstream out(dump, std::ios::binary);
        const double* p = Ex.data();
        out.write(reinterpret_cast<const char*>(p), sizeof(double) * method.cells());
        if (!out) {
            std::cerr << "ERROR: cannot write " << dump << std::endl;
            return 1;
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    int N = 32, iters = 100;  // kokkos_sample.cpp:168-176
    Precision prec = Precision::FP64;
    const char* dump = nullptr;
    int pos = 0;
    for (int a = 1; a < argc; ++a) {
        if (std::strcmp(argv[a], "--dump") == 0 && a + 1 < argc) {
            dump = argv[++a];
        } else if (std::strcmp(argv[a], "fp64") == 0) {
            prec = Precision::FP64;
        } else if (std::strcmp(argv[a], "fp32") == 0) {
            prec = Precision::FP32;
        } else if (std::strcmp(argv[a], "fp16") == 0) {
            prec = Precision::FP16;
        } else if (pos == 0) {
            N = std::atoi(argv[a]);
            ++pos;
        } else if (pos == 1) {
            iters = std::atoi(argv[a]);
            ++pos;
        } else {
            pos = 3;
        }
    }
    if (pos == 1 || pos > 2) {  // :177-180
        std::cout << "ERROR: Incorrect number of parameters" << std::endl;
        return 1;
    }
    try {
        return spherical_wave(N, iters, prec, dump);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
}
