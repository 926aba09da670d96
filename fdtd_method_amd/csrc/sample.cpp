// Counterpart of the reference's perf-test drivers
// (perf-tests/sample/sample.cpp:14-168, perf-tests/kokkos_sample/kokkos_sample.cpp:16-185)
// on the MI355X drop-in class: a spherical wave from the drivers' 8-cell
// cos^2 box current, switched on for the first period/dt = 40 steps.
//
//   sample                 N = 32, 100 iterations
//   sample N iters         as the reference
//   sample N iters [fp64|fp32|fp16] [--pml [PCT]] [--dump FILE] [--dump-pml FILE]
//   sample N iters 1       also writes the plane Nk/2 of Ex..Bz after every
//                          step t (0-based) to OutFiles_<1..6>/<t>.csv, the
//                          slices the reference's plotting script reads
//                          (python_script_legend/visualization.py:10-49 runs
//                          the driver with this third argument)
//
// Prints "Execution time: X s" and the 10x10 Ex slice around the centre of
// plane Nk/2 in the reference's format (setw(12), fixed, setprecision(5)).
// --pml also runs the same wave on the PML class with pml_percent PCT
// (default 0.2), as the reference driver's __PML_TEST__ build does
// (sample.cpp:93-123,136-147): "Execution time (PML): X s" after the first
// time, and "PML: " with its slice after the first slice.
// --dump / --dump-pml write Ex (N^3 doubles, index i + j*N + k*N*N) to FILE.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <functional>
#include <memory>
#include <string>
#include <sys/stat.h>

#include "FDTD_mi355x.hpp"

using namespace FDTD_mi355x;

static Parameters wave_parameters(int n) {
    const double d = FDTD_const::C;  // perf-tests/sample/sample.cpp:33-49
    const double boundary = static_cast<double>(n) / 2.0 * d;
    return Parameters{n, n, n, -boundary, boundary, -boundary, boundary, -boundary, boundary,
                      d, d, d};
}

static CurrentParameters wave_current() {
    CurrentParameters cur_param{8, 4, 0.2};  // sample.cpp:15-19,24
    cur_param.iterations = static_cast<int>(static_cast<double>(cur_param.period) / cur_param.dt);
    return cur_param;
}

// The drivers' time loop (perf-tests/sample/sample.cpp:55-87) on `method`:
// the cos^2 box current on Jx = Jy = Jz before each of the first
// min(iterations, it) steps, then zeroed_currents() and the remaining steps.
// Returns the elapsed seconds.
static double run_wave(FDTD& method, const Parameters& params, int it,
                       const std::function<void(int)>& after_step = nullptr) {
    const CurrentParameters cur_param = wave_current();
    const double T = cur_param.period;
    const double Tx = cur_param.period_x, Ty = cur_param.period_y, Tz = cur_param.period_z;
    auto current = [T, Tx, Ty, Tz](double x, double y, double z, double t) {
        return std::sin(2.0 * FDTD_const::PI * t / T) *
               std::pow(std::cos(2.0 * FDTD_const::PI * x / Tx), 2.0) *
               std::pow(std::cos(2.0 * FDTD_const::PI * y / Ty), 2.0) *
               std::pow(std::cos(2.0 * FDTD_const::PI * z / Tz), 2.0);
    };
    const int cur_time = std::min(cur_param.iterations, it);
    const int start_i = static_cast<int>(std::floor((-Tx / 4.0 - params.ax) / params.dx));
    const int start_j = static_cast<int>(std::floor((-Ty / 4.0 - params.ay) / params.dy));
    const int start_k = static_cast<int>(std::floor((-Tz / 4.0 - params.az) / params.dz));
    const int max_i = static_cast<int>(std::floor((Tx / 4.0 - params.ax) / params.dx));
    const int max_j = static_cast<int>(std::floor((Ty / 4.0 - params.ay) / params.dy));
    const int max_k = static_cast<int>(std::floor((Tz / 4.0 - params.az) / params.dz));

    auto start = std::chrono::high_resolution_clock::now();
    for (int t = 0; t < cur_time; ++t) {
        for (int k = start_k; k < max_k; ++k)
            for (int j = start_j; j < max_j; ++j)
                for (int i = start_i; i < max_i; ++i) {
                    const int index = i + j * params.Ni + k * params.Ni * params.Nj;
                    const double value =
                        current(static_cast<double>(i) * params.dx, static_cast<double>(j) * params.dy,
                                static_cast<double>(k) * params.dz,
                                static_cast<double>(t + 1) * cur_param.dt);
                    method.get_field(Component::JX)[index] = value;
                    method.get_field(Component::JY)[index] = value;
                    method.get_field(Component::JZ)[index] = value;
                }
        method.update_fields();
        if (after_step) after_step(t);
    }
    method.zeroed_currents();
    for (int t = cur_time; t < it; ++t) {
        method.update_fields();
        if (after_step) after_step(t);
    }
    method.synchronize();  // runs the deferred steps
    auto end = std::chrono::high_resolution_clock::now();
    std::chrono::duration<double> elapsed = end - start;
    return elapsed.count();
}

// 10x10 Ex slice of plane Nk/2 (sample.cpp:125-134).
static void print_slice(FDTD& method, const Parameters& params) {
    Field& Ex = method.get_field(Component::EX);
    const int k = params.Nk / 2;
    for (int j = params.Nj / 2 - 5; j < params.Nj / 2 + 5; j++) {
        for (int i = params.Ni / 2 - 5; i < params.Ni / 2 + 5; i++) {
            const int index = i + j * params.Ni + k * params.Ni * params.Nj;
            std::cout << std::setw(12) << std::fixed << std::setprecision(5) << (double)Ex[index];
        }
        std::cout << std::endl;
    }
    std::cout << std::endl;
}

static bool dump_ex(FDTD& method, const char* path) {
    if (!path) return true;
    std::ofstream out(path, std::ios::binary);
    const double* p = method.get_field(Component::EX).data();
    out.write(reinterpret_cast<const char*>(p), sizeof(double) * method.cells());
    if (!out) {
        std::cerr << "ERROR: cannot write " << path << std::endl;
        return false;
    }
    return true;
}

// The plotting script's input: plane Nk/2 of Ex..Bz into OutFiles_<1..6>/<t>.csv
// (component numbering of visualization.py:10).
static void write_slices(FDTD& method, const Parameters& params, int t) {
    for (int c = 0; c < 6; ++c) {
        const std::string dir = "OutFiles_" + std::to_string(c + 1);
        ::mkdir(dir.c_str(), 0755);
        write_csv_slice(method, static_cast<Component>(c), params.Nk / 2,
                        dir + "/" + std::to_string(t) + ".csv");
    }
}

static int spherical_wave(int n, int it, Precision prec, double pml, const char* dump,
                          const char* dump_pml, bool slices) {
    const Parameters params = wave_parameters(n);
    const double dt = wave_current().dt;
    FDTD method(params, dt, prec);
    std::function<void(int)> after_step;
    if (slices) after_step = [&](int t) { write_slices(method, params, t); };
    const double elapsed = run_wave(method, params, it, after_step);
    std::cout << "Execution time: " << elapsed << " s" << std::endl;

    std::unique_ptr<FDTD_PML> pml_method;
    if (pml >= 0) {  // sample.cpp:93-123
        pml_method.reset(new FDTD_PML(params, dt, pml, prec));
        const double elapsed_pml = run_wave(*pml_method, params, it);
        std::cout << "Execution time (PML): " << elapsed_pml << " s" << std::endl;
    }
    print_slice(method, params);
    if (pml_method) {  // sample.cpp:136-147
        std::cout << "PML: \n" << std::endl;
        print_slice(*pml_method, params);
    }
    if (!dump_ex(method, dump)) return 1;
    if (pml_method && !dump_ex(*pml_method, dump_pml)) return 1;
    return 0;
}

int main(int argc, char* argv[]) {
    int N = 32, iters = 100;
    Precision prec = Precision::FP64;
    double pml = -1.0;
    const char* dump = nullptr;
    const char* dump_pml = nullptr;
    bool slices = false;
    int pos = 0;
    for (int a = 1; a < argc; ++a) {
        if (std::strcmp(argv[a], "--dump") == 0 && a + 1 < argc) {
            dump = argv[++a];
        } else if (std::strcmp(argv[a], "--dump-pml") == 0 && a + 1 < argc) {
            dump_pml = argv[++a];
        } else if (std::strcmp(argv[a], "--pml") == 0) {
            pml = 0.2;  // sample.cpp:94
            if (a + 1 < argc && std::strchr(argv[a + 1], '.')) pml = std::atof(argv[++a]);
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
        } else if (pos == 2 && std::strcmp(argv[a], "1") == 0) {
            slices = true;  // visualization.py:48-49 passes str(int(True))
            ++pos;
        } else {
            pos = 4;
        }
    }
    if (pos == 1 || pos > 3) {  // perf-tests/sample/sample.cpp:163-166
        std::cout << "ERROR: Incorrect number of parameters" << std::endl;
        return 1;
    }
    try {
        return spherical_wave(N, iters, prec, pml, dump, dump_pml, slices);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
}
