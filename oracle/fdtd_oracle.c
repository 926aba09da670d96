/*
 * CPU oracle for the FDTD Yee time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is a from-scratch C restatement of the reference solver's hot
 * path, used only by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker.  The product path (fdtd_method_amd/) never
 * links, loads or calls it.
 *
 * Semantics followed (reference file:line):
 *   constants            include/Constants.h:6,11   C = 3e10, PI = 3.14159265358 (truncated)
 *   coefficients         src/FDTD/FDTD.cpp:43-53    (identical: src/FDTD_kokkos/FDTD_kokkos.cpp:45-55)
 *   periodic wrap        include/FDTD/FDTD.h:24-31  (identical: include/FDTD_kokkos/kokkos_functors.h:13-20)
 *   E update             include/FDTD_kokkos/kokkos_functors.h:64-90  (Jx->Ex, Jy->Ey, Jz->Ez)
 *   B update             include/FDTD_kokkos/kokkos_functors.h:128-151 (== src/FDTD/FDTD.cpp:99-130)
 *   one step             src/FDTD_kokkos/FDTD_kokkos.cpp:91-104        (B/2, E, B/2)
 *   source timing        perf-tests/sample/sample.cpp:74               J sampled at t = (n+1)*dt
 *
 * The E update follows the Kokkos functor, not src/FDTD/FDTD.cpp:85-93,
 * which adds Jx to all three E components.  The two agree whenever
 * Jx == Jy == Jz (every reference driver) or J == 0.
 *
 * Expression trees keep the reference's C++ left-to-right order, e.g.
 *   Ex = Ex + ((cur*Jx + cEy*(Bz - Bz[j-1])) - cEz*(By - By[k-1]))
 * and the file is compiled with -ffp-contract=off, so results are
 * bit-identical to the compiled reference class (no FMA).
 *
 * Pins (see tests/test_oracle.py): bit-exact against the compiled reference
 * OpenMP class and its golden dumps, known-answer single-cell source, the
 * reference's own convergence-order tests, shift equivariance, linearity and
 * discrete div(B) conservation.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define ORACLE_C 3e10            /* include/Constants.h:9  */
#define ORACLE_PI 3.14159265358  /* include/Constants.h:14 */

/* include/FDTD/FDTD.h:24-31: -1 -> N-1, N -> 0, otherwise unchanged.
 * Written with the same integer arithmetic as the reference. */
static inline int oracle_wrap(int i, int n) {
    int is_minus_one = (i < 0);
    int is_n = (i == n);
    return (n - 1) * is_minus_one + i * !(is_minus_one || is_n);
}

/* Coefficients in the reference's order, src/FDTD/FDTD.cpp:43-53.
 * out = {cEx, cEy, cEz, cBx, cBy, cBz, cur} */
void oracle_coefficients(double dx, double dy, double dz, double dt, double out[7]) {
    const double cdt = ORACLE_C * dt;
    out[0] = cdt / dx;
    out[1] = cdt / dy;
    out[2] = cdt / dz;
    out[3] = cdt / (2.0 * dx);
    out[4] = cdt / (2.0 * dy);
    out[5] = cdt / (2.0 * dz);
    out[6] = -4.0 * ORACLE_PI * dt;
}

/* One half-step B update, kokkos_functors.h:128-151 (B += c*dt/(2d) * (-curl E)). */
void oracle_update_B(int Ni, int Nj, int Nk, const double c[7],
                     const double* Ex, const double* Ey, const double* Ez,
                     double* Bx, double* By, double* Bz) {
    const double cBx = c[3], cBy = c[4], cBz = c[5];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Nk; ++k) {
        for (int j = 0; j < Nj; ++j) {
            for (int i = 0; i < Ni; ++i) {
                const int idx = i + j * Ni + k * Ni * Nj;
                const int ip = oracle_wrap(i + 1, Ni);
                const int jp = oracle_wrap(j + 1, Nj);
                const int kp = oracle_wrap(k + 1, Nk);
                const int i_next = ip + j * Ni + k * Ni * Nj;
                const int j_next = i + jp * Ni + k * Ni * Nj;
                const int k_next = i + j * Ni + kp * Ni * Nj;
                Bx[idx] += cBz * (Ey[k_next] - Ey[idx]) - cBy * (Ez[j_next] - Ez[idx]);
                By[idx] += cBx * (Ez[i_next] - Ez[idx]) - cBz * (Ex[k_next] - Ex[idx]);
                Bz[idx] += cBy * (Ex[j_next] - Ex[idx]) - cBx * (Ey[i_next] - Ey[idx]);
            }
        }
    }
}

/* One full-step E update, kokkos_functors.h:64-90 (E += c*dt/d * curl B - 4*pi*dt*J).
 * J pointers may be NULL, meaning J == 0 (adding cur*0 == -0 is exact). */
void oracle_update_E(int Ni, int Nj, int Nk, const double c[7],
                     double* Ex, double* Ey, double* Ez,
                     const double* Bx, const double* By, const double* Bz,
                     const double* Jx, const double* Jy, const double* Jz) {
    const double cEx = c[0], cEy = c[1], cEz = c[2], cur = c[6];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Nk; ++k) {
        for (int j = 0; j < Nj; ++j) {
            for (int i = 0; i < Ni; ++i) {
                const int idx = i + j * Ni + k * Ni * Nj;
                const int im = oracle_wrap(i - 1, Ni);
                const int jm = oracle_wrap(j - 1, Nj);
                const int km = oracle_wrap(k - 1, Nk);
                const int i_pred = im + j * Ni + k * Ni * Nj;
                const int j_pred = i + jm * Ni + k * Ni * Nj;
                const int k_pred = i + j * Ni + km * Ni * Nj;
                const double jx = Jx ? Jx[idx] : 0.0;
                const double jy = Jy ? Jy[idx] : 0.0;
                const double jz = Jz ? Jz[idx] : 0.0;
                Ex[idx] += cur * jx + cEy * (Bz[idx] - Bz[j_pred]) - cEz * (By[idx] - By[k_pred]);
                Ey[idx] += cur * jy + cEz * (Bx[idx] - Bx[k_pred]) - cEx * (Bz[idx] - Bz[i_pred]);
                Ez[idx] += cur * jz + cEx * (By[idx] - By[i_pred]) - cEy * (Bx[idx] - Bx[j_pred]);
            }
        }
    }
}

/* nsteps calls of update_fields(), FDTD_kokkos.cpp:91-104: B/2, E, B/2 with
 * a constant J (the reference's J persists between steps). */
void oracle_update_fields(int Ni, int Nj, int Nk, double dx, double dy, double dz,
                          double dt, int nsteps, double* E[3], double* B[3],
                          const double* J[3]) {
    double c[7];
    oracle_coefficients(dx, dy, dz, dt, c);
    for (int s = 0; s < nsteps; ++s) {
        oracle_update_B(Ni, Nj, Nk, c, E[0], E[1], E[2], B[0], B[1], B[2]);
        oracle_update_E(Ni, Nj, Nk, c, E[0], E[1], E[2], B[0], B[1], B[2],
                        J ? J[0] : NULL, J ? J[1] : NULL, J ? J[2] : NULL);
        oracle_update_B(Ni, Nj, Nk, c, E[0], E[1], E[2], B[0], B[1], B[2]);
    }
}

/* Point-source waveform, evaluated in FP64 at t = (n+1)*dt.
 *   kind 0: constant  amp
 *   kind 1: amp * sin(2*pi*(t - t0)/period + phase) * envelope,
 *           envelope = exp(-(t - t0)^2 / (2*sigma^2)) if sigma > 0, else 1.
 * Kind 1 with (period 20dt, sigma 40dt, t0 120dt) is the Gaussian-modulated
 * sinusoid of the 512^3 point-source workload.  The truncated PI of
 * include/Constants.h:14 is used for 2*pi, as the reference's own sources do
 * (perf-tests/sample/sample.cpp:27). */
double oracle_waveform(int kind, double amp, double period, double t0, double sigma,
                       double phase, double t) {
    if (kind == 0) return amp;
    const double tau = t - t0;
    double v = amp * sin(2.0 * ORACLE_PI * tau / period + phase);
    if (sigma > 0.0) v = v * exp(-(tau * tau) / (2.0 * sigma * sigma));
    return v;
}

/* nsteps steps driven by point sources: before step n every source cell's
 * J component is set to the waveform at t=(n+1)*dt (sample.cpp:66-82), all
 * other J is zero.  Sources that share a cell and component add up.
 * src layout per source: {i, j, k, comp(0..2), kind, amp, period, t0, sigma, phase}
 * given as doubles.  J arrays are caller-provided scratch (zeroed here). */
void oracle_run_point_sources(int Ni, int Nj, int Nk, double dx, double dy, double dz,
                              double dt, long long step0, int nsteps,
                              double* E[3], double* B[3], double* J[3],
                              int nsrc, const double* src) {
    double c[7];
    const size_t n = (size_t)Ni * Nj * Nk;
    oracle_coefficients(dx, dy, dz, dt, c);
    for (int d = 0; d < 3; ++d) memset(J[d], 0, n * sizeof(double));
    for (int s = 0; s < nsteps; ++s) {
        const double t = (double)(step0 + s + 1) * dt;
        for (int q = 0; q < nsrc; ++q) {
            const double* p = src + 10 * q;
            const size_t idx = (size_t)p[0] + (size_t)p[1] * Ni + (size_t)p[2] * Ni * Nj;
            J[(int)p[3]][idx] = 0.0;
        }
        for (int q = 0; q < nsrc; ++q) {
            const double* p = src + 10 * q;
            const size_t idx = (size_t)p[0] + (size_t)p[1] * Ni + (size_t)p[2] * Ni * Nj;
            J[(int)p[3]][idx] += oracle_waveform((int)p[4], p[5], p[6], p[7], p[8], p[9], t);
        }
        oracle_update_B(Ni, Nj, Nk, c, E[0], E[1], E[2], B[0], B[1], B[2]);
        oracle_update_E(Ni, Nj, Nk, c, E[0], E[1], E[2], B[0], B[1], B[2], J[0], J[1], J[2]);
        oracle_update_B(Ni, Nj, Nk, c, E[0], E[1], E[2], B[0], B[1], B[2]);
    }
}

/* Sum of squares of one field, accumulated in FP64 (energy diagnostic check). */
double oracle_sum_squares(const double* a, size_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long long i = 0; i < (long long)n; ++i) s += a[i] * a[i];
    return s;
}
