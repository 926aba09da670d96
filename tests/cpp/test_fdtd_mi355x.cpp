// Unit tests of the C++ drop-in class FDTD_mi355x::FDTD, restating the
// reference's gtest suite (unit-tests/test_FDTD_method.cpp:71-203 and
// unit-tests/test_FDTD_method_kokkos.cpp:71-201): second-order convergence
// of a plane wave along each axis, err(N)/err(2N) = 4 +- 0.1, plus the
// class's error behaviour (src/FDTD/FDTD.cpp:5-7, :149).
//
// Usage: test_fdtd_mi355x [fp64|fp32]   (exit status 0 = all passed)
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "FDTD_mi355x.hpp"

using namespace FDTD_mi355x;

namespace {

int g_failures = 0;
Precision g_precision = Precision::FP64;

#define EXPECT_TRUE(cond)                                                        \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);      \
            ++g_failures;                                                        \
        }                                                                        \
    } while (0)

#define EXPECT_NEAR(a, b, tol)                                                             \
    do {                                                                                   \
        const double a_ = (a), b_ = (b);                                                   \
        if (!(std::fabs(a_ - b_) <= (tol))) {                                              \
            std::printf("  FAILED %s:%d: %s = %.6f, expected %.6f +- %g\n", __FILE__,      \
                        __LINE__, #a, a_, b_, (double)(tol));                              \
            ++g_failures;                                                                  \
        }                                                                                  \
    } while (0)

const FP kTime = 5e-13;  // unit-tests/test_FDTD_method.cpp:7

// Initial profile and exact solution of the travelling wave
// (unit-tests/test_FDTD_method.cpp:9-16).
FP initial(FP x, FP a, FP b) { return std::sin(2.0 * FDTD_const::PI * (x - a) / (b - a)); }
FP exact(FP x, FP t, FP a, FP b) {
    return std::sin(2.0 * FDTD_const::PI * (x - a - FDTD_const::C * t) / (b - a));
}

// Propagation axis and sign of a (E, B) pair: the wave moves toward +axis
// when E = sign * f and B = f (src/FDTD/test_FDTD.cpp:53-80).
struct Wave {
    Axis axis;
    FP sign;
};

Wave wave_of(Component e, Component b) {
    if (e == Component::EY && b == Component::BZ) return {Axis::X, 1.0};
    if (e == Component::EZ && b == Component::BY) return {Axis::X, -1.0};
    if (e == Component::EX && b == Component::BZ) return {Axis::Y, -1.0};
    if (e == Component::EZ && b == Component::BX) return {Axis::Y, 1.0};
    if (e == Component::EX && b == Component::BY) return {Axis::Z, 1.0};
    if (e == Component::EY && b == Component::BX) return {Axis::Z, -1.0};
    throw std::logic_error("ERROR: invalid selected fields");
}

// Fill E at grid points and B half a cell further along the axis
// (src/FDTD/test_FDTD.cpp:5-51), writing through get_field()[index].
void fill(FDTD& m, const Parameters& p, SelectedFields sel) {
    const Wave w = wave_of(sel.selected_E, sel.selected_B);
    Field& E = m.get_field(sel.selected_E);
    Field& B = m.get_field(sel.selected_B);
    for (int k = 0; k < p.Nk; ++k)
        for (int j = 0; j < p.Nj; ++j)
            for (int i = 0; i < p.Ni; ++i) {
                const int index = i + j * p.Ni + k * p.Ni * p.Nj;
                FP x, d, a, b;
                if (w.axis == Axis::X) { x = i * p.dx; d = p.dx; a = p.ax; b = p.bx; }
                else if (w.axis == Axis::Y) { x = j * p.dy; d = p.dy; a = p.ay; b = p.by; }
                else { x = k * p.dz; d = p.dz; a = p.az; b = p.bz; }
                E[index] = w.sign * initial(x, a, b);
                B[index] = initial(x + d / 2.0, a, b);
            }
}

// Max |error| along the grid line through the origin
// (src/FDTD/test_FDTD.cpp:89-130); B is sampled half a cell along the axis.
FP line_error(FDTD& m, const Parameters& p, SelectedFields sel, Component c, FP t) {
    const Wave w = wave_of(sel.selected_E, sel.selected_B);
    const bool is_b = static_cast<int>(c) > static_cast<int>(Component::EZ);
    const FP sign = is_b ? 1.0 : w.sign;
    Field& f = m.get_field(c);
    FP err = 0.0;
    const int n = w.axis == Axis::X ? p.Ni : (w.axis == Axis::Y ? p.Nj : p.Nk);
    const FP d = w.axis == Axis::X ? p.dx : (w.axis == Axis::Y ? p.dy : p.dz);
    const FP a = w.axis == Axis::X ? p.ax : (w.axis == Axis::Y ? p.ay : p.az);
    const FP b = w.axis == Axis::X ? p.bx : (w.axis == Axis::Y ? p.by : p.bz);
    FP x = is_b ? d / 2.0 : 0.0;
    for (int q = 0; q < n; ++q, x += d) {
        const int index = w.axis == Axis::X ? q : (w.axis == Axis::Y ? q * p.Ni : q * p.Ni * p.Nj);
        err = std::max(err, std::fabs(sign * f[index] - exact(x, t, a, b)));
    }
    return err;
}

FP run_once(Component c, SelectedFields sel, int Ni, int Nj, int Nk, int iters) {
    Parameters p{Ni, Nj, Nk, 0.0, 1.0, 0.0, 2.0, 0.0, 3.0,
                 1.0 / Ni, 2.0 / Nj, 3.0 / Nk};
    FDTD m(p, kTime / iters, g_precision);
    fill(m, p, sel);
    for (int t = 0; t < iters; ++t) m.update_fields();
    return line_error(m, p, sel, c, kTime);
}

// unit-tests/test_FDTD_method.cpp:18-69: ratio of errors at N (16 steps)
// and 2N (32 steps).
FP convergence_ratio(Component c, SelectedFields sel, int Ni, int Nj, int Nk) {
    return run_once(c, sel, Ni, Nj, Nk, 16) / run_once(c, sel, 2 * Ni, 2 * Nj, 2 * Nk, 32);
}

struct Case {
    const char* name;
    Component test, e, b;
    int Ni, Nj, Nk;
};

const Case kCases[] = {
    {"x_axis_EY", Component::EY, Component::EY, Component::BZ, 16, 8, 4},
    {"x_axis_BZ", Component::BZ, Component::EY, Component::BZ, 16, 8, 4},
    {"x_axis_EZ", Component::EZ, Component::EZ, Component::BY, 16, 8, 4},
    {"x_axis_BY", Component::BY, Component::EZ, Component::BY, 16, 8, 4},
    {"y_axis_EX", Component::EX, Component::EX, Component::BZ, 8, 16, 4},
    {"y_axis_BZ", Component::BZ, Component::EX, Component::BZ, 8, 16, 4},
    {"y_axis_EZ", Component::EZ, Component::EZ, Component::BX, 8, 16, 4},
    {"y_axis_BX", Component::BX, Component::EZ, Component::BX, 8, 16, 4},
    {"z_axis_EX", Component::EX, Component::EX, Component::BY, 4, 8, 16},
    {"z_axis_BY", Component::BY, Component::EX, Component::BY, 4, 8, 16},
    {"z_axis_EY", Component::EY, Component::EY, Component::BX, 4, 8, 16},
    {"z_axis_BX", Component::BX, Component::EY, Component::BX, 4, 8, 16},
};

void test_errors() {
    Parameters bad{0, 4, 4, 0, 1, 0, 1, 0, 1, 1, 1, 1};
    bool thrown = false;
    try {
        FDTD m(bad, 1e-12, g_precision);
    } catch (const std::invalid_argument& e) {
        thrown = std::string(e.what()) == "ERROR: invalid parameters";
    }
    EXPECT_TRUE(thrown);
    Parameters ok{4, 4, 4, 0, 1, 0, 1, 0, 1, 1, 1, 1};
    thrown = false;
    try {
        FDTD m(ok, -1.0, g_precision);
    } catch (const std::invalid_argument&) {
        thrown = true;
    }
    EXPECT_TRUE(thrown);
    FDTD m(ok, 1e-12, g_precision);
    thrown = false;
    try {
        m.get_field(static_cast<Component>(9));
    } catch (const std::logic_error& e) {
        thrown = std::string(e.what()) == "ERROR: Invalid field component";
    }
    EXPECT_TRUE(thrown);
}

void test_currents() {
    // J writes read back, persist across steps, and zeroed_currents clears them.
    Parameters p{8, 6, 4, 0, 1, 0, 1, 0, 1, 1, 1, 1};
    FDTD m(p, 1e-12, g_precision);
    m.get_field(Component::JZ)[17] = 0.5;
    m.get_field(Component::JX)(3) = -2.0;
    EXPECT_TRUE(m.get_field(Component::JZ)[17] == 0.5);
    EXPECT_TRUE(m.get_field(Component::JX)[3] == -2.0);
    m.update_fields();
    EXPECT_TRUE(m.get_field(Component::JZ)[17] == 0.5);
    // Ez at the source cell after one step from zero: cur * J exactly.
    const FP cur = -4.0 * FDTD_const::PI * 1e-12;
    if (g_precision == Precision::FP64) EXPECT_TRUE(m.get_field(Component::EZ)[17] == cur * 0.5);
    m.zeroed_currents();
    EXPECT_TRUE(m.get_field(Component::JZ)[17] == 0.0);
    EXPECT_TRUE(m.get_field(Component::JX)[3] == 0.0);
}

void test_multi_step_equivalence() {
    // update_fields(n) == n calls of update_fields().
    Parameters p{16, 12, 10, 0, 1, 0, 1, 0, 1, 1.0, 1.1, 0.9};
    FDTD a(p, 5e-12, g_precision), b(p, 5e-12, g_precision);
    for (int c = 0; c < 6; ++c) {
        Field& fa = a.get_field(static_cast<Component>(c));
        Field& fb = b.get_field(static_cast<Component>(c));
        for (int i = 0; i < (int)a.cells(); ++i) {
            const double v = std::sin(0.37 * i + c) * (c < 3 ? 1.0 : 0.5);
            fa[i] = v;
            fb[i] = v;
        }
    }
    for (int s = 0; s < 5; ++s) a.update_fields();
    b.update_fields(5);
    bool same = true;
    for (int c = 0; c < 6; ++c)
        for (int i = 0; i < (int)a.cells(); ++i)
            same = same && (double)a.get_field(static_cast<Component>(c))[i] ==
                               (double)b.get_field(static_cast<Component>(c))[i];
    EXPECT_TRUE(same);
}

void test_time_blocking() {
    // Strip passes (4 steps each, plus remainders) and the blocking autotuner
    // keep the fields bitwise those of single steps.
    Parameters p{40, 24, 12, 0, 1, 0, 1, 0, 1, 1.0, 1.0, 1.0};
    const FP dt = 0.5 / (FDTD_const::C * std::sqrt(3.0));
    FDTD a(p, dt, g_precision), b(p, dt, g_precision), t(p, dt, g_precision);
    FDTD* all[3] = {&a, &b, &t};
    for (FDTD* g : all)
        for (int c = 0; c < 6; ++c) {
            Field& f = g->get_field(static_cast<Component>(c));
            for (int i = 0; i < (int)g->cells(); ++i) f[i] = std::cos(0.11 * i + 2 * c);
        }
    b.set_time_blocking_strips(4, 1, 2, 8);
    EXPECT_TRUE(t.autotune_blocking(2) > 0.0f);
    for (int n : {9, 4, 2}) {
        for (FDTD* g : all) {
            g->get_field(Component::JY)[777] = 0.5 * n;
            g->update_fields(n);
        }
    }
    bool same = true;
    for (int c = 0; c < 6; ++c)
        for (int i = 0; i < (int)a.cells(); ++i) {
            const double x = a.get_field(static_cast<Component>(c))[i];
            same = same && x == (double)b.get_field(static_cast<Component>(c))[i] &&
                   x == (double)t.get_field(static_cast<Component>(c))[i];
        }
    EXPECT_TRUE(same);
}

void test_deferred_steps() {
    // The reference drivers' loop (perf-tests/sample/sample.cpp:66-87): J
    // box writes, then update_fields(), one step at a time.  The class
    // defers the steps and runs them as one call (strip passes of 4 steps on
    // a grid this size); the fields equal stepping one call at a time.
    Parameters p{128, 64, 32, 0, 1, 0, 1, 0, 1, 1.0, 1.0, 1.0};
    const FP dt = 0.5 / (FDTD_const::C * std::sqrt(3.0));
    FDTD a(p, dt, g_precision), b(p, dt, g_precision);
    const int centre = 64 + 32 * 128 + 16 * 128 * 64;
    for (FDTD* g : {&a, &b}) {
        Field& ex = g->get_field(Component::EX);
        for (int i = 0; i < (int)g->cells(); i += 97) ex[i] = 1e-3 * std::cos(0.3 * i);
    }
    for (int t = 0; t < 14; ++t) {
        const double v = std::sin(2.0 * FDTD_const::PI * (t + 1) / 20.0);
        for (FDTD* g : {&a, &b})
            for (int d = 0; d < 2; ++d) {
                g->get_field(Component::JX)[centre + d] = v;
                g->get_field(Component::JY)[centre + d * 128] = 0.5 * v;
                g->get_field(Component::JZ)[centre + d * 128 * 64] = -v;
            }
        a.update_fields();     // deferred
        b.update_fields(1);    // one call per step
    }
    EXPECT_TRUE(a.pending_steps() == 14);
    a.zeroed_currents();
    b.zeroed_currents();
    for (int t = 0; t < 6; ++t) {
        a.update_fields();
        b.update_fields(1);
    }
    const auto sa = a.launch_stats(), sb = b.launch_stats();
    EXPECT_TRUE(sa[0] == 20 && sb[0] == 20);
    EXPECT_TRUE(sa[1] == 14 / 4 + 6 / 4);  // strip passes
    EXPECT_TRUE(sb[1] == 0);
    bool same = true;
    for (int c = 0; c < 6; ++c)
        for (int i = 0; i < (int)a.cells(); ++i)
            same = same && (double)a.get_field(static_cast<Component>(c))[i] ==
                               (double)b.get_field(static_cast<Component>(c))[i];
    EXPECT_TRUE(same);
}

void test_probes_and_energy() {
    // Probes record every deferred step of the drivers' loop on the device;
    // each row equals the fields read after stepping one call at a time,
    // and the energy records equal the sums of squares of the read fields.
    Parameters p{128, 64, 32, 0, 1, 0, 1, 0, 1, 1.0, 1.0, 1.0};
    const FP dt = 0.5 / (FDTD_const::C * std::sqrt(3.0));
    FDTD a(p, dt, g_precision), b(p, dt, g_precision);
    const int centre = 64 + 32 * 128 + 16 * 128 * 64;
    const std::vector<FDTD::ProbeCell> cells = {
        {64, 32, 16, Component::EZ}, {65, 32, 16, Component::BY}, {0, 0, 0, Component::EX},
        {127, 63, 31, Component::BZ}, {66, 30, 17, Component::EY}, {64, 33, 15, Component::BX}};
    const int steps = 10;
    a.set_probes(cells, steps);
    a.set_energy_interval(5, 2);
    std::vector<double> expect;
    std::vector<double> energy;
    for (int t = 0; t < steps; ++t) {
        const double v = std::sin(2.0 * FDTD_const::PI * (t + 1) / 20.0);
        for (FDTD* g : {&a, &b}) g->get_field(Component::JZ)[centre] = v;
        a.update_fields();   // deferred
        b.update_fields(1);
        for (const auto& c : cells)
            expect.push_back(b.get_field(c.c)[c.i + 128 * (c.j + 64 * c.k)]);
        if ((t + 1) % 5 == 0) {
            for (int c = 0; c < 6; ++c) {
                double e = 0;
                for (int i = 0; i < (int)b.cells(); ++i) {
                    const double x = b.get_field(static_cast<Component>(c))[i];
                    e += x * x;
                }
                energy.push_back(e);
            }
        }
    }
    EXPECT_TRUE(a.pending_steps() == steps);
    const std::vector<double> rows = a.read_probes();
    EXPECT_TRUE(rows.size() == expect.size());
    EXPECT_TRUE(rows == expect);
    const auto rec = a.energy_records();
    EXPECT_TRUE(rec.size() == 2 && rec[0].first == 5 && rec[1].first == 10);
    for (std::size_t r = 0; r < rec.size() && r < 2; ++r)
        for (int c = 0; c < 6; ++c)
            EXPECT_NEAR(rec[r].second[c], energy[6 * r + c], 1e-12 * energy[6 * r + c] + 1e-300);
    // a call beyond the rows left fails before any step runs
    bool thrown = false;
    try {
        a.update_fields(1);
    } catch (const std::runtime_error&) {
        thrown = true;
    }
    EXPECT_TRUE(thrown);
}

void test_pml_class() {
    // FDTD_PML (src/FDTD/FDTD_PML.cpp:205-365): layer widths, J ignored in
    // the layer, a pulse absorbed, pml_percent validated.
    Parameters p{24, 20, 16, 0, 1, 0, 1, 0, 1, 1.0, 1.0, 1.0};
    const FP dt = 0.5 / (FDTD_const::C * std::sqrt(3.0));
    FDTD_PML m(p, dt, 0.25, g_precision);
    const auto w = m.pml_widths();
    EXPECT_TRUE(w[0] == 6 && w[1] == 5 && w[2] == 4);
    m.get_field(Component::JZ)[0] = 1.0;  // a layer cell: no effect
    m.update_fields(3);
    bool zero = true;
    for (int c = 0; c < 6; ++c)
        for (int i = 0; i < (int)m.cells(); ++i)
            zero = zero && (double)m.get_field(static_cast<Component>(c))[i] == 0.0;
    EXPECT_TRUE(zero);
    m.zeroed_currents();
    const int centre = 12 + 10 * 24 + 8 * 24 * 20;
    for (int t = 0; t < 20; ++t) {
        m.get_field(Component::JZ)[centre] = std::sin(2.0 * FDTD_const::PI * (t + 1) / 20.0);
        m.update_fields();
    }
    m.zeroed_currents();
    auto energy = [&]() {
        double e = 0;
        for (int c = 0; c < 6; ++c)
            for (int i = 0; i < (int)m.cells(); ++i) {
                const double v = m.get_field(static_cast<Component>(c))[i];
                e += v * v;
            }
        return e;
    };
    const double e0 = energy();
    m.update_fields(200);
    EXPECT_TRUE(e0 > 0 && energy() < 1e-2 * e0);
    bool thrown = false;
    try {
        FDTD_PML bad(p, dt, 0.7, g_precision);
    } catch (const std::runtime_error&) {
        thrown = true;
    }
    EXPECT_TRUE(thrown);
}

// A subclass written against the reference's protected interface
// (include/FDTD/FDTD.h:10-34), composing update_fields() from update_B() and
// update_E() as FDTD_PML does (src/FDTD/FDTD_PML.cpp:343-365).
class ThreeSweep : public FDTD {
public:
    ThreeSweep(Parameters p, FP dt, Precision prec) : FDTD(p, dt, prec) {}
    void update_fields() override {
        update_B();
        update_E();
        update_B();
    }
    bool members_ok(const Parameters& p, FP time_step) {
        const FP cdt = FDTD_const::C * time_step;  // src/FDTD/FDTD.cpp:43-53
        int a = -1, b = Ni, c = 3;
        applyPeriodicBoundary(a, Ni);
        applyPeriodicBoundary(b, Ni);
        applyPeriodicBoundary(c, Ni);
        return Ni == p.Ni && Nj == p.Nj && Nk == p.Nk && dx == p.dx && dt == time_step &&
               coef_E_dy == cdt / p.dy && coef_B_dz == cdt / (2.0 * p.dz) &&
               cur_coef == -4.0 * FDTD_const::PI * time_step && begin_main_i == 0 &&
               end_main_k == p.Nk && a == Ni - 1 && b == 0 && c == 3 &&
               &Ex == &get_field(Component::EX) && &Jz == &get_field(Component::JZ) &&
               parameters.Nj == p.Nj;
    }
    // update_B() over a caller-chosen main box
    void half_b_in(int i0, int i1) {
        begin_main_i = i0;
        end_main_i = i1;
        update_B();
        begin_main_i = 0;
        end_main_i = Ni;
    }
};

class InteriorPml : public FDTD_PML {
public:
    InteriorPml(Parameters p, FP dt, Precision prec) : FDTD_PML(p, dt, 0.25, prec) {}
    std::array<int, 6> main_box() const {
        return {begin_main_i, end_main_i, begin_main_j, end_main_j, begin_main_k, end_main_k};
    }
};

void test_subclass_hooks() {
    // update_B(); update_E(); update_B() in a subclass equals the class's
    // own steps bit for bit, with single-cell J written between steps
    // through the protected Field references; the protected members hold
    // the reference's values; FDTD_PML's main box is its interior
    // (src/FDTD/FDTD_PML.cpp:259-268).
    Parameters p{40, 24, 16, 0, 1, 0, 1, 0, 1, 1.0, 1.3, 0.7};
    const FP dt = 0.4 / (FDTD_const::C * std::sqrt(3.0));
    ThreeSweep s(p, dt, g_precision);
    FDTD f(p, dt, g_precision);
    EXPECT_TRUE(s.members_ok(p, dt));
    for (FDTD* g : {static_cast<FDTD*>(&s), &f}) {
        Field& ex = g->get_field(Component::EX);
        Field& bz = g->get_field(Component::BZ);
        for (int i = 0; i < (int)g->cells(); i += 7) {
            ex[i] = 1e-3 * std::cos(0.1 * i);
            bz[i] = 2e-3 * std::sin(0.07 * i);
        }
    }
    const int centre = 20 + 12 * 40 + 8 * 40 * 24;
    for (int t = 0; t < 9; ++t) {
        for (FDTD* g : {static_cast<FDTD*>(&s), &f}) {
            g->get_field(Component::JX)[centre] = std::sin(0.3 * (t + 1));
            g->get_field(Component::JZ)[centre + 1] = std::cos(0.2 * (t + 1));
        }
        s.update_fields();
        f.update_fields();
    }
    bool same = true;
    for (int c = 0; c < 6; ++c)
        for (int i = 0; i < (int)f.cells(); ++i)
            same = same && (double)s.get_field(static_cast<Component>(c))[i] ==
                               (double)f.get_field(static_cast<Component>(c))[i];
    EXPECT_TRUE(same);
    // a box update leaves the cells outside the box alone
    const double outside = s.get_field(Component::BX)[0 + 3 * 40];
    const double inside = s.get_field(Component::BX)[20 + 3 * 40 + 8 * 40 * 24];
    s.half_b_in(10, 30);
    EXPECT_TRUE((double)s.get_field(Component::BX)[0 + 3 * 40] == outside);
    EXPECT_TRUE((double)s.get_field(Component::BX)[20 + 3 * 40 + 8 * 40 * 24] != inside);
    InteriorPml m(p, dt, g_precision);
    const auto b = m.main_box();
    const auto w = m.pml_widths();
    EXPECT_TRUE(b[0] == w[0] && b[1] == p.Ni - w[0] && b[2] == w[1] && b[3] == p.Nj - w[1] &&
                b[4] == w[2] && b[5] == p.Nk - w[2]);
}

void test_device_access() {
    // device_field(): typed access checks the element size; once device
    // addresses are out, update_fields() runs at once (no deferral) and
    // host-mirror J writes still reach the steps, so the results equal a
    // class that never handed out an address.
    Parameters p{40, 24, 16, 0, 1, 0, 1, 0, 1, 1.0, 1.0, 1.0};
    const FP dt = 0.4 / (FDTD_const::C * std::sqrt(3.0));
    FDTD a(p, dt, g_precision), b(p, dt, g_precision);
    bool thrown = false;
    try {
        if (g_precision == Precision::FP64) a.device_field_as<float>(Component::JX);
        else a.device_field_as<double>(Component::JX);
    } catch (const std::runtime_error&) {
        thrown = true;
    }
    EXPECT_TRUE(thrown);
    int eb = 0;
    EXPECT_TRUE(a.device_field(Component::JZ, &eb) != nullptr);
    EXPECT_TRUE(eb == (g_precision == Precision::FP64 ? 8 : 4));
    EXPECT_TRUE(a.stream() != nullptr);
    const int centre = 20 + 12 * 40 + 8 * 40 * 24;
    for (int t = 0; t < 6; ++t) {
        for (FDTD* g : {&a, &b}) g->get_field(Component::JZ)[centre] = std::sin(0.4 * (t + 1));
        a.update_fields();
        b.update_fields();
        EXPECT_TRUE(a.pending_steps() == 0);
    }
    a.zeroed_currents();
    b.zeroed_currents();
    a.update_fields(3);
    b.update_fields(3);
    bool same = true;
    for (int c = 0; c < 9; ++c)
        for (int i = 0; i < (int)a.cells(); ++i)
            same = same && (double)a.get_field(static_cast<Component>(c))[i] ==
                               (double)b.get_field(static_cast<Component>(c))[i];
    EXPECT_TRUE(same);
}

}  // namespace

int main(int argc, char** argv) {
    std::setvbuf(stdout, nullptr, _IOLBF, 0);  // each result line reaches a log at once
    if (argc > 1 && std::strcmp(argv[1], "fp32") == 0) g_precision = Precision::FP32;
    const char* prec_name = g_precision == Precision::FP64 ? "fp64" : "fp32";
    int ran = 0;
    for (const Case& c : kCases) {
        const int before = g_failures;
        const FP ratio = convergence_ratio(c.test, SelectedFields{c.e, c.b}, c.Ni, c.Nj, c.Nk);
        EXPECT_NEAR(ratio, 4.0, 0.1);
        std::printf("[%s] Convergence.%s ratio=%.4f %s\n", prec_name, c.name, ratio,
                    g_failures == before ? "OK" : "FAIL");
        ++ran;
    }
    test_errors();
    test_currents();
    test_multi_step_equivalence();
    test_pml_class();
    test_time_blocking();
    test_deferred_steps();
    test_probes_and_energy();
    test_subclass_hooks();
    test_device_access();
    ran += 9;
    std::printf("%d tests, %d failures\n", ran, g_failures);
    return g_failures == 0 ? 0 : 1;
}
