// C++ drop-in for the reference's FDTD class, on libfdtd_mi355x.so.
//
// Replaces (same public surface, same error types and messages):
//   FDTD_openmp::FDTD   include/FDTD/FDTD.h:8-41,  src/FDTD/FDTD.cpp:3-157
//   FDTD_kokkos::FDTD   include/FDTD_kokkos/FDTD_kokkos.h:7-36, src/FDTD_kokkos/FDTD_kokkos.cpp:5-104
// and the shared types of include/FP.h:3, include/Constants.h:4-12,
// include/Enums.h:4-7, include/Structures.h:7-39.
//
// A caller switches by including this header instead of "FDTD.h" /
// "FDTD_kokkos.h" and naming FDTD_mi355x::FDTD.  Differences a caller can see:
//  * get_field() returns a host mirror (FDTD_mi355x::Field) of the device
//    array.  f[idx] and f(idx) read (downloading lazily) and assign; E/B
//    writes are uploaded before the next step, single-cell J writes are
//    sent to the device as sparse currents (the drivers' box source,
//    perf-tests/sample/sample.cpp:67-81).
//  * update_fields() defers: steps accumulate, with the single-cell J
//    writes made between them, and run as one library call when a field is
//    read or written (other than a single J cell), at zeroed_currents(),
//    synchronize(), update_fields(n), destruction, or after 256 steps.  A
//    one-step loop therefore runs as temporally blocked passes, with the
//    results of stepping one call at a time, bit for bit.
//  * The constructor takes an optional storage precision (FP64 default).
//  * update_fields(int n) advances n steps in one call.
//  * device_field(Component) gives the device array (the Kokkos class's
//    get_field() returns device Views, filled by its driver on the device,
//    perf-tests/kokkos_sample/kokkos_sample.cpp:82-108).  From the first
//    such call on, update_fields() runs at once (device writes between
//    calls are invisible to the deferral) and host mirrors refresh after
//    every step, zeroed_currents() and synchronize().
//  * The reference's protected members are here for subclasses that write
//    their own update_fields() (as FDTD_PML does, src/FDTD/FDTD_PML.cpp:343-365):
//    update_E() / update_B() over begin_main_* .. end_main_*,
//    applyPeriodicBoundary, the coefficients, Ni..dt, parameters, and Field
//    references Jx..Bz to the host mirrors.  The accessors of the earlier
//    public surface are get_parameters() and get_dt().
//  * Only single-threaded use per object, like the reference.
//
// Header-only: it needs the C-ABI (fdtd_mi355x.h) and links against
// libfdtd_mi355x.so; no HIP header is required by callers.
#pragma once

#include <array>
#include <cstddef>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fdtd_mi355x.h"

typedef double FP;  // include/FP.h:3

namespace FDTD_const {  // include/Constants.h:4-12
const double C = 3e10;
const double R = 1e-12;
const double EPS0 = 1.0;
const double MU0 = EPS0;
const double N = 4.0;
const double PI = 3.14159265358;
}  // namespace FDTD_const

namespace FDTD_enums {  // include/Enums.h:4-7
enum class Component { EX, EY, EZ, BX, BY, BZ, JX, JY, JZ };
enum class Axis { X, Y, Z };
}  // namespace FDTD_enums

namespace FDTD_struct {  // include/Structures.h:7-39
struct SelectedFields {
    FDTD_enums::Component selected_E;
    FDTD_enums::Component selected_B;
};

struct CurrentParameters {
    int period;
    int m;
    FP dt;
    int iterations;
    FP period_x = static_cast<FP>(m) * FDTD_const::C;
    FP period_y = static_cast<FP>(m) * FDTD_const::C;
    FP period_z = static_cast<FP>(m) * FDTD_const::C;
};

struct Parameters {
    int Ni;
    int Nj;
    int Nk;
    FP ax;
    FP bx;
    FP ay;
    FP by;
    FP az;
    FP bz;
    FP dx;
    FP dy;
    FP dz;
};
}  // namespace FDTD_struct

namespace FDTD_mi355x {

using namespace FDTD_enums;
using namespace FDTD_struct;

enum class Precision { FP64 = FDTD_FP64, FP32 = FDTD_FP32, FP16 = FDTD_FP16 };

// Converts a C-ABI status into the reference's exception types.
inline void check_status(int rc) {
    if (rc == FDTD_OK) return;
    const std::string msg = fdtd_last_error();
    if (rc == FDTD_ERR_INVALID_PARAMETERS) throw std::invalid_argument(msg);
    if (rc == FDTD_ERR_INVALID_COMPONENT) throw std::logic_error(msg);
    throw std::runtime_error(msg);
}

class FDTD;

// Host mirror of one device field (the reference's Field&).
class Field {
public:
    class Ref {
    public:
        Ref(Field* f, std::size_t i) : f_(f), i_(i) {}
        operator double() const { return f_->get(i_); }
        Ref& operator=(double v) {
            f_->set(i_, v);
            return *this;
        }
        Ref& operator=(const Ref& o) { return *this = static_cast<double>(o); }
        Ref& operator+=(double v) { return *this = static_cast<double>(*this) + v; }
        Ref& operator-=(double v) { return *this = static_cast<double>(*this) - v; }
        Ref& operator*=(double v) { return *this = static_cast<double>(*this) * v; }

    private:
        Field* f_;
        std::size_t i_;
    };

    Field(FDTD* owner, int comp) : owner_(owner), comp_(comp) {}

    Ref operator[](std::size_t i) { return Ref(this, i); }
    Ref operator()(std::size_t i) { return Ref(this, i); }
    double operator[](std::size_t i) const { return const_cast<Field*>(this)->get(i); }
    double operator()(std::size_t i) const { return const_cast<Field*>(this)->get(i); }
    std::size_t size() const;

    // Mirror pointer for bulk access; the field is treated as modified.
    double* data() {
        pull();
        dirty_ = true;
        return host_.data();
    }

    inline double get(std::size_t i);
    inline void set(std::size_t i, double v);
    inline void pull();
    inline void push();
    void invalidate() {
        if (!dirty_) valid_ = false;
    }
    void drop() {
        valid_ = false;
        dirty_ = false;
    }

private:
    FDTD* owner_;
    int comp_;
    std::vector<double> host_;
    bool valid_ = false;
    bool dirty_ = false;
};

class FDTD {
public:
    // FDTD(Parameters, FP dt) -- src/FDTD/FDTD.cpp:3-61.
    FDTD(Parameters parameters, FP dt, Precision precision = Precision::FP64)
        : FDTD(parameters, dt, precision, -1.0) {}
    FDTD(const FDTD&) = delete;
    FDTD& operator=(const FDTD&) = delete;
    virtual ~FDTD() {
        try {
            run_pending();
        } catch (...) {
        }
        fdtd_destroy(grid_);
    }

    // Field& get_field(Component) -- src/FDTD/FDTD.cpp:138-151.
    Field& get_field(Component c) {
        const int id = static_cast<int>(c);
        if (id < 0 || id > 8) throw std::logic_error("ERROR: Invalid field component");
        return fields_[id];
    }

    // virtual void update_fields() -- src/FDTD/FDTD.cpp:153-157: one step,
    // deferred (see the header comment).
    virtual void update_fields() {
        if (device_mode_) {
            update_fields(1);
            return;
        }
        defer_step();
    }

    // n steps in one call (after any deferred ones).
    void update_fields(int nsteps) {
        flush();
        check_status(fdtd_update_fields(grid_, nsteps));
        invalidate_mirrors();
    }

    // void zeroed_currents() -- src/FDTD/FDTD.cpp:132-136.
    void zeroed_currents() {
        run_pending();
        for (int c = 6; c < 9; ++c) fields_[c].drop();
        check_status(fdtd_zero_currents(grid_));
    }

    // Device array of a component (see the header comment and
    // fdtd_device_field): J in double (FP64 grids) or float, at a fixed
    // address; E/B in the storage type, valid until the next step.  The
    // arrays belong to the object (a Kokkos View's reference count does not
    // exist here): no pointer outlives it.
    void* device_field(Component c, int* elem_bytes = nullptr, double* scale = nullptr) {
        const int id = static_cast<int>(c);
        if (id < 0 || id > 8) throw std::logic_error("ERROR: Invalid field component");
        flush();
        void* p = nullptr;
        check_status(fdtd_device_field(grid_, id, &p, elem_bytes, scale));
        device_mode_ = true;
        for (auto& f : fields_) f.invalidate();
        return p;
    }
    // Typed form: throws std::runtime_error unless T has the element size.
    template <typename T>
    T* device_field_as(Component c) {
        int eb = 0;
        void* p = device_field(c, &eb);
        if (static_cast<std::size_t>(eb) != sizeof(T))
            throw std::runtime_error("device_field: elements are " + std::to_string(eb) +
                                     " bytes, not " + std::to_string(sizeof(T)));
        return static_cast<T*>(p);
    }
    // The grid's HIP stream (hipStream_t): the library queues its device
    // work there in call order.
    void* stream() {
        void* s = nullptr;
        check_status(fdtd_get_stream(grid_, &s));
        return s;
    }

    // Runs the deferred steps now.
    void run_pending() {
        if (pending_ == 0) return;
        const int steps = pending_;
        pending_ = 0;
        // groups 0..steps-1 precede the steps; the last group follows them
        std::size_t before = 0;
        for (int s = 0; s < steps; ++s) before += static_cast<std::size_t>(w_counts_[s]);
        const int rc = fdtd_update_fields_with_currents(
            grid_, steps, w_counts_.data(), w_index_.data(), w_comp_.data(), w_value_.data());
        const int trailing = w_counts_[steps];
        int rc2 = FDTD_OK;
        if (rc == FDTD_OK && trailing > 0)
            rc2 = fdtd_set_current_values(grid_, trailing, w_index_.data() + before,
                                          w_comp_.data() + before, w_value_.data() + before);
        w_counts_.clear();
        w_index_.clear();
        w_comp_.clear();
        w_value_.clear();
        invalidate_mirrors();
        check_status(rc);
        check_status(rc2);
    }

    // Deferred steps not yet run.
    int pending_steps() const { return pending_; }

    // {steps, temporally blocked passes, single-step sweeps, B/2 prologues}
    // launched so far (after any deferred steps run).
    std::array<long long, 4> launch_stats() {
        run_pending();
        std::array<long long, 4> s{};
        check_status(fdtd_get_stats(grid_, s.data()));
        return s;
    }

    // Probes (no counterpart in the reference, whose drivers read single
    // values of get_field() after each update_fields()): the FP64 values of
    // the given cells after every later step, recorded on the device into
    // `capacity` rows, read back with read_probes().  Deferred steps run
    // first, so they are not recorded.
    struct ProbeCell {
        int i, j, k;
        Component c;
    };
    void set_probes(const std::vector<ProbeCell>& cells, long long capacity) {
        run_pending();
        std::vector<long long> index;
        std::vector<int> comp;
        for (const auto& p : cells) {
            index.push_back(p.i + static_cast<long long>(parameters.Ni) *
                                      (p.j + static_cast<long long>(parameters.Nj) * p.k));
            comp.push_back(static_cast<int>(p.c));
        }
        check_status(fdtd_set_probes(grid_, static_cast<int>(cells.size()), index.data(),
                                     comp.data(), capacity));
        nprobes_ = cells.size();
    }
    long long probe_rows() {
        run_pending();
        long long rows = 0;
        check_status(fdtd_probe_rows(grid_, &rows));
        return rows;
    }
    // All rows recorded so far, row-major (probe_rows() x cells).
    std::vector<double> read_probes() {
        const long long rows = probe_rows();
        std::vector<double> out(static_cast<std::size_t>(rows) * nprobes_);
        check_status(fdtd_read_probes(grid_, 0, rows, out.data()));
        return out;
    }
    // Energy records: the six sums of squares (Ex..Bz, physical units)
    // after every step that is a multiple of `every`, reduced on the device.
    void set_energy_interval(int every, long long capacity) {
        run_pending();
        check_status(fdtd_set_energy_interval(grid_, every, capacity));
    }
    std::vector<std::pair<long long, std::array<double, 6>>> energy_records() {
        run_pending();
        long long rows = 0;
        check_status(fdtd_energy_rows(grid_, &rows));
        std::vector<long long> steps(static_cast<std::size_t>(rows));
        std::vector<double> sums(static_cast<std::size_t>(rows) * 6);
        check_status(fdtd_read_energy(grid_, 0, rows, steps.data(), sums.data()));
        std::vector<std::pair<long long, std::array<double, 6>>> out(steps.size());
        for (std::size_t r = 0; r < steps.size(); ++r) {
            out[r].first = steps[r];
            for (int c = 0; c < 6; ++c) out[r].second[c] = sums[6 * r + c];
        }
        return out;
    }

    // Temporal blocking (no counterpart in the reference): L steps per pass
    // over HBM with the wavefront-strip kernel (0 = defaults; L = 0 turns it
    // off), or the fastest of single sweeps and blocked passes timed on
    // scratch copies of the state.  Bit-identical results either way.
    void set_time_blocking_strips(int L, int V = 0, int R = 0, int W = 0, int kchunk = 0) {
        check_status(fdtd_set_time_blocking_strips(grid_, L, V, R, W, kchunk));
    }
    float autotune_blocking(int reps = 3) {
        flush();
        int cfg[5];
        float ms = 0.0f;
        check_status(fdtd_autotune_blocking(grid_, reps, cfg, &ms));
        return ms;
    }

    std::size_t cells() const { return cells_; }
    fdtd_grid* handle() { return grid_; }
    const Parameters& get_parameters() const { return parameters; }
    FP get_dt() const { return dt; }

    void synchronize() {
        run_pending();
        check_status(fdtd_synchronize(grid_));
        if (device_mode_)
            for (auto& f : fields_) f.invalidate();
    }

    // FP64 copy of plane k of a component (Ni*Nj values, index i + j*Ni),
    // after pending host writes are uploaded.
    std::vector<double> get_plane(Component c, int k) {
        flush();
        std::vector<double> out(static_cast<std::size_t>(parameters.Ni) * parameters.Nj);
        check_status(fdtd_download_plane(grid_, static_cast<int>(c), k, out.data(), out.size()));
        return out;
    }

protected:
    // pml_percent < 0: periodic grid (fdtd_create); otherwise the PML
    // subclass's grid (fdtd_create_pml).
    FDTD(Parameters params, FP time_step, Precision precision, FP pml_percent)
        : fields_(make_fields(this)),
          Jx(fields_[6]), Jy(fields_[7]), Jz(fields_[8]),
          Ex(fields_[0]), Ey(fields_[1]), Ez(fields_[2]),
          Bx(fields_[3]), By(fields_[4]), Bz(fields_[5]),
          parameters(params) {
        fdtd_parameters p{params.Ni, params.Nj, params.Nk, params.ax, params.bx, params.ay,
                          params.by, params.az, params.bz, params.dx, params.dy, params.dz};
        if (pml_percent < 0)
            check_status(fdtd_create(&p, time_step, static_cast<int>(precision), &grid_));
        else
            check_status(
                fdtd_create_pml(&p, time_step, pml_percent, static_cast<int>(precision), &grid_));
        cells_ = static_cast<std::size_t>(params.Ni) * params.Nj * params.Nk;
        // the reference's protected members, src/FDTD/FDTD.cpp:35-60
        Ni = params.Ni;
        Nj = params.Nj;
        Nk = params.Nk;
        dx = params.dx;
        dy = params.dy;
        dz = params.dz;
        dt = time_step;
        const FP cdt = FDTD_const::C * dt;
        coef_E_dx = cdt / dx;
        coef_E_dy = cdt / dy;
        coef_E_dz = cdt / dz;
        coef_B_dx = cdt / (2.0 * dx);
        coef_B_dy = cdt / (2.0 * dy);
        coef_B_dz = cdt / (2.0 * dz);
        cur_coef = -4.0 * FDTD_const::PI * dt;
        begin_main_i = 0;
        begin_main_j = 0;
        begin_main_k = 0;
        end_main_i = Ni;
        end_main_j = Nj;
        end_main_k = Nk;
    }

    // include/FDTD/FDTD.h:24-31: -1 -> N-1, N -> 0, otherwise unchanged.
    static void applyPeriodicBoundary(int& i, const int& N) {
        if (i < 0) i = N - 1;
        else if (i == N) i = 0;
    }

    // void update_E() / void update_B() -- include/FDTD/FDTD.h:33-34,
    // src/FDTD/FDTD.cpp:63-130: one E step (J paired Jx->Ex, Jy->Ey, Jz->Ez
    // as include/FDTD_kokkos/kokkos_functors.h:64-90) / one B half step
    // over the main box, on the device and in place; deferred steps and
    // host-mirror writes go first.
    void update_E() { box_update(true); }
    void update_B() { box_update(false); }

private:
    friend class Field;

    static std::vector<Field> make_fields(FDTD* owner) {
        std::vector<Field> f;
        f.reserve(9);
        for (int c = 0; c < 9; ++c) f.emplace_back(owner, c);
        return f;
    }

    void box_update(bool e_step) {
        flush();
        const int lo[3] = {begin_main_i, begin_main_j, begin_main_k};
        const int hi[3] = {end_main_i, end_main_j, end_main_k};
        check_status(e_step ? fdtd_update_e_box(grid_, lo, hi) : fdtd_update_b_box(grid_, lo, hi));
        invalidate_mirrors();
    }

    // E/B mirrors after device work; J too once the caller holds device
    // addresses.
    void invalidate_mirrors() {
        for (int c = 0; c < 9; ++c)
            if (c < 6 || device_mode_) fields_[c].invalidate();
    }

    static constexpr int kMaxPending = 256;

    void flush() {
        run_pending();
        for (auto& f : fields_) f.push();
    }

    void defer_step() {
        for (int c = 0; c < 6; ++c) fields_[c].push();  // E/B writes precede the step
        if (pending_ == 0) w_counts_.assign(1, 0);
        ++pending_;
        w_counts_.push_back(0);
        if (pending_ >= kMaxPending) run_pending();
    }

    // A single J cell write: before the next deferred step when steps are
    // pending, else at once.
    void write_current(long long index, int comp, double v) {
        if (pending_ == 0) {
            check_status(fdtd_set_current_values(grid_, 1, &index, &comp, &v));
            return;
        }
        w_index_.push_back(index);
        w_comp_.push_back(comp);
        w_value_.push_back(v);
        w_counts_.back() += 1;
    }

    int pending_ = 0;
    bool device_mode_ = false;
    std::size_t nprobes_ = 0;
    std::vector<int> w_counts_;  // writes before deferred step s; the last: after them
    std::vector<long long> w_index_;
    std::vector<int> w_comp_;
    std::vector<double> w_value_;

    fdtd_grid* grid_ = nullptr;
    std::size_t cells_ = 0;
    std::vector<Field> fields_;  // before the references below, which bind to it

protected:
    // The reference's protected data, include/FDTD/FDTD.h:10-22.  Field
    // references are the host mirrors get_field() returns.
    Field &Jx, &Jy, &Jz;
    Field &Ex, &Ey, &Ez;
    Field &Bx, &By, &Bz;
    Parameters parameters;
    int Ni = 0, Nj = 0, Nk = 0;
    double dx = 0, dy = 0, dz = 0, dt = 0;
    double cur_coef = 0;
    double coef_E_dx = 0, coef_E_dy = 0, coef_E_dz = 0;
    double coef_B_dx = 0, coef_B_dy = 0, coef_B_dz = 0;
    int begin_main_i = 0, begin_main_j = 0, begin_main_k = 0;
    int end_main_i = 0, end_main_j = 0, end_main_k = 0;
};

// FDTD_PML(Parameters, FP dt, FP pml_percent) -- include/FDTD/FDTD_PML.h:10-44,
// src/FDTD/FDTD_PML.cpp:205-365: the outer (int)(N*pml_percent) cells of
// every axis form an absorbing layer; J acts on the main box only.
// pml_percent must lie in [0, 0.5] (std::runtime_error otherwise).
class FDTD_PML : public FDTD {
public:
    FDTD_PML(Parameters params, FP time_step, FP pml_percent,
             Precision precision = Precision::FP64)
        : FDTD(params, time_step, precision, pml_percent) {
        // main box = interior, src/FDTD/FDTD_PML.cpp:259-268
        const std::array<int, 3> p = pml_widths();
        begin_main_i = p[0];
        begin_main_j = p[1];
        begin_main_k = p[2];
        end_main_i = Ni - p[0];
        end_main_j = Nj - p[1];
        end_main_k = Nk - p[2];
    }

    using FDTD::update_fields;
    // void update_fields() override -- src/FDTD/FDTD_PML.cpp:343-365
    // (deferred like FDTD::update_fields()).
    void update_fields() override { FDTD::update_fields(); }

    // Layer widths {p_i, p_j, p_k} (FDTD_PML.cpp:254-256).
    std::array<int, 3> pml_widths() const {
        std::array<int, 3> w{};
        check_status(fdtd_get_pml(grid(), w.data(), nullptr));
        return w;
    }

private:
    const fdtd_grid* grid() const { return const_cast<FDTD_PML*>(this)->handle(); }
};

// Writes plane k of a component as the CSV slice read by the reference's
// plotting script (python_script_legend/visualization.py:10-45, which reads
// it with pandas.read_csv(sep=';') and turns decimal commas into points):
// a header line of column indices 0..Ni-1, then one line per row j with the
// Ni values of that row, ';'-separated, in %.9g with a decimal comma.
inline void write_csv_slice(FDTD& m, Component c, int k, const std::string& path) {
    const std::vector<double> v = m.get_plane(c, k);
    const int Ni = m.get_parameters().Ni, Nj = m.get_parameters().Nj;
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    for (int i = 0; i < Ni; ++i) std::fprintf(f, i ? ";%d" : "%d", i);
    std::fputc('\n', f);
    char num[64];
    for (int j = 0; j < Nj; ++j) {
        for (int i = 0; i < Ni; ++i) {
            std::snprintf(num, sizeof num, "%.9g", v[static_cast<std::size_t>(j) * Ni + i]);
            for (char* q = num; *q; ++q)
                if (*q == '.') *q = ',';
            std::fputs(num, f);
            std::fputc(i + 1 < Ni ? ';' : '\n', f);
        }
    }
    if (std::fclose(f) != 0) throw std::runtime_error("cannot write " + path);
}

inline std::size_t Field::size() const { return owner_->cells(); }

inline void Field::pull() {
    owner_->run_pending();
    if (valid_) return;
    host_.resize(owner_->cells());
    check_status(fdtd_download(owner_->handle(), comp_, host_.data(), host_.size()));
    valid_ = true;
}

inline void Field::push() {
    if (!dirty_) return;
    check_status(fdtd_upload(owner_->handle(), comp_, host_.data(), host_.size()));
    dirty_ = false;
}

inline double Field::get(std::size_t i) {
    pull();
    return host_.at(i);
}

inline void Field::set(std::size_t i, double v) {
    if (comp_ >= FDTD_JX && !dirty_) {
        if (i >= owner_->cells()) throw std::out_of_range("field index out of range");
        owner_->write_current(static_cast<long long>(i), comp_, v);
        valid_ = false;  // the mirror no longer reflects J (refreshed on read)
        return;
    }
    pull();
    host_.at(i) = v;
    dirty_ = true;
}

}  // namespace FDTD_mi355x
