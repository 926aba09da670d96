// C entry points over the reference's own FDTD_openmp::FDTD class -- TEST
// INFRASTRUCTURE ONLY.  Compiled together with the reference sources
// src/FDTD/FDTD.cpp (read in place from the reference checkout) by
// oracle/Makefile into oracle/_ref/libfdtd_ref.so, so that the oracle and the
// golden fixtures can be checked against the reference's actual machine code.
//
// Wrapped interface: include/FDTD/FDTD.h:8-41
//   FDTD(Parameters, FP dt)            -> ref_create
//   Field& get_field(Component)        -> ref_get / ref_set / ref_set_value
//                                         (component ordinals of include/Enums.h:5)
//   virtual void update_fields()       -> ref_update_fields (called nsteps times)
//   void zeroed_currents()             -> ref_zeroed_currents
// and the PML subclass, include/FDTD/FDTD_PML.h:10-44 (src/FDTD/FDTD_PML.cpp):
//   FDTD_PML(Parameters, FP dt, FP pml_percent) -> ref_pml_create
//   update_fields() override                    -> ref_pml_update_fields
// and the protected curl updates over a box (include/FDTD/FDTD.h:17-18,33-34),
// reached through a subclass that sets begin_main_* / end_main_*:
//   update_B() / update_E()                     -> ref_update_B_box / ref_update_E_box
// FDTD_PML::update_E_PML/update_B_PML keep their per-cell coefficients in
// variables shared by the OpenMP team (src/FDTD/FDTD_PML.cpp:69,138), so
// ref_pml_update_fields runs the class on one thread to get the values the
// code means.
#include <omp.h>

#include <cstring>
#include <exception>

#include "FDTD.h"
#include "FDTD_PML.h"

using FDTD_openmp::FDTD;
using FDTD_openmp::FDTD_PML;

namespace {
// The reference class with its protected box updates reachable, as a
// subclass of the reference's own would call them.
class BoxFDTD : public FDTD {
public:
    using FDTD::FDTD;
    void box_step(const int box[6], bool e_step) {
        const int keep[6] = {begin_main_i, end_main_i, begin_main_j,
                             end_main_j, begin_main_k, end_main_k};
        set_main(box);
        if (e_step) update_E();
        else update_B();
        set_main(keep);
    }

private:
    void set_main(const int box[6]) {
        begin_main_i = box[0];
        end_main_i = box[1];
        begin_main_j = box[2];
        end_main_j = box[3];
        begin_main_k = box[4];
        end_main_k = box[5];
    }
};
}  // namespace

extern "C" {

void* ref_create(int Ni, int Nj, int Nk, double ax, double bx, double ay, double by,
                 double az, double bz, double dx, double dy, double dz, double dt) {
    FDTD_struct::Parameters p{Ni, Nj, Nk, ax, bx, ay, by, az, bz, dx, dy, dz};
    try {
        return static_cast<FDTD*>(new BoxFDTD(p, dt));
    } catch (const std::exception&) {
        return nullptr;
    }
}

void ref_destroy(void* h) { delete static_cast<BoxFDTD*>(static_cast<FDTD*>(h)); }

// update_B() / update_E() of a ref_create grid over box = {i0, i1, j0, j1,
// k0, k1} (the main box is restored afterwards).  The
// reference's update_E adds Jx to all three E components
// (src/FDTD/FDTD.cpp:85-93).
void ref_update_B_box(void* h, const int box[6]) {
    static_cast<BoxFDTD*>(static_cast<FDTD*>(h))->box_step(box, false);
}

void ref_update_E_box(void* h, const int box[6]) {
    static_cast<BoxFDTD*>(static_cast<FDTD*>(h))->box_step(box, true);
}

int ref_set(void* h, int comp, const double* src, long long n) {
    try {
        auto& f = static_cast<FDTD*>(h)->get_field(static_cast<FDTD_enums::Component>(comp));
        if ((long long)f.size() != n) return -2;
        std::memcpy(f.data(), src, sizeof(double) * n);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_get(void* h, int comp, double* dst, long long n) {
    try {
        auto& f = static_cast<FDTD*>(h)->get_field(static_cast<FDTD_enums::Component>(comp));
        if ((long long)f.size() != n) return -2;
        std::memcpy(dst, f.data(), sizeof(double) * n);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_set_value(void* h, int comp, long long idx, double value) {
    try {
        auto& f = static_cast<FDTD*>(h)->get_field(static_cast<FDTD_enums::Component>(comp));
        if (idx < 0 || idx >= (long long)f.size()) return -2;
        f[idx] = value;
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

void ref_update_fields(void* h, int nsteps) {
    for (int s = 0; s < nsteps; ++s) static_cast<FDTD*>(h)->update_fields();
}

void ref_zeroed_currents(void* h) { static_cast<FDTD*>(h)->zeroed_currents(); }

// A PML grid; the other ref_* calls take it as an FDTD (get_field and
// zeroed_currents are inherited), except destruction and stepping.
void* ref_pml_create(int Ni, int Nj, int Nk, double ax, double bx, double ay, double by,
                     double az, double bz, double dx, double dy, double dz, double dt,
                     double pml_percent) {
    FDTD_struct::Parameters p{Ni, Nj, Nk, ax, bx, ay, by, az, bz, dx, dy, dz};
    try {
        FDTD_PML* g = new FDTD_PML(p, dt, pml_percent);
        return static_cast<FDTD*>(g);
    } catch (const std::exception&) {
        return nullptr;
    }
}

void ref_pml_destroy(void* h) { delete static_cast<FDTD_PML*>(static_cast<FDTD*>(h)); }

void ref_pml_update_fields(void* h, int nsteps) {
    const int threads = omp_get_max_threads();
    omp_set_num_threads(1);
    FDTD* g = static_cast<FDTD*>(h);
    for (int s = 0; s < nsteps; ++s) g->update_fields();
    omp_set_num_threads(threads);
}

}  // extern "C"
