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

extern "C" {

void* ref_create(int Ni, int Nj, int Nk, double ax, double bx, double ay, double by,
                 double az, double bz, double dx, double dy, double dz, double dt) {
    FDTD_struct::Parameters p{Ni, Nj, Nk, ax, bx, ay, by, az, bz, dx, dy, dz};
    try {
        return new FDTD(p, dt);
    } catch (const std::exception&) {
        return nullptr;
    }
}

void ref_destroy(void* h) { delete static_cast<FDTD*>(h); }

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
