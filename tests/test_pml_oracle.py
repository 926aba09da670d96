"""Pins of the oracle's PML layer (oracle/fdtd_oracle.c, oracle_pml_*) -- CPU.

The layer follows src/FDTD/FDTD_PML.cpp (OpenMP) and
src/FDTD_kokkos/FDTD_PML_kokkos.cpp + include/FDTD_kokkos/kokkos_functors.h:154-415.
Checks: bit-exact against the reference's own PML class compiled from its
sources (run on one thread, see oracle/ref_shim.cpp) and against golden dumps
it wrote; the conductivity profile's closed form; a zero-width layer is the
periodic solver; currents in the layer are ignored; the main box away from
the layer evolves exactly as the periodic solver until the layer's influence
arrives; and the layer absorbs (energy falls by orders of magnitude where the
periodic grid conserves it) -- a wrong sign in a decay or coefficient grows
instead.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from conftest import courant_dt, random_state
from oracle import oracle as O


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def pml_grid(shape, spacing, dt, pml, state, J=None):
    g = O.OraclePmlGrid(*shape, *spacing, dt, pml)
    for c in range(6):
        g.fields()[c][:] = state[c]
    if J is not None:
        for c in range(3):
            g.J[c][:] = J[c]
    return g


PML_CASES = [((16, 12, 10), 0.2), ((13, 9, 11), 0.3), ((10, 6, 4), 0.5), ((8, 8, 8), 0.0),
             ((7, 5, 3), 0.15), ((20, 6, 9), 0.25), ((5, 12, 1), 0.25)]


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("shape,pml", PML_CASES)
def test_pml_oracle_matches_compiled_reference(shape, pml):
    """Bit-exact against src/FDTD/FDTD_PML.cpp compiled with the reference's
    flags, J = Jx = Jy = Jz on cells of the main box and of the layer."""
    n = int(np.prod(shape))
    spacing = (1.0, 1.3, 0.7)
    dt = courant_dt(*spacing)
    st = random_state(n, seed=n + 7)
    Jc = np.zeros(n)
    Jc[:: max(1, n // 7)] = np.random.default_rng(2).standard_normal(Jc[:: max(1, n // 7)].size)
    g = pml_grid(shape, spacing, dt, pml, st, [Jc, Jc, Jc])
    r = O.RefPmlGrid(*shape, *spacing, dt, pml)
    for c in range(6):
        r.set(c, st[c])
    for c in range(6, 9):
        r.set(c, Jc)
    g.update_fields(7)
    r.update_fields(7)
    for c in range(6):
        assert np.array_equal(g.fields()[c], r.get(c)), f"component {c}"


def test_pml_oracle_matches_golden_random(golden_dir):
    f = np.load(os.path.join(golden_dir, "ref_pml_16x12x10_random_40steps.npz"))
    shape = tuple(int(x) for x in f["dims"])
    g = pml_grid(shape, tuple(f["spacing"]), float(f["dt"]), float(f["pml_percent"]),
                 list(f["E_in"]) + list(f["B_in"]), [f["J_in"]] * 3)
    g.update_fields(int(f["nsteps"]))
    for c in range(3):
        assert np.array_equal(g.E[c], f["E_out"][c])
        assert np.array_equal(g.B[c], f["B_out"][c])


def test_pml_oracle_matches_golden_sample(golden_dir):
    """The driver's PML block, spherical_wave(32, 100) with FDTD_PML(.., 0.2)."""
    with open(os.path.join(golden_dir, "ref_pml_sample_32_100.json")) as fh:
        ref = json.load(fh)
    out = O.spherical_wave(ref["n"], ref["iterations"], pml_percent=ref["pml_percent"])
    assert [sha(a) for a in out] == ref["sha256"]


def test_pml_sizes():
    """(int)(N * pml_percent), FDTD_PML.cpp:254-256."""
    assert O.pml_sizes(512, 512, 512, 0.2) == (102, 102, 102)
    assert O.pml_sizes(16, 12, 10, 0.25) == (4, 3, 2)
    assert O.pml_sizes(7, 5, 3, 0.15) == (1, 0, 0)
    assert O.pml_sizes(10, 6, 4, 0.5) == (5, 3, 2)


@pytest.mark.parametrize("n,p,d", [(20, 5, 1.3), (11, 3, 0.7), (64, 12, 3e10)])
def test_pml_sigma_profile_closed_form(n, p, d):
    """sigma = SGm (dist/p)^4 with SGm = (4+1)/2 ln(1/R)/(p d), R = 1e-12,
    dist = p at the outer cells and 1 next to the main box; zero inside."""
    s = O.pml_sigma(n, p, d)
    sgm = 2.5 * math.log(1e12) / (p * d)
    assert s[0] == pytest.approx(sgm, rel=1e-14)
    assert s[n - 1] == pytest.approx(sgm, rel=1e-14)
    assert s[p - 1] == pytest.approx(sgm * (1.0 / p) ** 4, rel=1e-14)
    for i in range(p):
        assert s[i] == pytest.approx(sgm * ((p - i) / p) ** 4, rel=1e-14)
        assert s[i] == s[n - 1 - i]  # mirror symmetric, bitwise
    assert np.all(np.diff(s[:p]) < 0)
    assert np.all(s[p:n - p] == 0.0)


def test_pml_zero_width_is_the_periodic_solver():
    """pml_percent 0 and one too small for a single cell: no layer, exactly the
    periodic FDTD (every layer slab of FDTD_PML.cpp:346-362 is empty)."""
    shape, spacing = (16, 16, 16), (1.0, 1.0, 1.0)
    dt = courant_dt(*spacing)
    n = int(np.prod(shape))
    st = random_state(n, seed=5)
    for pml in (0.0, 0.05):
        g = pml_grid(shape, spacing, dt, pml, st)
        h = O.OracleGrid(*shape, *spacing, dt)
        for c in range(6):
            h.fields()[c][:] = st[c]
        g.update_fields(6)
        h.update_fields(6)
        for c in range(6):
            assert np.array_equal(g.fields()[c], h.fields()[c])
        assert all(not a.any() for a in g.ES + g.BS)


def test_pml_layer_ignores_currents():
    """J acts through the main-box E update only (FDTD_PML.cpp:354 vs the
    layer's update_E_PML, which has no J term)."""
    shape, spacing, pml = (14, 12, 10), (1.0, 1.0, 1.0), 0.25
    dt = courant_dt(*spacing)
    n = int(np.prod(shape))
    st = random_state(n, seed=9)
    Ni, Nj = shape[0], shape[1]
    layer_cell = 1 + 1 * Ni + 1 * Ni * Nj
    main_cell = 7 + 6 * Ni + 5 * Ni * Nj
    base = pml_grid(shape, spacing, dt, pml, st)
    base.update_fields(5)
    J = np.zeros(n)
    J[layer_cell] = 3.0
    g = pml_grid(shape, spacing, dt, pml, st, [J, J, J])
    g.update_fields(5)
    for c in range(6):
        assert np.array_equal(g.fields()[c], base.fields()[c])
    J[main_cell] = 3.0
    g = pml_grid(shape, spacing, dt, pml, st, [J, J, J])
    g.update_fields(5)
    assert any(not np.array_equal(g.fields()[c], base.fields()[c]) for c in range(6))


def test_pml_main_box_interior_matches_periodic_until_the_layer_is_felt():
    """Per step a cell depends on cells at most 2 away along each axis
    (B reads +1 twice, E reads -1 once), so after s steps cells farther than
    2s from the layer hold exactly the periodic solver's values."""
    shape, spacing, pml, s = (28, 28, 28), (1.0, 1.0, 1.0), 0.15, 3
    dt = courant_dt(*spacing)
    n = int(np.prod(shape))
    st = random_state(n, seed=11)
    p = O.pml_sizes(*shape, pml)[0]
    g = pml_grid(shape, spacing, dt, pml, st)
    h = O.OracleGrid(*shape, *spacing, dt)
    for c in range(6):
        h.fields()[c][:] = st[c]
    g.update_fields(s)
    h.update_fields(s)
    lo, hi = p + 2 * s + 1, shape[0] - p - 2 * s - 1
    inner = np.zeros(shape[::-1], dtype=bool)
    inner[lo:hi, lo:hi, lo:hi] = True
    inner = inner.reshape(-1)
    for c in range(6):
        assert np.array_equal(g.fields()[c][inner], h.fields()[c][inner])
        assert not np.array_equal(g.fields()[c], h.fields()[c])


def test_pml_absorbs_an_outgoing_pulse():
    """A Gaussian-modulated Jz pulse at the centre: with the layer the field
    energy falls by orders of magnitude once the pulse has left the main box;
    the periodic grid keeps it."""
    N, d = 32, 1.0
    dt = 0.5 * d / (O.C * math.sqrt(3.0))
    src = [dict(i=N // 2, j=N // 2, k=N // 2, comp=2, kind=1, amp=1.0, period=20 * dt,
                t0=60 * dt, sigma=15 * dt)]
    energies = {}
    for pml in (None, 0.2):
        g = (O.OracleGrid(N, N, N, d, d, d, dt) if pml is None
             else O.OraclePmlGrid(N, N, N, d, d, d, dt, pml))
        g.run_point_sources(src, 130)
        e0 = sum(g.sum_squares())
        g.run_point_sources(src, 270)
        energies[pml] = (e0, sum(g.sum_squares()))
    e0, e1 = energies[None]
    assert e1 == pytest.approx(e0, rel=1e-2)
    e0, e1 = energies[0.2]
    assert e1 < 1e-3 * e0
