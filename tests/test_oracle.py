"""Pins of the CPU oracle (oracle/fdtd_oracle.c) -- runs on the CPU.

Each check is chosen so that a plausible mistake in the restatement fails
it: a wrong component or sign (known answer, convergence), a transposed or
mis-wrapped index (numpy-roll restatement, shift equivariance, compiled
reference), a dropped term (golden dumps of the reference class).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import courant_dt, random_state
from oracle import oracle as O


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def make(shape, spacing, dt, state, J=None):
    g = O.OracleGrid(*shape, *spacing, dt)
    for c in range(6):
        g.fields()[c][:] = state[c]
    if J is not None:
        for c in range(3):
            g.J[c][:] = J[c]
    return g


SHAPES = [(16, 12, 10), (7, 5, 3), (1, 4, 6), (5, 1, 2), (3, 2, 1), (2, 2, 2)]


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("shape", SHAPES)
def test_oracle_matches_compiled_reference(shape):
    """Bit-exact against src/FDTD/FDTD.cpp compiled with the reference's flags,
    with J = Jx = Jy = Jz (where OpenMP and Kokkos semantics coincide)."""
    n = int(np.prod(shape))
    spacing = (1.0, 1.3, 0.7)
    dt = courant_dt(*spacing)
    st = random_state(n, seed=n)
    Jc = np.zeros(n)
    Jc[:: max(1, n // 5)] = np.random.default_rng(1).standard_normal(Jc[:: max(1, n // 5)].size)
    g = make(shape, spacing, dt, st, [Jc, Jc, Jc])
    r = O.RefGrid(*shape, *spacing, dt)
    for c in range(6):
        r.set(c, st[c])
    for c in range(6, 9):
        r.set(c, Jc)
    g.update_fields(9)
    r.update_fields(9)
    for c in range(6):
        assert np.array_equal(g.fields()[c], r.get(c)), f"component {c}"


BOXES = [((16, 12, 10), (3, 2, 1), (13, 10, 9)), ((16, 12, 10), (0, 5, 0), (16, 12, 4)),
         ((7, 5, 3), (6, 0, 2), (7, 5, 3)), ((5, 1, 2), (1, 0, 0), (4, 1, 2))]


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("shape,lo,hi", BOXES)
def test_oracle_box_updates_match_compiled_reference(shape, lo, hi):
    """The box updates (the reference's protected update_B / update_E over
    begin_main_* .. end_main_*, reached through a subclass of the compiled
    class) bit-exact, cells outside the box untouched, J = Jx = Jy = Jz."""
    n = int(np.prod(shape))
    spacing = (1.0, 1.3, 0.7)
    dt = courant_dt(*spacing)
    st = random_state(n, seed=n + 7)
    Jc = np.random.default_rng(3).standard_normal(n)
    g = make(shape, spacing, dt, st, [Jc, Jc, Jc])
    r = O.RefGrid(*shape, *spacing, dt)
    for c in range(6):
        r.set(c, st[c])
    for c in range(6, 9):
        r.set(c, Jc)
    inside = np.zeros(shape[::-1], dtype=bool)
    inside[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = True
    inside = inside.reshape(-1)
    for step in ("B", "E", "B", "E"):
        before = [a.copy() for a in g.fields()[:6]]
        getattr(g, f"update_{step}_box")(lo, hi)
        getattr(r, f"update_{step}_box")(lo, hi)
        for c in range(6):
            assert np.array_equal(g.fields()[c], r.get(c)), f"{step} component {c}"
            assert np.array_equal(g.fields()[c][~inside], before[c][~inside])
        moved = [c for c in range(6) if not np.array_equal(g.fields()[c], before[c])]
        assert moved == ([3, 4, 5] if step == "B" else [0, 1, 2])


def test_oracle_matches_golden_random(golden_dir):
    f = np.load(os.path.join(golden_dir, "ref_16x12x10_random_50steps.npz"))
    shape = tuple(int(x) for x in f["dims"])
    g = make(shape, tuple(f["spacing"]), float(f["dt"]), list(f["E_in"]) + list(f["B_in"]),
             [f["J_in"]] * 3)
    g.update_fields(int(f["nsteps"]))
    for c in range(3):
        assert np.array_equal(g.E[c], f["E_out"][c])
        assert np.array_equal(g.B[c], f["B_out"][c])


def config1_state(N=64):
    n = N ** 3
    i = np.arange(n) % N
    ey = np.sin(2.0 * np.pi * i / N)
    bz = np.sin(2.0 * np.pi * (i + 0.5) / N) / O.C
    z = np.zeros(n)
    return [z.copy(), ey, z.copy(), z.copy(), z.copy(), bz]


def test_oracle_matches_golden_config1(golden_dir):
    rec = json.load(open(os.path.join(golden_dir, "ref_config1_64cube_200steps.json")))
    st = config1_state(rec["dims"][0])
    assert [sha(a) for a in st] == rec["init_sha256"]
    g = make(tuple(rec["dims"]), (rec["d"],) * 3, rec["dt"], st)
    g.update_fields(rec["nsteps"])
    assert [sha(a) for a in g.E + g.B] == rec["sha256"]


def test_oracle_matches_golden_sample(golden_dir):
    """The drivers' spherical wave (sample.cpp:14-91) on the oracle."""
    rec = json.load(open(os.path.join(golden_dir, "ref_sample_32_100.json")))
    out = O.spherical_wave(rec["n"], rec["iterations"])
    assert [sha(a) for a in out] == rec["sha256"]


@pytest.mark.parametrize("shape", SHAPES)
def test_oracle_matches_numpy_roll(shape):
    """Independent restatement with np.roll neighbours (all J components distinct)."""
    n = int(np.prod(shape))
    spacing = (0.9, 1.1, 1.3)
    dt = courant_dt(*spacing)
    st = random_state(n, seed=7 + n, count=9)
    g = make(shape, spacing, dt, st[:6], st[6:])
    E, B, J = [a.copy() for a in st[:3]], [a.copy() for a in st[3:6]], st[6:]
    g.update_fields(5)
    O.roll_update_fields(E, B, J, shape, O.coefficients(*spacing, dt), 5)
    for c in range(3):
        assert np.array_equal(g.E[c], E[c])
        assert np.array_equal(g.B[c], B[c])


def test_known_answer_single_source():
    """Zero fields, Jz = v at one cell, one step: Ez = cur*v exactly at the
    cell, B from the final half step in closed form (kokkos_functors.h:87,
    :143-150), everything else zero."""
    Ni, Nj, Nk = 6, 5, 4
    dx, dy, dz = 1.0, 2.0, 0.5
    dt = courant_dt(dx, dy, dz)
    g = O.OracleGrid(Ni, Nj, Nk, dx, dy, dz, dt)
    i, j, k, v = 2, 3, 1, 0.7
    idx = i + j * Ni + k * Ni * Nj
    g.J[2][idx] = v
    g.update_fields(1)
    cEx, cEy, cEz, cBx, cBy, cBz, cur = O.coefficients(dx, dy, dz, dt)
    assert cur == -4.0 * 3.14159265358 * dt
    ez = cur * v
    expect = [np.zeros(Ni * Nj * Nk) for _ in range(6)]
    expect[2][idx] = ez
    at = lambda ii, jj, kk: ii + jj * Ni + kk * Ni * Nj  # noqa: E731
    expect[3][at(i, j - 1, k)] = 0.0 - cBy * (ez - 0.0)     # Bx at j-1
    expect[3][idx] = 0.0 - cBy * (0.0 - ez)                  # Bx at j
    expect[4][at(i - 1, j, k)] = cBx * (ez - 0.0) - 0.0      # By at i-1
    expect[4][idx] = cBx * (0.0 - ez) - 0.0                  # By at i
    for c in range(6):
        assert np.array_equal(g.fields()[c], expect[c]), f"component {c}"


# unit-tests/test_FDTD_method.cpp:71-203: 12 convergence assertions.
CONVERGENCE = [
    ("x_axis_EY", 1, 1, 5, (16, 8, 4)), ("x_axis_BZ", 5, 1, 5, (16, 8, 4)),
    ("x_axis_EZ", 2, 2, 4, (16, 8, 4)), ("x_axis_BY", 4, 2, 4, (16, 8, 4)),
    ("y_axis_EX", 0, 0, 5, (8, 16, 4)), ("y_axis_BZ", 5, 0, 5, (8, 16, 4)),
    ("y_axis_EZ", 2, 2, 3, (8, 16, 4)), ("y_axis_BX", 3, 2, 3, (8, 16, 4)),
    ("z_axis_EX", 0, 0, 4, (4, 8, 16)), ("z_axis_BY", 4, 0, 4, (4, 8, 16)),
    ("z_axis_EY", 1, 1, 3, (4, 8, 16)), ("z_axis_BX", 3, 1, 3, (4, 8, 16)),
]
SIGN = {(0, 5): -1.0, (2, 4): -1.0, (1, 3): -1.0, (1, 5): 1.0, (2, 3): 1.0, (0, 4): 1.0}
AXIS = {(1, 5): 0, (2, 4): 0, (0, 5): 1, (2, 3): 1, (0, 4): 2, (1, 3): 2}
T_END = 5e-13


def plane_wave(shape, e, b):
    """E = sign*sin(2 pi x/L) at x = q*d, B = sin(2 pi (x + d/2)/L) along the
    pair's axis (src/FDTD/test_FDTD.cpp:5-80), domain [0,1]x[0,2]x[0,3]."""
    L = (1.0, 2.0, 3.0)
    d = tuple(L[a] / shape[a] for a in range(3))
    ax = AXIS[(e, b)]
    n = int(np.prod(shape))
    idx = np.arange(n)
    q = (idx % shape[0], (idx // shape[0]) % shape[1], idx // (shape[0] * shape[1]))[ax]
    x = q.astype(float) * d[ax]
    st = [np.zeros(n) for _ in range(6)]
    st[e] = SIGN[(e, b)] * np.sin(2.0 * O.PI * x / L[ax])
    st[b] = np.sin(2.0 * O.PI * (d[ax] / 2.0 + x) / L[ax])
    return st, d, ax, L


def line_error(field, comp, shape, e, b, d, ax, L):
    sign = 1.0 if comp >= 3 else SIGN[(e, b)]
    n = shape[ax]
    stride = (1, shape[0], shape[0] * shape[1])[ax]
    x = (d[ax] / 2.0 if comp >= 3 else 0.0) + np.arange(n) * d[ax]
    exact = np.sin(2.0 * O.PI * (x - O.C * T_END) / L[ax])
    return np.max(np.abs(sign * field[np.arange(n) * stride] - exact))


def run_convergence(stepper, comp, e, b, shape):
    errs = []
    for mult, iters in ((1, 16), (2, 32)):
        sh = tuple(s * mult for s in shape)
        st, d, ax, L = plane_wave(sh, e, b)
        out = stepper(sh, d, T_END / iters, st, iters)
        errs.append(line_error(out[comp], comp, sh, e, b, d, ax, L))
    return errs[0] / errs[1]


def oracle_stepper(shape, d, dt, st, iters):
    g = make(shape, d, dt, st)
    g.update_fields(iters)
    return g.E + g.B


@pytest.mark.parametrize("name,comp,e,b,shape", CONVERGENCE, ids=[c[0] for c in CONVERGENCE])
def test_oracle_convergence(name, comp, e, b, shape):
    assert abs(run_convergence(oracle_stepper, comp, e, b, shape) - 4.0) <= 0.1


def test_shift_equivariance():
    """Shifting the input by (1,2,3) cells shifts the output exactly."""
    shape = (9, 7, 5)
    n = int(np.prod(shape))
    spacing = (1.0, 1.0, 1.0)
    dt = courant_dt(*spacing)
    st = random_state(n, seed=11)
    shift = lambda a: np.roll(a.reshape(shape[::-1]), (3, 2, 1), axis=(0, 1, 2)).reshape(-1)  # noqa
    g1 = make(shape, spacing, dt, st)
    g2 = make(shape, spacing, dt, [shift(a) for a in st])
    g1.update_fields(6)
    g2.update_fields(6)
    for c in range(6):
        assert np.array_equal(shift(g1.fields()[c]), g2.fields()[c])


def test_linearity():
    shape = (8, 6, 5)
    n = int(np.prod(shape))
    spacing = (1.0, 1.0, 1.0)
    dt = courant_dt(*spacing)
    a, b = random_state(n, 1), random_state(n, 2)
    ga, gb = make(shape, spacing, dt, a), make(shape, spacing, dt, b)
    gs = make(shape, spacing, dt, [x + y for x, y in zip(a, b)])
    for g in (ga, gb, gs):
        g.update_fields(10)
    for c in range(6):
        np.testing.assert_allclose(gs.fields()[c], ga.fields()[c] + gb.fields()[c],
                                   rtol=0, atol=1e-12)


def test_divergence_of_B_is_preserved():
    """The B update is minus a forward-difference curl (kokkos_functors.h:143-150);
    the forward-difference divergence of such a curl vanishes identically, so
    that div B stays at its initial value up to rounding."""
    shape = (8, 7, 6)
    Ni, Nj, Nk = shape
    n = Ni * Nj * Nk
    spacing = (1.0, 1.5, 0.75)
    dt = courant_dt(*spacing)
    st = random_state(n, 5)

    def div(Bx, By, Bz):
        v = lambda a: a.reshape(Nk, Nj, Ni)  # noqa: E731
        return ((np.roll(v(Bx), -1, 2) - v(Bx)) / spacing[0] + (np.roll(v(By), -1, 1) - v(By))
                / spacing[1] + (np.roll(v(Bz), -1, 0) - v(Bz)) / spacing[2])

    g = make(shape, spacing, dt, st)
    d0 = div(*g.B)
    g.update_fields(20)
    assert np.max(np.abs(div(*g.B) - d0)) < 1e-11


def test_waveform_closed_form():
    t0, per, sig = 120.0, 20.0, 40.0
    for t in (1.0, 119.0, 120.0, 250.0):
        want = 0.5 * np.sin(2.0 * O.PI * (t - t0) / per) * np.exp(-(t - t0) ** 2 / (2 * sig ** 2))
        assert abs(O.waveform(1, 0.5, per, t0, sig, 0.0, t) - want) <= 1e-15
    assert O.waveform(0, 0.25, 1, 0, 0, 0, 3.0) == 0.25


def test_point_source_runner_matches_dense_J():
    """run_point_sources with constant (kind 0) sources equals a persistent J."""
    shape = (6, 5, 4)
    n = int(np.prod(shape))
    spacing = (1.0, 1.0, 1.0)
    dt = courant_dt(*spacing)
    st = random_state(n, 9)
    src = [dict(i=1, j=2, k=3, comp=2, kind=0, amp=0.3), dict(i=4, j=0, k=1, comp=0, kind=0, amp=-1.0)]
    g1 = make(shape, spacing, dt, st)
    g1.run_point_sources(src, 7)
    J = [np.zeros(n) for _ in range(3)]
    J[2][1 + 2 * 6 + 3 * 30] = 0.3
    J[0][4 + 0 * 6 + 1 * 30] = -1.0
    g2 = make(shape, spacing, dt, st, J)
    g2.update_fields(7)
    for c in range(6):
        assert np.array_equal(g1.fields()[c], g2.fields()[c])


def test_sum_squares_against_numpy_and_exact_cases():
    """oracle_sum_squares (the diagnostics checker) against numpy's float64
    dot product, and exact on values whose squares sum without rounding."""
    rng = np.random.default_rng(77)
    for n in (1, 7, 1000, 123457):
        a = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4)
        got = O.lib().oracle_sum_squares(O._ptr(a), a.size)
        assert abs(got - float(np.dot(a, a))) <= 1e-13 * float(np.dot(a, a))
    ints = np.arange(-50, 51, dtype=np.float64)            # squares sum exactly
    assert O.lib().oracle_sum_squares(O._ptr(ints), ints.size) == 2 * sum(k * k for k in range(51))
    assert O.lib().oracle_sum_squares(O._ptr(np.zeros(5)), 5) == 0.0
    g = O.OracleGrid(4, 3, 2, 1.0, 1.0, 1.0, 1e-12)
    for c in range(6):
        g.fields()[c][:] = c + 1.0
    assert g.sum_squares() == [24.0 * (c + 1.0) ** 2 for c in range(6)]
