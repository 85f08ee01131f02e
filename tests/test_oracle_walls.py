"""Pins of the oracle's second-order walls (BoundCondWallO2Cartesian).

An inviscid wall is a mirror: a case with a wall on its minus side must evolve
as the right half of the case mirrored about the wall (normal velocity negated)
with no wall at all.  The mirror image carries exactly the slopes the wall
supplies (zero for the scalars, 2 u_n / dx for the normal velocity through the
face-to-centre distance), its interface switch sees the cell's own volume
fraction, and the interior HLLC solver on mirrored states reduces to the wall
solver up to rounding.  A wrong sign of the wall distance or of the
extrapolation, a missing wall slope in the neighbouring faces, or the first-order
wall state fail the comparison by orders of magnitude.  All limiters are minmod
here (the wall's and the interior face's interface switch then agree).
A closed box conserves mass per phase and total energy to round-off."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from ecogen_amd.case import IG_AIR, SG_WATER

EPS = 1e-6


def _euler_case(n, lx, bc):
    return dict(name="w", nx=n, ny=1, nz=1, lx=lx, ly=0.1, lz=0.1, model="euler", eos=[dict(IG_AIR)], order=2,
                limiters=("minmod",) * 4, cfl=0.8, bc=bc, domains=[dict(kind="entire", rho=[1.0], p=[1.e5],
                                                                     vel=[0., 0., 0.])])


def _run(o, case, steps):
    dt = 1.e-15
    for _ in range(steps):
        dt = case["cfl"] * o.step(dt)
    return o.get_state()


def _rel(a, b):
    return np.sum(np.abs(a - b), axis=1) / np.maximum(np.sum(np.abs(b), axis=1), 1e-300)


@pytest.mark.parametrize("steps", [1, 25])
def test_euler_wall_is_a_mirror(steps):
    n = 60
    wall = _euler_case(n, 0.5, ("wall", "nonreflecting") + ("nonreflecting",) * 4)
    x = (np.arange(n) + 0.5) / n
    rho = 1.0 + 0.5 * np.exp(-((x - 0.15) / 0.05) ** 2)
    p = 1.e5 * (1.0 + 2.0 * np.exp(-((x - 0.2) / 0.04) ** 2))
    u = -80.0 * np.exp(-((x - 0.25) / 0.06) ** 2)          # flow toward the wall
    e = p / 0.4 / rho
    st = np.stack([rho, p, u, e + 0.5 * u * u])
    ow = Oracle(wall)
    ow.set_state(st)
    got = _run(ow, wall, steps)
    mirror = _euler_case(2 * n, 1.0, ("nonreflecting",) * 6)
    ms = np.concatenate([st[:, ::-1], st], axis=1)
    ms[2, :n] *= -1.0
    om = Oracle(mirror)
    om.set_state(ms)
    ref = _run(om, mirror, steps)[:, n:]
    assert np.all(_rel(got, ref) < 1e-11), _rel(got, ref)
    # the first-order wall state would be far off: the O2 wall state matters
    assert np.max(np.abs(got[2])) > 1.0


def _two_phase_2d(nx, ny, ly, model, bc):
    return dict(name="w2", nx=nx, ny=ny, nz=1, lx=0.4, ly=ly, lz=1.0, model=model,
                eos=[dict(SG_WATER), dict(IG_AIR)], order=2, limiters=("minmod",) * 4, cfl=0.6,
                alpha_null=False, relax_p=True, bc=bc,
                domains=[dict(kind="entire", alpha=[1 - EPS, EPS], rho=[1000., 1.2], p=[1e5, 1e5], vel=[0., 0., 0.])])


@pytest.mark.parametrize("model", ["pueq", "ueq"])
def test_two_phase_wall_on_y_is_a_mirror(model):
    nx, ny = 20, 24
    wall = _two_phase_2d(nx, ny, 0.3, model, ("nonreflecting", "nonreflecting", "wall") + ("nonreflecting",) * 3)
    x = ((np.arange(nx) + 0.5) / nx * 0.4)[None, :]
    y = ((np.arange(ny) + 0.5) / ny * 0.3)[:, None]
    r2 = (x - 0.2) ** 2 + (y - 0.08) ** 2
    aa = np.where(r2 < 0.05 ** 2, 1 - EPS, EPS) * np.ones_like(x * y)      # an air bubble near the wall
    aw = 1.0 - aa
    p = 1e5 + 2e8 * np.exp(-((y - 0.2) / 0.03) ** 2) * np.ones_like(x)     # a pressure pulse above it
    v = -50.0 * np.exp(-((y - 0.15) / 0.05) ** 2) * np.ones_like(x)
    u = 10.0 * np.sin(np.pi * x / 0.4) * np.ones_like(y)
    rw = 1000.0 * np.ones_like(p)
    ra = 1.2 * np.ones_like(p)
    o = Oracle(wall)
    st = o.get_state().reshape(-1, ny, nx)
    st[0], st[1], st[2] = aw, rw, p
    st[3], st[4], st[5] = aa, ra, p
    st[6], st[7] = u, v
    # E from the phases as the initial condition builds it
    e_w = (p + 4.4 * 6e8) / 3.4 / rw
    e_a = p / 0.4 / ra
    rho = aw * rw + aa * ra
    st[8] = (aw * rw * e_w + aa * ra * e_a) / rho + 0.5 * (u * u + v * v)
    st = st.reshape(-1, nx * ny)
    o.set_state(st)
    got = _run(o, wall, 12)
    mirror = _two_phase_2d(nx, 2 * ny, 0.6, model, ("nonreflecting",) * 6)
    s3 = st.reshape(-1, ny, nx)
    m3 = np.concatenate([s3[:, ::-1, :], s3], axis=1)
    m3[7, :ny] *= -1.0
    om = Oracle(mirror)
    om.set_state(m3.reshape(-1, 2 * ny * nx))
    ref = _run(om, mirror, 12).reshape(-1, 2 * ny, nx)[:, ny:, :].reshape(-1, ny * nx)
    rel = _rel(got, ref)
    assert np.all(rel < 1e-10), rel


@pytest.mark.parametrize("model", ["euler", "pueq", "ueq"])
def test_closed_box_conserves(model):
    n = 16
    if model == "euler":
        c = dict(name="box", nx=n, ny=n, nz=1, lx=1.0, ly=1.0, lz=1.0, model="euler", eos=[dict(IG_AIR)], order=2,
                 limiters=("minmod",) * 4, cfl=0.8, bc=("wall",) * 6,
                 domains=[dict(kind="entire", rho=[1.0], p=[1.e5], vel=[0., 0., 0.]),
                          dict(kind="disc", axis=0, axis2=1, center=(0.3, 0.6, 0.), radius=0.2, rho=[3.0],
                               p=[5.e5], vel=[30., -20., 0.])])
    else:
        c = _two_phase_2d(n, n, 0.4, model, ("wall",) * 6)
        c["domains"].append(dict(kind="disc", axis=0, axis2=1, center=(0.2, 0.15, 0.), radius=0.08,
                                 alpha=[EPS, 1 - EPS], rho=[1000., 5.0], p=[5e6, 5e6], vel=[20., 30., 0.]))
    o = Oracle(c)
    st0 = o.get_state()
    st = _run(o, c, 20)

    def tot(s):
        if model == "euler":
            return [s[0].sum(), (s[0] * s[4]).sum()]
        rho = s[0] * s[1] + s[3] * s[4]
        return [(s[0] * s[1]).sum(), (s[3] * s[4]).sum(), (rho * s[8]).sum()]
    a, b = tot(st0), tot(st)
    assert np.allclose(a, b, rtol=1e-13, atol=0), (a, b)
    assert not np.array_equal(st, st0)
