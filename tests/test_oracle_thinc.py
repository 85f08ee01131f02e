"""Pins of the oracle's THINC volume-fraction reconstruction
(CellInterfaceO2Cartesian.cpp:133-149, 165-180; LimiterTHINC.cpp).

THINC (Xiao, Honma and Kono 2005) replaces the linear volume-fraction profile
of a cell at an interface by qmin + qmax/2 (1 + s tanh(beta (x - x0))) on the
cell [0, 1], with x0 such that the cell average is the cell's volume fraction.
The reference evaluates the profile at the face in closed form; here the face
value is recomputed from that definition alone (x0 by bisection on the cell
average, the average by quadrature) for the left cell (face at x = 1) and the
right cell (face at x = 0).  The slope limiter THINC names is minmod.  An
advected interface stays sharper with THINC than with minmod."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle, lib
from ecogen_amd.case import IG_AIR, SG_WATER

BETA = 1.6


def _profile_face(a, qmin, qmax, s, at):
    def avg(x0):
        xs = (np.arange(4000) + 0.5) / 4000
        return float(np.mean(qmin + 0.5 * qmax * (1 + s * np.tanh(BETA * (xs - x0)))))
    lo, hi = -20.0, 21.0          # average decreases with x0 for s > 0, increases for s < 0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if (avg(mid) - a) * s > 0:
            lo = mid
        else:
            hi = mid
    x0 = 0.5 * (lo + hi)
    v = qmin + 0.5 * qmax * (1 + s * math.tanh(BETA * (at - x0)))
    return min(max(v, 1e-4), 1 - 1e-4)


@pytest.mark.parametrize("a,a_after,a_other", [(0.4, 0.05, 0.95), (0.7, 0.9, 0.1), (0.5, 1e-3, 0.999),
                                               (0.2, 0.01, 0.6), (0.93, 0.99, 0.2)])
def test_face_value_is_the_tanh_profile(a, a_after, a_other):
    L = lib()
    out = C.c_double()
    # left cell: the face is its right face (x = 1); the profile runs from a_after to a_other
    assert L.orc_thinc(0, a, a_after, a_other, C.byref(out)) == 1
    qmin, qmax = min(a_other, a_after), max(a_other, a_after) - min(a_other, a_after)
    s = 1.0 if a_other > a_after else -1.0
    assert abs(out.value - _profile_face(a, qmin, qmax, s, 1.0)) < 2e-4
    # right cell: the face is its left face (x = 0); the profile runs from a_other to a_after
    assert L.orc_thinc(1, a, a_after, a_other, C.byref(out)) == 1
    s = 1.0 if a_after > a_other else -1.0
    assert abs(out.value - _profile_face(a, qmin, qmax, s, 0.0)) < 2e-4


def test_no_reconstruction_outside_an_interface():
    L = lib()
    out = C.c_double(-1.0)
    assert L.orc_thinc(0, 0.5, 0.4, 0.45, C.byref(out)) == 0      # not monotone
    assert L.orc_thinc(0, 5e-5, 0.0, 0.1, C.byref(out)) == 0      # pure cell
    assert out.value == -1.0
    for s1, s2 in [(1.0, 2.0), (-3.0, -0.5), (1.0, -1.0), (0.0, 1.0), (1e-5, 1e-5)]:
        assert L.orc_limiter(6, s1, s2) == L.orc_limiter(1, s1, s2)


def _advected_interface(vf_limiter, steps=60):
    n = 100
    e = 1e-6
    c = dict(name="adv", nx=n, ny=1, nz=1, lx=1.0, ly=0.1, lz=0.1, model="pueq",
             eos=[dict(SG_WATER), dict(IG_AIR)], order=2, limiters=("minmod", "minmod", vf_limiter, vf_limiter),
             cfl=0.5, alpha_null=False, relax_p=True, bc=("nonreflecting",) * 6,
             domains=[dict(kind="entire", alpha=[1 - e, e], rho=[1000., 1.2], p=[1e5, 1e5], vel=[100., 0., 0.]),
                      dict(kind="halfspace", axis=0, origin=0.3, dir=1, alpha=[e, 1 - e], rho=[1000., 1.2],
                           p=[1e5, 1e5], vel=[100., 0., 0.])])
    o = Oracle(c)
    dt = 1e-15
    for _ in range(steps):
        dt = c["cfl"] * o.step(dt)
    a = o.get_state()[0]
    return int(np.sum((a > 0.01) & (a < 0.99))), a


def test_thinc_keeps_the_interface_sharper_than_minmod():
    n_thinc, a_t = _advected_interface("thinc")
    n_mm, a_m = _advected_interface("minmod")
    assert n_thinc < n_mm, (n_thinc, n_mm)
    assert np.all(np.isfinite(a_t)) and a_t.min() > -1e-12 and a_t.max() < 1 + 1e-12


def test_library_exp_changes_thinc_runs_only_in_the_last_bits():
    """The reference calls the C library's exp, which misrounds about 1 argument in
    1200; the oracle's correctly rounded mode (what the kernels compute) changes a
    THINC run by no more than such last-bit effects propagated through limiters."""
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from small_cases import all_cases
    L = lib()
    c = all_cases()["pueq_2d_o2_thinc_mix"]
    out = []
    for mode in (0, 1):
        L.orc_set_exact_exp(mode)
        o = Oracle(c)
        dt = 1e-15
        for _ in range(12):
            dt = c["cfl"] * o.step(dt)
        out.append(o.get_state())
    L.orc_set_exact_exp(0)
    a, b = out
    u_scale = np.sum(np.abs(b[6])) + np.sum(np.abs(b[7]))
    for v in range(a.shape[0]):
        den = u_scale if v in (6, 7) else np.sum(np.abs(b[v]))
        assert np.sum(np.abs(a[v] - b[v])) / den < 1e-9, v
