"""The oracle's OpenMP sweeps keep every cell's summation order: states, dtMax and
the relaxation counters must not depend on the thread count."""
import numpy as np
import pytest

from small_cases import all_cases
from oracle.pyoracle import Oracle, lib


def _run(case, threads, steps=5):
    lib().orc_set_threads(threads)
    o = Oracle(case)
    dt, dts = 1.e-15, []
    for _ in range(steps):
        d = o.step(dt)
        dts.append(d)
        dt = case["cfl"] * d
    return o.get_state(), dts, (o.stats.relax_calls, o.stats.newton_iters, o.stats.newton_fail)


@pytest.mark.parametrize("name", ["pueq_2d_o1", "ueq3_3d_o1", "pueq3_3d_tiles_o2", "pueq_2d_o2_wall_tank",
                                  "ueq_2d_o1_fin1e-07_tank", "euler_3d_o2_wall"])
def test_results_do_not_depend_on_threads(name):
    case = all_cases()[name]
    before = lib().orc_threads()
    try:
        a = _run(case, 1)
        b = _run(case, 4)
    finally:
        lib().orc_set_threads(before)
    assert np.array_equal(a[0], b[0])
    assert a[1] == b[1] and a[2] == b[2]
