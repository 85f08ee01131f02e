"""Parity of the HIP FP64 path with the oracle on small replicas (needs a GPU).

Both solvers start from the same initial-condition domains, take the same
sequence of time steps (dt0 = 1e-15, then cfl*dtMax of the oracle) and are
compared after every step:
  * dtMax within 1e-12 relative,
  * every persistent variable within 1e-12 relative L1 (the FP64 parity bar;
    the only non-shared arithmetic is the device pow() in the isentropic star
    pressures, which may differ from glibc's by an ulp).
Cases with a THINC limiter compare with the oracle's correctly rounded exp, the
one the kernels compute (the C library's exp is not correctly rounded for about
1 argument in 1200; tests/test_oracle_thinc.py bounds what that changes).
"""
import numpy as np
import pytest

from small_cases import all_cases

CASES = all_cases()
STEPS = 12
TOL = 1e-12
# The finite-rate relaxation decides per cell, with hard thresholds (pressures equal
# to 1e-5 relative, sub-step bounds), whether and how far to integrate; in this
# three-phase second-order case an ulp-level difference of the device pow (the one
# non-shared operation) flips that decision in a few cells after 6 steps (first
# differences at step 5 are 2e-15; the same case with the infinite relaxation stays
# at 1e-14), and the flipped cells then spread: 1e-6 relative L1 at step 6, 1e-5 at
# step 9.  Its bar only separates that from a wrong relaxation (the infinite
# relaxation differs from this one by O(1) in the air pressure); the three other
# finite-rate cases hold the 1e-12 bar.
TOL_CASE = {"ueq3_3d_o2_fin0.0001": 1e-3}


def rel_l1(a, b):
    den = np.sum(np.abs(b))
    return float(np.sum(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_step_parity(name):
    from ecogen_amd import Solver
    from ecogen_amd.case import variable_names
    from oracle.pyoracle import Oracle

    case = CASES[name]
    print(f"[parity] {name} {case['nx']}x{case['ny']}x{case['nz']} model={case['model']} order={case['order']}",
          flush=True)
    g = Solver(case)
    o = Oracle(case)
    from oracle.pyoracle import lib
    lib().orc_set_exact_exp(1 if "thinc" in case.get("limiters", ()) else 0)
    names = variable_names(case)
    s0g, s0o = g.get_state(), o.get_state()
    assert s0g.shape == s0o.shape
    np.testing.assert_array_equal(s0g, s0o, err_msg="initial condition differs")
    dt = 1.e-15
    worst = 0.0
    for it in range(STEPS):
        dtm_g, _ = g.step(dt)
        dtm_o = o.step(dt)
        assert abs(dtm_g - dtm_o) <= TOL_CASE.get(name, TOL) * abs(dtm_o), (it, dtm_g, dtm_o)
        sg, so = g.get_state(), o.get_state()
        assert np.all(np.isfinite(sg)), f"non-finite GPU state at step {it}"
        for v, nm in enumerate(names):
            err = rel_l1(sg[v], so[v])
            worst = max(worst, err)
            assert err <= TOL_CASE.get(name, TOL), f"step {it} variable {nm}: rel L1 {err:.3e}"
        dt = case["cfl"] * dtm_o
    exact = float(np.mean(sg == so))
    print(f"[parity] {name}: worst rel L1 {worst:.2e}, bit-identical fraction {exact:.6f}", flush=True)


@pytest.mark.gpu
def test_sod_time_loop_matches_reference_run():
    """Config-1 (1000 cells, t_end 5e-4) through the device-clocked loop.

    Golden values from a run of the reference binary on the same case (y/z
    extent 0.1): 433 iterations, t = 5.0106045125640768e-4,
    dt_next = 1.1542745158200889e-6 (infoCalcul.out, 17 significant digits)."""
    from ecogen_amd import Solver
    from ecogen_amd.case import sod_case

    g = Solver(sod_case(1000))
    st = g.run(t_end=5.e-4)
    assert st["iteration"] == 433
    assert st["t"] == float("0.00050106045125640768")
    assert abs(st["dt_next"] - float("1.1542745158200889e-06")) <= 1e-15 * 1.2e-6


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_non_converged_tank_newton_is_reported(precision):
    """A tank Newton that spends its 50 iterations ends the step with ECG_E_STATE and the
    reference's message (ModUEq.cpp:586), in a single step and in the device-clocked loop."""
    from ecogen_amd import EcgError, Solver
    from small_cases import tank_failure_case
    print(f"[parity] tank failure {precision}", flush=True)
    case = dict(tank_failure_case())
    case["precision"] = precision
    g = Solver(case)
    with pytest.raises(EcgError, match="solveRiemannInletTank not converged"):
        g.step(1.e-15)
    # Fresh contexts after one that failed: each must count against its own counters
    # (pinned host memory can come back from the earlier context with its contents).
    for n in (1, 3):
        g = Solver(case)
        with pytest.raises(EcgError, match="solveRiemannInletTank not converged"):
            g.run(max_iter=n)
