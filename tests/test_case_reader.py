"""The case reader (ecg_read_case, reference XML input format): CPU checks.

Cases written in the reference's format must come back as the case they
describe; the initial state the oracle builds from the read case must equal the
one built from the hand-written case dict; unsupported features and malformed
files must be reported, not guessed.
"""
import os

import numpy as np
import pytest

from case_files import write_eos, write_kapila_bubble, write_sod
from ecogen_amd import EcgError, read_case
from ecogen_amd.case import IG_AIR, SG_WATER, sod_case
from oracle.pyoracle import Oracle


def test_sod_case_reads_back(tmp_path):
    write_eos(str(tmp_path))
    d = os.path.join(str(tmp_path), "cases", "sod")
    write_sod(d, n=200)
    c = read_case(d)                     # libEOS found in a parent directory
    ref = sod_case(n=200)
    for k in ("nx", "ny", "nz", "lx", "ly", "lz", "model", "order", "cfl"):
        assert c[k] == ref[k], k
    assert c["control"]["iterations_mode"] is False and c["control"]["total_time"] == 5.e-4
    assert c["eos"][0]["gamma"] == 1.4 and c["eos"][0]["type"] == "IG"
    assert np.array_equal(Oracle(c).get_state(), Oracle(ref).get_state())


@pytest.mark.parametrize("model", ["PressureVelocityEq", "VelocityEq"])
def test_two_phase_case_reads_back(tmp_path, model):
    write_eos(str(tmp_path))
    d = os.path.join(str(tmp_path), "bubble")
    write_kapila_bubble(d, n=32, second_order=True, model=model)
    c = read_case(d, eos_dir=os.path.join(str(tmp_path), "libEOS"))
    assert c["model"] == ("pueq" if model == "PressureVelocityEq" else "ueq")
    assert c["relax_p"] is True
    assert c["limiters"] == ("minmod", "superbee", "vanleer", "superbee")   # defaults of Input.cpp:254-300
    assert [d_["kind"] for d_ in c["domains"]] == ["entire", "disc", "rectangle", "halfspace"]
    sh = c["domains"][3]
    assert sh["dir"] == -1 and sh["origin"] == 0.2 and sh["vel"][0] == 120.0
    # density of the shocked air from its temperature: p / ((gamma - 1) cv T)  (EosIG.cpp:85-88)
    assert sh["rho"][1] == 3.e8 / ((1.4 - 1.) * 800. * 400.)
    assert sh["p"] == [3.e8, 3.e8]
    rect = c["domains"][2]
    assert rect["lo"][:2] == (0.75, 0.05) and rect["len"][:2] == (0.1, 0.2)
    assert c["bc"] == ("nonreflecting",) * 6
    st = Oracle(c).get_state()
    assert np.all(np.isfinite(st))


def test_first_order_wall_and_iterations(tmp_path):
    write_eos(str(tmp_path))
    d = os.path.join(str(tmp_path), "b")
    write_kapila_bubble(d, n=24, iterations=7, second_order=False)
    c = read_case(d)
    assert c["order"] == 1 and c["limiters"] == ("none",) * 4
    assert c["bc"] == ("nonreflecting", "nonreflecting", "wall", "nonreflecting", "nonreflecting", "nonreflecting")
    assert c["control"] == dict(iterations_mode=True, iterations=7, iter_freq=7, total_time=0.0, time_freq=0.0)


def _edit(path, old, new):
    s = open(path).read()
    assert old in s
    open(path, "w").write(s.replace(old, new))


@pytest.mark.parametrize("old,new,code_word", [
    ('<AMR lvlMax="0"', '<AMR lvlMax="2"', "adaptive mesh refinement"),
    ("<globalLimiter>minmod</globalLimiter>", "<globalLimiter>thinc</globalLimiter>", "THINC"),
    ('structure="cartesian"', 'structure="unstructured"', "Cartesian"),
    ('name="PressureVelocityEq"', 'name="EulerKorteweg"', "flow model"),
])
def test_unsupported_features_are_reported(tmp_path, old, new, code_word):
    write_eos(str(tmp_path))
    d = os.path.join(str(tmp_path), "b")
    write_kapila_bubble(d, n=16)
    for f in ("mesh.xml", "main.xml", "model.xml"):
        p = os.path.join(d, f)
        if old in open(p).read():
            _edit(p, old, new)
    with pytest.raises(EcgError, match=code_word):
        read_case(d)


@pytest.mark.parametrize("file,old,new,msg", [
    ("initialConditions.xml", 'density="1000.0"/>', 'density="1000.0" temperature="300."/>', "only one of"),
    ("initialConditions.xml", 'state="air" type="disc"', 'state="vapour" type="disc"', "not found"),
    ("mesh.xml", '<numberCells x="16"', '<numberCells x="1"', "superior to 1"),
    ("main.xml", "</computationParam>", "", "unexpected end of file"),
    ("initialConditions.xml", 'EOS="SG_water.xml">', 'EOS="IG_air.xml">', "ordered"),
])
def test_malformed_cases_are_rejected(tmp_path, file, old, new, msg):
    write_eos(str(tmp_path))
    d = os.path.join(str(tmp_path), "b")
    write_kapila_bubble(d, n=16)
    p = os.path.join(d, file)
    s = open(p).read()
    assert old in s
    open(p, "w").write(s.replace(old, new, 1))
    with pytest.raises(EcgError, match=msg):
        read_case(d)


def test_missing_eos_directory(tmp_path):
    d = os.path.join(str(tmp_path), "b")
    write_kapila_bubble(d, n=16)
    with pytest.raises(EcgError, match="libEOS"):
        read_case(d)


REF = "/root/reference/libTests/referenceTestCases"


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference test cases not present")
def test_reference_cases_in_scope_are_read():
    """Every reference test case the library builds for is read; the others
    are rejected with a reason (never misread)."""
    import glob
    n_ok = 0
    for m in sorted(glob.glob(REF + "/**/main.xml", recursive=True)):
        d = os.path.dirname(m)
        try:
            c = read_case(d)
        except EcgError as e:
            assert "not built" in str(e) or "only" in str(e), (d, str(e))
            continue
        n_ok += 1
        assert c["nx"] > 1 and len(c["domains"]) >= 1
    assert n_ok >= 5


def test_thinc_volume_fraction_limiter(tmp_path):
    write_eos(str(tmp_path))
    d = os.path.join(str(tmp_path), "b")
    write_kapila_bubble(d, n=16, second_order=True)
    _edit(os.path.join(d, "main.xml"), "<globalVolumeFractionLimiter>vanleer</globalVolumeFractionLimiter>",
          "<globalVolumeFractionLimiter>THINC</globalVolumeFractionLimiter>")
    c = read_case(d)
    assert c["limiters"] == ("minmod", "superbee", "thinc", "superbee")
    _edit(os.path.join(d, "main.xml"), "<interfaceLimiter>superbee</interfaceLimiter>",
          "<interfaceLimiter>thinc</interfaceLimiter>")
    with pytest.raises(EcgError, match="only be a volume-fraction limiter"):
        read_case(d)


def test_finite_rate_relaxation(tmp_path):
    write_eos(str(tmp_path))
    d = os.path.join(str(tmp_path), "b")
    write_kapila_bubble(d, n=16, second_order=False, model="VelocityEq")
    p = os.path.join(d, "model.xml")
    _edit(p, '<relaxation type="P" speed="infinite"/>', '<relaxation type="P" speed="finite" rate="7.5e-4" solver="euler"/>')
    c = read_case(d)
    assert c["relax_p"] is True and c["relax_rate"] == 7.5e-4
    _edit(p, 'solver="euler"', 'solver="LSODA"')
    with pytest.raises(EcgError, match="LSODA"):
        read_case(d)
