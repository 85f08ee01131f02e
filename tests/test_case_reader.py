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


TANK_XM = '<boundCond name="CLXm" type="nonReflecting" number="1"/>'


def _tank_xml(fluids, data='<dataInletTank p0="3.e5" T0="300."/>'):
    inner = "".join(f'\n      <dataFluid EOS="{e}" {a}/>' for e, a in fluids)
    return (f'<boundCond name="CLXm" type="inletTank" number="1">\n    {data}\n    <fluidsProp>{inner}\n'
            f'    </fluidsProp>\n  </boundCond>')


def test_inlet_tank_boundary(tmp_path):
    """<boundCond type="inletTank"> (BoundCondInletTank.cpp:37-165): volume fractions
    matched to the phases by EOS name and normalised by their sum; mass fractions
    turned into volume fractions through the phase densities at (p0, T0)."""
    write_eos(str(tmp_path))
    d = os.path.join(str(tmp_path), "b")
    write_kapila_bubble(d, n=16, second_order=True)
    p = os.path.join(d, "initialConditions.xml")
    base = open(p).read()
    open(p, "w").write(base.replace(TANK_XM, _tank_xml([("IG_air.xml", 'alpha="0.3"'),
                                                        ("SG_water.xml", 'alpha="0.7000001"')])))
    c = read_case(d)
    assert c["bc"][0] == "inletTank" and c["bc"][1] == "nonreflecting"
    t = c["tanks"][0]
    assert t["p0"] == 3.e5 and t["T0"] == 300.
    np.testing.assert_allclose(t["alpha"], [0.7000001 / 1.0000001, 0.3 / 1.0000001], rtol=1e-15)
    o = Oracle(c)
    for _ in range(3):
        o.step(1.e-7)
    assert np.all(np.isfinite(o.get_state()))
    # mass fractions
    open(p, "w").write(base.replace(TANK_XM, _tank_xml([("SG_water.xml", 'massFrac="0.5"'),
                                                        ("IG_air.xml", 'massFrac="0.5"')])))
    c = read_case(d)
    rw = (3.e5 + 6.e8) / ((4.4 - 1.) * 1000. * 300.)
    ra = 3.e5 / ((1.4 - 1.) * 800. * 300.)
    vw, va = 0.5 / rw, 0.5 / ra
    np.testing.assert_allclose(c["tanks"][0]["alpha"], [vw / (vw + va), va / (vw + va)], rtol=1e-14)
    # a single phase takes alpha 1 without <fluidsProp>
    ds = os.path.join(str(tmp_path), "sod")
    write_sod(ds, n=50)
    ps = os.path.join(ds, "initialConditions.xml")
    _edit(ps, TANK_XM, '<boundCond name="CLXm" type="inletTank" number="1"><dataInletTank p0="2.e5" T0="350."/>'
                       '</boundCond>')
    c = read_case(ds)
    assert c["bc"][0] == "inletTank" and c["tanks"][0]["p0"] == 2.e5 and c["tanks"][0]["alpha"] == [1.0]


@pytest.mark.parametrize("fluids,data,msg", [
    ([("IG_air.xml", 'alpha="0.3"'), ("SG_water.xml", 'massFrac="0.7"')], None, "only one of"),
    ([("IG_air.xml", 'alpha="1."')], None, "per phase"),
    ([("IG_air.xml", 'alpha="0.5"'), ("SG_water.xml", 'alpha="0.7"')], None, "sum of alpha"),
    ([("IG_air.xml", 'alpha="-0.5"'), ("SG_water.xml", 'alpha="1.5"')], None, r"in \[0,1\]"),
    ([("IG_air.xml", 'alpha="0.3"'), ("SG_vapor.xml", 'alpha="0.7"')], None, "unknown"),
    ([("IG_air.xml", ""), ("SG_water.xml", "")], None, "One of following"),
    ([("IG_air.xml", 'alpha="0.3"'), ("SG_water.xml", 'alpha="0.7"')], '<dataInletTank p0="3.e5"/>', "T0"),
    ([("IG_air.xml", 'alpha="0.3"'), ("SG_water.xml", 'alpha="0.7"')], "<other/>", "dataInletTank"),
])
def test_malformed_inlet_tanks_are_rejected(tmp_path, fluids, data, msg):
    write_eos(str(tmp_path))
    d = os.path.join(str(tmp_path), "b")
    write_kapila_bubble(d, n=16)
    p = os.path.join(d, "initialConditions.xml")
    kw = {} if data is None else dict(data=data)
    _edit(p, TANK_XM, _tank_xml(fluids, **kw))
    with pytest.raises(EcgError, match=msg):
        read_case(d)
