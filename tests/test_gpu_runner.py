"""The command-line runner (ecogen_run) on cases written in the reference's XML
input format: its final state must match the oracle run of the case the reader
produces (needs a GPU)."""
import os
import subprocess

import numpy as np
import pytest

from case_files import write_eos, write_kapila_bubble, write_sod

RUNNER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ecogen_amd", "ecogen_run")
TOL = 1e-12


def _read_out(path):
    with open(path, "rb") as f:
        head = f.readline().split()
        nx, ny, nz, nv = int(head[0]), int(head[1]), int(head[2]), int(head[3])
        t, it = float(head[4]), int(head[5])
        data = np.frombuffer(f.read(), dtype="<f8").reshape(nv, nx * ny * nz)
    return data, t, it


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["sod_time", "pueq_o2", "ueq_o1_wall"])
def test_runner_matches_oracle(tmp_path, which):
    from ecogen_amd import read_case
    from oracle.pyoracle import Oracle
    print(f"[runner] {which}", flush=True)
    write_eos(str(tmp_path))
    d = os.path.join(str(tmp_path), which)
    if which == "sod_time":
        write_sod(d, n=300, t_end=2.e-4)
    elif which == "pueq_o2":
        write_kapila_bubble(d, n=40, iterations=15, second_order=True)
    else:
        write_kapila_bubble(d, n=40, iterations=15, second_order=False, model="VelocityEq")
    out = os.path.join(str(tmp_path), "state.bin")
    r = subprocess.run([RUNNER, d, "--out", out], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    got, t, it = _read_out(out)
    case = read_case(d)
    o = Oracle(case)
    ctl = case["control"]
    if ctl["iterations_mode"]:
        t_o, _, it_o = o.run(max_iter=ctl["iterations"])
    else:
        t_o, _, it_o = o.run(t_end=ctl["total_time"])
    assert it == it_o and abs(t - t_o) <= 1e-12 * t_o
    ref = o.get_state()
    for v in range(ref.shape[0]):
        den = max(float(np.sum(np.abs(ref[v]))), 1e-300)
        err = float(np.sum(np.abs(got[v] - ref[v]))) / den
        assert err <= TOL or float(np.sum(np.abs(got[v] - ref[v]))) <= 1e-300, (v, err)
