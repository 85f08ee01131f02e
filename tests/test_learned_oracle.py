"""Pins of the learned warm start's CPU checker (oracle/mlp_oracle.py), CPU only."""
import ctypes as C

import numpy as np

import oracle.pyoracle as po
from oracle import mlp_oracle as mo
from ecogen_amd.case import IG_AIR, SG_EPOXY, SG_WATER

DP = C.POINTER(C.c_double)
EOS3 = [SG_WATER, SG_EPOXY, IG_AIR]


def _random_records(rng, n):
    rec = np.zeros((n, mo.SAMPLE_FLOATS))
    for i in range(n):
        k = rng.integers(2, 4)
        a = rng.dirichlet(np.ones(k))
        p = 10 ** rng.uniform(4.5, 9, size=k)
        f = mo.features(a, p, EOS3[:k])
        rec[i, :12] = f[0]
        u = rng.normal(0, 1.5)
        rec[i, 15] = f[5]
        rec[i, 14] = u
        rec[i, 12] = f[3] + f[5] * u
    return rec


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(11)
    prm = rng.normal(0, 0.3, mo.N_PARAM)
    rec = _random_records(rng, 40)
    loss, g = mo.loss_and_grad(prm, rec)
    idx = np.concatenate([rng.choice(np.arange(mo.OFF_B1), 20, replace=False),
                          mo.OFF_B1 + np.arange(0, 64, 9),
                          rng.choice(np.arange(mo.OFF_W2, mo.OFF_B2), 30, replace=False),
                          mo.OFF_B2 + np.arange(0, 64, 7), mo.OFF_W3 + np.arange(0, 64, 5), [mo.OFF_B3]])
    for i in idx:
        h = 1e-6 * max(1.0, abs(prm[i]))
        pp, pm = prm.copy(), prm.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_and_grad(pp, rec)[0] - mo.loss_and_grad(pm, rec)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (abs(g[i]) + 1e-3 * np.abs(g).max()), (i, fd, g[i])


def test_forward_agrees_with_loss_path():
    rng = np.random.default_rng(3)
    prm = rng.normal(0, 0.2, mo.N_PARAM)
    rec = _random_records(rng, 25)
    y = mo.forward(prm, rec[:, :12])
    r = mo.target(rec)
    assert abs(np.mean((y - r) ** 2) - mo.loss_and_grad(prm, rec)[0]) <= 1e-14 * (1 + np.mean(r * r))


def test_tlin_is_the_linearised_root():
    """sum_k alpha_k (p_lin - p_k) / Z_k = 0 at p_lin = p_cold + s t_lin."""
    rng = np.random.default_rng(5)
    for _ in range(300):
        k = rng.integers(2, 4)
        a = rng.dirichlet(np.ones(k))
        p = 10 ** rng.uniform(4, 9.5, size=k)
        x, pc, s, tl, _, _ = mo.features(a, p, EOS3[:k])
        Z = [e["gamma"] * (p[j] + e["pinf"]) for j, e in enumerate(EOS3[:k])]
        pl = pc + s * tl
        res = sum(a[j] * (pl - p[j]) / Z[j] for j in range(k))
        scale = sum(a[j] * abs(p[j]) / Z[j] for j in range(k)) + sum(a[j] * abs(pl) / Z[j] for j in range(k))
        assert abs(res) <= 1e-12 * scale
        assert np.isclose(sum(x[4 * j + 2] for j in range(k)), 1.0, rtol=0, atol=1e-14)
        assert x[3 * 4:].sum() == 0.0 or k == 3


def test_relaxed_root_tends_to_tlin_for_small_spreads():
    """The normalised Newton root of the oracle relaxation approaches t_lin as s/Z -> 0."""
    L = po.lib()
    eos = (po.OrcEos * 3)(*[po.make_eos(e) for e in EOS3])
    rng = np.random.default_rng(9)
    rho0 = np.array([1000., 1185., 1.2])
    for _ in range(100):
        k = 3
        a = rng.dirichlet(np.ones(k))
        a = np.clip(a, 1e-3, None)
        a /= a.sum()
        base = 10 ** rng.uniform(5, 8)
        eps_rel = 10 ** rng.uniform(-6, -3)
        p = base * (1 + eps_rel * rng.uniform(-1, 1, size=k))
        x, pc, s, tl, _, sg = mo.features(a, p, EOS3)
        aa, rr, pp, ps = a.copy(), rho0.copy(), p.copy(), np.zeros(1)
        it = L.orc_relax_cell(k, eos, 0.0, 0, aa.ctypes.data_as(DP), rr.ctypes.data_as(DP),
                              pp.ctypes.data_as(DP), ps.ctypes.data_as(DP))
        assert 0 < it < 100
        t = (ps[0] - pc) / s
        Zmin = min(e["gamma"] * (p[j] + e["pinf"]) for j, e in enumerate(EOS3))
        # Newton tolerance in pressure is ~1e-10 * Z; the nonlinearity is O(s / Z)
        assert abs(t - tl) <= 20 * s / Zmin + 2e-10 * max(e["gamma"] * (p[j] + e["pinf"])
                                                           for j, e in enumerate(EOS3)) / s


def test_sigma_is_the_second_order_remainder():
    """For small spreads t - t_lin approaches the second-order term of the
    rational relaxation equation, sum_k w_k (t_lin - x_k1)^2 s / (p_k + pinf_k),
    which is sigma (computed with Z_k = gamma_k (p_k + pinf_k)) times gamma-weighted
    factors in [min gamma, max gamma]."""
    L = po.lib()
    eos = (po.OrcEos * 3)(*[po.make_eos(e) for e in EOS3])
    rng = np.random.default_rng(13)
    rho0 = np.array([1000., 1185., 1.2])
    g = np.array([e["gamma"] for e in EOS3])
    for _ in range(100):
        a = rng.dirichlet(np.ones(3))
        a = np.clip(a, 1e-2, None)
        a /= a.sum()
        base = 10 ** rng.uniform(6, 8)
        p = base * (1 + 1e-3 * rng.uniform(-1, 1, size=3))
        x, pc, s, tl, _, sg = mo.features(a, p, EOS3)
        aa, rr, pp, ps = a.copy(), rho0.copy(), p.copy(), np.zeros(1)
        it = L.orc_relax_cell(3, eos, 0.0, 0, aa.ctypes.data_as(DP), rr.ctypes.data_as(DP),
                              pp.ctypes.data_as(DP), ps.ctypes.data_as(DP))
        assert 0 < it < 100
        rem = (ps[0] - pc) / s - tl
        # second-order term with the exact s/(p_k + pinf_k) = gamma_k s / Z_k
        r2 = sum(x[4 * k + 2] * (tl - x[4 * k + 1]) ** 2 * g[k] * (s / (EOS3[k]["gamma"] * (p[k] + EOS3[k]["pinf"])))
                 for k in range(3))
        tol = 1e-10 * max(e["gamma"] * (p[k] + e["pinf"]) for k, e in enumerate(EOS3)) / s
        assert abs(rem - r2) <= 0.05 * abs(r2) + tol
        assert g.min() * sg * 0.9 - tol <= abs(r2) <= g.max() * sg * 1.1 + tol
