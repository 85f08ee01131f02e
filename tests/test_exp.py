"""The exponential of the kernels (ecg_exp, the same code on host and device) is
correctly rounded: it equals the exact exponential rounded once to double,
computed here with 60-digit decimal arithmetic."""
import decimal

import numpy as np

from ecogen_amd import library


def _cr(x):
    with decimal.localcontext() as ctx:
        ctx.prec = 60
        return float(decimal.Decimal(x).exp())


def test_exp_is_correctly_rounded():
    L = library()
    rng = np.random.default_rng(11)
    xs = np.concatenate([rng.uniform(-1.6, 1.6, 60000), rng.uniform(-700., 709., 5000),
                         rng.uniform(-1e-8, 1e-8, 2000), [0.0, 1.0, -1.0, 1.6, -1.6, 0.5 * np.log(2.0)]])
    bad = [x for x in xs if L.ecg_exp(float(x)) != _cr(float(x))]
    assert not bad, (len(bad), bad[:5])
    assert L.ecg_exp(float("nan")) != L.ecg_exp(float("nan"))
