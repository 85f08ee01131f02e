"""Write tests/golden/eos_vectors.npz: inputs and outputs of the reference's own
EOS (EosSG.cpp, EosIG.cpp) and limiter (Order2/Limiter*.cpp) code.

Run in a container where the reference tree is mounted, after
`make -C oracle -f ref.mk` has compiled those sources into oracle/_ref/ref_eos.
The vectors let the CPU test-suite pin the oracle's restatement bit for bit
where the reference tree is absent.
"""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_eos")

REC = np.dtype([("op", "<i4"), ("kind", "<i4"), ("v", "<f8", (8,))])

EOS_SETS = [  # kind, gamma, pinf, cv, eref
    (1, 4.4, 6.e8, 1000.0, 0.0),      # SG_water
    (1, 2.43, 5.3e9, 1800.0, 0.0),    # SG_epoxy
    (0, 1.4, 0.0, 800.0, 0.0),        # IG_air
    (1, 2.35, 1.e9, 1816.0, -1.167e6),  # SG with a non-zero reference energy
    (0, 1.67, 0.0, 3116.0, 1.e5),
]


def records(seed=7):
    rng = np.random.default_rng(seed)
    recs = []

    def add(op, kind, g, pinf, cv, eref, eps, a, b, c=0.0):
        recs.append((op, kind, (g, pinf, cv, eref, eps, a, b, c)))

    for kind, g, pinf, cv, eref in EOS_SETS:
        for eps in (0.0, 1.0):
            for _ in range(80):
                rho = 10 ** rng.uniform(-3, 3.5)
                p = 10 ** rng.uniform(2, 10) - (0.9 * pinf * rng.uniform() if kind == 1 else 0.0)
                e = 10 ** rng.uniform(3, 7) + eref
                rho2 = rho * 10 ** rng.uniform(-0.5, 0.5)
                pf = 10 ** rng.uniform(2, 10)
                add(0, kind, g, pinf, cv, eref, eps, rho, p)
                add(1, kind, g, pinf, cv, eref, eps, rho, e)
                add(2, kind, g, pinf, cv, eref, eps, rho, p + pinf * 0.95)
                add(3, kind, g, pinf, cv, eref, eps, rho, p)
                add(4, kind, g, pinf, cv, eref, eps, p, rho, rho2)
                add(5, kind, g, pinf, cv, eref, eps, p, rho, pf)
                add(6, kind, g, pinf, cv, eref, eps, p, rho, pf)
            # pressure floors around the thresholds
            thr = -(1. - 1.e-15) * pinf + 1.e-15 if kind == 1 else 1.e-8
            for x in (thr, np.nextafter(thr, -np.inf), np.nextafter(thr, np.inf), thr * 2, -1e3, 0.0, 1e5):
                add(7, kind, g, pinf, cv, eref, eps, x, 0.0)
            # zero density with the alphaNull guard
            add(0, kind, g, pinf, cv, eref, eps, 0.0, 1e5)
            add(2, kind, g, pinf, cv, eref, eps, 0.0, 1e5 + pinf)
    for lim in range(1, 6):
        for _ in range(200):
            s1 = rng.choice([-1, 1]) * 10 ** rng.uniform(-7, 3)
            s2 = rng.choice([-1, 1]) * 10 ** rng.uniform(-7, 3)
            if rng.uniform() < 0.1:
                s2 = 0.0
            recs.append((8, lim, (0, 0, 0, 0, 0, s1, s2, 0)))
        for s1, s2 in ((1e-5, 1e-4), (3e-5, 3.4e-5), (-1e-3, -1e-3), (1e-3, 1e-3), (0.0, 1.0), (2.0, 1.0)):
            recs.append((8, lim, (0, 0, 0, 0, 0, s1, s2, 0)))
    arr = np.zeros(len(recs), dtype=REC)
    for i, (op, kind, v) in enumerate(recs):
        arr[i]["op"], arr[i]["kind"], arr[i]["v"] = op, kind, v
    return arr


def run_reference(arr):
    out = subprocess.run([REF], input=arr.tobytes(), capture_output=True, check=True).stdout
    return np.frombuffer(out, dtype="<f8").reshape(-1, 2)


if __name__ == "__main__":
    arr = records()
    res = run_reference(arr)
    np.savez_compressed(os.path.join(HERE, "eos_vectors.npz"), op=arr["op"], kind=arr["kind"], v=arr["v"], out=res)
    print(f"wrote {len(arr)} records")
