"""Case descriptions of the benchmark configurations and their small replicas.

A case is a dict with the mesh (``nx, ny, nz, lx, ly, lz``), the flow model
(``"euler"``, ``"pueq"`` = PRESSUREVELOCITYEQ (Kapila), ``"ueq"`` =
VELOCITYEQ), one EOS dict per phase, ``order``, ``limiters`` (global,
interface, global volume fraction, interface volume fraction), ``bc`` (X-, X+,
Y-, Y+, Z-, Z+), ``cfl``, ``alpha_null``, ``relax_p`` and the ordered list of
initial-condition ``domains`` (later domains override earlier ones, as
``Cell::fill`` does).  Random bubble / sphere lists come from
``numpy.random.default_rng(seed)`` and are stored literally in the case, so
the XML writer and every solver see the same inclusions.
"""
import math

import numpy as np

SEED = 20261202

# libEOS parameter sets (SG_water.xml, IG_air.xml, SG_epoxy.xml)
SG_WATER = dict(type="SG", gamma=4.4, pinf=6.e8, cv=1000.0, eref=0.0, name="SG_water.xml")
IG_AIR = dict(type="IG", gamma=1.4, pinf=0.0, cv=800.0, eref=0.0, name="IG_air.xml")
SG_EPOXY = dict(type="SG", gamma=2.43, pinf=5.3e9, cv=1800.0, eref=0.0, name="SG_epoxy.xml")

EPS_PURE = 1.e-6   # volume fraction of an absent phase


def hugoniot_density(eos, p0, rho0, p1):
    """EosSG/EosIG::computeDensityHugoniot (EosSG.cpp:152-159, EosIG.cpp:150-157)."""
    g = eos["gamma"]
    pinf = eos.get("pinf", 0.0) if eos["type"] == "SG" else 0.0
    num = (g + 1.) * (p1 + pinf) + (g - 1.) * (p0 + pinf)
    den = (g - 1.) * (p1 + pinf) + (g + 1.) * (p0 + pinf)
    return rho0 * num / den


def shocked_state(eos, p0, rho0, p1):
    """Post-shock density and velocity of a right-running shock into fluid at rest."""
    rho1 = hugoniot_density(eos, p0, rho0, p1)
    u1 = math.sqrt((p1 - p0) * (1.0 / rho0 - 1.0 / rho1))
    return rho1, u1


def sod_case(n=1000, ly=0.1, lz=0.1):
    """1D Euler, ideal gas, Sod-like data (rho, p) = (1, 1e5) | (0.125, 1e4) at x = 0.5."""
    return dict(name=f"euler_sod_{n}", nx=n, ny=1, nz=1, lx=1.0, ly=ly, lz=lz, model="euler",
                eos=[dict(IG_AIR)], order=1, limiters=("none",) * 4, cfl=0.8, bc=("nonreflecting",) * 6,
                t_end=5.e-4,
                domains=[dict(kind="entire", rho=[0.125], p=[1.e4], vel=[0., 0., 0.]),
                         dict(kind="halfspace", axis=0, origin=0.5, dir=-1, rho=[1.0], p=[1.e5], vel=[0., 0., 0.])])


def _two_phase_states(p, rho_w=1000.0, rho_a=1.2, eps=EPS_PURE):
    water = dict(alpha=[1. - eps, eps], rho=[rho_w, rho_a], p=[p, p])
    air = dict(alpha=[eps, 1. - eps], rho=[rho_w, rho_a], p=[p, p])
    return water, air


def kapila_bubbles_2d(n=2048, nbub=64, order=1, seed=SEED, p0=1.e5, p_shock=1.e9, x_shock=0.05, model="pueq"):
    """2D two-phase shock-bubble case: water with `nbub` air discs hit by a planar water shock."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0.0, 1.0, size=(nbub, 2))
    radii = rng.uniform(0.01, 0.04, size=nbub)
    water, air = _two_phase_states(p0)
    rho1, u1 = shocked_state(SG_WATER, p0, 1000.0, p_shock)
    shocked = dict(alpha=[1. - EPS_PURE, EPS_PURE], rho=[rho1, 1.2], p=[p_shock, p_shock])
    doms = [dict(kind="entire", vel=[0., 0., 0.], **water)]
    for (cx, cy), r in zip(centers, radii):
        doms.append(dict(kind="disc", axis=0, axis2=1, center=(float(cx), float(cy), 0.0), radius=float(r),
                         vel=[0., 0., 0.], **air))
    doms.append(dict(kind="halfspace", axis=0, origin=x_shock, dir=-1, vel=[u1, 0., 0.], **shocked))
    lim = ("minmod",) * 4 if order == 2 else ("none",) * 4
    return dict(name=f"{model}_2d_bubbles_{n}_o{order}", nx=n, ny=n, nz=1, lx=1.0, ly=1.0, lz=1.0,
                model=model, eos=[dict(SG_WATER), dict(IG_AIR)], order=order, limiters=lim, cfl=0.8,
                alpha_null=False, relax_p=True, bc=("nonreflecting",) * 6, domains=doms)


def ueq3_spheres_3d(n=256, nsph=64, seed=SEED, p0=1.e5, p_shock=1.e9, x_shock=0.05, order=1):
    """3D three-phase (water, epoxy, air) pressure-non-equilibrium case with random inclusions."""
    rng = np.random.default_rng(seed)
    e = EPS_PURE
    water = dict(alpha=[1. - 2 * e, e, e], rho=[1000.0, 1185.0, 1.2], p=[p0] * 3)
    epoxy = dict(alpha=[e, 1. - 2 * e, e], rho=[1000.0, 1185.0, 1.2], p=[p0] * 3)
    air = dict(alpha=[e, e, 1. - 2 * e], rho=[1000.0, 1185.0, 1.2], p=[p0] * 3)
    rho1, u1 = shocked_state(SG_WATER, p0, 1000.0, p_shock)
    shocked = dict(alpha=[1. - 2 * e, e, e], rho=[rho1, 1185.0, 1.2], p=[p_shock] * 3)
    doms = [dict(kind="entire", vel=[0., 0., 0.], **water)]
    for q in range(nsph):
        st = epoxy if q % 2 == 0 else air
        c = rng.uniform(0.1, 1.0, size=3)
        r = rng.uniform(4.0, 16.0) / n
        doms.append(dict(kind="sphere", center=tuple(float(v) for v in c), radius=float(r), vel=[0., 0., 0.], **st))
    doms.append(dict(kind="halfspace", axis=0, origin=x_shock, dir=-1, vel=[u1, 0., 0.], **shocked))
    lim = ("minmod",) * 4 if order == 2 else ("none",) * 4
    return dict(name=f"ueq3_3d_spheres_{n}", nx=n, ny=n, nz=n, lx=1.0, ly=1.0, lz=1.0, model="ueq",
                eos=[dict(SG_WATER), dict(SG_EPOXY), dict(IG_AIR)], order=order, limiters=lim, cfl=0.8,
                alpha_null=False, relax_p=True, bc=("nonreflecting",) * 6, domains=doms)


def drive_tank(eos, p_drive, rho1, u1, eps=EPS_PURE):
    """Inlet tank (BoundCondInletTank) that feeds a stream of pressure p_drive, density rho1 and
    speed u1: stagnation pressure p_drive + rho1 u1^2 / 2, and the temperature at which the EOS
    gives the stream's isentrope density there."""
    g, pinf, cv = eos["gamma"], eos.get("pinf", 0.0), eos["cv"]
    p0 = p_drive + 0.5 * rho1 * u1 * u1
    rho0 = rho1 * ((p0 + pinf) / (p_drive + pinf)) ** (1.0 / g)
    return dict(p0=p0, T0=(p0 + pinf) / ((g - 1.0) * cv * rho0), alpha=[1. - eps, eps])


def ueq2_pockets_3d(n=512, npockets=4096, seed=SEED, p0=1.e5, p_drive=5.e8, x_drive=0.04, order=2,
                    inflow_tank=False):
    """3D two-phase water/air pressure-non-equilibrium case: air pockets, high-pressure slab drive;
    with inflow_tank the slab is fed through the x- side by an inlet tank (drive_tank)."""
    rng = np.random.default_rng(seed)
    water, air = _two_phase_states(p0)
    rho1, u1 = shocked_state(SG_WATER, p0, 1000.0, p_drive)
    driven = dict(alpha=[1. - EPS_PURE, EPS_PURE], rho=[rho1, 1.2], p=[p_drive, p_drive])
    centers = rng.uniform(0.0, 1.0, size=(npockets, 3))
    radii = np.exp(rng.uniform(math.log(2.0), math.log(10.0), size=npockets)) / n
    doms = [dict(kind="entire", vel=[0., 0., 0.], **water)]
    for c, r in zip(centers, radii):
        doms.append(dict(kind="sphere", center=tuple(float(v) for v in c), radius=float(r), vel=[0., 0., 0.], **air))
    doms.append(dict(kind="halfspace", axis=0, origin=x_drive, dir=-1, vel=[u1, 0., 0.], **driven))
    lim = ("minmod",) * 4 if order == 2 else ("none",) * 4
    case = dict(name=f"ueq2_3d_pockets_{n}_o{order}", nx=n, ny=n, nz=n, lx=1.0, ly=1.0, lz=1.0, model="ueq",
                eos=[dict(SG_WATER), dict(IG_AIR)], order=order, limiters=lim, cfl=0.8, alpha_null=False,
                relax_p=True, bc=("nonreflecting",) * 6, domains=doms)
    if inflow_tank:
        case["bc"] = ("inletTank",) + ("nonreflecting",) * 5
        case["tanks"] = {0: drive_tank(SG_WATER, p_drive, rho1, u1)}
        case["name"] += "_tank"
    return case


def variable_names(case):
    """Names of the persistent SoA arrays, in storage order."""
    D = (case["nx"] > 1) + (case["ny"] > 1) + (case["nz"] > 1)
    names = []
    if case["model"] == "euler":
        names += ["rho", "p"]
    else:
        for k in range(len(case["eos"])):
            names += [f"alpha{k}", f"rho{k}", f"p{k}"]
    names += ["u", "v", "w"][:D]
    names.append("E")
    return names
