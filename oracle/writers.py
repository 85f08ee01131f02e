"""Oracle of the reference's result files (test infrastructure only).

A plain restatement of what the reference writes when it prints a Cartesian
case, used by the tests to check the library's writers byte for byte:

  printed variables   Phase*/Mix*::returnScalar / returnVector
                      (src/Models/UEq/PhaseUEq.cpp:115-164, PUEq/PhasePUEq.cpp:87-129,
                      Euler/PhaseEuler.cpp:131-185, UEq/MixUEq.cpp:136-259),
                      reducedOutput lists of src/InputOutput/Input.cpp:601-631
  density gradient    GradientFiniteDifference::computeGradients
                      (src/Gradients/GradientFiniteDifference.cpp:45-122) over the
                      faces of MeshCartesian::initializeGeometrieMonoCpu (Meshes/MeshCartesian.cpp:212-557)
  GNU files           OutputGNU::writeResults, Cell::printGnuplotAMR, writeScriptGnuplot,
                      printBlocGnuplot (src/InputOutput/OutputGNU.cpp:61-226, Order1/Cell.cpp:759-792)
  VTK XML files       OutputXML::writeResultsXML, writeMeshRectilinearXML, writePhysicalDataXML,
                      writeCollectionXML, initializeSpecificOutput (src/InputOutput/OutputXML.cpp:58-783),
                      Output::writeDataset + IO::writeb64Chaine (Output.cpp:376-428, IO.cpp:44-88)
  infoCalcul, infoMesh Output::saveInfos, saveInfoCells (Output.cpp:361-372, 494-519)
  global quantities   OutputGlobalGNU (src/InputOutput/OutputGlobalGNU.cpp:68-117), Cell::computeTotalMass /
                      computeTotalEnergy (Order1/Cell.cpp:863-895)
  restart             OutputXML::ReadPhysicalDataXML (OutputXML.cpp:194-270), ModPUEq::fulfillStateRestart
                      (Models/PUEq/ModPUEq.cpp:76-82), Cell::completeFulfillState (Order1/Cell.cpp:344-353)

Numbers are formatted as C++ streams format a double under the default float
field: printf's %g with the stream precision (6 unless <outputMode precision> is set).

Checked against: the C oracle's EOS temperature (itself pinned bit for bit
against the reference-compiled EOS), closed-form gradients of linear fields,
XML well-formedness and round trips of every ASCII and base64 array
(tests/test_output_oracle.py).
"""
import base64 as _base64
import math
import os
import xml.etree.ElementTree as ET

import numpy as np


def fmt(x, prec):
    """C++ `stream << double` under the default float field."""
    return "%.*g" % (prec if prec else 6, float(x))


def _axis(n, l):
    d = l / float(n)
    return d, [(i + 0.5) * d for i in range(n)]


def dimension(case):
    return 1 + (case["ny"] > 1) + (case["nz"] > 1)


def _eps(case):
    return 1.e-15 if (case["model"] != "euler" and case.get("alpha_null", False)) else 0.0


def temperature(e, eps, rho, p):
    rho = np.maximum(rho, eps)
    if e["type"].upper() == "SG":
        return (p + e["pinf"]) / (e["gamma"] - 1.) / rho / e["cv"]
    return p / (e["gamma"] - 1.) / rho / e["cv"]


def energy(e, eps, rho, p):
    rho = np.maximum(rho, eps)
    if e["type"].upper() == "SG":
        return (p + e["gamma"] * e["pinf"]) / (e["gamma"] - 1.) / rho + e.get("eref", 0.0)
    return p / (e["gamma"] - 1.) / rho + e.get("eref", 0.0)


def mixture_density(case, st):
    if case["model"] == "euler":
        return st[0].copy()
    r = np.zeros(st.shape[1])
    for k in range(len(case["eos"])):
        r = r + st[3 * k] * st[3 * k + 1]
    return r


def printed(case, st, reduced=False):
    """[(name, owner, is_vector, values)] in print order; vectors as (ncell, 3)."""
    D = dimension(case)
    n = st.shape[1]
    eps = _eps(case)
    out = []

    def vel(off):
        v = np.zeros((n, 3))
        for a in range(D):
            v[:, a] = st[off + a]
        return v

    if case["model"] == "euler":
        e = case["eos"][0]
        out.append(("Density", 0, False, st[0].copy()))
        out.append(("Pressure", 0, False, st[1].copy()))
        out.append(("Temperature", 0, False, temperature(e, eps, st[0], st[1])))
        out.append(("Velocity", 0, True, vel(2)))
        return out
    N = len(case["eos"])
    rho = mixture_density(case, st)
    pm = np.zeros(n)
    for k in range(N):
        pm = pm + st[3 * k] * st[3 * k + 2]
    ueq = case["model"] == "ueq"
    for k in range(N):
        a, r, p = st[3 * k], st[3 * k + 1], st[3 * k + 2]
        T = temperature(case["eos"][k], eps, r, p)
        Y = a * r / np.maximum(rho, eps)
        out.append(("Alpha", k, False, np.where(a < 1e-20, 0.0, a)))
        out.append(("Density", k, False, np.where(r < 1e-20, 1e-20, r)))
        if ueq:
            out.append(("Pressure", k, False, p.copy()))
        if not reduced:
            out.append(("Temperature", k, False, np.where(T < 1e-20, 1e-20, T)))
            out.append(("Mass fraction", k, False, np.where(Y < 1e-20, 0.0, Y)))
    out.append(("Density_Mixture", -1, False, rho))
    out.append(("Pressure_Mixture", -1, False, pm))
    if not reduced:
        out.append(("Total_energy_Mixture", -1, False, st[3 * N + D].copy()))
    out.append(("Velocity_Mixture", -1, True, vel(3 * N)))
    return out


def grad_rho(case, rho):
    """|grad rho| of the finite-difference gradient, cell by cell."""
    nx, ny, nz = case["nx"], case["ny"], case["nz"]
    dx, px = _axis(nx, case["lx"])
    dy, py = _axis(ny, case["ly"])
    dz, pz = _axis(nz, case["lz"])
    out = np.zeros(nx * ny * nz)
    idx = lambda i, j, k: i + nx * (j + ny * k)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                g = [0.0, 0.0, 0.0]
                s = [0.0, 0.0, 0.0]
                faces = []
                if i > 0: faces.append(((i - 1, j, k), (i, j, k), 0))
                if i < nx - 1: faces.append(((i, j, k), (i + 1, j, k), 0))
                if j > 0: faces.append(((i, j - 1, k), (i, j, k), 1))
                if j < ny - 1: faces.append(((i, j, k), (i, j + 1, k), 1))
                if k > 0: faces.append(((i, j, k - 1), (i, j, k), 2))
                if k < nz - 1: faces.append(((i, j, k), (i, j, k + 1), 2))
                for (l, r, a) in faces:
                    pl = (px[l[0]], py[l[1]], pz[l[2]])
                    pr = (px[r[0]], py[r[1]], pz[r[2]])
                    v = [pl[q] - pr[q] for q in range(3)]
                    d = [abs(x) for x in v]
                    for q in range(3):
                        s[q] += d[q]
                    dist = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
                    gf = (rho[idx(*r)] - rho[idx(*l)]) / dist
                    for q in range(3):
                        nq = 1.0 if q == a else 0.0
                        g[q] = g[q] + nq * gf * d[q]
                c = (px[i], py[j], pz[k])
                bfaces = []
                if i == 0: bfaces.append((0.0, c[1], c[2]))
                if i == nx - 1: bfaces.append((c[0] + 0.5 * dx, c[1], c[2]))
                if ny > 1:
                    if j == 0: bfaces.append((c[0], 0.0, c[2]))
                    if j == ny - 1: bfaces.append((c[0], c[1] + 0.5 * dy, c[2]))
                if nz > 1:
                    if k == 0: bfaces.append((c[0], c[1], 0.0))
                    if k == nz - 1: bfaces.append((c[0], c[1], c[2] + 0.5 * dz))
                for f in bfaces:
                    for q in range(3):
                        s[q] += abs(c[q] - f[q]) * 2.
                for q in range(3):
                    if s[q] <= 1e-12:
                        s[q] = 1.0
                    g[q] = g[q] / s[q]
                out[idx(i, j, k)] = math.sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2])
    return out


def _stem(name):
    return name[:-4] if name.endswith(".xml") else name


def _gnu_style(s):
    s = s.replace("_", "\\_")
    return _stem(s)


def _b64(b):
    return _base64.b64encode(b).decode()


class Writer:
    """The reference's result files of one case, written into `folder` (= results/<run>/)."""

    def __init__(self, case, folder, fmt_="XML", binary=False, precision=0, reduced=False, eos_names=None,
                 global_quantities=()):
        self.case, self.folder = case, folder
        self.format, self.binary, self.prec, self.reduced = fmt_.upper(), binary, precision, reduced
        self.eos_names = eos_names
        self.gq = list(global_quantities)
        self.num = 0
        self.files = {}          # relative path -> text

    def _dataset(self, values):
        if not self.binary:
            return "".join(fmt(x, self.prec) + " " for x in values)
        data = np.ascontiguousarray(values, dtype="<f8").tobytes()
        return _b64(np.int32(len(data)).tobytes()) + _b64(data)

    # --- start-of-run files
    def open(self):
        b = "B64" if self.binary else ""
        if self.format == "XML":
            self.files[f"collectionParaview{b}.pvd"] = (
                '<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" byte_order="LittleEndian" '
                'compressor="vtkZLibDataCompressor">\n  <Collection>\n\n  </Collection>\n</VTKFile>\n')
            self.files[f"collectionVisIt{b}.visit"] = "!NBLOCKS 1\n"
        for q in self.gq:
            self.files[f"globalQuantities/{q}.out"] = ""
            self.files[f"globalQuantities/plot_{q}.gnu"] = (
                "reset\nset style data lines\nset nokey\n\nset xlabel 't (s)'\n"
                f"set title '{_gnu_style(q)}'\nplot \"{q}.out\" u 1:2\npause(-1)\n\n")
        self.files["infoCalcul.out"] = ""
        self.files["infoMesh/infoMesh_CPU0.out"] = ""
        self.files["errorsAndWarnings/error_CPU0.out"] = ""
        self.files["errorsAndWarnings/warning_CPU0.out"] = ""

    # --- one print
    def write(self, st, iteration, t, dt_next, seconds=0.0):
        c = self.case
        line = ""
        if self.num == 0:
            line += "1\n"
        line += " ".join([str(self.num), str(int(iteration)), fmt(t, self.prec), fmt(dt_next, self.prec),
                          fmt(seconds, self.prec), "0", "0"]) + "\n"
        self.files["infoCalcul.out"] += line
        self.files["infoMesh/infoMesh_CPU0.out"] += f"{c['nx'] * c['ny'] * c['nz']}\n"
        for q in self.gq:
            self.files[f"globalQuantities/{q}.out"] += f"{fmt(t, self.prec)} {fmt(global_quantity(c, st, q), self.prec)}\n"
        pv = printed(c, st, self.reduced)
        if self.format == "GNU":
            self._gnu(pv)
        else:
            self._xml(pv, st)
        self.num += 1

    def _gnu(self, pv):
        c = self.case
        D = dimension(c)
        _, px = _axis(c["nx"], c["lx"])
        _, py = _axis(c["ny"], c["ly"])
        _, pz = _axis(c["nz"], c["lz"])
        rows = []
        n = c["nx"] * c["ny"] * c["nz"]
        for cell in range(n):
            i = cell % c["nx"]
            j = (cell // c["nx"]) % c["ny"]
            k = cell // (c["nx"] * c["ny"])
            r = [fmt(px[i], self.prec)]
            if D >= 2: r.append(fmt(py[j], self.prec))
            if D == 3: r.append(fmt(pz[k], self.prec))
            for (_, _, vec, v) in pv:
                if vec:
                    x, y, z = v[cell]
                    r.append(fmt(math.sqrt(x * x + y * y + z * z), self.prec))
                else:
                    r.append(fmt(v[cell], self.prec))
            r += ["0", "0"]
            rows.append(" ".join(r) + " \n")
        self.files[f"datasets/result_CPU0_TIME{self.num}.out"] = "".join(rows) + "\n"
        s = "reset\nset style data lines\nset nokey\n\nset xlabel 'x (m)'\n"
        if D == 2:
            s += "set surface\nset dgrid3d 50,50\nset contour base\nset cntrparam levels 25\nshow contour\n"
        index = 1 + D
        titles = []
        for (name, owner, _, _) in pv:
            titles.append(_gnu_style(name) + ("\\_" + _gnu_style(self.eos_names[owner]) if owner >= 0 else ""))
        titles += ["AMR level", "Xi"]
        for title in titles:
            s += f"set title '{title}'\n" + ("plot" if D <= 1 else "splot")
            for t in range(self.num + 1):
                s += f' "./datasets/result_CPU0_TIME{t}.out"'
                s += f" u 1:{index}" if D <= 1 else f" u 1:2:{index}"
                if t != self.num:
                    s += ",\\"
                s += "\n"
            s += "pause(-1)\n\n"
            index += 1
        self.files["visualization.gnu"] = s

    def dataset_name(self, num):
        return f"result{'B64' if self.binary else ''}_CPU0_TIME{num}.vtr"

    def _xml(self, pv, st):
        c = self.case
        f = self.binary and "binary" or "ascii"
        ext = f"0 {c['nx']} 0 {c['ny']} 0 {c['nz']}"
        s = ('<?xml version="1.0"?>\n<VTKFile type="RectilinearGrid" version="0.1" byte_order="LittleEndian">\n'
             f'  <RectilinearGrid WholeExtent="{ext}">\n    <Piece Extent="{ext}">\n      <Coordinates>\n')
        for (n, l) in ((c["nx"], c["lx"]), (c["ny"], c["ly"]), (c["nz"], c["lz"])):
            d, p = _axis(n, l)
            x = [pi - 0.5 * d for pi in p] + [p[-1] + 0.5 * d]
            s += '        <DataArray type="Float64" '
            s += 'format="ascii">\n          ' if not self.binary else 'format="binary">\n'
            s += self._dataset(x) + "\n        </DataArray>\n"
        s += "      </Coordinates>\n      <CellData>\n"
        multiphase = c["model"] != "euler"
        for (name, owner, vec, v) in pv:
            if owner < 0 and not multiphase:
                continue
            nm = f"F{owner}_{name}_{_stem(self.eos_names[owner])}" if owner >= 0 else name
            s += f'        <DataArray type="Float64" Name="{nm}"'
            if vec:
                s += ' NumberOfComponents="3"'
            s += f' format="{f}">\n'
            s += self._dataset(np.asarray(v).reshape(-1)) + "\n        </DataArray>\n"
        s += f'        <DataArray type="Float64" Name="gradRho" format="{f}">\n'
        s += self._dataset(grad_rho(c, mixture_density(c, st))) + "\n        </DataArray>\n"
        s += "      </CellData>\n    </Piece>\n  </RectilinearGrid>\n</VTKFile>\n"
        self.files["datasets/" + self.dataset_name(self.num)] = s
        b = "B64" if self.binary else ""
        pvd = ('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" byte_order="LittleEndian" '
               'compressor="vtkZLibDataCompressor">\n    <Collection>\n')
        vis = "!NBLOCKS 1\n"
        for t in range(self.num + 1):
            pvd += f'        <DataSet timestep="{t}" part="0" file="datasets/{self.dataset_name(t)}"/>\n'
            vis += f"datasets/{self.dataset_name(t)}\n"
        pvd += "    </Collection>\n</VTKFile>\n"
        self.files[f"collectionParaview{b}.pvd"] = pvd
        self.files[f"collectionVisIt{b}.visit"] = vis


def global_quantity(case, st, q):
    vol = (case["lx"] / float(case["nx"])) * (case["ly"] / float(case["ny"])) * (case["lz"] / float(case["nz"]))
    rho = mixture_density(case, st)
    D = dimension(case)
    E = st[2 + D] if case["model"] == "euler" else st[3 * len(case["eos"]) + D]
    acc = 0.0
    if q == "mass":
        for i in range(rho.size):
            acc += rho[i] * vol
    else:
        for i in range(rho.size):
            acc += rho[i] * E[i] * vol
    return acc


def read_restart(case, path, reduced=False):
    """State a restart rebuilds from an ASCII .vtr file (arrays in print order)."""
    root = ET.parse(path).getroot()
    arrays = root.find("RectilinearGrid").find("Piece").find("CellData").findall("DataArray")
    n = case["nx"] * case["ny"] * case["nz"]
    names = [(nm, ow, vec) for (nm, ow, vec, _) in printed(case, np.ones((nvar(case), n)), reduced)
             if not (ow < 0 and case["model"] == "euler")]
    vals = {}
    for q, (nm, ow, vec) in enumerate(names):
        v = np.array([float(x) for x in arrays[q].text.split()])
        vals[(nm, ow)] = v.reshape(n, 3) if vec else v[:n]
    D = dimension(case)
    eps = _eps(case)
    if case["model"] == "euler":
        st = np.zeros((D + 3, n))
        st[0], st[1] = vals[("Density", 0)], vals[("Pressure", 0)]
        u = vals[("Velocity", 0)]
        for a in range(D):
            st[2 + a] = u[:, a]
        st[2 + D] = energy(case["eos"][0], eps, st[0], st[1]) + 0.5 * (u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1] + u[:, 2] * u[:, 2])
        return st
    N = len(case["eos"])
    st = np.zeros((3 * N + D + 1, n))
    pm = vals[("Pressure_Mixture", -1)]
    for k in range(N):
        st[3 * k] = vals[("Alpha", k)]
        st[3 * k + 1] = vals[("Density", k)]
        st[3 * k + 2] = vals[("Pressure", k)] if case["model"] == "ueq" else pm
    if case["model"] == "pueq":
        for k in range(N):
            a, r, p = st[3 * k], st[3 * k + 1], st[3 * k + 2]
            if eps > 1e-20:
                a[a < 0.] = 0.
                a[a > 1.] = 1.
                r[r <= 1e-15] = 1e-15
            else:
                a[a < 1e-15] = 1e-14
                a[a > 1. - 1e-15] = 1. - 1e-14
                r[r < 1e-15] = 1e-14
            e = case["eos"][k]
            if e["type"].upper() == "SG":
                floor = -(1. - 1.e-15) * e["pinf"] + 1.e-15
                p[p <= floor] = floor
            else:
                p[p < 1e-8] = 1e-8
    u = vals[("Velocity_Mixture", -1)]
    for a in range(D):
        st[3 * N + a] = u[:, a]
    rho = np.zeros(n)
    for k in range(N):
        rho = rho + st[3 * k] * st[3 * k + 1]
    em = np.zeros(n)
    for k in range(N):
        Y = st[3 * k] * st[3 * k + 1] / np.maximum(rho, eps)
        em = em + Y * energy(case["eos"][k], eps, st[3 * k + 1], st[3 * k + 2])
    st[3 * N + D] = em + 0.5 * (u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1] + u[:, 2] * u[:, 2])
    return st


def nvar(case):
    D = dimension(case)
    return D + 3 if case["model"] == "euler" else 3 * len(case["eos"]) + D + 1


def compare_tree(folder, files, skip_timing=True):
    """Differences between the files under `folder` and the oracle's `files` (relative paths)."""
    bad = []
    for rel, text in files.items():
        p = os.path.join(folder, rel)
        if not os.path.exists(p):
            bad.append(f"{rel}: missing")
            continue
        got = open(p).read()
        if rel == "infoCalcul.out" and skip_timing:
            got = "\n".join(" ".join(l.split()[:4]) for l in got.splitlines())
            text = "\n".join(" ".join(l.split()[:4]) for l in text.splitlines())
        if got != text:
            for ln, (a, b) in enumerate(zip(got.splitlines(), text.splitlines())):
                if a != b:
                    bad.append(f"{rel}:{ln + 1}: got {a[:160]!r} expected {b[:160]!r}")
                    break
            else:
                bad.append(f"{rel}: lengths {len(got)} vs {len(text)}")
    return bad
