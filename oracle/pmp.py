"""Restated reader for pampa's `.pmp` keyword input format.

TEST INFRASTRUCTURE ONLY (part of the oracle): used by tests/, the smoke check
and the cpu_baseline leg of bench.py to build the checker's view of an input.
The product has its own C++ reader.

Follows the reference's observable parsing rules:
  - line cleaning, including the tab quirk: a tab and the character after it
    become one space (input.cxx:4-38);
  - tokenising on single spaces, skipping empty and '#' lines (input.cxx:41-60);
  - keyword dispatch of the main file (Parser.cxx:4-260), materials
    (Material.cxx:18-136), constant nuclear data (ConstantNuclearData.cxx:4-126),
    precursor data (PrecursorData.cxx:4-50), Cartesian meshes
    (CartesianMesh.cxx:4-143), unstructured extruded meshes
    (UnstructuredExtrudedMesh.cxx:4-149) and the neutronic solver blocks
    (DiffusionSolver.cxx:4-85, SNSolver.cxx:8-124);
  - derived nuclear data (ConstantNuclearData.cxx:129-177).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

BC_NONE, BC_VACUUM, BC_REFLECTIVE, BC_ROBIN, BC_DIRICHLET, BC_ADIABATIC, BC_CONVECTION = range(7)
_BC_NAMES = {"vacuum": BC_VACUUM, "reflective": BC_REFLECTIVE, "robin": BC_ROBIN,
             "dirichlet": BC_DIRICHLET, "adiabatic": BC_ADIABATIC, "convection": BC_CONVECTION}


class PmpError(RuntimeError):
    pass


def clean(s: str) -> str:
    """input::clean (input.cxx:4-38)."""
    if not s:
        return s
    while True:
        i = s.find("\t")
        if i < 0:
            break
        s = s[:i] + " " + s[i + 2:]
    while True:
        i = s.find("  ")
        if i < 0:
            break
        s = s[:i] + s[i + 1:]
    if s == " ":
        return ""
    s = s.strip(" ")
    if s.startswith("#"):
        return ""
    return s


class Lines:
    """Sequential reader with input::get_next_line semantics (input.cxx:41-60)."""

    def __init__(self, path: str):
        with open(path, "r") as fh:
            self._lines = fh.read().split("\n")
        if self._lines and self._lines[-1] == "":
            self._lines.pop()
        self._k = 0

    def next(self) -> list[str]:
        while self._k < len(self._lines):
            s = clean(self._lines[self._k])
            self._k += 1
            if s:
                return s.split(" ")
        return []

    def floats(self, n: int) -> list[float]:
        out: list[float] = []
        while len(out) < n:
            line = self.next()
            if not line:
                raise PmpError("missing data")
            for t in line:
                if len(out) >= n:
                    raise PmpError("out-of-bounds data")
                out.append(float(t))
        return out

    def ints(self, n: int) -> list[int]:
        return [int(v) for v in self._ints(n)]

    def _ints(self, n):
        out = []
        while len(out) < n:
            line = self.next()
            if not line:
                raise PmpError("missing data")
            for t in line:
                if len(out) >= n:
                    raise PmpError("out-of-bounds data")
                out.append(int(t))
        return out


def _read_bc(line: list[str], i: int, rd: Lines):
    """input::read(BoundaryCondition&) (input.cxx:367-404); Robin/Dirichlet values
    given as a constant (time tables are read and their first value kept)."""
    kind = line[i]
    if kind not in _BC_NAMES:
        raise PmpError("wrong boundary-condition type")
    t = _BC_NAMES[kind]
    val = 0.0
    if t in (BC_ROBIN, BC_DIRICHLET):
        rest = line[i + 1:]
        if len(rest) == 1:
            val = float(rest[0])
        elif len(rest) == 2 and rest[1] == "{":
            n = int(rest[0])
            rd.floats(n)
            val = rd.floats(n)[0]
            rd.next()
        else:
            raise PmpError("wrong number of function arguments")
    return t, val


@dataclass
class NuclearData:
    G: int = -1
    sigma_total: np.ndarray | None = None
    nu_sigma_fission: np.ndarray | None = None
    kappa_sigma_fission: np.ndarray | None = None
    sigma_transport: np.ndarray | None = None
    sigma_scattering: np.ndarray | None = None  # [from, to]
    diffusion_coefficient: np.ndarray | None = None
    chi_prompt: np.ndarray | None = None
    chi_delayed: np.ndarray | None = None
    chi_effective: np.ndarray | None = None
    velocity: np.ndarray | None = None

    def check(self, beta_total: float):
        """ConstantNuclearData::check(beta) (ConstantNuclearData.cxx:129-177)."""
        G = self.G
        if self.nu_sigma_fission is not None or self.chi_prompt is not None:
            if self.nu_sigma_fission is None:
                raise PmpError("missing nu-fission cross sections")
            if self.chi_prompt is None:
                raise PmpError("missing fission spectrum")
        if self.nu_sigma_fission is None:
            self.nu_sigma_fission = np.zeros(G)
        if self.kappa_sigma_fission is None:
            f = 3.2e-11 / 2.4355
            self.kappa_sigma_fission = np.array([v * f for v in self.nu_sigma_fission])
        if self.chi_prompt is None:
            self.chi_prompt = np.zeros(G)
        if self.chi_delayed is None:
            self.chi_delayed = self.chi_prompt.copy()
        if beta_total > 0.0:
            self.chi_effective = np.array([(1.0 - beta_total) * p + beta_total * d
                                           for p, d in zip(self.chi_prompt, self.chi_delayed)])
        else:
            self.chi_effective = self.chi_prompt.copy()
        if self.sigma_transport is not None:
            if self.diffusion_coefficient is not None:
                raise PmpError("transport cross sections can only be defined if diffusion "
                               "coefficients are not")
            self.diffusion_coefficient = np.array([1.0 / (3.0 * s) for s in self.sigma_transport])


def _read_nuclear_data(rd: Lines) -> NuclearData:
    nd = NuclearData()
    keys_vec = {"sigma-total": "sigma_total", "nu-sigma-fission": "nu_sigma_fission",
                "kappa-sigma-fission": "kappa_sigma_fission", "sigma-transport": "sigma_transport",
                "diffusion-coefficient": "diffusion_coefficient",
                "fission-spectrum": "chi_prompt", "fission-spectrum-prompt": "chi_prompt",
                "fission-spectrum-delayed": "chi_delayed", "neutron-velocity": "velocity"}
    while True:
        line = rd.next()
        if not line or line[0] == "}":
            break
        k = line[0]
        if k == "energy-groups":
            nd.G = int(line[1])
        elif k in keys_vec:
            setattr(nd, keys_vec[k], np.array(rd.floats(nd.G), dtype=np.float64))
        elif k == "sigma-scattering":
            vals = []
            for _ in range(nd.G):
                row = rd.next()
                if len(row) != nd.G:
                    raise PmpError("wrong scattering cross sections")
                vals.append([float(t) for t in row])
            nd.sigma_scattering = np.array(vals, dtype=np.float64)
        else:
            raise PmpError(f"unrecognized keyword '{k}'")
    return nd


@dataclass
class Material:
    name: str
    nd: NuclearData | None = None
    beta_total: float = 0.0
    fuel: bool = False
    bc: bool = False
    split: bool = False
    # temperature-dependent data (nuclear-data-set, FeedbackNuclearData.cxx:4-53):
    # one data set per temperature; nd is the first
    temperatures: list = field(default_factory=list)
    sets: list = field(default_factory=list)


def _read_material_block(rd: Lines, mat: Material):
    """Material::read (Material.cxx:18-136)."""
    while True:
        line = rd.next()
        if not line or line[0] == "}":
            break
        k = line[0]
        if k == "nuclear-data":
            mat.nd = _read_nuclear_data(rd)
        elif k == "nuclear-data-set":
            while True:
                ln = rd.next()
                if not ln or ln[0] == "}":
                    break
                if ln[0] == "temperature":
                    mat.temperatures = rd.floats(int(ln[1]))
                elif ln[0] == "nuclear-data":
                    mat.sets.append(_read_nuclear_data(rd))
                else:
                    raise PmpError(f"unrecognized keyword '{ln[0]}'")
            if not mat.sets or len(mat.sets) != len(mat.temperatures):
                raise PmpError("one nuclear-data block per temperature is needed")
            mat.nd = mat.sets[0]
        elif k == "precursor-data":
            beta_total = 0.0
            ng = 0
            while True:
                pl = rd.next()
                if not pl or pl[0] == "}":
                    break
                if pl[0] == "precursor-groups":
                    ng = int(pl[1])
                elif pl[0] == "lambda":
                    rd.floats(ng)
                elif pl[0] == "beta":
                    for b in rd.floats(ng):
                        beta_total += b
                else:
                    raise PmpError(f"unrecognized keyword '{pl[0]}'")
            mat.beta_total = beta_total
        elif k == "thermal-properties":
            pass
        elif k == "fuel":
            mat.fuel = bool(int(line[1]))
        elif k == "bc":
            mat.bc = bool(int(line[1]))
        elif k == "split":
            mat.split = bool(int(line[1]))
        else:
            raise PmpError(f"unrecognized keyword '{k}'")
    for nd in (mat.sets if mat.sets else [mat.nd] if mat.nd is not None else []):
        nd.check(mat.beta_total)


@dataclass
class CartesianInput:
    nx: int = 1
    ny: int = 0
    nz: int = 0
    dx: list = field(default_factory=lambda: [1.0])
    dy: list = field(default_factory=lambda: [0.0])
    dz: list = field(default_factory=lambda: [0.0])
    boundaries: list = field(default_factory=list)
    bcs: dict = field(default_factory=dict)  # boundary index -> (type, value)
    materials: np.ndarray | None = None  # 0-based, -1 void
    num_dims: int = 0


@dataclass
class UnstructuredInput:
    xy: np.ndarray | None = None
    cptr: np.ndarray | None = None
    cpts: np.ndarray | None = None
    nz: int = 0
    dz: list = field(default_factory=lambda: [0.0])
    boundaries: list = field(default_factory=list)
    xy_bnames: list = field(default_factory=list)
    xy_bpoints: list = field(default_factory=list)
    default_b: int = -1
    bcs: dict = field(default_factory=dict)
    materials: np.ndarray | None = None
    num_dims: int = 0


def _axis(rd, line, store, name):
    n = int(line[1])
    if n > 0:
        vals = rd.floats(n)
    else:
        vals = rd.floats(1) * (-n)
        n = -n
    return n, vals


def read_cartesian(path: str) -> CartesianInput:
    """CartesianMesh::read (CartesianMesh.cxx:4-143)."""
    rd = Lines(path)
    m = CartesianInput()
    while True:
        line = rd.next()
        if not line:
            break
        k = line[0]
        if k == "dx":
            m.nx, m.dx = _axis(rd, line, m, "x")
            m.num_dims += 1
            m.boundaries += ["-x", "+x"]
        elif k == "dy":
            m.ny, m.dy = _axis(rd, line, m, "y")
            if m.ny > 1:
                m.num_dims += 1
            m.boundaries += ["-y", "+y"]
        elif k == "dz":
            m.nz, m.dz = _axis(rd, line, m, "z")
            if m.nz > 1:
                m.num_dims += 1
            m.boundaries += ["-z", "+z"]
        elif k == "bc":
            name = line[1]
            if name not in m.boundaries:
                raise PmpError("wrong boundary name")
            m.bcs[m.boundaries.index(name)] = _read_bc(line, 2, rd)
        elif k == "materials":
            n = m.nx * max(m.ny, 1) * max(m.nz, 1)
            if int(line[1]) != n:
                raise PmpError("wrong number of materials")
            m.materials = np.array(rd.ints(n), dtype=np.int32) - 1
        elif k == "nodal-indices":
            n = m.nx * max(m.ny, 1) * max(m.nz, 1)
            rd.ints(n)
        else:
            raise PmpError(f"unrecognized keyword '{k}'")
    return m


def read_unstructured(path: str) -> UnstructuredInput:
    """UnstructuredExtrudedMesh::read (UnstructuredExtrudedMesh.cxx:4-149)."""
    rd = Lines(path)
    m = UnstructuredInput()
    ncells = 0
    while True:
        line = rd.next()
        if not line:
            break
        k = line[0]
        if k == "points":
            n = int(line[1])
            pts = []
            for _ in range(n):
                row = rd.next()
                if len(row) != 2:
                    raise PmpError("wrong point data")
                pts.append([float(row[0]), float(row[1])])
            m.xy = np.array(pts, dtype=np.float64)
        elif k == "cells":
            ncells, nt = int(line[1]), int(line[2])
            ptr, flat = [0], []
            for _ in range(ncells):
                row = rd.next()
                if not row:
                    raise PmpError("missing data")
                flat += [int(t) for t in row]
                ptr.append(len(flat))
            if len(flat) > nt:
                raise PmpError("out-of-bounds data")
            m.cptr = np.array(ptr, dtype=np.int32)
            m.cpts = np.array(flat, dtype=np.int32)
            m.num_dims += 2
        elif k == "dz":
            m.nz, m.dz = _axis(rd, line, m, "z")
            if m.nz > 1:
                m.num_dims += 1
            m.boundaries += ["-z", "+z"]
        elif k == "boundary":
            name = line[1]
            m.boundaries.append(name)
            m.xy_bnames.append(name)
            n = int(line[2])
            if n > 0:
                m.xy_bpoints.append(rd.ints(n))
            else:
                m.default_b = len(m.xy_bpoints)
                m.xy_bpoints.append([])
        elif k == "bc":
            name = line[1]
            if name not in m.boundaries:
                raise PmpError("wrong boundary name")
            m.bcs[m.boundaries.index(name)] = _read_bc(line, 2, rd)
        elif k == "materials":
            n = ncells * max(m.nz, 1)
            if int(line[1]) != n:
                raise PmpError("wrong number of materials")
            m.materials = np.array(rd.ints(n), dtype=np.int32) - 1
        elif k == "nodal-indices":
            rd.ints(ncells * max(m.nz, 1))
        else:
            raise PmpError(f"unrecognized keyword '{k}'")
    return m


@dataclass
class SolverInput:
    kind: str
    G: int = -1
    power: float = 1.0
    bcs: dict = field(default_factory=dict)
    order: int = -1
    delta: float = 0.1
    ls: bool = False


def _read_solver_block(rd: Lines, s: SolverInput, boundaries: list):
    """DiffusionSolver::read / SNSolver::read (DiffusionSolver.cxx:4-85, SNSolver.cxx:8-124)."""
    while True:
        line = rd.next()
        if not line or line[0] == "}":
            break
        k = line[0]
        if k == "energy-groups":
            s.G = int(line[1])
        elif k == "bc":
            name = line[1]
            if name not in boundaries:
                raise PmpError("wrong boundary name")
            s.bcs[boundaries.index(name)] = _read_bc(line, 2, rd)
        elif k == "power":
            s.power = float(line[1])
        elif k == "convergence":
            pass
        elif k == "order" and s.kind == "sn":
            s.order = int(line[1])
        elif k == "mixed-face-interpolation" and s.kind == "sn":
            s.delta = float(line[1])
        elif k == "least-squares-boundary-interpolation" and s.kind == "sn":
            s.ls = bool(int(line[1]))
        else:
            raise PmpError(f"unrecognized keyword '{k}'")


@dataclass
class PartitionedInput:
    """Mesh data of one domain (PartitionedMesh.cxx:4-263)."""
    C: int = 0
    ghost: int = 0
    points: np.ndarray | None = None     # (np, 3)
    cpoints: list = field(default_factory=list)
    vol: np.ndarray | None = None
    cen: np.ndarray | None = None        # (C + ghost, 3)
    materials: np.ndarray | None = None  # 0-based, C + ghost
    nfaces: np.ndarray | None = None
    farea: list = field(default_factory=list)     # ragged per cell
    fcen: np.ndarray | None = None       # (F, 3)
    fnorm: np.ndarray | None = None      # (F, 3)
    fnb: list = field(default_factory=list)       # ragged per cell
    boundaries: list = field(default_factory=list)
    bcs: dict = field(default_factory=dict)
    num_dims: int = 1


def read_partitioned(path: str) -> PartitionedInput:
    """PartitionedMesh::read (PartitionedMesh.cxx:4-263)."""
    rd = Lines(path)
    m = PartitionedInput()
    F = 0

    def ragged(n, conv):
        rows = []
        for _ in range(n):
            line = rd.next()
            if not line:
                raise PmpError("missing data")
            rows.append([conv(v) for v in line])
        return rows

    while True:
        line = rd.next()
        if not line:
            break
        k = line[0]
        if k == "points":
            n = int(line[1])
            m.points = np.array(ragged(n, float))
        elif k == "cells":
            m.C, m.ghost = int(line[1]), int(line[2])
            if m.ghost < 1:
                raise PmpError("wrong number of ghost cells")
        elif k == "cell-points":
            m.cpoints = ragged(m.C, int)
        elif k == "cell-volumes":
            m.vol = np.array(rd.floats(m.C))
        elif k == "cell-centroids":
            m.cen = np.array(ragged(m.C + m.ghost, float))
            c = m.cen[:m.C]
            m.num_dims = 1 + int(np.any(c[1:, 1] != c[0, 1])) + int(np.any(c[1:, 2] != c[0, 2]))
        elif k == "cell-materials":
            m.materials = np.array(rd.ints(m.C + m.ghost)) - 1
        elif k in ("cell-nodal-indices",):
            rd.ints(m.C)
        elif k == "cell-global-indices":
            rd.ints(m.C + m.ghost)
        elif k == "faces":
            m.nfaces = np.array(rd.ints(m.C), np.int32)
            F = int(m.nfaces.sum())
        elif k == "face-areas":
            if int(line[2]) != F:
                raise PmpError("wrong number of face areas")
            m.farea = ragged(m.C, float)
        elif k in ("face-centroids", "face-normals"):
            if int(line[2]) != 3 * F:
                raise PmpError("wrong number of face coordinates")
            a = np.array(ragged(F, float))
            if k == "face-centroids":
                m.fcen = a
            else:
                m.fnorm = a
        elif k == "face-neighbors":
            if int(line[2]) != F:
                raise PmpError("wrong number of face neighbors")
            m.fnb = ragged(m.C, int)
        elif k == "boundary":
            m.boundaries.append(line[1])
        elif k == "bc":
            m.bcs[m.boundaries.index(line[1])] = _read_bc(line, 2, rd)
        else:
            raise PmpError(f"unrecognized keyword '{k}'")
    return m


@dataclass
class Case:
    mesh_type: str
    mesh: object
    materials: list
    solver: SolverInput
    dt: list


def read_input(path: str) -> Case:
    """Parser::read for the steady neutronics inputs (Parser.cxx:4-260).
    File names inside are resolved against the working directory first, as
    the reference does, and then against the input file's directory."""
    base = os.path.dirname(os.path.abspath(path))

    def resolve(p):
        if os.path.exists(p):
            return p
        return os.path.join(base, p)

    rd = Lines(path)
    mesh_type, mesh, materials, solver, dt = None, None, [], None, []
    while True:
        line = rd.next()
        if not line:
            break
        k = line[0]
        if k == "mesh":
            mesh_type = line[1]
            if mesh_type == "cartesian":
                mesh = read_cartesian(resolve(line[2]))
            elif mesh_type == "unstructured":
                mesh = read_unstructured(resolve(line[2]))
            elif mesh_type == "partitioned":
                mesh = read_partitioned(resolve(line[2]))
            else:
                raise PmpError("wrong mesh type")
        elif k == "material":
            mat = Material(line[1])
            if line[2] == "{":
                _read_material_block(rd, mat)
            else:
                _read_material_block(Lines(resolve(line[2])), mat)
            if mat.bc:
                # a boundary named after the material (Parser.cxx:93)
                mesh.boundaries.append(mat.name)
            materials.append(mat)
        elif k == "solver":
            solver = SolverInput(line[1])
            if line[2] == "{":
                _read_solver_block(rd, solver, mesh.boundaries)
            else:
                _read_solver_block(Lines(resolve(line[2])), solver, mesh.boundaries)
        elif k == "dt":
            n = int(line[1])
            dt = rd.floats(n) if n > 0 else rd.floats(1) * (-n)
        elif k in ("vtk", "petsc", "mesh-nodal"):
            pass
        else:
            raise PmpError(f"unrecognized keyword '{k}'")
    return Case(mesh_type, mesh, materials, solver, dt)
