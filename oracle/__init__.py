"""FP64 CPU oracle for pampa's neutronics hot path.

TEST INFRASTRUCTURE ONLY.  This package is the parity checker and the CPU
baseline: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import it.  The product (pampa_amd) never imports, links or calls it.

Layers:
  - oracle.pmp       restated `.pmp` reader (input format of the reference);
  - oracle.c         C restatement of mesh building, operator assembly with
                     PETSc-AIJ row semantics, SN quadrature support, upwind
                     sweeps (compiled by oracle/Makefile into _build/liboracle.so);
  - this module      glue: geometry/XS tables, assembly wrappers, and the
                     reference's eigen-solver restated -- shift-and-invert with
                     sigma = 0 and an exact LU inner solve around an implicitly
                     restarted Arnoldi method (SLEPc Krylov-Schur + STSINVERT +
                     PCLU, petsc.cxx:175-204, DiffusionSolver.cxx:323-350).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
from dataclasses import dataclass

import numpy as np

from . import pmp
from .pmp import (BC_NONE, BC_REFLECTIVE, BC_ROBIN, BC_VACUUM)  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    """Compile oracle.c with gcc (the recipe is oracle/Makefile)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ct.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_I = ct.c_int
_D = ct.c_double
_P = ct.c_void_p


def _declare(L):
    L.orc_cart_count.argtypes = [_I, _I, _I, _P, _P]
    L.orc_cart_build.argtypes = [_I, _I, _I] + [_P] * 14
    L.orc_unst_build.argtypes = ([_I, _P, _I, _P, _P, _I, _P, _I, _P, _P, _P, _I, _P, _I]
                                 + [_P] * 8)
    L.orc_diff_assemble.argtypes = [_I, _I] + [_P] * 8 + [_I] + [_P] * 13
    L.orc_diff_assemble_sets.argtypes = [_I, _I] + [_P] * 8 + [_I] + [_P] * 17
    L.orc_diff_assemble_transient.argtypes = ([_I, _I] + [_P] * 8 + [_I] + [_P] * 7 + [_D, _P, _P, _D]
                                              + [_P] * 3)
    L.orc_diff_rhs.argtypes = [_I, _I, _P, _P, _P, _P, _D, _P, _P, _P]
    L.orc_diff_rhs.restype = None
    L.orc_sn_gauss.argtypes = [_I, _I, _P, _P, _P, _P, _D, _P]
    L.orc_sn_ls.argtypes = [_I, _I, _P, _P, _P, _P, _I, _P, _P]
    L.orc_sn_assemble.argtypes = ([_I, _I] + [_P] * 8 + [_I, _I] + [_P] * 4 + [_I]
                                  + [_P] * 13)
    L.orc_sn_order.argtypes = [_I, _I, _P, _P, _P, _I, _P, _P, _P, _P]
    L.orc_sn_sweep.argtypes = [_I, _I] + [_P] * 6 + [_I, _I] + [_P] * 8
    L.orc_sn_source.argtypes = [_I, _I, _P, _P, _P, _P, _P, _D, _P, _P]
    L.orc_sn_eigen.argtypes = ([_I, _I] + [_P] * 6 + [_I, _I] + [_P] * 8 + [_I, _I, _D, _D]
                               + [_P] * 3)


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------

@dataclass
class Mesh:
    """Single-domain mesh in the oracle's padded face layout."""
    C: int
    fmax: int
    num_dims: int
    vol: np.ndarray
    cen: np.ndarray      # (C, 3)
    cmat: np.ndarray     # (C,)
    nfaces: np.ndarray   # (C,)
    farea: np.ndarray    # (C, fmax)
    fcen: np.ndarray     # (C, fmax, 3)
    fnorm: np.ndarray    # (C, fmax, 3)
    fnb: np.ndarray      # (C, fmax)
    boundaries: list
    bcs: dict            # boundary index -> (type, value)


def _alloc(C, fmax):
    return dict(vol=np.zeros(C), cen=np.zeros((C, 3)), cmat=np.zeros(C, np.int32),
                nfaces=np.zeros(C, np.int32), farea=np.zeros((C, fmax)),
                fcen=np.zeros((C, fmax, 3)), fnorm=np.zeros((C, fmax, 3)),
                fnb=np.zeros((C, fmax), np.int32))


def build_cartesian(m: pmp.CartesianInput) -> Mesh:
    """CartesianMesh::build (CartesianMesh.cxx:146-414) via orc_cart_build."""
    L = lib()
    mats = np.ascontiguousarray(m.materials, dtype=np.int32)
    C = ct.c_int(0)
    if L.orc_cart_count(m.nx, m.ny, m.nz, _p(mats), ct.byref(C)):
        raise pmp.PmpError("wrong material definition")
    C = C.value
    fmax = 6 if m.nz > 0 else (4 if m.ny > 0 else 2)
    a = _alloc(C, fmax)
    bidx = np.array([m.boundaries.index(n) if n in m.boundaries else -1
                     for n in ("-x", "+x", "-y", "+y", "-z", "+z")], dtype=np.int32)
    dx = np.array(m.dx, dtype=np.float64)
    dy = np.array(m.dy, dtype=np.float64)
    dz = np.array(m.dz, dtype=np.float64)
    fm = ct.c_int(0)
    L.orc_cart_build(m.nx, m.ny, m.nz, _p(dx), _p(dy), _p(dz), _p(mats), _p(bidx), ct.byref(fm),
                     _p(a["vol"]), _p(a["cen"]), _p(a["cmat"]), _p(a["nfaces"]), _p(a["farea"]),
                     _p(a["fcen"]), _p(a["fnorm"]), _p(a["fnb"]))
    return Mesh(C=C, fmax=fmax, num_dims=m.num_dims, boundaries=list(m.boundaries),
                bcs=dict(m.bcs), **a)


def build_unstructured(m: pmp.UnstructuredInput) -> Mesh:
    """UnstructuredExtrudedMesh::build (UnstructuredExtrudedMesh.cxx:153-364)."""
    L = lib()
    ncxy = len(m.cptr) - 1
    nz = m.nz
    C = ncxy * max(nz, 1)
    nmax = int(np.max(np.diff(m.cptr)))
    fmax = nmax + 2 if nz > 0 else nmax
    a = _alloc(C, fmax)
    bptr = np.zeros(len(m.xy_bpoints) + 1, np.int32)
    for b, pts in enumerate(m.xy_bpoints):
        bptr[b + 1] = bptr[b] + len(pts)
    bpts = np.array([p for pts in m.xy_bpoints for p in pts] + [0], dtype=np.int32)
    bglob = np.array([m.boundaries.index(n) for n in m.xy_bnames] + [0], dtype=np.int32)
    bz = np.array([m.boundaries.index(n) if n in m.boundaries else -1 for n in ("-z", "+z")],
                  dtype=np.int32)
    xy = np.ascontiguousarray(m.xy, dtype=np.float64)
    dz = np.array(m.dz, dtype=np.float64)
    err = L.orc_unst_build(len(xy), _p(xy), ncxy, _p(m.cptr), _p(m.cpts), nz, _p(dz),
                           len(m.xy_bpoints), _p(bptr), _p(bpts), _p(bglob), m.default_b, _p(bz),
                           fmax, _p(a["vol"]), _p(a["cen"]), _p(a["cmat"]), _p(a["nfaces"]),
                           _p(a["farea"]), _p(a["fcen"]), _p(a["fnorm"]), _p(a["fnb"]))
    if err:
        raise pmp.PmpError("wrong mesh connectivity")
    a["cmat"] = np.ascontiguousarray(m.materials, dtype=np.int32)
    return Mesh(C=C, fmax=fmax, num_dims=m.num_dims, boundaries=list(m.boundaries),
                bcs=dict(m.bcs), **a)


def build_partitioned(m: pmp.PartitionedInput) -> Mesh:
    """The mesh data as read (PartitionedMesh.hxx: nothing to build), one
    domain: no face may lead to a ghost cell."""
    C = m.C
    fmax = int(m.nfaces.max())
    a = _alloc(C, fmax)
    a["vol"][:] = m.vol
    a["cen"][:] = m.cen[:C]
    a["cmat"][:] = m.materials[:C]
    a["nfaces"][:] = m.nfaces
    f = 0
    for i in range(C):
        n = int(m.nfaces[i])
        if len(m.farea[i]) != n or len(m.fnb[i]) != n:
            raise pmp.PmpError("wrong number of face values for a cell")
        a["farea"][i, :n] = m.farea[i]
        a["fnb"][i, :n] = m.fnb[i]
        a["fcen"][i, :n] = m.fcen[f:f + n]
        a["fnorm"][i, :n] = m.fnorm[f:f + n]
        f += n
    if np.any(a["fnb"] >= C):
        raise pmp.PmpError("a face leads to a ghost cell")
    mesh = Mesh(C=C, fmax=fmax, num_dims=m.num_dims, boundaries=list(m.boundaries), bcs=dict(m.bcs), **a)
    mesh.partitioned = True
    return mesh


def split_materials(mesh: Mesh, materials: list):
    """Mesh::splitMaterials (Mesh.cxx:9-60): every connected region of a split
    material becomes a material of its own, <name>-<k>, appended to the list;
    a boundary-condition material's regions name boundaries of their own."""
    import copy
    subs = {}
    for i in range(mesh.C):
        im = int(mesh.cmat[i])
        if not materials[im].split:
            continue
        im2 = len(materials)
        mesh.cmat[i] = im2
        stack = [i]
        while stack:
            c = stack.pop()
            for f in range(mesh.nfaces[c]):
                c2 = int(mesh.fnb[c, f])
                if c2 >= 0 and mesh.cmat[c2] == im:
                    mesh.cmat[c2] = im2
                    stack.append(c2)
        sub = copy.copy(materials[im])
        sub.name = f"{materials[im].name}-{subs.get(im, 0)}"
        subs[im] = subs.get(im, 0) + 1
        sub.split = False
        if sub.bc:
            mesh.boundaries.append(sub.name)
        materials.append(sub)


def remove_bc_materials(mesh: Mesh, materials: list) -> Mesh:
    """Mesh::removeBCMaterials (Mesh.cxx:63-204): cells of boundary-condition
    materials leave the mesh and their faces become faces of the boundary named
    after the material."""
    if getattr(mesh, "partitioned", False):
        raise pmp.PmpError("unable to remove materials from a partitioned mesh")
    keep = np.array([not materials[int(m)].bc for m in mesh.cmat])
    new = np.full(mesh.C, -1, np.int64)
    new[keep] = np.arange(int(keep.sum()))
    fnb = mesh.fnb.copy()
    for i in range(mesh.C):
        for f in range(mesh.nfaces[i]):
            c2 = int(fnb[i, f])
            if c2 >= 0:
                m2 = materials[int(mesh.cmat[c2])]
                fnb[i, f] = new[c2] if keep[c2] else -mesh.boundaries.index(m2.name) - 1
    sel = lambda a: np.ascontiguousarray(a[keep])  # noqa: E731
    nf = sel(mesh.nfaces)
    return Mesh(C=int(keep.sum()), fmax=int(nf.max()), num_dims=mesh.num_dims, vol=sel(mesh.vol),
                cen=sel(mesh.cen), cmat=sel(mesh.cmat), nfaces=nf, farea=sel(mesh.farea), fcen=sel(mesh.fcen),
                fnorm=sel(mesh.fnorm), fnb=np.ascontiguousarray(fnb[keep], np.int32),
                boundaries=list(mesh.boundaries), bcs=dict(mesh.bcs))


def build_mesh(case: pmp.Case) -> Mesh:
    """The case's mesh, with the edits the parser makes before the first solver
    (Parser.cxx:107-130): split materials, then boundary-condition materials
    removed.  Split materials are appended to case.materials (once)."""
    if case.mesh_type == "cartesian":
        mesh = build_cartesian(case.mesh)
    elif case.mesh_type == "unstructured":
        mesh = build_unstructured(case.mesh)
    elif case.mesh_type == "partitioned":
        mesh = build_partitioned(case.mesh)
    else:
        raise pmp.PmpError("wrong mesh type")
    if any(m.split for m in case.materials):
        if getattr(case, "_split_done", None) is None:
            case._split_done = len(case.materials)
        else:
            del case.materials[case._split_done:]
        split_materials(mesh, case.materials)
    if any(m.bc for m in case.materials):
        mesh = remove_bc_materials(mesh, case.materials)
    return mesh


# --------------------------------------------------------------------------
# cross-section tables and boundary table
# --------------------------------------------------------------------------

@dataclass
class XS:
    G: int
    sigt: np.ndarray   # (nm, G)
    sigs: np.ndarray   # (nm, G, G) [from, to]
    nusf: np.ndarray
    kapf: np.ndarray
    chi: np.ndarray    # effective spectrum
    D: np.ndarray
    beta: np.ndarray   # (nm,)
    # temperature sets (None: one set per material, tables indexed by material)
    set0: np.ndarray | None = None
    nset: np.ndarray | None = None
    tref: np.ndarray | None = None


def xs_tables(materials, G: int, diffusion: bool) -> XS:
    """Per-material tables; with temperature-dependent materials the tables
    are per set (material m's sets from set0[m], nset[m] of them, at
    temperatures tref)."""
    nm = len(materials)
    sets = [mat.sets if mat.sets else [mat.nd] for mat in materials]
    ns = sum(len(x) for x in sets)
    t = {k: np.zeros((ns, G)) for k in ("sigt", "nusf", "kapf", "chi", "D")}
    sigs = np.zeros((ns, G, G))
    beta = np.zeros(nm)
    set0, nset, tref = np.zeros(nm, np.int32), np.zeros(nm, np.int32), np.zeros(ns)
    k = 0
    for i, mat in enumerate(materials):
        set0[i], nset[i] = k, len(sets[i])
        beta[i] = mat.beta_total
        for j, nd in enumerate(sets[i]):
            if nd is None:
                raise pmp.PmpError("missing nuclear data")
            if nd.G != G:
                raise pmp.PmpError("wrong number of energy groups")
            t["sigt"][k] = nd.sigma_total
            t["nusf"][k] = nd.nu_sigma_fission
            t["kapf"][k] = nd.kappa_sigma_fission
            t["chi"][k] = nd.chi_effective
            if diffusion:
                if nd.diffusion_coefficient is None:
                    raise pmp.PmpError("missing diffusion coefficients")
                t["D"][k] = nd.diffusion_coefficient
            sigs[k] = nd.sigma_scattering
            tref[k] = mat.temperatures[j] if mat.sets else 0.0
            k += 1
    if ns == nm:
        return XS(G=G, sigs=sigs, beta=beta, **t)
    return XS(G=G, sigs=sigs, beta=beta, set0=set0, nset=nset, tref=tref, **t)


def cell_values(xs: XS, mesh, name: str, T=None) -> np.ndarray:
    """(C, G) values of table `name` for every cell, at the cells'
    temperatures T (0 when None) for temperature-dependent materials
    (the interpolation of orc_sets in oracle.c)."""
    tab = getattr(xs, name)
    if xs.nset is None:
        return tab[mesh.cmat]
    T = np.zeros(mesh.C) if T is None else np.asarray(T, np.float64)
    out = np.zeros((mesh.C, xs.G))
    for i in range(mesh.C):
        m = int(mesh.cmat[i])
        s0, n = int(xs.set0[m]), int(xs.nset[m])
        Tr = xs.tref[s0:s0 + n]
        if n == 1 or T[i] <= Tr[0]:
            out[i] = tab[s0]
        elif T[i] > Tr[-1]:
            out[i] = tab[s0 + n - 1]
        else:
            k = int(np.searchsorted(Tr, T[i], side="left"))
            f = (T[i] - Tr[k - 1]) / (Tr[k] - Tr[k - 1])
            out[i] = (1.0 - f) * tab[s0 + k - 1] + f * tab[s0 + k]
    return out


def bc_table(mesh: Mesh, solver_bcs: dict):
    """Boundary table indexed by bc id = boundary index + 1.  A solver block with
    any `bc` line replaces the mesh's table (DiffusionSolver.cxx:91, :27-44)."""
    src = solver_bcs if solver_bcs else mesh.bcs
    n = len(mesh.boundaries) + 1
    bct = np.zeros(n, np.int32)
    bcv = np.zeros(n)
    for b, (t, v) in src.items():
        bct[b + 1] = t
        bcv[b + 1] = v
    return bct, bcv


# --------------------------------------------------------------------------
# diffusion
# --------------------------------------------------------------------------

@dataclass
class CSR:
    n: int
    ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray

    def scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.ptr), shape=(self.n, self.n))


def assemble_diffusion(mesh: Mesh, xs: XS, bct, bcv, T=None):
    """R and F of DiffusionSolver::buildMatrices (DiffusionSolver.cxx:88-320), n = 0;
    temperature-dependent materials evaluated at the cell temperatures T
    (0 when None)."""
    L = lib()
    G = xs.G
    N = mesh.C * G
    r_ptr = np.zeros(N + 1, np.int32)
    f_ptr = np.zeros(N + 1, np.int32)
    args = [mesh.C, mesh.fmax, _p(mesh.nfaces), _p(mesh.vol), _p(mesh.cen), _p(mesh.cmat),
            _p(mesh.farea), _p(mesh.fcen), _p(mesh.fnorm), _p(mesh.fnb), G, _p(xs.sigt),
            _p(xs.sigs), _p(xs.nusf), _p(xs.chi), _p(xs.D), _p(bct), _p(bcv)]
    fn = L.orc_diff_assemble
    if xs.nset is not None:
        Tc = np.zeros(mesh.C) if T is None else np.ascontiguousarray(T, np.float64)
        args += [_p(xs.set0), _p(xs.nset), _p(xs.tref), _p(Tc)]
        fn = L.orc_diff_assemble_sets
    if fn(*args, _p(r_ptr), None, None, _p(f_ptr), None, None):
        raise pmp.PmpError("boundary condition not implemented")
    r_col = np.zeros(r_ptr[-1], np.int32)
    r_val = np.zeros(r_ptr[-1])
    f_col = np.zeros(f_ptr[-1], np.int32)
    f_val = np.zeros(f_ptr[-1])
    fn(*args, _p(r_ptr), _p(r_col), _p(r_val), _p(f_ptr), _p(f_col), _p(f_val))
    return CSR(N, r_ptr, r_col, r_val), CSR(N, f_ptr, f_col, f_val)


@dataclass
class TransientXS:
    """Per-material time-step data: neutron velocity, prompt and delayed
    spectra [nm][G], total delayed fraction [nm]."""
    vel: np.ndarray
    chip: np.ndarray
    chid: np.ndarray
    beta: np.ndarray


def transient_tables(materials, G: int) -> TransientXS:
    nm = len(materials)
    t = TransientXS(np.zeros((nm, G)), np.zeros((nm, G)), np.zeros((nm, G)), np.zeros(nm))
    for i, mat in enumerate(materials):
        nd = mat.nd
        if nd.velocity is None:
            raise pmp.PmpError("missing neutron velocities")
        t.vel[i] = nd.velocity
        t.chip[i] = nd.chi_prompt
        t.chid[i] = nd.chi_delayed
        t.beta[i] = mat.beta_total
    return t


def assemble_diffusion_transient(mesh: Mesh, xs: XS, bct, bcv, tr: TransientXS, dt: float, keff: float):
    """R of a time step n > 0 of DiffusionSolver::buildMatrices
    (DiffusionSolver.cxx:134-148,156-175): the steady rows plus V/(v dt) on the
    diagonal and the prompt fission -(1 - beta) chip(g) nusf(g2) V / keff on the
    cell block."""
    L = lib()
    G = xs.G
    N = mesh.C * G
    r_ptr = np.zeros(N + 1, np.int32)
    args = [mesh.C, mesh.fmax, _p(mesh.nfaces), _p(mesh.vol), _p(mesh.cen), _p(mesh.cmat),
            _p(mesh.farea), _p(mesh.fcen), _p(mesh.fnorm), _p(mesh.fnb), G, _p(xs.sigt),
            _p(xs.sigs), _p(xs.nusf), _p(xs.D), _p(bct), _p(bcv), _p(tr.vel), float(dt), _p(tr.chip),
            _p(tr.beta), float(keff)]
    if L.orc_diff_assemble_transient(*args, _p(r_ptr), None, None):
        raise pmp.PmpError("boundary condition not implemented")
    r_col = np.zeros(r_ptr[-1], np.int32)
    r_val = np.zeros(r_ptr[-1])
    L.orc_diff_assemble_transient(*args, _p(r_ptr), _p(r_col), _p(r_val))
    return CSR(N, r_ptr, r_col, r_val)


def diffusion_rhs(mesh: Mesh, tr: TransientXS, dt: float, S: np.ndarray, phi0: np.ndarray):
    """b of a time step (DiffusionSolver.cxx:134-143): chid S + V/(v dt) phi0."""
    G = tr.vel.shape[1]
    b = np.zeros(mesh.C * G)
    S = np.ascontiguousarray(S, np.float64)
    phi0 = np.ascontiguousarray(phi0, np.float64)
    lib().orc_diff_rhs(mesh.C, G, _p(mesh.cmat), _p(mesh.vol), _p(tr.chid), _p(tr.vel), float(dt), _p(S),
                       _p(phi0), _p(b))
    return b


def linear_solve(A: CSR, b: np.ndarray) -> np.ndarray:
    """The time step's linear system R phi = b (petsc::solve(KSP),
    petsc.cxx:382-425, with the KSP of petsc.cxx:147-172) solved exactly by a
    sparse LU: the solution the reference's iterative solve converges to."""
    import scipy.sparse.linalg as sla
    lu = sla.splu(A.scipy().tocsc(), permc_spec="MMD_AT_PLUS_A")
    return lu.solve(np.asarray(b, np.float64))


def eigen_sinvert(R: CSR, F: CSR, tol=1e-13, v0=None, permc="MMD_AT_PLUS_A"):
    """Largest 1/lambda of R x = lambda F x by shift-and-invert at sigma = 0:
    an exact sparse LU of R and an implicitly restarted Arnoldi iteration on
    R^-1 F, the algorithm the reference configures in SLEPc
    (petsc.cxx:193-197).  Returns (keff, x) with x real and of positive sum.
    `permc` is SuperLU's column ordering (COLAMD keeps the fill of S_N
    operators with downwind couplings, delta < 1, far smaller)."""
    import scipy.sparse.linalg as sla
    Rs = R.scipy().tocsc()
    Fs = F.scipy()
    lu = sla.splu(Rs, permc_spec=permc)
    op = sla.LinearOperator(Rs.shape, matvec=lambda v: lu.solve(Fs @ v), dtype=np.float64)
    if v0 is None:
        v0 = np.ones(R.n)
    vals, vecs = sla.eigs(op, k=1, which="LM", tol=tol, v0=v0, ncv=min(R.n - 1, 20))
    k = float(vals[0].real)
    x = vecs[:, 0].real.copy()
    if x.sum() < 0:
        x = -x
    return k, x


def normalize_flux(mesh: Mesh, xs: XS, phi: np.ndarray, power: float, T=None):
    """NeutronicSolver::normalizeScalarFlux + calculatePowerAndProductionRate
    (NeutronicSolver.cxx:46-116): phi scaled so sum(phi kappaSf V) = power;
    returns (phi, q, P, total_power)."""
    G = xs.G
    ph = phi.reshape(mesh.C, G)
    kap = cell_values(xs, mesh, "kapf", T)
    p0 = float(np.sum(ph * kap * mesh.vol[:, None]))
    ph = ph * (power / p0)
    if np.any(ph < 0.0):
        raise pmp.PmpError("negative values in the scalar-flux solution")
    q = np.sum(ph * kap * mesh.vol[:, None], axis=1)
    return ph.reshape(-1), q, float(q.sum())


@dataclass
class Result:
    keff: float
    phi: np.ndarray
    q: np.ndarray
    power: float
    R: CSR | None = None
    F: CSR | None = None


def solve_diffusion_case(case: pmp.Case, return_ops=False, T=None) -> Result:
    mesh = build_mesh(case)
    s = case.solver
    xs = xs_tables(case.materials, s.G, diffusion=True)
    bct, bcv = bc_table(mesh, s.bcs)
    R, F = assemble_diffusion(mesh, xs, bct, bcv, T)
    k, x = eigen_sinvert(R, F)
    phi, q, p = normalize_flux(mesh, xs, x, s.power, T)
    return Result(k, phi, q, p, R if return_ops else None, F if return_ops else None)


def solve_input(path: str, **kw) -> Result:
    case = pmp.read_input(path)
    if case.solver.kind == "diffusion":
        return solve_diffusion_case(case, **kw)
    if case.solver.kind == "sn":
        return solve_sn_case(case, **kw)
    raise pmp.PmpError("wrong solver type")


# --------------------------------------------------------------------------
# SN
# --------------------------------------------------------------------------

def quadrature(order: int):
    """Level-symmetric first-octant tables with the reference's 7-digit constants
    and octant layout l = o*(M/8) + m with sign bits x:1, y:2, z:4, plus the
    reflection table (AngularQuadratureSet.cxx:4-209)."""
    s3 = 1.0 / np.sqrt(3.0)
    if order == 2:
        mu = [[s3, s3, s3]]
        w = [1.0 / 8.0]
    elif order == 4:
        a, b = 0.3500212, 0.8688903
        mu = [[a, a, b], [a, b, a], [b, a, a]]
        w = [(1.0 / 3.0) / 8.0] * 3
    elif order == 6:
        a, b, c = 0.2666355, 0.6815076, 0.9261808
        w0, w1 = 0.1761263, 0.1572071
        mu = [[a, a, c], [a, c, a], [c, a, a], [a, b, b], [b, a, b], [b, b, a]]
        w = [w0 / 8.0] * 3 + [w1 / 8.0] * 3
    elif order == 8:
        a, b, c, d = 0.2182179, 0.5773503, 0.7867958, 0.9511897
        w0, w1, w2 = 0.1209877, 0.0907407, 0.0925926
        mu = [[a, a, d], [a, d, a], [d, a, a], [a, b, c], [a, c, b], [b, a, c], [b, c, a],
              [c, a, b], [c, b, a], [b, b, b]]
        w = [w0 / 8.0] * 3 + [w1 / 8.0] * 6 + [w2 / 8.0]
    else:
        raise pmp.PmpError("SN order not implemented")
    M = order * (order + 2)
    q = M // 8
    dirs = np.zeros((M, 3))
    wts = np.zeros(M)
    for o in range(8):
        for m in range(q):
            l = o * q + m
            dirs[l, 0] = -mu[m][0] if (o & 1) else mu[m][0]
            dirs[l, 1] = -mu[m][1] if (o & 2) else mu[m][1]
            dirs[l, 2] = -mu[m][2] if (o & 4) else mu[m][2]
            wts[l] = w[m]
    refl = np.full((M, 3), -1, np.int32)
    for ax in range(3):
        e = np.zeros(3)
        e[ax] = 1.0
        for m in range(M):
            d = dirs[m] + (-2.0 * float(dirs[m] @ e)) * e
            hits = [m2 for m2 in range(M) if float(dirs[m2] @ d) > 1.0 - 1.0e-6]
            if len(hits) != 1:
                raise pmp.PmpError("reflected direction not found")
            refl[m, ax] = hits[0]
    return dirs, wts, refl


def boundary_cells(mesh: Mesh):
    """SNSolver::getBoundaryCells (SNSolver.cxx:127-156)."""
    ic2bc = np.full(mesh.C, -1, np.int32)
    nb = 0
    for i in range(mesh.C):
        if np.any(mesh.fnb[i, :mesh.nfaces[i]] < 0):
            ic2bc[i] = nb
            nb += 1
    return ic2bc, nb


def sn_setup(mesh: Mesh, s: pmp.SolverInput):
    dirs, wts, refl = quadrature(s.order)
    L = lib()
    coef = np.zeros((mesh.C, mesh.fmax, 4))
    L.orc_sn_gauss(mesh.C, mesh.fmax, _p(mesh.nfaces), _p(mesh.cen), _p(mesh.fcen), _p(mesh.fnb),
                   s.delta, _p(coef))
    ic2bc, nb = boundary_cells(mesh)
    lsc = np.zeros((max(nb, 1), mesh.fmax, 3))
    if s.ls:
        L.orc_sn_ls(mesh.C, mesh.fmax, _p(mesh.nfaces), _p(mesh.cen), _p(mesh.fcen), _p(mesh.fnb),
                    mesh.num_dims, _p(ic2bc), _p(lsc))
    return dirs, wts, refl, coef, ic2bc, lsc


def assemble_sn(mesh: Mesh, xs: XS, bct, s: pmp.SolverInput):
    """R and F of SNSolver::buildMatrices in reference form (SNSolver.cxx:344-631)."""
    L = lib()
    dirs, wts, refl, coef, ic2bc, lsc = sn_setup(mesh, s)
    G, M = xs.G, len(wts)
    N = mesh.C * G * M
    r_ptr = np.zeros(N + 1, np.int32)
    f_ptr = np.zeros(N + 1, np.int32)
    args = [mesh.C, mesh.fmax, _p(mesh.nfaces), _p(mesh.vol), _p(mesh.cen), _p(mesh.cmat),
            _p(mesh.farea), _p(mesh.fcen), _p(mesh.fnorm), _p(mesh.fnb), G, M, _p(dirs), _p(wts),
            _p(refl), _p(coef), int(s.ls), _p(ic2bc), _p(lsc), _p(xs.sigt), _p(xs.sigs),
            _p(xs.nusf), _p(xs.chi), _p(bct)]
    if L.orc_sn_assemble(*args, _p(r_ptr), None, None, _p(f_ptr), None, None):
        raise pmp.PmpError("boundary condition not implemented")
    r_col = np.zeros(r_ptr[-1], np.int32)
    r_val = np.zeros(r_ptr[-1])
    f_col = np.zeros(f_ptr[-1], np.int32)
    f_val = np.zeros(f_ptr[-1])
    L.orc_sn_assemble(*args, _p(r_ptr), _p(r_col), _p(r_val), _p(f_ptr), _p(f_col), _p(f_val))
    return CSR(N, r_ptr, r_col, r_val), CSR(N, f_ptr, f_col, f_val), (dirs, wts)


def sn_normalize(mesh: Mesh, xs: XS, psi: np.ndarray, wts: np.ndarray, power: float):
    """calculateScalarFlux, normalizeScalarFlux, normalizeAngularFlux
    (SNSolver.cxx:272-341, NeutronicSolver.cxx:46-78): phi = 4 pi sum_m w psi,
    phi scaled to the power; psi scaled with sum w psi kappaSf V (no 4 pi)."""
    G, M = xs.G, len(wts)
    ps = psi.reshape(mesh.C, G, M)
    phi = 4.0 * np.pi * np.einsum("igm,m->ig", ps, wts)
    kap = xs.kapf[mesh.cmat]
    p0 = float(np.sum(phi * kap * mesh.vol[:, None]))
    phi = phi * (power / p0)
    if np.any(phi < 0):
        raise pmp.PmpError("negative values in the scalar-flux solution")
    pa = float(np.sum(np.einsum("igm,m->ig", ps, wts) * kap * mesh.vol[:, None]))
    ps = ps * (power / pa)
    if np.any(ps < 0):
        raise pmp.PmpError("negative values in the angular-flux solution")
    q = np.sum(phi * kap * mesh.vol[:, None], axis=1)
    return phi.reshape(-1), ps.reshape(-1), q, float(q.sum())


def solve_sn_case(case: pmp.Case, return_ops=False) -> Result:
    """Reference-form SN eigen solve (assembled operator + shift-invert)."""
    mesh = build_mesh(case)
    s = case.solver
    xs = xs_tables(case.materials, s.G, diffusion=False)
    bct, _ = bc_table(mesh, s.bcs)
    R, F, (dirs, wts) = assemble_sn(mesh, xs, bct, s)
    k, x = eigen_sinvert(R, F)
    phi, psi, q, p = sn_normalize(mesh, xs, x, wts, s.power)
    res = Result(k, phi, q, p, R if return_ops else None, F if return_ops else None)
    res.psi = psi
    return res


def sn_order(mesh: Mesh, dirs: np.ndarray):
    """Per-direction upwind level schedule (orc_sn_order)."""
    L = lib()
    M = len(dirs)
    order = np.zeros((M, mesh.C), np.int32)
    level = np.zeros((M, mesh.C), np.int32)
    nlev = np.zeros(M, np.int32)
    d = np.ascontiguousarray(dirs)
    if L.orc_sn_order(mesh.C, mesh.fmax, _p(mesh.nfaces), _p(mesh.fnorm), _p(mesh.fnb), M, _p(d),
                      _p(order), _p(level), _p(nlev)):
        raise pmp.PmpError("cyclic sweep dependencies")
    return order, level, nlev


def sn_sweep_eigen(mesh: Mesh, xs: XS, s: pmp.SolverInput, inner=4, max_outer=5000,
                   tol_k=1e-11, tol_phi=1e-10):
    """Sweep-based eigen solve (source iteration inside power iteration) for the
    sweepable regime: mixed-face-interpolation 1, LS off, vacuum boundaries."""
    L = lib()
    dirs, wts, refl, coef, _, _ = sn_setup(mesh, s)
    order, _, _ = sn_order(mesh, dirs)
    G, M = xs.G, len(wts)
    psi = np.zeros(mesh.C * G * M)
    phiw = np.ones(mesh.C * G)
    k = ct.c_double(1.0)
    its = L.orc_sn_eigen(mesh.C, mesh.fmax, _p(mesh.nfaces), _p(mesh.vol), _p(mesh.cmat),
                         _p(mesh.farea), _p(mesh.fnorm), _p(mesh.fnb), G, M, _p(dirs), _p(wts),
                         _p(coef), _p(xs.sigt), _p(xs.sigs), _p(xs.nusf), _p(xs.chi), _p(order),
                         inner, max_outer, tol_k, tol_phi, ct.byref(k), _p(psi), _p(phiw))
    return k.value, psi, phiw, its


def sn_sweep_eigen_anderson(mesh: Mesh, xs: XS, s: pmp.SolverInput, depth=5, max_outer=5000,
                            tol_k=1e-12, tol_phi=1e-10):
    """The fixed point of orc_sn_eigen's iteration (inner = 1: source from the
    moment, one sweep, k from the production ratio, the moment rescaled to the
    first production) reached with Anderson mixing of depth `depth` (Walker and
    Ni 2011, least squares over the last residual differences), for lattices
    whose plain iteration needs thousands of sweeps.  The same orc_sn_source
    and orc_sn_sweep do the work.  Stops when |dk| < tol_k and the relative L2
    change of the moment < tol_phi.  Returns (keff, phiw, sweeps)."""
    L = lib()
    dirs, wts, refl, coef, _, _ = sn_setup(mesh, s)
    order, _, _ = sn_order(mesh, dirs)
    G, M = xs.G, len(wts)
    n = mesh.C * G
    nus = xs.nusf[mesh.cmat].reshape(-1) * np.repeat(mesh.vol, G)
    psi = np.zeros(n * M)
    q = np.zeros(n)
    phi = np.ones(n)
    p0 = float(nus @ phi)
    k = 1.0
    dF, dG = [], []
    f_prev = g_prev = None
    for it in range(1, max_outer + 1):
        L.orc_sn_source(mesh.C, G, _p(mesh.vol), _p(mesh.cmat), _p(xs.sigs), _p(xs.nusf), _p(xs.chi),
                        1.0 / k, _p(phi), _p(q))
        g = np.zeros(n)
        L.orc_sn_sweep(mesh.C, mesh.fmax, _p(mesh.nfaces), _p(mesh.vol), _p(mesh.cmat), _p(mesh.farea),
                       _p(mesh.fnorm), _p(mesh.fnb), G, M, _p(dirs), _p(wts), _p(coef), _p(xs.sigt),
                       _p(order), _p(q), _p(psi), _p(g))
        p = float(nus @ g)
        knew = k * p / p0
        g *= p0 / p
        f = g - phi
        if abs(knew - k) < tol_k and np.linalg.norm(f) < tol_phi * np.linalg.norm(g):
            return knew, g, it
        k = knew
        if f_prev is not None and depth > 0:
            dF.append(f - f_prev)
            dG.append(g - g_prev)
            if len(dF) > depth:
                dF.pop(0)
                dG.pop(0)
        f_prev, g_prev = f, g
        if dF:
            A = np.stack(dF, axis=1)
            gam, *_ = np.linalg.lstsq(A, f, rcond=None)
            phi = g - np.stack(dG, axis=1) @ gam
        else:
            phi = g
        phi *= p0 / float(nus @ phi)
    raise RuntimeError("Anderson-mixed source iteration did not converge")
