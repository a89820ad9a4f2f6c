"""Writes the oracle fixtures of the workloads too large for the oracle to solve
inside a test run (tests/golden/*.npz), in this build container:

  batched_classes.npz  the batched family's largest size classes (128x128x16,
                       256x256x8, 192x192x16, 256x256x16 groups; one instance
                       each): keff and every 64th entry of the power-normalised
                       scalar flux of the FP64 oracle's eigen solve;
  hex_s6_lattice.npz   the hexagonal S6 lattice of 61,009,920 angular unknowns
                       (331 columns x 24 prisms x 40 layers, 4 groups): keff and
                       every 64th entry of the moment sum_m w_m psi (production
                       normalised) of the oracle's FP64 source iteration, driven
                       to its fixed point by Anderson mixing (the plain
                       iteration needs thousands of sweeps at this size).

The eigen solves are the oracle's (oracle/__init__.py): exact sparse LU plus
restarted Arnoldi (shift-invert) for the diffusion instances; for 256x256x16
(1,048,576 rows), whose LU with SuperLU's own column orderings does not
finish here, the same shift-invert Arnoldi on an LU of R with its cells in
nested-dissection order (recursive bisection of the cell grid, separator
lines last, a cell's groups together).  The file is rewritten after every
instance.  Run:
    python tests/golden/make_fixtures.py [batched] [lattice]
"""
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from pampa_amd import workloads as W  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
STRIDE = 64
BIG_CLASSES = [(128, 16), (256, 8), (192, 16), (256, 16)]


def _instance_of(side, G):
    for i in range(4096):
        rng = W.SplitMix64(3_000_000 + i)
        s = W.BATCH_SIDES[rng.choice(len(W.BATCH_SIDES))]
        g = W.BATCH_GROUPS[rng.choice(len(W.BATCH_GROUPS))]
        if (s, g) == (side, G):
            return i
    raise RuntimeError("class not drawn")


def _nd_cells(side):
    """Nested-dissection order of a side x side cell grid (cell = y side + x):
    halves first, then the separator line, down to blocks of 64 cells."""
    order = []

    def rec(x0, x1, y0, y1):
        nx, ny = x1 - x0, y1 - y0
        if nx * ny <= 64:
            order.extend(y * side + x for y in range(y0, y1) for x in range(x0, x1))
            return
        if nx >= ny:
            m = (x0 + x1) // 2
            rec(x0, m, y0, y1)
            rec(m + 1, x1, y0, y1)
            order.extend(y * side + m for y in range(y0, y1))
        else:
            m = (y0 + y1) // 2
            rec(x0, x1, y0, m)
            rec(x0, x1, m + 1, y1)
            order.extend(m * side + x for x in range(x0, x1))

    rec(0, side, 0, side)
    return np.array(order)


def _sinvert_nd(R, F, side, G, tol=1e-13):
    """eigen_sinvert with the LU of R taken in nested-dissection order of the
    cells (symmetric permutation, diagonal pivots)."""
    import scipy.sparse.linalg as sla
    cells = _nd_cells(side)
    assert len(cells) == side * side and len(np.unique(cells)) == side * side
    p = (cells[:, None] * G + np.arange(G)[None, :]).reshape(-1)
    inv = np.empty_like(p)
    inv[p] = np.arange(len(p))
    Rs = R.scipy().tocsr()
    Fs = F.scipy().tocsr()
    lu = sla.splu(Rs[p][:, p].tocsc(), permc_spec="NATURAL", diag_pivot_thresh=0.0,
                  options=dict(SymmetricMode=True))
    op = sla.LinearOperator(Rs.shape, matvec=lambda v: lu.solve((Fs @ v)[p])[inv], dtype=np.float64)
    vals, vecs = sla.eigs(op, k=1, which="LM", tol=tol, v0=np.ones(R.n), ncv=20)
    x = vecs[:, 0].real.copy()
    return float(vals[0].real), (x if x.sum() > 0 else -x)


def batched():
    out = {}
    for side, G in BIG_CLASSES:
        i = _instance_of(side, G)
        d = tempfile.mkdtemp(prefix="fixture_")
        W.batched_instance(d, i)
        case = oracle.pmp.read_input(os.path.join(d, "input.pmp"))
        mesh = oracle.build_mesh(case)
        s = case.solver
        xs = oracle.xs_tables(case.materials, s.G, diffusion=True)
        bct, bcv = oracle.bc_table(mesh, s.bcs)
        R, F = oracle.assemble_diffusion(mesh, xs, bct, bcv)
        t0 = time.time()
        if side * side * G > 600_000:
            k, x = _sinvert_nd(R, F, side, G)
        else:
            k, x = oracle.eigen_sinvert(R, F, permc="COLAMD")
        phi, _, _ = oracle.normalize_flux(mesh, xs, x, s.power)
        print(f"instance {i} ({side}x{side}, {G} groups, {R.n} rows): keff {k:.10f} in {time.time() - t0:.0f} s",
              flush=True)
        out[f"inst{i}_keff"] = np.array(k)
        out[f"inst{i}_phi"] = phi[::STRIDE].copy()
        out["stride"] = np.array(STRIDE)
        np.savez_compressed(os.path.join(HERE, "batched_classes.npz"), **out)


def lattice():
    d = tempfile.mkdtemp(prefix="fixture_")
    W.hex_sn_3d(d)
    case = oracle.pmp.read_input(os.path.join(d, "input.pmp"))
    mesh = oracle.build_mesh(case)
    s = case.solver
    xs = oracle.xs_tables(case.materials, s.G, diffusion=False)
    t0 = time.time()
    k, phiw, its = oracle.sn_sweep_eigen_anderson(mesh, xs, s, tol_k=1e-12, tol_phi=1e-10)
    kap = xs.kapf[mesh.cmat]
    # production normalisation of the moment: sum phiw nusf V = 1
    nus = xs.nusf[mesh.cmat]
    phiw = phiw / float(np.sum(phiw.reshape(mesh.C, -1) * nus * mesh.vol[:, None]))
    print(f"lattice: keff {k:.10f} after {its} sweeps in {time.time() - t0:.0f} s", flush=True)
    assert np.all(kap >= 0.0)
    np.savez_compressed(os.path.join(HERE, "hex_s6_lattice.npz"), keff=np.array(k), phiw=phiw[::STRIDE].copy(),
                        stride=np.array(STRIDE), outer=np.array(its))


if __name__ == "__main__":
    what = sys.argv[1:] or ["batched", "lattice"]
    oracle.build()
    if "batched" in what:
        batched()
    if "lattice" in what:
        lattice()
