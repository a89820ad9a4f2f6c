"""The S_N scattering + fission source as a per-material 16 x 16 contraction on
the matrix cores (FP64 MFMA), against the oracle's source (orc_sn_source,
the right-hand side form of SNSolver.cxx:417-439) and through eigen solves
against the oracle's source iteration."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DK_PCM = 1.0
FLUX_L2 = 1e-5


def _hex16(oracle_mod, tmp_path, **kw):
    from pampa_amd import workloads as W
    d = W.hex_sn_3d(str(tmp_path / "hex16"), G=16, **kw)
    case = oracle_mod.pmp.read_input(os.path.join(d, "input.pmp"))
    mesh = oracle_mod.build_mesh(case)
    xs = oracle_mod.xs_tables(case.materials, case.solver.G, False)
    return d, case, mesh, xs


def _sweep(pk, oracle_mod, mesh, xs, case, precision, cmat=None):
    fptr, fa, fc, fn, fb = pk.ragged_faces(mesh.nfaces, mesh.fmax, mesh.farea, mesh.fcen, mesh.fnorm,
                                           mesh.fnb)
    dm = pk.DeviceMesh(fptr, mesh.vol, mesh.cen, mesh.cmat if cmat is None else cmat, fa, fc, fn, fb)
    dx = pk.DeviceXS(xs.sigt, xs.sigs.reshape(len(xs.sigt), -1), xs.nusf, xs.kapf, xs.chi)
    dirs, wts, _ = oracle_mod.quadrature(case.solver.order)
    bct, _ = oracle_mod.bc_table(mesh, case.solver.bcs)
    return pk.Sweep(dm, dx, dirs, wts, bct, precision)


def _oracle_source(oracle_mod, mesh, xs, cmat, phiw, k):
    q = np.zeros(mesh.C * xs.G)
    P = oracle_mod._p
    oracle_mod.lib().orc_sn_source(mesh.C, xs.G, P(mesh.vol), P(cmat), P(xs.sigs), P(xs.nusf), P(xs.chi),
                                   1.0 / k, P(phiw), P(q))
    return q


@pytest.mark.parametrize("materials", ["mesh", "random-per-cell"])
def test_source_on_matrix_cores_matches_oracle(pk, oracle_mod, tmp_path, materials):
    """1,368 cells (85 full tiles of 16 and a tail of 8), 16 groups, 4
    materials: as laid out by the lattice (tiles straddling columns of two
    materials) and drawn at random per cell (up to 4 materials per tile).  The
    vector kernel reproduces the oracle's operation order (bit-identical); the
    matrix cores sum in another order (relative difference at FP64 round-off
    of a 16-term positive sum)."""
    d, case, mesh, xs = _hex16(oracle_mod, tmp_path, rings=2, layers=3, order=2)
    assert mesh.C % 16 != 0 and len(xs.sigt) == 4
    rng = np.random.default_rng(7)
    cmat = np.ascontiguousarray(mesh.cmat if materials == "mesh" else rng.integers(0, 4, mesh.C), np.int32)
    sw = _sweep(pk, oracle_mod, mesh, xs, case, pk.PK_PREC_FP64, cmat)
    phiw = rng.uniform(0.05, 3.0, mesh.C * xs.G)
    k = 1.137
    q_o = _oracle_source(oracle_mod, mesh, xs, cmat, phiw, k)
    tiles = cmat[:mesh.C // 16 * 16].reshape(-1, 16)
    print(f"[case] source contraction, materials {materials}: C={mesh.C}, tiles with more than one "
          f"material {int(np.sum(tiles.min(1) != tiles.max(1)))}", flush=True)
    sw.set_contraction(1)
    q_v = sw.source(phiw, k)
    assert np.array_equal(q_v, q_o)
    sw.set_contraction(2)
    q_m = sw.source(phiw, k)
    rel = np.max(np.abs(q_m - q_o) / np.abs(q_o))
    print(f"   matrix cores vs oracle: max relative difference {rel:.2e}", flush=True)
    assert rel < 1e-14
    # the default picks the matrix cores at 16 groups
    sw.set_contraction(0)
    assert np.array_equal(sw.source(phiw, k), q_m)


def test_matrix_core_contraction_rejects_other_group_counts(pk, oracle_mod, tmp_path):
    from pampa_amd import workloads as W
    d = W.hex_sn_3d(str(tmp_path / "hex4"), rings=1, layers=2, order=2, G=4)
    case = oracle_mod.pmp.read_input(os.path.join(d, "input.pmp"))
    mesh = oracle_mod.build_mesh(case)
    xs = oracle_mod.xs_tables(case.materials, case.solver.G, False)
    sw = _sweep(pk, oracle_mod, mesh, xs, case, pk.PK_PREC_FP64)
    with pytest.raises(pk.PkError):
        sw.set_contraction(2)
    sw.set_contraction(1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sn_eigen_16_groups_on_matrix_cores(pk, oracle_mod, tmp_path, precision):
    """A 16-group lattice (336 cells, S2) solved with the source on the matrix
    cores and on the vector kernel: both eigenpairs within 1 pcm and 1e-5 of
    the oracle's source iteration."""
    d, case, mesh, xs = _hex16(oracle_mod, tmp_path, rings=1, layers=2, order=2)
    k_o, _, phiw_o, _ = oracle_mod.sn_sweep_eigen(mesh, xs, case.solver, inner=1, tol_k=1e-12, tol_phi=1e-11)
    prec = pk.PK_PREC_FP64 if precision == "fp64" else pk.PK_PREC_MIXED
    shape = lambda v: np.asarray(v) / np.sum(v)  # noqa: E731
    for mode in (2, 1):
        sw = _sweep(pk, oracle_mod, mesh, xs, case, prec)
        sw.set_contraction(mode)
        k, phiw, outer, sweeps = sw.solve()
        dk = abs(k - k_o) * 1e5
        l2 = np.linalg.norm(shape(phiw) - shape(phiw_o)) / np.linalg.norm(shape(phiw_o))
        print(f"[case] 16-group eigen {precision} contraction {mode}: dk={dk:.2e} pcm L2={l2:.2e} "
              f"sweeps {sweeps}", flush=True)
        assert dk <= DK_PCM and l2 <= FLUX_L2
