"""Parity of the HIP path with the FP64 oracle on identical inputs:
operators bit-exact, eigenpairs within |dk| <= 1 pcm and scalar-flux relative
L2 <= 1e-5 (FP64 drop-in and FP32-inner mixed precision), sweeps and their
schedules against the oracle's restatement."""
import os

import numpy as np
import pytest

from conftest import REF

pytestmark = pytest.mark.gpu

DK_PCM = 1.0          # |dk| tolerance in pcm (1e-5)
FLUX_L2 = 1e-5        # relative L2 of the power-normalised scalar flux


def _case(tmp_path, name):
    from pampa_amd import workloads as W
    if name == "slab-1d":
        return W.slab_1d(str(tmp_path / name))
    if name == "quarter-core-170":
        return W.quarter_core_2d(str(tmp_path / name))
    if name == "quarter-core-g4":
        return W.quarter_core_2d(str(tmp_path / name), seed=11, side=48, G=4)
    if name == "quarter-core-300":
        # 180,000 rows = 704 row units: reductions in groups of 32 workgroups
        return W.quarter_core_2d(str(tmp_path / name), seed=5, side=300)
    return os.path.join(REF, name)


def _oracle(oracle_mod, d):
    cwd = os.getcwd()
    os.chdir(d)
    try:
        return oracle_mod.solve_input("input.pmp", return_ops=True)
    finally:
        os.chdir(cwd)


def _product(d, *opts):
    import pampa_amd
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with pampa_amd.DropIn("input.pmp", "-silent", *opts) as run:
            run.solve()
            k, p = run.result()
            phi = run.field("scalar-flux")
            R = run.operator("R")
            F = run.operator("F")
            st = run.stats()
        return k, phi, R, F, st
    finally:
        os.chdir(cwd)


DIFF = ["slab-1d", "quarter-core-170", "quarter-core-g4", "quarter-core-300", "slabs/reflected-diffusion",
        "pwr-iaea-benchmark/cartesian-diffusion", "pwr-iaea-benchmark/unstructured-diffusion"]


@pytest.mark.parametrize("name", DIFF)
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_diffusion_parity(oracle_mod, tmp_path, name, precision):
    print(f"[case] {name} {precision}", flush=True)
    d = _case(tmp_path, name)
    ref = _oracle(oracle_mod, d)
    k, phi, R, F, st = _product(d, "-pk_precision", precision)
    # operators: pattern and FP64 values bit-exact
    for (p, c, v), O in ((R, ref.R), (F, ref.F)):
        assert np.array_equal(p, O.ptr) and np.array_equal(c, O.col)
        assert np.array_equal(v, O.val)
    assert abs(k - ref.keff) * 1e5 <= DK_PCM
    assert np.linalg.norm(phi - ref.phi) / np.linalg.norm(ref.phi) <= FLUX_L2
    print(f"   dk={abs(k - ref.keff) * 1e5:.2e} pcm outer={st['outer']} inner={st['inner']}", flush=True)


# inner-solver variants: plain block-Jacobi preconditioning, and restart lengths
# whose cycles end by exhausting the basis (the norm-only step after the last
# column) as well as by the reduction target
@pytest.mark.parametrize("opts", [("-pk_preconditioner", "bj"), ("-ksp_gmres_restart", "4"),
                                  ("-ksp_gmres_restart", "60"),
                                  ("-pk_preconditioner", "bj", "-ksp_gmres_restart", "7")])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_diffusion_parity_inner_variants(oracle_mod, tmp_path, opts, precision):
    print(f"[case] quarter-core-g4 {precision} {' '.join(opts)}", flush=True)
    d = _case(tmp_path, "quarter-core-g4")
    ref = _oracle(oracle_mod, d)
    k, phi, _, _, st = _product(d, "-pk_precision", precision, *opts)
    assert abs(k - ref.keff) * 1e5 <= DK_PCM
    assert np.linalg.norm(phi - ref.phi) / np.linalg.norm(ref.phi) <= FLUX_L2
    print(f"   dk={abs(k - ref.keff) * 1e5:.2e} pcm outer={st['outer']} inner={st['inner']}", flush=True)


def _hex(oracle_mod, tmp_path, **kw):
    from pampa_amd import workloads as W
    d = W.hex_sn_3d(str(tmp_path / "hex"), **kw)
    case = oracle_mod.pmp.read_input(os.path.join(d, "input.pmp"))
    mesh = oracle_mod.build_mesh(case)
    xs = oracle_mod.xs_tables(case.materials, case.solver.G, False)
    return d, case, mesh, xs


def _device_sweep(pk, oracle_mod, mesh, xs, case, precision):
    fptr, fa, fc, fn, fb = pk.ragged_faces(mesh.nfaces, mesh.fmax, mesh.farea, mesh.fcen, mesh.fnorm,
                                           mesh.fnb)
    dm = pk.DeviceMesh(fptr, mesh.vol, mesh.cen, mesh.cmat, fa, fc, fn, fb)
    dx = pk.DeviceXS(xs.sigt, xs.sigs.reshape(len(xs.sigt), -1), xs.nusf, xs.kapf, xs.chi)
    dirs, wts, _ = oracle_mod.quadrature(case.solver.order)
    bct, _ = oracle_mod.bc_table(mesh, case.solver.bcs)
    return pk.Sweep(dm, dx, dirs, wts, bct, precision), dirs, wts


@pytest.mark.parametrize("G,gt,pd,step", [(2, None, None, None), (3, None, None, None), (4, None, None, None),
                                          (4, 2, None, None), (4, 1, None, None), (4, 1, (1, 4), 64),
                                          (4, 1, (2, 1), 96), (4, 4, (1, 1), 50)])
def test_record_sweep_matches_geometry_sweep(pk, oracle_mod, tmp_path, monkeypatch, G, gt, pd, step):
    """Sweeps of the per-direction kernel (workgroups walking one direction's
    steps for G / gt groups each, upwind values in local-memory slots, records
    prefetched D steps ahead, levels split into steps of at most `step` items)
    and sweeps of the level-step kernel that reads the geometry compute the
    same items in the same arithmetic: bit-identical angular fluxes in FP64 and
    FP32, for every configuration; a repeated sweep gives the same bits."""
    d, case, mesh, xs = _hex(oracle_mod, tmp_path, rings=3, layers=4, order=6, G=G)
    rng = np.random.default_rng(3)
    q = rng.uniform(0.1, 2.0, mesh.C * xs.G)
    if gt:
        monkeypatch.setenv("PK_SWEEP_GT", str(gt))
    if pd:
        monkeypatch.setenv("PK_SWEEP_P", str(pd[0]))
        monkeypatch.setenv("PK_SWEEP_D", str(pd[1]))
    if step:
        monkeypatch.setenv("PK_SWEEP_STEP", str(step))
    for prec in (pk.PK_PREC_FP64, pk.PK_PREC_MIXED):
        print(f"[case] records vs geometry, G {G}, groups per workgroup {gt}, (P, D) {pd}, step {step}, "
              f"precision {prec}", flush=True)
        rec, _, _ = _device_sweep(pk, oracle_mod, mesh, xs, case, prec)
        psi_r, phi_r = rec.sweep(q)
        psi_r2, _ = rec.sweep(q)
        monkeypatch.setenv("PK_SWEEP_GEOMETRY", "1")
        geo, _, _ = _device_sweep(pk, oracle_mod, mesh, xs, case, prec)
        monkeypatch.delenv("PK_SWEEP_GEOMETRY")
        psi_g, phi_g = geo.sweep(q)
        assert np.array_equal(psi_r, psi_g) and np.array_equal(phi_r, phi_g)
        assert np.array_equal(psi_r, psi_r2)


def test_sweep_schedule_and_single_sweep(pk, oracle_mod, tmp_path):
    d, case, mesh, xs = _hex(oracle_mod, tmp_path, rings=2, layers=3, order=4, G=4)
    sw, dirs, wts = _device_sweep(pk, oracle_mod, mesh, xs, case, pk.PK_PREC_FP64)
    # schedules: identical levels (integer work, bit-exact)
    order, level, nlev = oracle_mod.sn_order(mesh, dirs)
    nl, lv = sw.schedule()
    assert np.array_equal(nl, nlev) and np.array_equal(lv, level)
    # one FP64 sweep against the oracle's sweep on the same source
    rng = np.random.default_rng(1)
    q = rng.uniform(0.1, 2.0, mesh.C * xs.G)
    psi, phiw = sw.sweep(q)
    _, _, _, coef, _, _ = oracle_mod.sn_setup(mesh, case.solver)
    psi_o = np.zeros_like(psi)
    phiw_o = np.zeros_like(phiw)
    L = oracle_mod.lib()
    P = oracle_mod._p
    L.orc_sn_sweep(mesh.C, mesh.fmax, P(mesh.nfaces), P(mesh.vol), P(mesh.cmat), P(mesh.farea),
                   P(mesh.fnorm), P(mesh.fnb), xs.G, len(wts), P(dirs), P(wts), P(coef), P(xs.sigt),
                   P(order), P(q), P(psi_o), P(phiw_o))
    assert np.max(np.abs(psi - psi_o) / np.abs(psi_o)) < 1e-13
    assert np.max(np.abs(phiw - phiw_o) / np.abs(phiw_o)) < 1e-13
    # FP32 sweep of the same source: relative error at FP32 round-off
    sw32, _, _ = _device_sweep(pk, oracle_mod, mesh, xs, case, pk.PK_PREC_MIXED)
    psi32, _ = sw32.sweep(q)
    assert np.max(np.abs(psi32 - psi_o) / np.abs(psi_o)) < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sn_anderson_same_eigenpair_fewer_sweeps(pk, oracle_mod, tmp_path, precision):
    """Anderson-accelerated source iteration converges to the eigenpair of the
    plain iteration (and of the FP64 oracle) in fewer sweeps."""
    d, case, mesh, xs = _hex(oracle_mod, tmp_path, rings=3, layers=4, order=4, G=4)
    prec = pk.PK_PREC_FP64 if precision == "fp64" else pk.PK_PREC_MIXED
    sw, _, _ = _device_sweep(pk, oracle_mod, mesh, xs, case, prec)
    k0, phi0, o0, s0 = sw.solve(accel=0)
    k5, phi5, o5, s5 = sw.solve(accel=5)
    k_o, _, phiw_o, _ = oracle_mod.sn_sweep_eigen(mesh, xs, case.solver, inner=1, tol_k=1e-12, tol_phi=1e-11)
    shape = lambda v: np.asarray(v) / np.sum(v)
    print(f"[case] S_N Anderson {precision}: plain {s0} sweeps, accelerated {s5} sweeps; "
          f"dk {abs(k5 - k_o) * 1e5:.2e} pcm", flush=True)
    for k, phi in ((k0, phi0), (k5, phi5)):
        assert abs(k - k_o) * 1e5 <= DK_PCM
        assert np.linalg.norm(shape(phi) - shape(phiw_o)) / np.linalg.norm(shape(phiw_o)) <= FLUX_L2
    assert s5 < s0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sn_eigen_parity(oracle_mod, tmp_path, precision):
    d, case, mesh, xs = _hex(oracle_mod, tmp_path, rings=2, layers=3, order=4, G=4)
    k_o, psi_o, phiw_o, its = oracle_mod.sn_sweep_eigen(mesh, xs, case.solver, inner=1,
                                                        tol_k=1e-12, tol_phi=1e-11)
    phi_o = 4 * np.pi * phiw_o
    kap = xs.kapf[mesh.cmat]
    phi_o = phi_o * (1.0 / np.sum(phi_o.reshape(mesh.C, -1) * kap * mesh.vol[:, None]))
    import pampa_amd
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with pampa_amd.DropIn("input.pmp", "-silent", "-pk_precision", precision) as run:
            run.solve()
            k, p = run.result()
            phi = run.field("scalar-flux")
            psi = run.field("angular-flux")
    finally:
        os.chdir(cwd)
    assert abs(k - k_o) * 1e5 <= DK_PCM
    assert np.linalg.norm(phi - phi_o) / np.linalg.norm(phi_o) <= FLUX_L2
    assert np.all(psi >= 0.0)


def test_sn_tiny_matches_assembled_operator(oracle_mod, tmp_path):
    """Tiny lattice: the GPU's swept eigenpair equals the eigenpair of the
    reference-form assembled S_N operator (direct shift-invert solve)."""
    d, case, mesh, xs = _hex(oracle_mod, tmp_path, rings=1, layers=2, order=2, G=2)
    bct, _ = oracle_mod.bc_table(mesh, case.solver.bcs)
    R, F, (dirs, wts) = oracle_mod.assemble_sn(mesh, xs, bct, case.solver)
    k_as, x = oracle_mod.eigen_sinvert(R, F)
    import pampa_amd
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with pampa_amd.DropIn("input.pmp", "-silent") as run:
            run.solve()
            k, _ = run.result()
    finally:
        os.chdir(cwd)
    assert abs(k - k_as) * 1e5 <= DK_PCM


LATTICE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hex_s6_lattice.npz")


@pytest.mark.skipif(not os.path.exists(LATTICE), reason="oracle fixture of the S6 lattice not generated")
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_full_s6_lattice_against_oracle_fixture(tmp_path, precision):
    """The full hexagonal S6 lattice (317,760 prisms, 4 groups, 61,009,920 angular
    unknowns) through the drop-in against the FP64 oracle's source iteration:
    keff and every 64th entry of the scalar-flux shape
    (tests/golden/make_fixtures.py)."""
    import pampa_amd
    from pampa_amd import workloads as W
    fx = np.load(LATTICE)
    st = int(fx["stride"])
    d = W.hex_sn_3d(str(tmp_path / "hex-s6"))
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with pampa_amd.DropIn("input.pmp", "-silent", "-pk_precision", precision) as run:
            run.solve()
            k, _ = run.result()
            phi = run.field("scalar-flux")
    finally:
        os.chdir(cwd)
    shape = lambda v: np.asarray(v, np.float64) / np.sum(v)  # noqa: E731
    xs, xo = shape(phi[::st]), shape(fx["phiw"])
    dk = abs(k - float(fx["keff"])) * 1e5
    l2 = np.linalg.norm(xs - xo) / np.linalg.norm(xo)
    print(f"[case] S6 lattice {precision}: dk={dk:.2e} pcm flux L2 (every {st}th entry) {l2:.2e}", flush=True)
    assert dk <= DK_PCM
    assert l2 <= FLUX_L2
