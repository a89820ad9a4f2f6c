"""Time steps (n > 0) of the diffusion drop-in against the oracle: the
time-step operator R_t and right-hand side b of DiffusionSolver.cxx:134-175
bit-exact, and the step's flux against an exact solve of R_t phi = b, from a
flux and a delayed source that are not in equilibrium (so a step that left
the flux unchanged would fail)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(oracle_mod, d):
    case = oracle_mod.pmp.read_input(os.path.join(d, "input.pmp"))
    mesh = oracle_mod.build_mesh(case)
    xs = oracle_mod.xs_tables(case.materials, case.solver.G, True)
    bct, bcv = oracle_mod.bc_table(mesh, case.solver.bcs)
    tr = oracle_mod.transient_tables(case.materials, case.solver.G)
    return mesh, xs, bct, bcv, tr


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_transient_steps_match_oracle(pk, oracle_mod, tmp_path, precision):
    from pampa_amd import workloads as W
    d = W.transient_core_2d(str(tmp_path / "t"), side=40, G=3, steps=2)
    mesh, xs, bct, bcv, tr = _setup(oracle_mod, d)
    rng = np.random.default_rng(11)
    cwd = os.getcwd()
    os.chdir(d)
    try:
        # ksp_rtol 1e-11: about 5x the residual an exact FP64 solve leaves on this
        # operator (condition number ~4e4), so the flux is pinned to ~1e-7
        with pk.DropIn("input.pmp", "-silent", "-ksp_rtol", "1e-11", "-pk_precision", precision) as run:
            C, G, _ = run.sizes()
            run.solve()
            k, _ = run.result()
            phi = run.field("scalar-flux")
            dts = run.dt()
            assert len(dts) == 2
            # step 1 from a perturbed flux and a random delayed source; step 2
            # from the flux and delayed source step 1 produced
            phi0 = phi * rng.uniform(0.8, 1.2, C * G)
            # delayed source around its equilibrium value beta P_i
            S = tr.beta[mesh.cmat] * np.sum(phi.reshape(C, G) * xs.nusf[mesh.cmat], axis=1) * mesh.vol / k
            S *= rng.uniform(0.5, 1.5, C)
            run.set_field("scalar-flux", phi0)
            run.set_field("delayed-source", S)
            t = 0.0
            for n in (1, 2):
                print(f"[case] transient {precision} step {n} dt={dts[n - 1]}", flush=True)
                t += dts[n - 1]
                run.solve(n, dts[n - 1], t)
                p, c, v = run.operator("R")
                b = run.operator("b")
                phi1 = run.field("scalar-flux")
                Rt = oracle_mod.assemble_diffusion_transient(mesh, xs, bct, bcv, tr, dts[n - 1], k)
                bo = oracle_mod.diffusion_rhs(mesh, tr, dts[n - 1], S, phi0)
                assert np.array_equal(p, Rt.ptr) and np.array_equal(c, Rt.col)
                assert np.array_equal(v, Rt.val)
                assert np.array_equal(b, bo)
                ref = oracle_mod.linear_solve(Rt, bo)
                err = np.linalg.norm(phi1 - ref) / np.linalg.norm(ref)
                moved = np.linalg.norm(phi1 - phi0) / np.linalg.norm(phi0)
                print(f"   flux L2 vs oracle {err:.2e}, change over the step {moved:.2e}", flush=True)
                assert err <= 1e-6
                # the first step starts away from equilibrium and moves the flux;
                # the second starts from the first's result
                assert moved > (1e-4 if n == 1 else 0.0)
                phi0 = phi1
                # each solve leaves the equilibrium delayed source S = beta P of its
                # flux for the next step (NeutronicSolver.cxx:95-102)
                P = run.field("production-rate")
                S = tr.beta[mesh.cmat] * P
                assert np.array_equal(run.field("delayed-source"), S)
    finally:
        os.chdir(cwd)
