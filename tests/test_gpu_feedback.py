"""Temperature-dependent cross sections (FeedbackNuclearData.hxx:8-150) as a
coupling hook: a coupled code sets the cells' temperatures through
pampa_set_field("temperature") between solves, and each steady solve's
operators, evaluated on the device at those temperatures, must equal the
oracle's bit for bit, with the eigenpair within |dk| <= 1 pcm and flux L2 <=
1e-5 -- over a few rounds of a power-to-temperature feedback loop."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_feedback_loop_against_oracle(pk, oracle_mod, tmp_path, precision):
    from pampa_amd import workloads as W
    d = W.feedback_core_2d(str(tmp_path / "fb"), side=30)
    case = oracle_mod.pmp.read_input(os.path.join(d, "input.pmp"))
    mesh = oracle_mod.build_mesh(case)
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with pk.DropIn("input.pmp", "-silent", "-pk_precision", precision) as run:
            C, G, _ = run.sizes()
            T = np.full(C, 300.0)
            for it in range(3):
                if it < 2:
                    run.set_field("temperature", T)
                else:
                    # the last round writes the field in device memory directly
                    import ctypes as ct
                    dev, n = run.field_device("temperature")
                    assert n == C
                    hip = ct.CDLL("libamdhip64.so")
                    hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
                    Tc = np.ascontiguousarray(T)
                    assert hip.hipMemcpy(dev, Tc.ctypes.data, 8 * C, 1) == 0  # host to device
                    assert np.array_equal(run.field("temperature"), T)
                run.solve()
                k, _ = run.result()
                phi = run.field("scalar-flux")
                q = run.field("power")
                R, F = run.operator("R"), run.operator("F")
                ref = oracle_mod.solve_diffusion_case(case, return_ops=True, T=T)
                for (p, c, v), O in ((R, ref.R), (F, ref.F)):
                    assert np.array_equal(p, O.ptr) and np.array_equal(c, O.col)
                    assert np.array_equal(v, O.val)
                dk = abs(k - ref.keff) * 1e5
                l2 = np.linalg.norm(phi - ref.phi) / np.linalg.norm(ref.phi)
                print(f"[case] feedback {precision} round {it}: keff {k:.6f} dk={dk:.2e} pcm L2={l2:.2e} "
                      f"T {T.min():.0f}..{T.max():.0f} K", flush=True)
                assert dk <= 1.0 and l2 <= 1e-5
                # a crude thermal model: temperature rises with the cell's power density
                T = 300.0 + 700.0 * q / mesh.vol / np.max(q / mesh.vol)
    finally:
        os.chdir(cwd)
