"""Drop-in behaviour on the GPU: the pampa CLI and C-ABI on the reference's own
inputs reproduce its golden stdout (test/check_ref.txt)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REF

pytestmark = pytest.mark.gpu

CASES = ["slabs/reflected-diffusion", "pwr-iaea-benchmark/cartesian-diffusion",
         "pwr-iaea-benchmark/unstructured-diffusion", "pwr-iaea-benchmark/cartesian-diffusion-3d",
         "pwr-iaea-benchmark/unstructured-diffusion-3d", "slabs/reflected-s2", "slabs/reflected-s4"]


def golden_stdout(case):
    """The case's first run as printed by check.sh (between 'Run <case>...' and the
    end of the program's own output)."""
    text = open(os.path.join(REF, "check_ref.txt")).read()
    block = text.split(f"\nRun {case}...\n", 1)[1]
    end = block.index("Finalize...\nDone.\n\n") + len("Finalize...\nDone.\n\n")
    return block[:end]


@pytest.mark.parametrize("case", CASES)
def test_cli_stdout_matches_reference(case):
    import pampa_amd
    print(f"[case] {case}", flush=True)
    r = subprocess.run([pampa_amd.CLI_PATH, "input.pmp"], cwd=os.path.join(REF, case),
                       capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout == golden_stdout(case)


def test_c_abi_fields_and_power():
    import pampa_amd
    d = os.path.join(REF, "pwr-iaea-benchmark", "cartesian-diffusion")
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with pampa_amd.DropIn("input.pmp", "-silent") as run:
            assert run.dt() == []
            run.solve()
            k, p = run.result()
            C, G, M = run.sizes()
            q = run.field("power")
            phi = run.field("scalar-flux")
            P = run.field("production-rate")
            assert (C, G) == (3856, 2)
            assert f"{k:.6f}" == "1.029406"
            assert abs(q.sum() - 1.0) < 1e-12 and abs(p - 1.0) < 1e-12
            assert np.all(phi > 0)
            # production rate P = sum_g phi nuSf V / keff and power q = sum phi kSf V share
            # the kappa/nu = 3.2e-11/2.4355 ratio (kappa-fission defaulted from nu-fission)
            assert np.allclose(q, P * k * 3.2e-11 / 2.4355, rtol=1e-12)
            # set_field round trip on an input field
            T = np.full(C, 600.0)
            run.set_field("temperature", T)
            assert np.array_equal(run.field("temperature"), T)
            with pytest.raises(pampa_amd.PkError):
                run.field("no-such-field")
    finally:
        os.chdir(cwd)


def test_petsc_dump(tmp_path):
    """-petsc_dump writes the solution vector as VecView writes it to a PETSc
    binary viewer (petsc.cxx:491-511): scalar_flux_<n>.ptc in the working
    directory, big-endian int32 class id 1211214 and length, then the values
    as big-endian doubles, beside an (empty) .info file."""
    import shutil
    import pampa_amd
    from pampa_amd import workloads as W
    d = W.slab_1d(str(tmp_path / "slab"))
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with pampa_amd.DropIn("input.pmp", "-silent", "-petsc_dump") as run:
            run.solve()
            phi = run.field("scalar-flux")
        # written by the solve's output step
        raw = open(os.path.join(d, "scalar_flux_0.ptc"), "rb").read()
    finally:
        os.chdir(cwd)
    hdr = np.frombuffer(raw[:8], dtype=">i4")
    assert hdr[0] == 1211214 and hdr[1] == len(phi)
    vals = np.frombuffer(raw[8:], dtype=">f8")
    assert np.array_equal(vals, phi)
    assert os.path.exists(os.path.join(d, "scalar_flux_0.ptc.info"))
    shutil.rmtree(d)
