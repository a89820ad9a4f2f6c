"""Input formats beyond Cartesian and unstructured-extruded meshes, through the
drop-in against the oracle: partitioned mesh data of one domain
(PartitionedMesh.cxx:4-263), boundary-condition materials (Mesh.cxx:63-204) and
split materials (Mesh.cxx:9-60).  Operators bit-exact, eigenpairs within
|dk| <= 1 pcm and flux L2 <= 1e-5, and the CLI's stdout on the reference's
partitioned rerun of cartesian-diffusion equal to its golden transcript."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REF

pytestmark = pytest.mark.gpu


def _product(d):
    import pampa_amd
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with pampa_amd.DropIn("input.pmp", "-silent") as run:
            run.solve()
            k, _ = run.result()
            return k, run.field("scalar-flux"), run.operator("R"), run.operator("F")
    finally:
        os.chdir(cwd)


def _oracle(oracle_mod, d):
    cwd = os.getcwd()
    os.chdir(d)
    try:
        return oracle_mod.solve_input("input.pmp", return_ops=True)
    finally:
        os.chdir(cwd)


def _cases(oracle_mod, tmp_path, name):
    import meshdata
    import test_oracle
    from pampa_amd import workloads as W
    if name == "partitioned-exact":
        return test_oracle._partitioned_case(oracle_mod, tmp_path, repr)[0]
    if name == "bc-material-wall":
        core = W.quarter_core_2d(str(tmp_path / "core"), seed=4, side=24)
        return meshdata.with_wall(core, str(tmp_path / "wall"))
    if name == "split-material":
        d = W.quarter_core_2d(str(tmp_path / "core"), seed=4, side=34)
        with open(os.path.join(d, "fuel-absorber.pmp"), "a") as fh:
            fh.write("split 1\n")
        return d
    raise KeyError(name)


@pytest.mark.parametrize("name", ["partitioned-exact", "bc-material-wall", "split-material"])
def test_format_parity(oracle_mod, tmp_path, name):
    print(f"[case] {name}", flush=True)
    d = _cases(oracle_mod, tmp_path, name)
    ref = _oracle(oracle_mod, d)
    k, phi, R, F = _product(d)
    for (p, c, v), O in ((R, ref.R), (F, ref.F)):
        assert np.array_equal(p, O.ptr) and np.array_equal(c, O.col)
        assert np.array_equal(v, O.val)
    assert abs(k - ref.keff) * 1e5 <= 1.0
    assert np.linalg.norm(phi - ref.phi) / np.linalg.norm(ref.phi) <= 1e-5


def test_partitioned_rerun_stdout(oracle_mod, tmp_path):
    """The mesh written as the reference's writer writes it (3 decimals) and read
    back as partitioned mesh data: the CLI prints the golden transcript of the
    reference's partitioned rerun (check_ref.txt:86-100)."""
    import meshdata
    import pampa_amd
    import test_gpu_dropin
    import test_oracle
    d, _ = test_oracle._partitioned_case(oracle_mod, tmp_path, meshdata.fixed3)
    r = subprocess.run([pampa_amd.CLI_PATH, "input.pmp"], cwd=d, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout == test_gpu_dropin.golden_stdout("pwr-iaea-benchmark/cartesian-diffusion")
