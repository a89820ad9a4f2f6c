"""CPU checks of the product's boundary: the C-ABI library loads and exports
every entry point its headers declare, the workload generators are
deterministic, and the drop-in CLI reports input errors the reference's way
(no device call happens before the input is valid)."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import REF, ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in ("pk.h", "pampa.h")]


def _declared(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\**\s+\**\s*([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return [n for n in names if n.startswith(("pk_", "pampa_"))]


def test_library_exports_declared_symbols():
    import pampa_amd
    lib = ctypes.CDLL(pampa_amd.LIB_PATH)
    names = [n for h in HEADERS for n in _declared(h)]
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_pk_last_error_without_device_calls():
    import pampa_amd
    lib = pampa_amd.lib()
    assert isinstance(lib.pk_last_error(), bytes)


def test_splitmix64_reference_values():
    from pampa_amd.workloads import SplitMix64
    r = SplitMix64(0)
    # published SplitMix64 outputs for seed 0
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_workloads_deterministic(tmp_path):
    from pampa_amd import workloads as W
    a = W.quarter_core_2d(str(tmp_path / "a"), seed=7, side=34)
    b = W.quarter_core_2d(str(tmp_path / "b"), seed=7, side=34)
    for f in sorted(os.listdir(a)):
        assert open(os.path.join(a, f)).read() == open(os.path.join(b, f)).read()


def test_hex_lattice_size(tmp_path, oracle_mod):
    from pampa_amd import workloads as W
    d = W.hex_sn_3d(str(tmp_path / "h"), rings=3, layers=2)
    m = oracle_mod.pmp.read_unstructured(os.path.join(d, "mesh.pmp"))
    assert len(m.cptr) - 1 == (1 + 3 * 3 * 4) * 24
    mesh = oracle_mod.build_unstructured(m)
    assert mesh.C == 2 * 37 * 24
    # boundary slots: the lattice perimeter has 6 (2 r + 1) hexagon edges, each
    # split into 2 triangle edges, per layer; plus one z face per cell (2 layers)
    assert (mesh.fnb < 0).sum() == 6 * (2 * 3 + 1) * 2 * 2 + mesh.C


def _cli(args, cwd):
    import pampa_amd
    return subprocess.run([pampa_amd.CLI_PATH] + args, cwd=cwd, capture_output=True, text=True,
                          timeout=60)


@pytest.mark.parametrize("text,msg", [
    ("mesh cartesian mesh.pmp\nfoo 1\n", "unrecognized keyword 'foo'"),
    ("mesh cartesian nowhere.pmp\n", "unable to open nowhere.pmp"),
    ("mesh cartesian mesh.pmp\nsolver heat {\n}\n", "is not part of this build"),
])
def test_cli_input_errors(tmp_path, text, msg):
    (tmp_path / "mesh.pmp").write_text("dx 2\n1 1\nbc -x vacuum\nbc +x vacuum\nmaterials 2\n1 1\n")
    (tmp_path / "input.pmp").write_text(text)
    r = _cli(["input.pmp"], str(tmp_path))
    assert r.returncode == 1
    assert msg in r.stdout
    assert "error:" in r.stdout and "Error in pampa.initialize()." in r.stdout


@pytest.mark.parametrize("opt,value", [("-pk_outer_acceleration", "4"), ("-pk_outer_acceleration", "9"),
                                       ("-pk_inner_reduction", "1.5")])
def test_cli_option_errors(tmp_path, opt, value):
    """Out-of-range build options are refused while the command line is read,
    before any device call."""
    (tmp_path / "mesh.pmp").write_text("dx 2\n1 1\nbc -x vacuum\nbc +x vacuum\nmaterials 2\n1 1\n")
    (tmp_path / "input.pmp").write_text("mesh cartesian mesh.pmp\n")
    r = _cli(["input.pmp", opt, value], str(tmp_path))
    assert r.returncode == 1
    assert f"wrong {opt[1:]}" in r.stdout and "wrong command-line option" in r.stdout


def test_cli_missing_input(tmp_path):
    r = _cli([], str(tmp_path))
    assert r.returncode == 1 and "missing input file" in r.stdout


def test_product_reader_accepts_reference_inputs(tmp_path):
    """The product's parser reads every reference neutronics input up to the
    first device call; an invalid group count is caught at that point."""
    d = os.path.join(REF, "slabs", "reflected-diffusion")
    text = open(os.path.join(d, "input.pmp")).read().replace("energy-groups 2", "energy-groups 3")
    (tmp_path / "input.pmp").write_text(text.replace("../materials", os.path.join(REF, "slabs", "materials"))
                                        .replace("mesh.pmp", os.path.join(d, "mesh.pmp")))
    r = _cli(["input.pmp"], str(tmp_path))
    assert r.returncode == 1
    assert "wrong number of energy groups" in r.stdout
