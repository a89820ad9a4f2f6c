"""S_N inputs outside the sweepable regime, through the drop-in: the
reference-form operator R and F (SNSolver.cxx:344-631) assembled on the device
bit-exact against the oracle, and the eigenpair against the oracle's direct
solve of that operator, for
  - the default face interpolation delta = 0.1 (SNSolver.hxx:16) and other
    delta < 1, LS off and on (24 x 24 cores: an exact LU of these operators is
    the oracle's eigen solve, and its fill grows fast with delta < 1),
  - the 2-D least-squares boundary correction with its flat-index reads
    (SNSolver.cxx:212-268) on the reference's cartesian-sn input, and that
    input with LS off,
  - reflective boundaries (SNSolver.cxx:528-557): the quarter core's -x/-y
    planes and a half domain reflected about x,
  - the 3-D least-squares correction on a hexagonal-prism lattice."""
import os
import shutil

import numpy as np
import pytest

from conftest import REF

pytestmark = pytest.mark.gpu

DK_PCM = 1.0
FLUX_L2 = 1e-5


def _cartesian_sn(tmp_path, name, delta, ls):
    """The reference's pwr-iaea-benchmark/cartesian-sn (3,856 cells, S2, 2 groups,
    reflective -x/-y, vacuum +x/+y) with its solver block rewritten."""
    src = os.path.join(REF, "pwr-iaea-benchmark")
    root = tmp_path / name
    shutil.copytree(os.path.join(src, "materials"), root / "materials")
    d = root / "case"
    d.mkdir()
    shutil.copy(os.path.join(src, "cartesian-sn", "mesh.pmp"), d / "mesh.pmp")
    text = open(os.path.join(src, "cartesian-sn", "input.pmp")).read()
    block = text[text.index("solver sn {"):]
    lines = ["solver sn {", "   energy-groups 2", "   order 2"]
    if delta is not None:
        lines.append(f"   mixed-face-interpolation {delta}")
    lines.append(f"   least-squares-boundary-interpolation {1 if ls else 0}")
    (d / "input.pmp").write_text(text.replace(block, "\n".join(lines) + "\n}\n"))
    return str(d)


def _square_sn(tmp_path, name, side, order, delta, ls):
    """A side x side quarter core of 1 cm cells (the 2-group materials and block
    map of the quarter-core workload; reflective -x/-y, vacuum +x/+y) solved
    with S_N of this order, face interpolation delta (None: the default) and
    the least-squares switch."""
    from pampa_amd import workloads as W
    d = W.quarter_core_2d(str(tmp_path / name), seed=3, side=side, G=2)
    p = os.path.join(d, "input.pmp")
    text = open(p).read()
    block = text[text.index("solver diffusion {"):]
    lines = ["solver sn {", "   energy-groups 2", f"   order {order}"]
    if delta is not None:
        lines.append(f"   mixed-face-interpolation {delta}")
    lines.append(f"   least-squares-boundary-interpolation {1 if ls else 0}")
    open(p, "w").write(text.replace(block, "\n".join(lines) + "\n}\n"))
    return d


def _hex(tmp_path, name, order, delta, ls):
    """Hexagonal-prism lattice (2 rings, 2 layers: 14,592 rows at S2), 2 groups,
    vacuum, face interpolation delta (None: the default) and the 3-D
    least-squares switch."""
    from pampa_amd import workloads as W
    d = W.hex_sn_3d(str(tmp_path / name), rings=2, layers=2, order=order, G=2)
    p = os.path.join(d, "input.pmp")
    text = open(p).read()
    if delta is None:
        text = text.replace("   mixed-face-interpolation 1.0\n", "")
    text = text.replace("least-squares-boundary-interpolation 0", f"least-squares-boundary-interpolation {ls}")
    open(p, "w").write(text)
    return d


def _mirror_half(tmp_path):
    import test_oracle
    return str(test_oracle._mirror_case(tmp_path, "half", "sn", True))


# With delta < 1 the eigenvector has negative angular fluxes (the downwind
# face interpolation oscillates), and the reference then stops with
# "negative values in the angular-flux solution" (SNSolver.cxx:329); those
# cases check the operator, keff, the eigenvector's shape and that error.
CASES = {
    "cartesian-sn": lambda t: os.path.join(REF, "pwr-iaea-benchmark", "cartesian-sn"),
    "cartesian-sn-ls-off": lambda t: _cartesian_sn(t, "lo", 1.0, False),
    "hex-s2-ls3d": lambda t: _hex(t, "h1", 2, 1.0, 1),
    "half-domain-reflective-s4": _mirror_half,
    "core-16-s2-delta-default": lambda t: _square_sn(t, "dd", 16, 2, None, False),
    "core-12-s4-delta-0.1-ls": lambda t: _square_sn(t, "dl", 12, 4, 0.1, True),
    "core-12-s4-delta-0.5-ls": lambda t: _square_sn(t, "dh", 12, 4, 0.5, True),
    "hex-s2-delta-default-ls3d": lambda t: _hex(t, "h2", 2, None, 1),
}


def _oracle(oracle_mod):
    case = oracle_mod.pmp.read_input("input.pmp")
    mesh = oracle_mod.build_mesh(case)
    s = case.solver
    xs = oracle_mod.xs_tables(case.materials, s.G, False)
    bct, _ = oracle_mod.bc_table(mesh, s.bcs)
    R, F, (dirs, wts) = oracle_mod.assemble_sn(mesh, xs, bct, s)
    k, x = oracle_mod.eigen_sinvert(R, F, permc="COLAMD" if s.delta < 1.0 else "MMD_AT_PLUS_A")
    try:
        phi, _, _, _ = oracle_mod.sn_normalize(mesh, xs, x, wts, s.power)
        err = None
    except oracle_mod.pmp.PmpError as e:
        phi, err = None, str(e)
    return R, F, k, x, phi, err


@pytest.mark.parametrize("name", list(CASES))
def test_sn_reference_form_operator_and_eigenpair(pk, oracle_mod, tmp_path, name):
    print(f"[case] {name}", flush=True)
    d = CASES[name](tmp_path)
    cwd = os.getcwd()
    os.chdir(d)
    try:
        R_o, F_o, k_o, x_o, phi_o, err_o = _oracle(oracle_mod)
        with pk.DropIn("input.pmp", "-silent") as run:
            failed = None
            try:
                run.solve()
            except pk.PkError as e:
                failed = str(e)
            k, _ = run.result()
            phi = run.field("scalar-flux")
            psi = run.field("angular-flux")
            R = run.operator("R")
            F = run.operator("F")
            st = run.stats()
    finally:
        os.chdir(cwd)
    for (p, c, v), O in ((R, R_o), (F, F_o)):
        assert np.array_equal(p, O.ptr) and np.array_equal(c, O.col)
        assert np.array_equal(v, O.val)
    dk = abs(k - k_o) * 1e5
    shape = lambda v: v / np.sum(v)  # noqa: E731
    l2x = np.linalg.norm(shape(psi) - shape(x_o)) / np.linalg.norm(shape(x_o))
    print(f"   rows {len(R[0]) - 1} keff {k:.8f} oracle {k_o:.8f} dk={dk:.2e} pcm eigenvector L2={l2x:.2e} "
          f"outer={st['outer']} inner={st['inner']} oracle error: {err_o}", flush=True)
    assert dk <= DK_PCM
    assert l2x <= FLUX_L2
    # the drop-in fails exactly when the reference's normalisation would
    assert (failed is not None) == (err_o is not None)
    if err_o is None:
        l2 = np.linalg.norm(phi - phi_o) / np.linalg.norm(phi_o)
        assert l2 <= FLUX_L2
