"""The allocation-free power normalisation (pk_normalize_flux_static: reduction
scratch in a fixed device array) against pk_normalize_flux (library scratch),
through the C-ABI: two calls in a row, the second on a smaller mesh, each
giving the same bits as the existing entry point on the same input."""
import ctypes as ct
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _device_problem(pk, oracle_mod, d):
    case = oracle_mod.pmp.read_input(os.path.join(d, "input.pmp"))
    mesh = oracle_mod.build_mesh(case)
    xs = oracle_mod.xs_tables(case.materials, case.solver.G, True)
    fptr, fa, fc, fn, fb = pk.ragged_faces(mesh.nfaces, mesh.fmax, mesh.farea, mesh.fcen, mesh.fnorm,
                                           mesh.fnb)
    dm = pk.DeviceMesh(fptr, mesh.vol, mesh.cen, mesh.cmat, fa, fc, fn, fb)
    dx = pk.DeviceXS(xs.sigt, xs.sigs.reshape(len(xs.sigt), -1), xs.nusf, xs.kapf, xs.chi, xs.D)
    return dm, dx, mesh.C, case.solver.G


def _normalize(pk, fn, dm, dx, phi0, C, power, keff):
    L = pk.lib()
    f = getattr(L, fn)
    f.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_double, ct.c_void_p,
                  ct.c_void_p, ct.c_void_p]
    f.restype = ct.c_int
    phi, q, P = pk.Vec(len(phi0), phi0), pk.Vec(C), pk.Vec(C)
    tot = ct.c_double()
    pk.check(f(dm.h, dx.h, phi.h, power, keff, q.h, P.h, ct.byref(tot)))
    return phi.get(), q.get(), P.get(), tot.value


def test_static_scratch_normalisation_matches(pk, oracle_mod, tmp_path):
    from pampa_amd import workloads as W
    cases = [W.quarter_core_2d(str(tmp_path / "core"), seed=5, side=300),   # 180,000 rows: 704 workgroups
             W.slab_1d(str(tmp_path / "slab"))]                             # 400 rows: 2 workgroups
    rng = np.random.default_rng(17)
    for d in cases:
        dm, dx, C, G = _device_problem(pk, oracle_mod, d)
        phi0 = rng.uniform(0.1, 3.0, C * G)
        print(f"[case] {os.path.basename(d)} C={C} G={G}", flush=True)
        for power, keff in ((1.0, 1.02), (-1.0, 0.97)):
            a = _normalize(pk, "pk_normalize_flux", dm, dx, phi0, C, power, keff)
            b = _normalize(pk, "pk_normalize_flux_static", dm, dx, phi0, C, power, keff)
            for x, y in zip(a[:3], b[:3]):
                assert np.array_equal(x, y)
            assert a[3] == b[3]
            if power > 0:
                assert abs(b[3] - power) <= 1e-12 * power
            # the same call again reuses the scratch it left: same bits
            c = _normalize(pk, "pk_normalize_flux_static", dm, dx, phi0, C, power, keff)
            assert all(np.array_equal(x, y) for x, y in zip(b[:3], c[:3])) and b[3] == c[3]


def test_static_scratch_negative_flux_reported(pk, oracle_mod, tmp_path):
    from pampa_amd import workloads as W
    dm, dx, C, G = _device_problem(pk, oracle_mod, W.slab_1d(str(tmp_path / "slab")))
    phi0 = np.ones(C * G)
    phi0[7] = -1.0
    with pytest.raises(pk.PkError, match="negative"):
        _normalize(pk, "pk_normalize_flux_static", dm, dx, phi0, C, 1.0, 1.0)
    # the next call starts from cleared counters and succeeds
    a = _normalize(pk, "pk_normalize_flux", dm, dx, np.ones(C * G), C, 1.0, 1.0)
    b = _normalize(pk, "pk_normalize_flux_static", dm, dx, np.ones(C * G), C, 1.0, 1.0)
    assert all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3]))
