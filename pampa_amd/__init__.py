"""pampa_amd -- MI355X-native neutronics hot path of pampa.

The product is the in-tree shared library libpampa_amd.so (HIP kernels for
gfx950, the pk_* numerics C-ABI of include/pk.h, and the C++ host that keeps
pampa's parser, meshes, solvers and the pampa_* C-ABI of include/pampa.h).
This module binds that library with ctypes for Python callers (tests,
benchmarks).  It fails loudly when the library has not been built: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpampa_amd.so")
CLI_PATH = os.path.join(HERE, "bin", "pampa")

PK_INSERT_VALUES, PK_ADD_VALUES = 0, 1
PK_PREC_FP64, PK_PREC_MIXED = 0, 1
PK_BC_NONE, PK_BC_VACUUM, PK_BC_REFLECTIVE, PK_BC_ROBIN = 0, 1, 2, 3

_lib = None


class PkError(RuntimeError):
    pass


def lib():
    """Load libpampa_amd.so (raises if it is missing: build it first)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PkError(f"{LIB_PATH} is not built; run `python -m pampa_amd.build`")
        _lib = ct.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = ct.c_void_p
_PP = ct.POINTER(ct.c_void_p)
_I = ct.c_int
_L = ct.c_long
_D = ct.c_double


def _declare(L):
    sig = {
        "pk_initialize": [_I], "pk_finalize": [], "pk_synchronize": [],
        "pk_vec_create": [_L, _PP], "pk_vec_destroy": [_PP], "pk_vec_set": [_P, _D],
        "pk_vec_set_values": [_P, _P], "pk_vec_get_values": [_P, _P], "pk_vec_copy": [_P, _P],
        "pk_vec_scale": [_P, _D], "pk_vec_axpy": [_P, _D, _P], "pk_vec_dot": [_P, _P, _P],
        "pk_vec_norm2": [_P, _P], "pk_vec_device_ptr": [_P, _PP], "pk_vec_pointwise_mult": [_P, _P, _P],
        "pk_mat_create": [_L, _I, _PP], "pk_mat_set_values": [_P, _I, _P, _I, _P, _P, _I],
        "pk_mat_assemble": [_P], "pk_mat_destroy": [_PP], "pk_mat_get_size": [_P, _P, _P],
        "pk_mat_get_csr": [_P, _P, _P, _P], "pk_mat_mult": [_P, _P, _P],
        "pk_mat_set_block_size": [_P, _I], "pk_mat_time_mult": [_P, _P, _P, _I, _I, _P],
        "pk_mesh_create": [_I] + [_P] * 8 + [_PP], "pk_mesh_destroy": [_PP],
        "pk_xs_create": [_I, _I] + [_P] * 6 + [_PP], "pk_xs_destroy": [_PP],
        "pk_diffusion_assemble": [_P, _P, _I, _P, _P, _PP, _PP],
        "pk_normalize_flux": [_P, _P, _P, _D, _D, _P, _P, _P],
        "pk_eig_create": [_P, _P, _PP], "pk_eig_set_tolerances": [_P, _D, _I],
        "pk_eig_set_precision": [_P, _I], "pk_eig_set_inner": [_P, _I, _D, _I],
        "pk_eig_set_shift": [_P, _D], "pk_eig_set_acceleration": [_P, _I], "pk_eig_set_initial": [_P, _P], "pk_eig_solve": [_P],
        "pk_eig_get_eigenpair": [_P, _P, _P], "pk_eig_get_stats": [_P, _P, _P, _P, _P],
        "pk_eig_destroy": [_PP],
        "pk_lin_create": [_P, _PP], "pk_lin_set_tolerances": [_P, _D, _D, _I],
        "pk_lin_solve": [_P, _P, _P], "pk_lin_get_stats": [_P, _P, _P], "pk_lin_destroy": [_PP],
        "pk_batch_create": [_I, _P, _P, _PP], "pk_batch_set_tolerances": [_P, _D, _I],
        "pk_batch_set_precision": [_P, _I], "pk_batch_set_inner": [_P, _I, _D, _I],
        "pk_batch_set_shift": [_P, _D], "pk_batch_set_acceleration": [_P, _I], "pk_batch_set_initial": [_P, _I, _P], "pk_batch_solve": [_P],
        "pk_batch_get_eigenpair": [_P, _I, _P, _P], "pk_batch_get_stats": [_P, _I, _P, _P, _P, _P],
        "pk_batch_destroy": [_PP],
        "pk_tune_count": [_P], "pk_tune_get": [_I, _P, _P, _P, _P, _P, _P, _P], "pk_tune_reset": [],
        "pk_pool_create": [_L, _L, _L, _L, _I, _I, _I, _PP],
        "pk_pool_set_options": [_P, _D, _I, _D, _I, _D], "pk_pool_admit": [_P, _P, _P, _P, _P],
        "pk_pool_step": [_P, _I, _P, _P], "pk_pool_collect": [_P, _I, _P, _P],
        "pk_pool_result": [_P, _I, _P, _P, _P, _P, _P, _P], "pk_pool_release": [_P, _I],
        "pk_pool_capacity": [_P, _P, _P, _P, _P],
    }
    for name, args in sig.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _I
    _IP = ct.POINTER(ct.c_int)
    drop = {
        "pampa_initialize": [_I, ct.POINTER(ct.c_char_p), ct.POINTER(ct.POINTER(ct.c_double)), _IP, _IP],
        "pampa_initialize_steady_state": [_I, ct.POINTER(ct.c_char_p), _IP],
        "pampa_solve": [_I, _D, _D, _IP], "pampa_solve_steady_state": [_IP],
        "pampa_finalize": [ct.POINTER(ct.POINTER(ct.c_double)), _IP],
        "pampa_finalize_steady_state": [_IP],
        "pampa_get_field": [_P, ct.c_char_p, _IP], "pampa_set_field": [_P, ct.c_char_p, _IP],
        "pampa_ext_get_result": [_P, _P, _IP], "pampa_ext_get_operator": [ct.c_char_p, _P, _P, _P, _P, _P, _IP],
        "pampa_ext_get_stats": [_P, _P, _P, _P, _IP], "pampa_ext_get_sizes": [_P, _P, _P, _IP],
        "pampa_ext_time_kernel": [_I, _I, _P, _P, _IP], "pampa_ext_cold_start": [_IP],
        "pampa_ext_time_source": [_I, _I, _I, _P, _P, _IP],
        "pampa_ext_field_device": [ct.c_char_p, _PP, ct.POINTER(ct.c_long), _IP],
    }
    for name, args in {"pk_profile_enable": [_I], "pk_profile_reset": [],
                       "pk_profile_get": [ct.c_char_p, _P, _P, _P]}.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _I
    _ULLP = ct.POINTER(ct.c_ulonglong)
    drop.update({
        "pampa_serve_create": [_D, _I, _I, _PP, _IP], "pampa_serve_set_options": [_P, _D, _D, _D, _IP],
        "pampa_serve_add_case": [_P, ct.c_char_p, _IP, _P, _IP, _IP],
        "pampa_serve_case_tables": [_P, _I, ct.c_ulonglong, _D, _P, _P, _P, _P, _IP],
        "pampa_serve_footprint": [_P, _I, _P, _P, _IP],
        "pampa_serve_run": [_P, _I, _P, _P, _ULLP, _D, _I, _P, _P, _P, _P, _P, _P, _IP],
        "pampa_serve_forget_warm": [_P, _IP], "pampa_serve_set_progress": [_P, _D, _IP], "pampa_serve_set_sn_workers": [_P, _I, _IP], "pampa_serve_time_spmv": [_P, _I, _P, _P, _P, _IP], "pampa_serve_stats": [_P, _P, _P, _P, _P, _IP],
        "pampa_serve_destroy": [_PP, _IP],
    })
    for name, args in drop.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = None
    L.pk_sn_create.argtypes = [_P, _P, _I, _P, _P, _D, _I, _P, _PP]
    L.pk_sn_create.restype = _I
    for name, args in {"pk_sn_get_schedule": [_P, _P, _P], "pk_sn_set_precision": [_P, _I],
                       "pk_sn_set_tolerances": [_P, _D, _D, _I], "pk_sn_sweep_host": [_P, _P, _P, _P],
                       "pk_sn_solve": [_P, _P, _P, _P], "pk_sn_get_stats": [_P, _P, _P, _P],
                       "pk_sn_time_sweep": [_P, _I, _P, _P], "pk_sn_destroy": [_PP],
                       "pk_sn_set_acceleration": [_P, _I], "pk_sn_reset": [_P], "pk_sn_get_moment": [_P, _P],
                       "pk_sn_prepare": [_P], "pk_sn_set_contraction": [_P, _I],
                       "pk_sn_source_host": [_P, _P, _D, _P], "pk_sn_time_source": [_P, _I, _P, _P],
                       "pk_sn_set_xs": [_P, _P, _P, _P, _P, _P]}.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _I
    L.pk_stream_copy.argtypes = [_L, _I, _P]
    L.pk_stream_copy.restype = _I
    L.pk_pool_bytes.argtypes = [_L, _L, _L, _L, _I, _I, _I]
    L.pk_pool_bytes.restype = _D
    L.pk_last_error.restype = ct.c_char_p
    L.pk_last_error.argtypes = []
    L.pk_vec_size.restype = _L
    L.pk_vec_size.argtypes = [_P]


def check(rc: int):
    if rc != 0:
        raise PkError(lib().pk_last_error().decode())


def ptr(a: np.ndarray):
    return a.ctypes.data_as(ct.c_void_p)


class Handle:
    """Owns one pk_* handle and destroys it with the matching function."""

    def __init__(self, kind: str, h: ct.c_void_p):
        self.kind = kind
        self.h = h

    @property
    def value(self):
        return self.h

    def close(self):
        if self.h is not None and self.h.value and _lib is not None:
            getattr(_lib, f"pk_{self.kind}_destroy")(ct.byref(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _new(kind, fn, *args):
    h = ct.c_void_p()
    check(fn(*args, ct.byref(h)))
    return Handle(kind, h)


def initialize(device: int = 0):
    check(lib().pk_initialize(device))


def profile(on: bool):
    check(lib().pk_profile_enable(int(on)))
    if on:
        check(lib().pk_profile_reset())


def stream_copy(n_floats: int, reps: int = 20) -> float:
    """Average seconds of one STREAM-style copy of n_floats floats (8 n bytes)."""
    t = ct.c_double()
    check(lib().pk_stream_copy(n_floats, reps, ct.byref(t)))
    return t.value


def profile_get(name: str):
    n, s, b = ct.c_long(), ct.c_double(), ct.c_double()
    check(lib().pk_profile_get(name.encode(), ct.byref(n), ct.byref(s), ct.byref(b)))
    return n.value, s.value, b.value


class Vec(Handle):
    def __init__(self, n: int, values=None):
        L = lib()
        h = ct.c_void_p()
        check(L.pk_vec_create(n, ct.byref(h)))
        super().__init__("vec", h)
        self.n = n
        if values is not None:
            self.set(values)

    def set(self, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        assert a.size == self.n
        check(lib().pk_vec_set_values(self.h, ptr(a)))

    def get(self) -> np.ndarray:
        a = np.empty(self.n)
        check(lib().pk_vec_get_values(self.h, ptr(a)))
        return a


class Mat(Handle):
    def __init__(self, h: ct.c_void_p):
        super().__init__("mat", h)

    @classmethod
    def from_csr(cls, n, row_ptr, col, val, block_size=1):
        L = lib()
        h = ct.c_void_p()
        check(L.pk_mat_create(n, 0, ct.byref(h)))
        m = cls(h)
        row_ptr = np.asarray(row_ptr)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        for r in range(n):
            a, b = int(row_ptr[r]), int(row_ptr[r + 1])
            if b > a:
                rows = np.array([r], np.int32)
                check(L.pk_mat_set_values(h, 1, ptr(rows), b - a, ptr(col[a:b]), ptr(val[a:b]),
                                          PK_INSERT_VALUES))
        check(L.pk_mat_assemble(h))
        if block_size > 1:
            check(L.pk_mat_set_block_size(h, block_size))
        return m

    def size(self):
        n, nnz = ct.c_long(), ct.c_long()
        check(lib().pk_mat_get_size(self.h, ct.byref(n), ct.byref(nnz)))
        return n.value, nnz.value

    def csr(self):
        n, nnz = self.size()
        p = np.zeros(n + 1, np.int32)
        c = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1))
        check(lib().pk_mat_get_csr(self.h, ptr(p), ptr(c), ptr(v)))
        return p, c[:nnz], v[:nnz]

    def mult(self, x: Vec, y: Vec):
        check(lib().pk_mat_mult(self.h, x.h, y.h))

    def time_mult(self, x: Vec, y: Vec, reps=100, fp32=False) -> float:
        t = ct.c_double()
        check(lib().pk_mat_time_mult(self.h, x.h, y.h, reps, int(fp32), ct.byref(t)))
        return t.value


class DeviceMesh(Handle):
    """pk_mesh from ragged face arrays (face_ptr/face data)."""

    def __init__(self, face_ptr, vol, cen, mat, farea, fcen, fnorm, fnb):
        arrs = [np.ascontiguousarray(face_ptr, np.int32), np.ascontiguousarray(vol, np.float64),
                np.ascontiguousarray(cen, np.float64), np.ascontiguousarray(mat, np.int32),
                np.ascontiguousarray(farea, np.float64), np.ascontiguousarray(fcen, np.float64),
                np.ascontiguousarray(fnorm, np.float64), np.ascontiguousarray(fnb, np.int32)]
        self._keep = arrs
        h = ct.c_void_p()
        check(lib().pk_mesh_create(len(arrs[1]), *[ptr(a) for a in arrs], ct.byref(h)))
        super().__init__("mesh", h)
        self.C = len(arrs[1])


class DeviceXS(Handle):
    def __init__(self, sigt, sigs, nusf, kapf, chi, D=None):
        sigt = np.ascontiguousarray(sigt, np.float64)
        nm, G = sigt.shape
        arrs = [sigt, np.ascontiguousarray(sigs, np.float64), np.ascontiguousarray(nusf, np.float64),
                np.ascontiguousarray(kapf, np.float64), np.ascontiguousarray(chi, np.float64)]
        Dp = None if D is None else np.ascontiguousarray(D, np.float64)
        self._keep = arrs + [Dp]
        h = ct.c_void_p()
        check(lib().pk_xs_create(nm, G, *[ptr(a) for a in arrs], None if Dp is None else ptr(Dp),
                                 ct.byref(h)))
        super().__init__("xs", h)
        self.G = G


def diffusion_assemble(mesh: DeviceMesh, xs: DeviceXS, bc_type, bc_value):
    bt = np.ascontiguousarray(bc_type, np.int32)
    bv = np.ascontiguousarray(bc_value, np.float64)
    R, F = ct.c_void_p(), ct.c_void_p()
    check(lib().pk_diffusion_assemble(mesh.h, xs.h, len(bt), ptr(bt), ptr(bv), ct.byref(R),
                                      ct.byref(F)))
    return Mat(R), Mat(F)


class Eigen(Handle):
    def __init__(self, R: Mat, F: Mat, tol=1e-8, precision=PK_PREC_FP64, shift=0.1, restart=60,
                 reduction=0.1, max_inner=2000, max_outer=20000, acceleration=3):
        L = lib()
        h = ct.c_void_p()
        check(L.pk_eig_create(R.h, F.h, ct.byref(h)))
        super().__init__("eig", h)
        self._mats = (R, F)
        check(L.pk_eig_set_tolerances(h, tol, max_outer))
        check(L.pk_eig_set_precision(h, precision))
        check(L.pk_eig_set_inner(h, restart, reduction, max_inner))
        check(L.pk_eig_set_shift(h, shift))
        check(L.pk_eig_set_acceleration(h, acceleration))

    def solve(self, x0: Vec | None = None):
        L = lib()
        if x0 is not None:
            check(L.pk_eig_set_initial(self.h, x0.h))
        check(L.pk_eig_solve(self.h))

    def eigenpair(self, n):
        lam = ct.c_double()
        x = Vec(n)
        check(lib().pk_eig_get_eigenpair(self.h, ct.byref(lam), x.h))
        return 1.0 / lam.value, x

    def stats(self):
        o, i, r, s = ct.c_int(), ct.c_long(), ct.c_double(), ct.c_double()
        check(lib().pk_eig_get_stats(self.h, ct.byref(o), ct.byref(i), ct.byref(r), ct.byref(s)))
        return dict(outer=o.value, inner=i.value, residual=r.value, seconds=s.value)


class Batch(Handle):
    """Batched eigen solver over pairs (R_b, F_b): one launch per step for all."""

    def __init__(self, Rs, Fs, tol=1e-8, precision=PK_PREC_FP64, shift=0.1, restart=60,
                 reduction=0.1, max_inner=2000, max_outer=20000, acceleration=3):
        L = lib()
        n = len(Rs)
        assert n == len(Fs) and n > 0
        ra = (ct.c_void_p * n)(*[r.h.value for r in Rs])
        fa = (ct.c_void_p * n)(*[f.h.value for f in Fs])
        h = ct.c_void_p()
        check(L.pk_batch_create(n, ra, fa, ct.byref(h)))
        super().__init__("batch", h)
        self._mats = (list(Rs), list(Fs))
        self.sizes = [r.size()[0] for r in Rs]
        check(L.pk_batch_set_tolerances(h, tol, max_outer))
        check(L.pk_batch_set_precision(h, precision))
        check(L.pk_batch_set_inner(h, restart, reduction, max_inner))
        check(L.pk_batch_set_shift(h, shift))
        check(L.pk_batch_set_acceleration(h, acceleration))

    def set_initial(self, index: int, x: "Vec"):
        check(lib().pk_batch_set_initial(self.h, index, x.h))

    def solve(self):
        check(lib().pk_batch_solve(self.h))

    def eigenpair(self, index: int):
        lam = ct.c_double()
        x = Vec(self.sizes[index])
        check(lib().pk_batch_get_eigenpair(self.h, index, ct.byref(lam), x.h))
        return 1.0 / lam.value, x

    def stats(self, index: int):
        o, i, r, s = ct.c_int(), ct.c_long(), ct.c_double(), ct.c_double()
        check(lib().pk_batch_get_stats(self.h, index, ct.byref(o), ct.byref(i), ct.byref(r), ct.byref(s)))
        return dict(outer=o.value, inner=i.value, residual=r.value, seconds=s.value)


PK_POOL_FULL = 2


def tune_choices():
    """The autotuner's choices so far, one dict per size class in the order
    tuned: rows, cell block, nonzeros, precision, slice height, SpMV workgroup
    and the timings (us) of every (height, workgroup) candidate."""
    L = lib()
    n = ct.c_int()
    check(L.pk_tune_count(ct.byref(n)))
    out = []
    for i in range(n.value):
        rows, nnz = ct.c_long(), ct.c_long()
        bs, prec, h, blk = ct.c_int(), ct.c_int(), ct.c_int(), ct.c_int()
        us = np.zeros(12)
        check(L.pk_tune_get(i, ct.byref(rows), ct.byref(bs), ct.byref(nnz), ct.byref(prec), ct.byref(h),
                            ct.byref(blk), ptr(us)))
        out.append({"rows": rows.value, "block": bs.value, "nnz": nnz.value, "precision": prec.value,
                    "slice_height": h.value, "spmv_block": blk.value,
                    "us": {f"{hh}x{bb}": float(us[a * 4 + q]) for a, hh in enumerate((32, 64, 128))
                           for q, bb in enumerate((128, 256, 512, 1024))}})
    return out


def tune_reset():
    check(lib().pk_tune_reset())


class Pool(Handle):
    """Eigen pool for continuous batching: admit (R, F[, x0]) problems while
    others iterate, step, collect finished slots, release them."""

    def __init__(self, max_rows, max_entries, max_fentries, max_binv, max_slots,
                 precision=PK_PREC_FP64, restart=60):
        h = ct.c_void_p()
        check(lib().pk_pool_create(max_rows, max_entries, max_fentries, max_binv, max_slots, precision,
                                   restart, ct.byref(h)))
        super().__init__("batch", h)  # a pool is a pk_batch handle
        self.sizes = {}

    @staticmethod
    def bytes(max_rows, max_entries, max_fentries, max_binv, max_slots, precision=PK_PREC_FP64,
              restart=60):
        return lib().pk_pool_bytes(max_rows, max_entries, max_fentries, max_binv, max_slots, precision,
                                   restart)

    def options(self, tol=1e-8, max_outer=10000, reduction=0.1, max_inner=2000, shift=0.1):
        check(lib().pk_pool_set_options(self.h, tol, max_outer, reduction, max_inner, shift))

    def admit(self, R: "Mat", F: "Mat", x0: "Vec | None" = None):
        """Slot index, or None when the pool has no room now."""
        s = ct.c_int(-1)
        rc = lib().pk_pool_admit(self.h, R.h, F.h, x0.h if x0 is not None else None, ct.byref(s))
        if rc == PK_POOL_FULL:
            return None
        check(rc)
        self.sizes[s.value] = R.size()[0]
        return s.value

    def step(self, ticks=1):
        live, done = ct.c_int(), ct.c_int()
        check(lib().pk_pool_step(self.h, ticks, ct.byref(live), ct.byref(done)))
        return live.value, done.value

    def collect(self, max_n=4096):
        arr = (ct.c_int * max_n)()
        n = ct.c_int()
        check(lib().pk_pool_collect(self.h, max_n, arr, ct.byref(n)))
        return [arr[i] for i in range(n.value)]

    def result(self, slot, with_vector=True):
        k, o, i, r, c = ct.c_double(), ct.c_int(), ct.c_long(), ct.c_double(), ct.c_int()
        x = Vec(self.sizes[slot]) if with_vector else None
        check(lib().pk_pool_result(self.h, slot, ct.byref(k), x.h if x is not None else None, ct.byref(o),
                                   ct.byref(i), ct.byref(r), ct.byref(c)))
        return dict(keff=k.value, x=x, outer=o.value, inner=i.value, residual=r.value,
                    converged=bool(c.value))

    def release(self, slot):
        check(lib().pk_pool_release(self.h, slot))
        self.sizes.pop(slot, None)

    def capacity(self):
        fr, fe, fs, t = ct.c_long(), ct.c_long(), ct.c_int(), ct.c_long()
        check(lib().pk_pool_capacity(self.h, ct.byref(fr), ct.byref(fe), ct.byref(fs), ct.byref(t)))
        return dict(free_rows=fr.value, free_entries=fe.value, free_slots=fs.value, ticks=t.value)


class Server:
    """Serving over the pampa_serve_* C-ABI: resident base cases, requests
    continuously batched into one device eigen pool."""

    def __init__(self, budget_bytes=0.0, max_slots=1024, precision=PK_PREC_MIXED):
        self.L = lib()
        self.h = ct.c_void_p()
        err = ct.c_int(0)
        self.L.pampa_serve_create(budget_bytes, max_slots, precision, ct.byref(self.h), ct.byref(err))
        if err.value:
            raise PkError("pampa_serve_create failed")
        self.rows = []

    def _call(self, name, *args):
        err = ct.c_int(0)
        getattr(self.L, name)(*args, ct.byref(err))
        if err.value:
            raise PkError(f"{name} failed")

    def options(self, tol=1e-8, reduction=0.1, shift=0.1):
        self._call("pampa_serve_set_options", self.h, tol, reduction, shift)

    def add_case(self, input_path: str):
        cid, rows, groups = ct.c_int(), ct.c_long(), ct.c_int()
        self._call("pampa_serve_add_case", self.h, input_path.encode(), ct.byref(cid), ct.byref(rows),
                   ct.byref(groups))
        self.rows.append(rows.value)
        return cid.value, rows.value, groups.value

    def case_tables(self, case_id, nm, G, seed, rel):
        st, ss = np.zeros((nm, G)), np.zeros((nm, G, G))
        D, nf = np.zeros((nm, G)), np.zeros((nm, G))
        self._call("pampa_serve_case_tables", self.h, case_id, seed, rel, ptr(st), ptr(ss), ptr(D), ptr(nf))
        return st, ss, D, nf

    def footprint(self, case_id):
        b, pb = ct.c_double(), ct.c_double()
        self._call("pampa_serve_footprint", self.h, case_id, ct.byref(b), ct.byref(pb))
        return b.value, pb.value

    def run(self, arrival, case_id, seed, rel=0.02, warm=True, keep_flux=None):
        """Open-loop stream; returns per-request arrays (keff, latency, outer,
        inner, converged) and the fluxes of the requests in keep_flux."""
        n = len(arrival)
        arr = np.ascontiguousarray(arrival, dtype=np.float64)
        cid = np.ascontiguousarray(case_id, dtype=np.int32)
        sd = np.ascontiguousarray(seed, dtype=np.uint64)
        keff, lat = np.zeros(n), np.zeros(n)
        outer, conv = np.zeros(n, np.int32), np.zeros(n, np.int32)
        inner = np.zeros(n, np.int64)
        fl = {}
        fp = (ct.c_void_p * max(n, 1))()
        for r in (keep_flux or []):
            fl[r] = np.zeros(self.rows[cid[r]])
            fp[r] = fl[r].ctypes.data
        self._call("pampa_serve_run", self.h, n, ptr(arr), ptr(cid),
                   sd.ctypes.data_as(ct.POINTER(ct.c_ulonglong)), rel, int(warm), ptr(keff), ptr(lat),
                   ptr(outer), ptr(inner), ptr(conv), fp)
        return dict(keff=keff, latency=lat, outer=outer, inner=inner, converged=conv.astype(bool), flux=fl)

    def time_spmv(self, reps=20):
        """(seconds per launch, algorithmic bytes, rows) of one SpMV over every case."""
        t, b, r = ct.c_double(), ct.c_double(), ct.c_double()
        self._call("pampa_serve_time_spmv", self.h, reps, ct.byref(t), ct.byref(b), ct.byref(r))
        return t.value, b.value, r.value

    def progress(self, seconds: float):
        """Print a progress line on stderr every `seconds` during run()."""
        self._call("pampa_serve_set_progress", self.h, seconds)

    def sn_workers(self, n: int):
        """Workers (sweep handles) per S_N case added after this call."""
        self._call("pampa_serve_set_sn_workers", self.h, int(n))

    def forget_warm(self):
        self._call("pampa_serve_forget_warm", self.h)

    def stats(self):
        s, t, m, pb = ct.c_double(), ct.c_long(), ct.c_int(), ct.c_double()
        self._call("pampa_serve_stats", self.h, ct.byref(s), ct.byref(t), ct.byref(m), ct.byref(pb))
        return dict(seconds=s.value, ticks=t.value, max_live=m.value, pool_bytes=pb.value)

    def close(self):
        if self.h is not None and self.h.value:
            err = ct.c_int(0)
            self.L.pampa_serve_destroy(ct.byref(self.h), ct.byref(err))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DropIn:
    """In-process use of the drop-in C-ABI (pampa_* entry points): one
    calculation at a time per process, as in the reference."""

    def __init__(self, input_path: str, *options: str):
        self.L = lib()
        args = [b"pampa", input_path.encode()] + [o.encode() for o in options]
        self._argv = (ct.c_char_p * len(args))(*args)
        self._dt = ct.POINTER(ct.c_double)()
        self.ndt = ct.c_int(0)
        err = ct.c_int(0)
        self.L.pampa_initialize(len(args), self._argv, ct.byref(self._dt), ct.byref(self.ndt),
                                ct.byref(err))
        if err.value:
            raise PkError("pampa_initialize failed")
        self.open = True

    def dt(self):
        return [self._dt[i] for i in range(self.ndt.value)]

    def solve(self, n=0, dt=0.0, t=0.0):
        err = ct.c_int(0)
        self.L.pampa_solve(n, dt, t, ct.byref(err))
        if err.value:
            raise PkError("pampa_solve failed")

    def sizes(self):
        c, g, m, err = ct.c_int(), ct.c_int(), ct.c_int(), ct.c_int()
        self.L.pampa_ext_get_sizes(ct.byref(c), ct.byref(g), ct.byref(m), ct.byref(err))
        return c.value, g.value, m.value

    def field(self, name: str) -> np.ndarray:
        C, G, M = self.sizes()
        n = {"scalar-flux": C * G, "angular-flux": C * G * M}.get(name, C)
        a = np.empty(n)
        err = ct.c_int(0)
        self.L.pampa_get_field(ptr(a), name.encode(), ct.byref(err))
        if err.value:
            raise PkError(f"pampa_get_field({name}) failed")
        return a

    def set_field(self, name: str, values):
        a = np.ascontiguousarray(values, np.float64)
        err = ct.c_int(0)
        self.L.pampa_set_field(ptr(a), name.encode(), ct.byref(err))
        if err.value:
            raise PkError(f"pampa_set_field({name}) failed")

    def result(self):
        k, p, err = ct.c_double(), ct.c_double(), ct.c_int()
        self.L.pampa_ext_get_result(ct.byref(k), ct.byref(p), ct.byref(err))
        return k.value, p.value

    def stats(self):
        o, i, r, s, err = ct.c_int(), ct.c_long(), ct.c_double(), ct.c_double(), ct.c_int()
        self.L.pampa_ext_get_stats(ct.byref(o), ct.byref(i), ct.byref(r), ct.byref(s), ct.byref(err))
        return dict(outer=o.value, inner=i.value, residual=r.value, seconds=s.value)

    def field_device(self, name: str):
        """(device address, length) of a field (pampa_ext_field_device)."""
        dev, n, err = ct.c_void_p(), ct.c_long(), ct.c_int()
        self.L.pampa_ext_field_device(name.encode(), ct.byref(dev), ct.byref(n), ct.byref(err))
        if err.value:
            raise PkError(f"pampa_ext_field_device({name}) failed")
        return dev.value, n.value

    def cold_start(self):
        """The next solve starts cold (pampa_ext_cold_start)."""
        err = ct.c_int()
        self.L.pampa_ext_cold_start(ct.byref(err))
        if err.value:
            raise PkError("pampa_ext_cold_start failed")

    def time_kernel(self, reps=50, fp32=False):
        s, b, err = ct.c_double(), ct.c_double(), ct.c_int()
        self.L.pampa_ext_time_kernel(reps, int(fp32), ct.byref(s), ct.byref(b), ct.byref(err))
        if err.value:
            raise PkError("pampa_ext_time_kernel failed")
        return s.value, b.value

    def time_source(self, reps=20, fp32=False, mode=1):
        """(seconds, algorithmic bytes) of the S_N source kernel; mode 1
        vector kernel, 2 matrix cores (16 groups)."""
        s, b, err = ct.c_double(), ct.c_double(), ct.c_int()
        self.L.pampa_ext_time_source(reps, int(fp32), int(mode), ct.byref(s), ct.byref(b), ct.byref(err))
        if err.value:
            raise PkError("pampa_ext_time_source failed")
        return s.value, b.value

    def operator(self, name: str):
        n, nnz, err = ct.c_long(), ct.c_long(), ct.c_int()
        self.L.pampa_ext_get_operator(name.encode(), ct.byref(n), ct.byref(nnz), None, None, None,
                                      ct.byref(err))
        if err.value:
            raise PkError(f"operator {name} not available")
        if name == "b":
            v = np.zeros(n.value)
            self.L.pampa_ext_get_operator(b"b", ct.byref(n), ct.byref(nnz), None, None, ptr(v), ct.byref(err))
            if err.value:
                raise PkError("right-hand side not available")
            return v
        p = np.zeros(n.value + 1, np.int32)
        c = np.zeros(max(nnz.value, 1), np.int32)
        v = np.zeros(max(nnz.value, 1))
        self.L.pampa_ext_get_operator(name.encode(), ct.byref(n), ct.byref(nnz), ptr(p), ptr(c), ptr(v),
                                      ct.byref(err))
        return p, c[:nnz.value], v[:nnz.value]

    def close(self):
        if self.open:
            err = ct.c_int(0)
            self.L.pampa_finalize(ct.byref(self._dt), ct.byref(err))
            self.open = False

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def ragged_faces(nfaces, fmax, *padded):
    """Padded [C][fmax] face arrays -> ragged (face_ptr, arrays...)."""
    nfaces = np.asarray(nfaces)
    C = len(nfaces)
    fptr = np.zeros(C + 1, np.int32)
    fptr[1:] = np.cumsum(nfaces)
    sel = np.concatenate([np.arange(i * fmax, i * fmax + nfaces[i]) for i in range(C)]) if C else np.zeros(0, int)
    out = [fptr]
    for a in padded:
        a = np.asarray(a)
        flat = a.reshape(C * fmax, -1) if a.ndim == 3 else a.reshape(C * fmax)
        out.append(np.ascontiguousarray(flat[sel]))
    return out


class Sweep(Handle):
    """pk_sn handle (matrix-free S_N operator)."""

    def __init__(self, mesh: DeviceMesh, xs: DeviceXS, dirs, wts, bc_type, precision=PK_PREC_FP64):
        d = np.ascontiguousarray(dirs, np.float64)
        w = np.ascontiguousarray(wts, np.float64)
        bt = np.ascontiguousarray(bc_type, np.int32)
        self._keep = (mesh, xs, d, w, bt)
        h = ct.c_void_p()
        check(lib().pk_sn_create(mesh.h, xs.h, len(w), ptr(d), ptr(w), 1.0, len(bt), ptr(bt), ct.byref(h)))
        super().__init__("sn", h)
        self.C, self.G, self.M = mesh.C, xs.G, len(w)
        check(lib().pk_sn_set_precision(h, precision))

    def schedule(self):
        nl = np.zeros(self.M, np.int32)
        lv = np.zeros((self.M, self.C), np.int32)
        check(lib().pk_sn_get_schedule(self.h, ptr(nl), ptr(lv)))
        return nl, lv

    def sweep(self, q):
        q = np.ascontiguousarray(q, np.float64)
        psi = np.zeros(self.C * self.G * self.M)
        phiw = np.zeros(self.C * self.G)
        check(lib().pk_sn_sweep_host(self.h, ptr(q), ptr(psi), ptr(phiw)))
        return psi, phiw

    def time_sweep(self, reps=10):
        t, b = ct.c_double(), ct.c_double()
        check(lib().pk_sn_time_sweep(self.h, reps, ct.byref(t), ct.byref(b)))
        return t.value, b.value

    def set_contraction(self, mode):
        """Source contraction: 0 by group count (matrix cores at 16 groups),
        1 vector kernel, 2 matrix cores (16 groups only)."""
        check(lib().pk_sn_set_contraction(self.h, int(mode)))

    def source(self, phiw, keff):
        """The scattering + fission source q[C*G] (volume included) of a
        moment phiw[C*G] at k, in FP64 through the handle's contraction."""
        phiw = np.ascontiguousarray(phiw, np.float64)
        q = np.zeros(self.C * self.G)
        check(lib().pk_sn_source_host(self.h, ptr(phiw), float(keff), ptr(q)))
        return q

    def time_source(self, reps=20):
        """(seconds per source kernel, algorithmic bytes per call)."""
        t, b = ct.c_double(), ct.c_double()
        check(lib().pk_sn_time_source(self.h, reps, ct.byref(t), ct.byref(b)))
        return t.value, b.value

    def solve(self, accel=5, tol_k=1e-9, tol_phi=1e-8, max_outer=100000, cold=True):
        """k-eigenvalue solve (source iteration, Anderson depth `accel`):
        (keff, moment sum_m w_m psi [C*G], outer iterations, sweeps)."""
        L = lib()
        check(L.pk_sn_set_acceleration(self.h, int(accel)))
        check(L.pk_sn_set_tolerances(self.h, tol_k, tol_phi, max_outer))
        if cold:
            check(L.pk_sn_reset(self.h))
        k = ct.c_double()
        check(L.pk_sn_solve(self.h, ct.byref(k), None, None))
        phiw = np.zeros(self.C * self.G)
        check(L.pk_sn_get_moment(self.h, ptr(phiw)))
        o, sw = ct.c_int(), ct.c_long()
        check(L.pk_sn_get_stats(self.h, ct.byref(o), ct.byref(sw), None))
        return k.value, phiw, o.value, sw.value
