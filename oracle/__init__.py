"""CPU oracle for mkFit's Matriplex Kalman hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, as the checker and as the timed CPU baseline.  The product
package (mkfit_amd) never imports or links it.

The C++ restatement lives in oracle_core.h / oracle.cc (reference citations
there); this module is a ctypes wrapper over liborc_avx512.so / liborc_avx2.so.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

FIT_CHI2 = 1
FIT_CHARGE_FLIP = 2


class OrcCtx(ctypes.Structure):
    _fields_ = [("bfield", ctypes.c_float), ("rl_grid", ctypes.c_void_p), ("xi_grid", ctypes.c_void_p)]


class _Batch32(ctypes.Structure):
    _fields_ = [("n_tracks", ctypes.c_int), ("n_slots", ctypes.c_int),
                ("par", ctypes.c_void_p), ("err", ctypes.c_void_p), ("chg", ctypes.c_void_p),
                ("chi2", ctypes.c_void_p), ("hit_pos", ctypes.c_void_p), ("hit_err", ctypes.c_void_p),
                ("n_hits", ctypes.c_void_p), ("hit_flags", ctypes.c_void_p),
                ("pflags", ctypes.c_int), ("fit_flags", ctypes.c_int)]


_Batch64 = _Batch32  # same field layout (pointers differ only in pointee type)


def _cpu_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            return "avx512f" in f.read()
    except OSError:
        return False


def default_lib_path():
    name = "liborc_avx512.so" if _cpu_has_avx512() else "liborc_avx2.so"
    return os.path.join(_HERE, name)


def lib(path=None):
    """The oracle shared library (the ISA-matched build unless `path` is given)."""
    p = path or default_lib_path()
    if p not in _LIBS:
        if not os.path.exists(p):
            raise RuntimeError("%s is not built (run __graft_entry__.build())" % p)
        L = ctypes.CDLL(p)
        L.orc_simd_width.restype = ctypes.c_int
        _LIBS[p] = L
    return _LIBS[p]


def simd_width():
    return lib().orc_simd_width()


def isa_name():
    return "avx512" if simd_width() == 16 else "avx2"


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """Holds the B field and material grid (Config:: globals of the reference)."""

    def __init__(self, bfield=3.8112, rl_grid=None, xi_grid=None, lib_path=None):
        self.L = lib(lib_path)
        self._rl = None if rl_grid is None else np.ascontiguousarray(rl_grid, np.float32)
        self._xi = None if xi_grid is None else np.ascontiguousarray(xi_grid, np.float32)
        self.ctx = OrcCtx(float(bfield), _p(self._rl), _p(self._xi))

    def _dt(self, dtype):
        return ("f64", np.float64) if np.dtype(dtype) == np.float64 else ("f32", np.float32)

    def propagate_to_r(self, in_err, in_par, in_chg, ms_rad, pflags=0, no_mat=None, dtype=np.float32):
        sfx, T = self._dt(dtype)
        n = in_par.shape[-1]
        ie = np.ascontiguousarray(in_err, T); ip = np.ascontiguousarray(in_par, T)
        ch = np.ascontiguousarray(in_chg, np.int32); r = np.ascontiguousarray(ms_rad, T)
        oe = np.zeros((21, n), T); op = np.zeros((6, n), T); fl = np.zeros(n, np.int32)
        nm = None if no_mat is None else np.ascontiguousarray(no_mat, np.int32)
        getattr(self.L, "orc_propagate_to_r_" + sfx)(ctypes.byref(self.ctx), n, _p(ie), _p(ip), _p(ch), _p(r),
                                                    _p(oe), _p(op), int(pflags), _p(nm), _p(fl))
        return oe, op, fl

    def propagate_to_z(self, in_err, in_par, in_chg, ms_z, pflags=0, no_mat=None, dtype=np.float32):
        sfx, T = self._dt(dtype)
        n = in_par.shape[-1]
        ie = np.ascontiguousarray(in_err, T); ip = np.ascontiguousarray(in_par, T)
        ch = np.ascontiguousarray(in_chg, np.int32); z = np.ascontiguousarray(ms_z, T)
        oe = np.zeros((21, n), T); op = np.zeros((6, n), T)
        nm = None if no_mat is None else np.ascontiguousarray(no_mat, np.int32)
        getattr(self.L, "orc_propagate_to_z_" + sfx)(ctypes.byref(self.ctx), n, _p(ie), _p(ip), _p(ch), _p(z),
                                                    _p(oe), _p(op), int(pflags), _p(nm))
        return oe, op

    def helix_at_r(self, in_par, in_chg, ms_rad, pflags=0, dtype=np.float32):
        sfx, T = self._dt(dtype)
        n = in_par.shape[-1]
        ip = np.ascontiguousarray(in_par, T); ch = np.ascontiguousarray(in_chg, np.int32)
        r = np.ascontiguousarray(ms_rad, T)
        op = np.zeros((6, n), T); J = np.zeros((36, n), T); fl = np.zeros(n, np.int32)
        getattr(self.L, "orc_helix_at_r_" + sfx)(ctypes.byref(self.ctx), n, _p(ip), _p(ch), _p(r), _p(op), _p(J),
                                                _p(fl), int(pflags))
        return op, J.reshape(6, 6, n), fl

    def helix_at_z(self, in_par, in_chg, ms_z, pflags=0, dtype=np.float32):
        sfx, T = self._dt(dtype)
        n = in_par.shape[-1]
        ip = np.ascontiguousarray(in_par, T); ch = np.ascontiguousarray(in_chg, np.int32)
        z = np.ascontiguousarray(ms_z, T)
        op = np.zeros((6, n), T); J = np.zeros((36, n), T)
        getattr(self.L, "orc_helix_at_z_" + sfx)(ctypes.byref(self.ctx), n, _p(ip), _p(ch), _p(z), _p(op), _p(J),
                                                int(pflags))
        return op, J.reshape(6, 6, n)

    def kalman_operation(self, op, ps_err, ps_par, ms_err, ms_par, endcap=False, dtype=np.float32):
        sfx, T = self._dt(dtype)
        n = ps_par.shape[-1]
        pe = np.ascontiguousarray(ps_err, T); pp = np.ascontiguousarray(ps_par, T)
        me = np.ascontiguousarray(ms_err, T); mp = np.ascontiguousarray(ms_par, T)
        oe = np.zeros((21, n), T); opar = np.zeros((6, n), T); c2 = np.zeros(n, T)
        getattr(self.L, "orc_kalman_op_" + sfx)(int(op), int(bool(endcap)), n, _p(pe), _p(pp), _p(me), _p(mp),
                                               _p(oe), _p(opar), _p(c2))
        return oe, opar, c2

    def fit(self, ev, pflags=0, fit_flags=FIT_CHI2, dtype=np.float32, n_threads=1, with_flags=True):
        """Whole fit of a toy-MC event dict; returns (par, err, chg, chi2)."""
        sfx, T = self._dt(dtype)
        n = ev["seed_par"].shape[1]
        S = ev["hit_pos"].shape[0]
        par = np.array(ev["seed_par"], T); err = np.array(ev["seed_err"], T)
        chg = np.array(ev["chg"], np.int32); chi2 = np.zeros(n, T)
        hp = np.ascontiguousarray(ev["hit_pos"], np.float32); he = np.ascontiguousarray(ev["hit_err"], np.float32)
        nh = np.ascontiguousarray(ev["n_hits"], np.int32)
        hf = np.ascontiguousarray(ev["hit_flags"], np.uint8) if with_flags else None
        b = _Batch32(n, S, _p(par), _p(err), _p(chg), _p(chi2), _p(hp), _p(he), _p(nh), _p(hf), int(pflags),
                     int(fit_flags))
        rc = getattr(self.L, "orc_fit_" + sfx)(ctypes.byref(self.ctx), ctypes.byref(b), int(n_threads))
        if rc != 0:
            raise RuntimeError("oracle fit failed (%d)" % rc)
        return par, err, chg, chi2
