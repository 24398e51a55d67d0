"""ctypes binding of the C oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.  The product (mkpp_amd) never does.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libmkpp_oracle.so")
TABLES = os.path.join(HERE, "tables")
METHODS = {"ros2": 0, "ros3": 1, "ros4": 2, "rodas3": 3, "rodas4": 4}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, dp, lp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_long)
        L.oracle_load.restype = vp
        L.oracle_load.argtypes = [C.c_char_p]
        L.oracle_free.argtypes = [vp]
        for nm in ("oracle_n", "oracle_nphoto", "oracle_nj", "oracle_nw", "oracle_nlu", "oracle_ngroups"):
            getattr(L, nm).argtypes = [vp]
            getattr(L, nm).restype = C.c_int
        L.oracle_rates.argtypes = [vp, C.c_double, C.c_double, dp, dp, dp, C.c_double]
        L.oracle_jacobian.argtypes = [vp, C.c_double, C.c_double, dp, dp, dp, C.c_double]
        L.oracle_rate_constants.argtypes = [vp, C.c_double, C.c_double, dp, dp, dp, dp, C.c_double]
        L.oracle_factorize.argtypes = [vp, dp, dp]
        L.oracle_solve.argtypes = [vp, dp, dp, dp, dp]
        L.oracle_integrate.restype = C.c_int
        L.oracle_integrate.argtypes = [vp, C.c_int, C.c_long, dp, dp, dp, dp, C.c_double, C.c_int, C.c_double,
                                       dp, dp, lp, lp, C.c_int, C.c_long, C.c_int, dp]
        L.oracle_integrate_trace.restype = C.c_long
        L.oracle_integrate_trace.argtypes = [vp, C.c_int, dp, C.c_double, C.c_double, dp, C.c_double, C.c_double,
                                             dp, dp, C.c_long, dp, C.c_long, lp, C.c_double]
        L.oracle_tableau.restype = C.c_int
        L.oracle_tableau.argtypes = [C.c_int, dp]
        L.oracle_integrate_reduced.restype = C.c_int
        L.oracle_integrate_reduced.argtypes = [vp, C.c_int, C.c_long, dp, dp, dp, dp, C.c_double, C.c_int,
                                               C.c_double, dp, dp, lp, lp, C.c_int, C.c_long, C.POINTER(C.c_int),
                                               C.c_int, dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _cells(v, ncell):
    return None if v is None else np.ascontiguousarray(np.broadcast_to(v, (ncell,)), dtype=np.float64)


def _l(a):
    return a.ctypes.data_as(C.POINTER(C.c_long)) if a is not None else None


class Oracle:
    """One mechanism's CPU reference integrator."""

    def __init__(self, mech: str, path: str | None = None):
        self.name = mech
        path = path or os.path.join(TABLES, f"{mech}.tab")
        self._h = lib().oracle_load(path.encode())
        if not self._h:
            raise RuntimeError(f"oracle: cannot load {path}")
        L = lib()
        self.n = L.oracle_n(self._h)
        self.nphoto = L.oracle_nphoto(self._h)
        self.nj = L.oracle_nj(self._h)
        self.nw = L.oracle_nw(self._h)
        self.nlu = L.oracle_nlu(self._h)
        self.ngroups = L.oracle_ngroups(self._h)

    def __del__(self):
        try:
            if self._h:
                lib().oracle_free(self._h)
        except Exception:
            pass

    def rates(self, y, jvals=(), T=300.0, air_density=2.4476e19, clw=0.0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        jv = np.ascontiguousarray(np.asarray(jvals, dtype=np.float64).reshape(-1) if len(jvals) else np.zeros(1))
        f = np.zeros(self.n)
        lib().oracle_rates(self._h, T, air_density, _d(y), _d(jv), _d(f), float(clw))
        return f

    def jacobian_slots(self, y, jvals=(), T=300.0, air_density=2.4476e19, clw=0.0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        jv = np.ascontiguousarray(np.asarray(jvals, dtype=np.float64).reshape(-1) if len(jvals) else np.zeros(1))
        out = np.zeros(self.nj)
        lib().oracle_jacobian(self._h, T, air_density, _d(y), _d(jv), _d(out), float(clw))
        return out

    def rate_constants(self, y, jvals=(), T=300.0, air_density=2.4476e19, clw=0.0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        jv = np.ascontiguousarray(np.asarray(jvals, dtype=np.float64).reshape(-1) if len(jvals) else np.zeros(1))
        k = np.zeros(max(self.ngroups, 1))
        dk = np.zeros(max(self.ngroups, 1))
        lib().oracle_rate_constants(self._h, T, air_density, _d(y), _d(jv), _d(k), _d(dk), float(clw))
        return k[: self.ngroups], dk[: self.ngroups]

    def factorize(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        lu = np.zeros(self.nlu)
        lib().oracle_factorize(self._h, _d(w), _d(lu))
        return lu

    def solve(self, lu, rhs):
        lu = np.ascontiguousarray(lu, dtype=np.float64)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros(self.n)
        work = np.zeros(self.n)
        lib().oracle_solve(self._h, _d(lu), _d(rhs), _d(x), _d(work))
        return x

    def integrate(self, conc, dt, steps=1, method="rodas3", temp=None, air_density=None, jvals=None,
                  atol=None, rtol=None, seed=1.0e-6, hoist=True, max_attempts=0, threads=0, cloud_water=None):
        """Integrate cells in place.  conc: (num_species, num_cells) float64,
        i.e. the LayoutLeft buffer conc[cell + num_cells*species]; cloud_water:
        per-cell cloud liquid water (None: 0).
        Returns (accepted, attempted, stalled_cells)."""
        conc = np.asarray(conc)
        assert conc.dtype == np.float64 and conc.flags.c_contiguous and conc.shape[0] == self.n
        ncell = conc.shape[1]
        T = None if temp is None else np.ascontiguousarray(np.broadcast_to(temp, (ncell,)), dtype=np.float64)
        rho = None if air_density is None else np.ascontiguousarray(np.broadcast_to(air_density, (ncell,)), dtype=np.float64)
        jv = None
        if self.nphoto:
            jv = np.zeros((self.nphoto, ncell)) if jvals is None else np.ascontiguousarray(jvals, dtype=np.float64)
            assert jv.shape == (self.nphoto, ncell)
        at = np.ascontiguousarray(np.broadcast_to(1.0e-3 if atol is None else atol, (self.n,)), dtype=np.float64)
        rt = np.ascontiguousarray(np.broadcast_to(1.0e-6 if rtol is None else rtol, (self.n,)), dtype=np.float64)
        acc = np.zeros(ncell, dtype=np.int64)
        att = np.zeros(ncell, dtype=np.int64)
        stalled = lib().oracle_integrate(self._h, METHODS[method], ncell, _d(conc), _d(T), _d(rho), _d(jv), float(dt),
                                         int(steps), float(seed), _d(at), _d(rt), _l(acc), _l(att), int(bool(hoist)),
                                         int(max_attempts), int(threads), _d(_cells(cloud_water, ncell)))
        return acc, att, stalled

    def integrate_trace(self, y, dt, method="rodas3", temp=300.0, air_density=2.4476e19, jvals=(), atol=1.0e-3,
                        rtol=1.0e-6, seed=1.0e-6, max_attempts=0, cap=100000, clw=0.0):
        """One cell's integrate(dt) with its attempt trace.  y: physical-order
        state (not modified).  Returns (y_end, trace) with trace an
        (attempts, 4) array of [t, dt, error norm, accepted] per attempt."""
        y = np.array(y, dtype=np.float64).reshape(-1).copy()
        assert y.shape[0] == self.n
        jv = np.ascontiguousarray(np.asarray(jvals, dtype=np.float64).reshape(-1) if len(jvals) else np.zeros(1))
        at = np.ascontiguousarray(np.broadcast_to(atol, (self.n,)), dtype=np.float64)
        rt = np.ascontiguousarray(np.broadcast_to(rtol, (self.n,)), dtype=np.float64)
        tr = np.zeros((cap, 4))
        acc = np.zeros(1, dtype=np.int64)
        att = lib().oracle_integrate_trace(self._h, METHODS[method], _d(y), float(temp), float(air_density), _d(jv),
                                           float(dt), float(seed), _d(at), _d(rt), int(max_attempts), _d(tr), int(cap),
                                           _l(acc), float(clw))
        assert att <= cap, "trace capacity exceeded"
        return y, tr[:att].copy()

    def integrate_reduced(self, conc, dt, steps=1, method="rodas3", temp=None, air_density=None, jvals=None,
                          atol=None, rtol=None, threshold=0.0, hoist=True, max_attempts=0, threads=0,
                          cloud_water=None):
        """integrate_with_reduction over cells in place (layout as integrate).
        atol/rtol are indexed by the solver's permuted position, as the
        reference's reduction variant indexes them.  Returns (accepted,
        attempted, stalled_cells, active) with active (num_species, num_cells)
        the flags of each cell's last attempted step in permuted order."""
        conc = np.asarray(conc)
        assert conc.dtype == np.float64 and conc.flags.c_contiguous and conc.shape[0] == self.n
        ncell = conc.shape[1]
        T = None if temp is None else np.ascontiguousarray(np.broadcast_to(temp, (ncell,)), dtype=np.float64)
        rho = None if air_density is None else np.ascontiguousarray(np.broadcast_to(air_density, (ncell,)), dtype=np.float64)
        jv = None
        if self.nphoto:
            jv = np.zeros((self.nphoto, ncell)) if jvals is None else np.ascontiguousarray(jvals, dtype=np.float64)
            assert jv.shape == (self.nphoto, ncell)
        at = np.ascontiguousarray(np.broadcast_to(1.0e-3 if atol is None else atol, (self.n,)), dtype=np.float64)
        rt = np.ascontiguousarray(np.broadcast_to(1.0e-6 if rtol is None else rtol, (self.n,)), dtype=np.float64)
        acc = np.zeros(ncell, dtype=np.int64)
        att = np.zeros(ncell, dtype=np.int64)
        active = np.zeros((self.n, ncell), dtype=np.int32)
        stalled = lib().oracle_integrate_reduced(self._h, METHODS[method], ncell, _d(conc), _d(T), _d(rho), _d(jv),
                                                 float(dt), int(steps), float(threshold), _d(at), _d(rt), _l(acc),
                                                 _l(att), int(bool(hoist)), int(max_attempts),
                                                 active.ctypes.data_as(C.POINTER(C.c_int)), int(threads),
                                                 _d(_cells(cloud_water, ncell)))
        return acc, att, stalled, active.astype(bool)


def tableau(method: str) -> dict:
    """The coefficients the oracle integrates with (oracle_tableau)."""
    out = np.zeros(51)
    assert lib().oracle_tableau(METHODS[method], _d(out)) == 0
    s = int(out[0])
    ntri = s * (s - 1) // 2
    return {"stages": s, "gamma1": out[1], "elo": out[2], "A": list(out[3:3 + ntri]), "C": list(out[18:18 + ntri]),
            "M": list(out[33:33 + s]), "E": list(out[39:39 + s]), "NewF": [bool(v) for v in out[45:45 + s]]}
