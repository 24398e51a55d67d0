"""ctypes binding of libmkpp_amd.so and the per-mechanism drop-in libraries.

Nothing here falls back to Python or CPU code: if the native library is
missing the import of a function raises.  Arrays use the C ABI's LayoutLeft
order, i.e. numpy/torch arrays shaped (num_species, num_cells), C-contiguous,
so that element [species, cell] sits at cell + num_cells * species.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

MKPP_SUCCESS = 0
ERRORS = {
    -101: "MKPP_ERROR_NULL_POINTER",
    -102: "MKPP_ERROR_INVALID_HANDLE",
    -103: "MKPP_ERROR_INVALID_ARGUMENT",
    -104: "MKPP_ERROR_INVALID_STATE",
    -105: "MKPP_ERROR_CONVERGENCE",
    -106: "MKPP_ERROR_NOT_FOUND",
}
METHODS = {"ros2": 0, "ros3": 1, "ros4": 2, "rodas3": 3, "rodas4": 4}


class MkppError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_vp, _dp, _ip, _sz = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_size_t

_SIGS = {
    "mkpp_amd_mechanism_count": ([], C.c_int),
    "mkpp_amd_mechanism_name": ([C.c_int], C.c_char_p),
    "mkpp_amd_has_mechanism": ([C.c_char_p], C.c_int),
    "mkpp_amd_mechanism_info": ([C.c_char_p, _ip, _ip, _ip, _ip, _ip], C.c_int),
    "mkpp_amd_species_name": ([C.c_char_p, C.c_int, C.POINTER(C.c_char_p)], C.c_int),
    "mkpp_amd_create_handle": ([C.c_char_p, C.POINTER(_vp), _sz], C.c_int),
    "mkpp_amd_destroy_handle": ([_vp], C.c_int),
    "mkpp_amd_initialize": ([_vp], C.c_int),
    "mkpp_amd_get_species_count": ([_vp, C.POINTER(_sz)], C.c_int),
    "mkpp_amd_get_reaction_count": ([_vp, C.POINTER(_sz)], C.c_int),
    "mkpp_amd_get_photolysis_count": ([_vp, C.POINTER(_sz)], C.c_int),
    "mkpp_amd_get_species_index": ([_vp, C.c_char_p, _ip], C.c_int),
    "mkpp_amd_get_photolysis_index": ([_vp, C.c_char_p, _ip], C.c_int),
    "mkpp_amd_set_state_ptrs": ([_vp, _dp, _dp, _dp, _dp], C.c_int),
    "mkpp_amd_set_rate_parameter_ptrs": ([_vp, _dp, _sz], C.c_int),
    "mkpp_amd_set_photolysis_ptrs": ([_vp, _dp], C.c_int),
    "mkpp_amd_set_cloud_water_ptr": ([_vp, _dp], C.c_int),
    "mkpp_amd_integrate": ([_vp, C.c_double], C.c_int),
    "mkpp_amd_integrate_steps": ([_vp, C.c_double, C.c_int], C.c_int),
    "mkpp_amd_evaluate_rates": ([_vp, _dp], C.c_int),
    "mkpp_amd_evaluate_jacobian": ([_vp, _dp], C.c_int),
    "mkpp_amd_get_last_error": ([_vp, C.c_char_p, _sz], C.c_int),
    "mkpp_amd_set_solver_options": ([_vp, C.c_int, _dp, _dp, C.c_double], C.c_int),
    "mkpp_amd_set_max_attempts": ([_vp, C.c_long], C.c_int),
    "mkpp_amd_get_solver_options": ([_vp, _ip, _dp, _dp, _dp], C.c_int),
    "mkpp_amd_execute": ([_vp, _vp, C.c_long, C.c_long, _vp, _vp], C.c_int),
    "mkpp_amd_get_cell_status": ([_vp, _ip, _ip, _ip], C.c_int),
    "mkpp_amd_device_integrate": ([C.c_char_p, _vp, C.c_long, C.c_long, _vp, _vp, _vp, C.c_long, C.c_double, C.c_int,
                                   C.c_int, C.c_double, _vp, _vp, _vp, _vp, _vp, C.c_long, _vp], C.c_int),
    "mkpp_amd_device_integrate_unrolled": ([C.c_char_p, _vp, C.c_long, C.c_long, _vp, _vp, _vp, C.c_long, C.c_double,
                                            C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, _vp, _vp, C.c_long, _vp],
                                           C.c_int),
    "mkpp_amd_device_evaluate": ([C.c_char_p, C.c_int, _vp, C.c_long, C.c_long, _vp, _vp, _vp, C.c_long, _vp, C.c_long,
                                  _vp], C.c_int),
    "mkpp_amd_device_integrate_reduced": ([C.c_char_p, _vp, C.c_long, C.c_long, _vp, _vp, _vp, C.c_long, C.c_double,
                                           C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, _vp, _vp, C.c_long, _vp],
                                          C.c_int),
}



class BatchSegment(C.Structure):
    """mkpp_amd_batch_segment (include/mkpp_amd.h)."""

    _fields_ = [("mechanism", C.c_char_p), ("conc", _vp), ("num_cells", C.c_long), ("stride", C.c_long),
                ("temp", _vp), ("air_density", _vp), ("rate_parameters", _vp), ("parameter_stride", C.c_long),
                ("atol", _vp), ("rtol", _vp), ("status", _vp), ("accepted", _vp), ("attempted", _vp)]


BATCH_MAX_SEGMENTS = 8
BATCH_WORKSPACE_BYTES = 32

_SIGS["mkpp_amd_device_integrate_batch"] = ([C.POINTER(BatchSegment), C.c_int, C.c_double, C.c_int, C.c_int,
                                             C.c_double, C.c_long, C.c_int, C.c_int, _vp, _vp], C.c_int)
_SIGS["mkpp_amd_device_integrate_batch_static"] = ([C.POINTER(BatchSegment), C.c_int, C.c_double, C.c_int, C.c_int,
                                                    C.c_double, C.c_long, C.c_int, C.c_int, _vp], C.c_int)
_SIGS["mkpp_amd_batch_capacity"] = ([C.c_int, _ip, _ip], C.c_int)
_SIGS["mkpp_amd_diff_info"] = ([C.c_char_p, _ip, _ip], C.c_int)
_SIGS["mkpp_amd_device_rate_constants"] = ([C.c_char_p, _vp, C.c_long, C.c_long, _vp, _vp, _vp, C.c_long, _vp,
                                            C.c_long, _vp], C.c_int)
_SIGS["mkpp_amd_device_integrate_checkpointed"] = ([C.c_char_p, _vp, C.c_long, C.c_long, _vp, _vp, _vp, C.c_long, _vp,
                                                    C.c_long, C.c_double, C.c_int, C.c_double, _vp, _vp, _vp, _vp,
                                                    _vp, C.c_long, _vp, _vp, _vp, C.c_int, C.c_int, _vp], C.c_int)
_SIGS["mkpp_amd_device_tangent_linear"] = ([C.c_char_p, _vp, C.c_long, C.c_long, _vp, _vp, _vp, C.c_long, _vp, C.c_long,
                                            C.c_int, _vp, _vp, C.c_int, _vp, _vp, _vp, _vp], C.c_int)
_SIGS["mkpp_amd_device_adjoint"] = ([C.c_char_p, _vp, C.c_long, C.c_long, _vp, _vp, _vp, C.c_long, _vp, C.c_long,
                                     C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp], C.c_int)

_lib = None


def library_path() -> str:
    return os.path.join(LIBDIR, "libmkpp_amd.so")


def lib():
    """Load libmkpp_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python -m mkpp_amd.build`")
        L = C.CDLL(path, mode=C.RTLD_GLOBAL)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def dropin_path(mechanism: str) -> str:
    return os.path.join(LIBDIR, f"libmkpp_{mechanism}.so")


def load_dropin(mechanism: str):
    """The reference-named C ABI (mkpp_*) of one mechanism."""
    lib()  # resolve libmkpp_amd.so first
    path = dropin_path(mechanism)
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m mkpp_amd.build`")
    return C.CDLL(path)


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def mechanisms() -> list[str]:
    L = lib()
    return [L.mkpp_amd_mechanism_name(i).decode() for i in range(L.mkpp_amd_mechanism_count())]


def mechanism_info(name: str) -> dict:
    vals = [C.c_int() for _ in range(5)]
    rc = lib().mkpp_amd_mechanism_info(name.encode(), *[C.byref(v) for v in vals])
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, name)
    keys = ("num_species", "num_reactions", "num_photolysis", "num_rate_parameters", "supports_runtime_environment")
    info = {k: v.value for k, v in zip(keys, vals)}
    names = []
    for i in range(info["num_species"]):
        p = C.c_char_p()
        lib().mkpp_amd_species_name(name.encode(), i, C.byref(p))
        names.append(p.value.decode())
    info["species"] = names
    return info


class ExecutionParams(C.Structure):
    """mkpp_amd_execution_params (include/mkpp_amd.h)."""

    _fields_ = [("dt", C.c_double), ("steps", C.c_int), ("jvals", C.c_void_p), ("jvals_extent", C.c_long),
                ("temp", C.c_double), ("air_density", C.c_double), ("rh", C.c_double), ("clw", C.c_double),
                ("cell_temperature", C.c_void_p), ("cell_humidity", C.c_void_p), ("cell_density", C.c_void_p),
                ("cell_cloud_water", C.c_void_p), ("cell_jvals", C.c_void_p), ("cell_jvals_extent0", C.c_long),
                ("cell_jvals_extent1", C.c_long), ("cell_jvals_stride0", C.c_long), ("cell_jvals_stride1", C.c_long)]


class Handle:
    """Handle API over host numpy buffers (borrowed, like the reference's C ABI)."""

    def __init__(self, mechanism: str, num_cells: int):
        self.mechanism = mechanism
        self.num_cells = num_cells
        self._h = _vp()
        self._check(lib().mkpp_amd_create_handle(mechanism.encode(), C.byref(self._h), num_cells), "create")
        self._check(lib().mkpp_amd_initialize(self._h), "initialize")
        self.info = mechanism_info(mechanism)
        self._keep = []

    def _check(self, rc, what):
        if rc != MKPP_SUCCESS:
            msg = ""
            if getattr(self, "_h", None) and self._h.value:
                buf = C.create_string_buffer(512)
                lib().mkpp_amd_get_last_error(self._h, buf, 512)
                msg = buf.value.decode()
            raise MkppError(rc, f"{what}: {msg}")

    def close(self):
        if self._h.value:
            lib().mkpp_amd_destroy_handle(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bind(self, conc: np.ndarray, temp, pressure, air_density, rate_parameters=None):
        n = self.info["num_species"]
        assert conc.dtype == np.float64 and conc.flags.c_contiguous and conc.shape == (n, self.num_cells)
        arrs = [np.ascontiguousarray(np.broadcast_to(x, (self.num_cells,)), dtype=np.float64)
                for x in (temp, pressure, air_density)]
        self._keep = [conc, *arrs]
        self._check(lib().mkpp_amd_set_state_ptrs(self._h, _d(conc), *[_d(a) for a in arrs]), "set_state_ptrs")
        if rate_parameters is not None:
            rp = np.ascontiguousarray(rate_parameters, dtype=np.float64)
            assert rp.shape == (self.info["num_rate_parameters"], self.num_cells)
            self._keep.append(rp)
            self._check(lib().mkpp_amd_set_rate_parameter_ptrs(self._h, _d(rp), rp.shape[0]), "set_rate_parameter_ptrs")

    def set_cloud_water(self, cloud_water):
        """Per-cell cloud liquid water (the reference's mkpp_set_cloud_water_ptr);
        cells read 0 until it is set."""
        cw = np.ascontiguousarray(np.broadcast_to(cloud_water, (self.num_cells,)), dtype=np.float64)
        self._keep_cw = cw
        self._check(lib().mkpp_amd_set_cloud_water_ptr(self._h, _d(cw)), "set_cloud_water_ptr")

    def set_solver(self, method="ros3", atol=None, rtol=None, seed=1.0e-6):
        n = self.info["num_species"]
        at = None if atol is None else np.ascontiguousarray(np.broadcast_to(atol, (n,)), dtype=np.float64)
        rt = None if rtol is None else np.ascontiguousarray(np.broadcast_to(rtol, (n,)), dtype=np.float64)
        self._check(lib().mkpp_amd_set_solver_options(self._h, METHODS[method], _d(at), _d(rt), float(seed)), "set_solver_options")

    def set_max_attempts(self, n: int):
        self._check(lib().mkpp_amd_set_max_attempts(self._h, int(n)), "set_max_attempts")

    def integrate(self, dt: float, steps: int = 1):
        self._check(lib().mkpp_amd_integrate_steps(self._h, float(dt), int(steps)), "integrate")

    def execute(self, conc, dt: float, steps: int = 1, temp=None, air_density=None, cloud_water=None,
                rate_parameters=None, stream=None):
        """mkpp_amd_execute on device-resident cells (torch CUDA tensors,
        float64): conc (N, cells) contiguous; temp, air_density, cloud_water
        per cell or None (the execution defaults 288.15 K, 2.4476e19, 0);
        rate_parameters (P, cells) contiguous or None.  Enqueued on `stream`
        (default: torch's current stream) and synchronous, with the handle's
        solver options; per-cell outcomes via cell_status()."""
        import torch

        n = self.num_cells
        assert tuple(conc.shape) == (self.info["num_species"], n) and conc.is_contiguous()
        p = ExecutionParams(dt=float(dt), steps=int(steps), temp=288.15, air_density=2.4476e19, rh=0.5, clw=0.0)
        for field, t in (("cell_temperature", temp), ("cell_density", air_density), ("cell_cloud_water", cloud_water)):
            if t is not None:
                assert tuple(t.shape) == (n,) and t.is_contiguous()
                setattr(p, field, t.data_ptr())
        if rate_parameters is not None:
            assert rate_parameters.shape[1] == n and rate_parameters.is_contiguous()
            p.cell_jvals = rate_parameters.data_ptr()
            p.cell_jvals_extent0, p.cell_jvals_extent1 = n, rate_parameters.shape[0]
            p.cell_jvals_stride0, p.cell_jvals_stride1 = 1, n
        s = (torch.cuda.current_stream() if stream is None else stream).cuda_stream
        self._check(lib().mkpp_amd_execute(self._h, conc.data_ptr(), n, n, C.addressof(p), s), "execute")

    def cell_status(self):
        st = np.zeros(self.num_cells, np.int32)
        acc = np.zeros(self.num_cells, np.int32)
        att = np.zeros(self.num_cells, np.int32)
        ip = C.POINTER(C.c_int)
        self._check(lib().mkpp_amd_get_cell_status(self._h, st.ctypes.data_as(ip), acc.ctypes.data_as(ip),
                                                     att.ctypes.data_as(ip)), "get_cell_status")
        return st, acc, att

    def evaluate_rates(self):
        out = np.zeros((self.info["num_species"], self.num_cells))
        self._check(lib().mkpp_amd_evaluate_rates(self._h, _d(out)), "evaluate_rates")
        return out

    def evaluate_jacobian(self):
        n = self.info["num_species"]
        out = np.zeros((n, n, self.num_cells))  # [j][i][cell] == cell + ncells*(i + n*j)
        self._check(lib().mkpp_amd_evaluate_jacobian(self._h, _d(out)), "evaluate_jacobian")
        return out.transpose(1, 0, 2)  # -> [i][j][cell]


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


_INFO_CACHE: dict = {}


def _info(mechanism: str) -> dict:
    if mechanism not in _INFO_CACHE:
        _INFO_CACHE[mechanism] = mechanism_info(mechanism)
    return _INFO_CACHE[mechanism]


def validate(mechanism: str, conc, *, temp=None, air_density=None, rate_parameters=None, atol=None, rtol=None,
             status=None, accepted=None, attempted=None, k=None, ckpt=None, hs=None, nsteps=None, lam=None,
             out=None, require_tolerances: bool = True) -> int:
    """Check every tensor an entry point hands to the kernels against the
    mechanism's metadata before anything is launched: the kernels address
    conc[cell + cells*species] for every species, rate parameters for every
    slot, per-cell arrays for every cell, so a short or mistyped tensor would
    be read or written out of bounds.  Raises ValueError; returns the cell
    count.  Shapes and dtypes are checked first, then that every tensor is a
    contiguous CUDA tensor on conc's device."""
    import torch

    info = _info(mechanism)
    N, P = info["num_species"], info["num_rate_parameters"]
    f64, i32 = torch.float64, torch.int32

    def need(t, name, shape, dtype):
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{mechanism}: {name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
        if t.dtype != dtype:
            raise ValueError(f"{mechanism}: {name} has dtype {t.dtype}, expected {dtype}")

    if conc.dim() != 2 or conc.shape[0] != N:
        raise ValueError(f"{mechanism}: conc has shape {tuple(conc.shape)}, expected ({N}, num_cells)")
    ncell = int(conc.shape[1])
    need(conc, "conc", (N, ncell), f64)
    checks = [("conc", conc)]
    for name, t, shape, dtype in (("temp", temp, (ncell,), f64), ("air_density", air_density, (ncell,), f64),
                                  ("atol", atol, (N,), f64), ("rtol", rtol, (N,), f64),
                                  ("status", status, (ncell,), i32), ("accepted", accepted, (ncell,), i32),
                                  ("attempted", attempted, (ncell,), i32), ("lam", lam, (N, ncell), f64)):
        if t is not None:
            need(t, name, shape, dtype)
            checks.append((name, t))
    if require_tolerances and (atol is None or rtol is None):
        raise ValueError(f"{mechanism}: atol and rtol ({N},) float64 tensors are required")
    if rate_parameters is not None:
        if P == 0:
            raise ValueError(f"{mechanism}: takes no rate parameters")
        need(rate_parameters, "rate_parameters", (P, ncell), f64)
        checks.append(("rate_parameters", rate_parameters))
    if k is not None or ckpt is not None or hs is not None or nsteps is not None:
        ng, nv = diff_info(mechanism)
        if k is not None:
            need(k, "k", (ng, ncell), f64)
            checks.append(("k", k))
        if hs is not None:
            if hs.dim() != 2 or hs.shape[0] != ncell or hs.shape[1] < 1:
                raise ValueError(f"{mechanism}: hs has shape {tuple(hs.shape)}, expected ({ncell}, max_steps)")
            need(hs, "hs", (ncell, hs.shape[1]), f64)
            checks.append(("hs", hs))
            if ckpt is not None:
                need(ckpt, "ckpt", (ncell, hs.shape[1], nv), f64)
                checks.append(("ckpt", ckpt))
        elif ckpt is not None:
            raise ValueError(f"{mechanism}: ckpt needs hs")
        if nsteps is not None:
            need(nsteps, "nsteps", (ncell,), i32)
            checks.append(("nsteps", nsteps))
    if out is not None:
        checks.append(("out", out))
    dev = conc.device
    for name, t in checks:
        if not t.is_cuda or t.device != dev:
            raise ValueError(f"{mechanism}: {name} must be a CUDA tensor on {dev}, got {t.device}")
        if not t.is_contiguous():
            raise ValueError(f"{mechanism}: {name} must be contiguous")
    return ncell


def device_integrate(mechanism: str, conc, dt: float, steps: int = 1, method: str = "rodas3", temp=None,
                     air_density=None, rate_parameters=None, atol=None, rtol=None, seed: float = 1.0e-6,
                     status=None, accepted=None, attempted=None, max_attempts: int = 100000, stream=None):
    """Integrate device-resident cells in place (torch CUDA tensors, float64).

    conc: (num_species, num_cells) contiguous; rate_parameters: (P, num_cells);
    atol/rtol: (num_species,) device tensors.  Enqueued on `stream` (default:
    torch's current stream); returns immediately.
    """
    import torch

    ncell = validate(mechanism, conc, temp=temp, air_density=air_density, rate_parameters=rate_parameters,
                     atol=atol, rtol=rtol, status=status, accepted=accepted, attempted=attempted)
    s = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().mkpp_amd_device_integrate(
        mechanism.encode(), _ptr(conc), ncell, ncell, _ptr(temp), _ptr(air_density), _ptr(rate_parameters),
        ncell if rate_parameters is not None else 0, float(dt), int(steps), METHODS[method], float(seed),
        _ptr(atol), _ptr(rtol), _ptr(status), _ptr(accepted), _ptr(attempted), int(max_attempts),
        C.c_void_p(s.cuda_stream))
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, "device_integrate")


def device_integrate_unrolled(mechanism: str, conc, dt: float, steps: int = 1, method: str = "rodas3", temp=None,
                              air_density=None, rate_parameters=None, atol=None, rtol=None, seed: float = 1.0e-6,
                              status=None, accepted=None, attempted=None, max_attempts: int = 100000, stream=None):
    """device_integrate through the validation backend: dense unrolled LU and
    solves in the solver permutation (the reference's
    MKPP_USE_UNROLLED_REFERENCE build).  Lane-per-cell mechanisms only."""
    import torch

    ncell = validate(mechanism, conc, temp=temp, air_density=air_density, rate_parameters=rate_parameters,
                     atol=atol, rtol=rtol, status=status, accepted=accepted, attempted=attempted)
    s = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().mkpp_amd_device_integrate_unrolled(
        mechanism.encode(), _ptr(conc), ncell, ncell, _ptr(temp), _ptr(air_density), _ptr(rate_parameters),
        ncell if rate_parameters is not None else 0, float(dt), int(steps), METHODS[method], float(seed),
        _ptr(atol), _ptr(rtol), _ptr(status), _ptr(accepted), _ptr(attempted), int(max_attempts),
        C.c_void_p(s.cuda_stream))
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, "device_integrate_unrolled")


def device_integrate_reduced(mechanism: str, conc, dt: float, steps: int = 1, method: str = "rodas3", temp=None,
                             air_density=None, rate_parameters=None, atol=None, rtol=None,
                             importance_threshold: float = 0.0, status=None, accepted=None, attempted=None,
                             max_attempts: int = 100000, stream=None):
    """The reference's integrate_with_reduction on device-resident cells, in
    place (arguments as device_integrate; atol/rtol are indexed by the
    species' position in the solver permutation, as the reference does)."""
    import torch

    ncell = validate(mechanism, conc, temp=temp, air_density=air_density, rate_parameters=rate_parameters,
                     atol=atol, rtol=rtol, status=status, accepted=accepted, attempted=attempted)
    s = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().mkpp_amd_device_integrate_reduced(
        mechanism.encode(), _ptr(conc), ncell, ncell, _ptr(temp), _ptr(air_density), _ptr(rate_parameters),
        ncell if rate_parameters is not None else 0, float(dt), int(steps), METHODS[method],
        float(importance_threshold), _ptr(atol), _ptr(rtol), _ptr(status), _ptr(accepted), _ptr(attempted),
        int(max_attempts), C.c_void_p(s.cuda_stream))
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, "device_integrate_reduced")


def device_evaluate(mechanism: str, what: str, conc, temp=None, air_density=None, rate_parameters=None, stream=None):
    import torch

    if what not in ("rates", "jacobian"):
        raise ValueError(f"what must be 'rates' or 'jacobian', got {what!r}")
    ncell = validate(mechanism, conc, temp=temp, air_density=air_density, rate_parameters=rate_parameters,
                     require_tolerances=False)
    n = conc.shape[0]
    if what == "rates":
        out = torch.empty((n, ncell), dtype=torch.float64, device=conc.device)
    else:
        out = torch.empty((n, n, ncell), dtype=torch.float64, device=conc.device)
    s = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().mkpp_amd_device_evaluate(mechanism.encode(), 0 if what == "rates" else 1, _ptr(conc), ncell, ncell,
                                        _ptr(temp), _ptr(air_density), _ptr(rate_parameters),
                                        ncell if rate_parameters is not None else 0, _ptr(out), ncell,
                                        C.c_void_p(s.cuda_stream))
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, "device_evaluate")
    return out if what == "rates" else out.permute(1, 0, 2)


def batch_capacity(method: str = "rodas3"):
    """(workgroups one CU holds, CU count) for the persistent batch kernel."""
    per_cu, cus = C.c_int(), C.c_int()
    rc = lib().mkpp_amd_batch_capacity(METHODS[method], C.byref(per_cu), C.byref(cus))
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, "batch_capacity")
    return per_cu.value, cus.value


def device_integrate_batch(segments, dt: float, steps: int = 1, method: str = "rodas3", seed: float = 1.0e-6,
                           max_attempts: int = 100000, cells_per_lane: int = 1, workgroups: int = 0,
                           workspace=None, stream=None):
    """One persistent launch over several mechanisms' device-resident cells.

    segments: list of dicts with keys mechanism, conc (species, cells) and the
    optional temp, air_density, rate_parameters, atol, rtol (required),
    status, accepted, attempted -- torch CUDA tensors as for device_integrate.
    workspace: int32 CUDA tensor of >= 8 elements (allocated here if None).
    Enqueued on `stream`; returns immediately.
    """
    import torch

    keep = []
    arr = (BatchSegment * len(segments))()
    for i, g in enumerate(segments):
        conc = g["conc"]
        ncell = validate(g["mechanism"], conc, temp=g.get("temp"), air_density=g.get("air_density"),
                         rate_parameters=g.get("rate_parameters"), atol=g.get("atol"), rtol=g.get("rtol"),
                         status=g.get("status"), accepted=g.get("accepted"), attempted=g.get("attempted"))
        name = g["mechanism"].encode()
        keep.append(name)
        rp = g.get("rate_parameters")
        arr[i] = BatchSegment(name, conc.data_ptr(), ncell, ncell, _addr(g.get("temp")), _addr(g.get("air_density")),
                              _addr(rp), ncell if rp is not None else 0, _addr(g.get("atol")), _addr(g.get("rtol")),
                              _addr(g.get("status")), _addr(g.get("accepted")), _addr(g.get("attempted")))
    if workspace is None:
        workspace = torch.zeros(BATCH_WORKSPACE_BYTES // 4, dtype=torch.int32, device=segments[0]["conc"].device)
    assert workspace.is_cuda and workspace.numel() * workspace.element_size() >= BATCH_WORKSPACE_BYTES
    s = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().mkpp_amd_device_integrate_batch(arr, len(segments), float(dt), int(steps), METHODS[method],
                                               float(seed), int(max_attempts), int(cells_per_lane), int(workgroups),
                                               C.c_void_p(workspace.data_ptr()), C.c_void_p(s.cuda_stream))
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, "device_integrate_batch")
    return workspace


def _addr(t):
    return None if t is None else t.data_ptr()


def diff_info(mechanism: str):
    """(flux groups, variable species) of a mechanism with a device adjoint."""
    ng, nv = C.c_int(), C.c_int()
    rc = lib().mkpp_amd_diff_info(mechanism.encode(), C.byref(ng), C.byref(nv))
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, "diff_info")
    return ng.value, nv.value


def _stream(stream):
    import torch

    return C.c_void_p((stream if stream is not None else torch.cuda.current_stream()).cuda_stream)


def device_rate_constants(mechanism: str, conc, temp=None, air_density=None, rate_parameters=None, stream=None):
    """Rate constants k (groups, cells) from the rate laws (device tensors)."""
    import torch

    ncell = validate(mechanism, conc, temp=temp, air_density=air_density, rate_parameters=rate_parameters,
                     require_tolerances=False)
    ng, _ = diff_info(mechanism)
    k = torch.empty((ng, ncell), dtype=torch.float64, device=conc.device)
    rc = lib().mkpp_amd_device_rate_constants(mechanism.encode(), _ptr(conc), ncell, ncell, _ptr(temp),
                                              _ptr(air_density), _ptr(rate_parameters),
                                              ncell if rate_parameters is not None else 0, _ptr(k), ncell,
                                              _stream(stream))
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, "device_rate_constants")
    return k


def device_integrate_checkpointed(mechanism: str, conc, dt: float, method: str, temp, air_density, rate_parameters,
                                  k, atol, rtol, ckpt, hs, nsteps, seed: float = 1.0e-6, status=None, accepted=None,
                                  attempted=None, max_attempts: int = 100000, replay: bool = False, stream=None):
    """Forward integration of device-resident cells that records (replay=False)
    or replays (replay=True) the accepted steps.  ckpt: (cells, max_steps, NV)
    float64, hs: (cells, max_steps) float64, nsteps: (cells,) int32."""
    if ckpt is None or hs is None or nsteps is None:
        raise ValueError(f"{mechanism}: ckpt, hs and nsteps are required")
    ncell = validate(mechanism, conc, temp=temp, air_density=air_density, rate_parameters=rate_parameters,
                     atol=atol, rtol=rtol, status=status, accepted=accepted, attempted=attempted, k=k, ckpt=ckpt,
                     hs=hs, nsteps=nsteps)
    max_steps = hs.shape[1]
    rc = lib().mkpp_amd_device_integrate_checkpointed(
        mechanism.encode(), _ptr(conc), ncell, ncell, _ptr(temp), _ptr(air_density), _ptr(rate_parameters),
        ncell if rate_parameters is not None else 0, _ptr(k), ncell if k is not None else 0, float(dt),
        METHODS[method], float(seed), _ptr(atol), _ptr(rtol), _ptr(status), _ptr(accepted), _ptr(attempted),
        int(max_attempts), _ptr(ckpt), _ptr(hs), _ptr(nsteps), int(max_steps), 1 if replay else 0, _stream(stream))
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, "device_integrate_checkpointed")


def device_adjoint(mechanism: str, conc, method: str, temp, air_density, rate_parameters, k, ckpt, hs, nsteps,
                   lam, stream=None):
    """Adjoint sweep over recorded steps.  lam (species, cells): dL/dy(T) in,
    dL/dy(0) out (in place).  Returns kbar (groups, cells) = dL/dk."""
    import torch

    if ckpt is None or hs is None or nsteps is None or lam is None:
        raise ValueError(f"{mechanism}: ckpt, hs, nsteps and lam are required")
    ncell = validate(mechanism, conc, temp=temp, air_density=air_density, rate_parameters=rate_parameters, k=k,
                     ckpt=ckpt, hs=hs, nsteps=nsteps, lam=lam, require_tolerances=False)
    ng, _ = diff_info(mechanism)
    kbar = torch.empty((ng, ncell), dtype=torch.float64, device=conc.device)
    rc = lib().mkpp_amd_device_adjoint(
        mechanism.encode(), _ptr(conc), ncell, ncell, _ptr(temp), _ptr(air_density), _ptr(rate_parameters),
        ncell if rate_parameters is not None else 0, _ptr(k), ncell, METHODS[method], _ptr(ckpt), _ptr(hs),
        _ptr(nsteps), int(hs.shape[1]), _ptr(lam), _ptr(kbar), _stream(stream))
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, "device_adjoint")
    return kbar


def device_tangent_linear(mechanism: str, conc, method: str, temp, air_density, rate_parameters, k, hs, nsteps, dy,
                          dk=None, status=None, stream=None):
    """Tangent-linear replay of recorded steps (device tensors, in place):
    conc (species, cells) y(0) -> y(T); dy (species, cells) dy(0) -> dy(T);
    dk (groups, cells) or None the rate-constant direction; k as for the
    adjoint (None: rate laws)."""
    if hs is None or nsteps is None or dy is None:
        raise ValueError(f"{mechanism}: hs, nsteps and dy are required")
    ncell = validate(mechanism, conc, temp=temp, air_density=air_density, rate_parameters=rate_parameters, k=k,
                     hs=hs, nsteps=nsteps, lam=dy, status=status, require_tolerances=False)
    if dk is not None:
        validate(mechanism, conc, k=dk, require_tolerances=False)
    rc = lib().mkpp_amd_device_tangent_linear(
        mechanism.encode(), _ptr(conc), ncell, ncell, _ptr(temp), _ptr(air_density), _ptr(rate_parameters),
        ncell if rate_parameters is not None else 0, _ptr(k), ncell, METHODS[method], _ptr(hs), _ptr(nsteps),
        int(hs.shape[1]), _ptr(dy), _ptr(dk), _ptr(status), _stream(stream))
    if rc != MKPP_SUCCESS:
        raise MkppError(rc, "device_tangent_linear")
