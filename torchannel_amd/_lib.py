"""ctypes binding of libtorchannel_hip.so (the C ABI declared in include/torchannel.h).

Every product entry point in this package goes through here.  There is no CPU path: a tensor
that is not on a HIP device, or a missing library, raises.
"""
from __future__ import annotations

import ctypes
import os
import threading
import weakref

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtorchannel_hip.so")

TC_TOP_DIRICHLET, TC_TOP_NEUMANN = 0, 1
TC_BRANCH_GPU, TC_BRANCH_CPU = 0, 1
TC_F64, TC_F32 = 0, 1
TC_NSCAL = 8
# scalar-block slots (tc_internal.h)
S_DT, S_F, S_UBULK, S_TIME, S_CFL, S_UTARGET, S_STEP, S_NONFINITE = 0, 1, 2, 3, 4, 5, 6, 7
TUNE_K1_JB = 1
TUNE_THOMAS_SB = 2
TUNE_K1_TILE = 3
TUNE_THOMAS_LDS = 4

_vp, _i, _d, _i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
_pd = ctypes.POINTER(ctypes.c_double)

# name -> argtypes (every entry point returns int)
SIGNATURES = {
    "tc_version": [],
    "tc_grid_create": [_i, _i, _i, _d, _d, _d, _i, _i, _pd, _pd, ctypes.POINTER(_vp)],
    "tc_grid_destroy": [_vp],
    "tc_grid_set_tuning": [_vp, _i, _i],
    "tc_grid_set_pitch": [_vp, _i],
    "tc_grid_get_pitch": [_vp, ctypes.POINTER(_i)],
    "tc_solver_set_newest": [_vp, _i],
    "tc_adjoint_create": [_vp, _d, ctypes.POINTER(_vp)],
    "tc_adjoint_destroy": [_vp],
    "tc_adjoint_objective": [_vp, _vp, _d, _vp, _vp],
    "tc_adjoint_begin": [_vp, _d, _d, _vp, _vp],
    "tc_adjoint_step": [_vp, _vp, _vp, _vp, _i, _d, _vp, _vp],
    "tc_ensemble_create": [ctypes.POINTER(_vp), _i, ctypes.POINTER(_vp)],
    "tc_ensemble_destroy": [_vp],
    "tc_ensemble_step": [_vp, _d, _i, _vp],
    "tc_ensemble_capture": [_vp, _i, _vp],
    "tc_ensemble_replay": [_vp, _vp],
    "tc_ensemble_set_form": [_vp, _i],
    "tc_ensemble_check": [_vp, _vp],
    "tc_grid_get_tuning": [_vp, _i, ctypes.POINTER(_i)],
    "tc_apply_bc": [_vp, _vp, _vp, _vp, _vp],
    "tc_rhs_imex": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tc_rhs_fused_v2": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tc_cn_solve": [_vp, _i, _d, _vp, _d, _vp, _vp, _vp],
    "tc_pcr_batched": [_i, _i, _i, _i, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    "tc_thomas_batched": [_i, _i, _i, _i, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp],
    "tc_thomas_batched_static": [_i, _i, _i, _i, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    "tc_divergence": [_vp, _vp, _vp, _vp, _vp, _vp],
    "tc_project": [_vp, _vp, _vp, _vp, _vp, _d, _vp, _vp],
    "tc_bulk_velocity": [_vp, _vp, _vp, _i64, _i64, _i64, _d, _vp, _vp],
    "tc_cfl_rate": [_vp, _vp, _vp, _vp, _vp, _vp],
    "tc_u_tau": [_vp, _vp, _d, _vp, _vp],
    "tc_poisson_create": [_vp, ctypes.POINTER(_vp)],
    "tc_poisson_destroy": [_vp],
    "tc_poisson_solve": [_vp, _vp, _vp, _vp],
    "tc_stats_accumulate": [_vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp],
    "tc_solver_workspace_size": [_vp, ctypes.POINTER(ctypes.c_size_t)],
    "tc_solver_create": [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp, _d,
                         _i, _vp, ctypes.c_size_t, ctypes.POINTER(_vp)],
    "tc_solver_destroy": [_vp],
    "tc_solver_step": [_vp, _d, _i, _vp, ctypes.POINTER(_i), _vp],
    "tc_solver_project_initial": [_vp, _i, _vp],
    "tc_solver_rescale": [_vp, _vp],
    "tc_solver_update_dt": [_vp, _d, _d, _d, _vp],
    "tc_solver_diagnostics": [_vp, _d, _vp, _i, _vp],
    "tc_solver_capture": [_vp, _i, _i, _d, _d, _d, _vp],
    "tc_solver_replay": [_vp, _vp],
    "tc_solver_phase": [_vp, _i, _i, _vp],
    "tc_calib_columns": [_i, _i, _i, _i, _i, _i, _i, ctypes.POINTER(_vp), _vp],
}

_lib = None
_lock = threading.Lock()


def load():
    """Load the HIP library once; raise if it was never built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"torchannel_amd: {LIB_PATH} is missing; build it with "
                                   "`python -m torchannel_amd.build` (HIP/gfx950)")
            lib = ctypes.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = ctypes.c_int
            lib.tc_strerror.argtypes = [ctypes.c_int]
            lib.tc_strerror.restype = ctypes.c_char_p
            _lib = lib
    return _lib


def check(rc: int):
    if rc != 0:
        msg = load().tc_strerror(rc).decode()
        raise RuntimeError(f"{msg} (code {rc})")


def call(name, *args):
    check(getattr(load(), name)(*args))


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def stream_of(t: torch.Tensor):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def require_device(*tensors, dense=True):
    """The product path runs only on HIP devices, on float32/float64 tensors, contiguous unless
    `dense` is False (field arrays, whose layout require_fields checks)."""
    dt = None
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
            raise RuntimeError("torchannel_amd: tensors must live on a HIP device (no CPU path)")
        if dense and not t.is_contiguous():
            raise RuntimeError("torchannel_amd: tensors must be contiguous (reference C-order layout)")
        if t.dtype not in (torch.float64, torch.float32):
            raise RuntimeError(f"torchannel_amd: unsupported dtype {t.dtype}")
        if dt is None:
            dt = t.dtype
        elif t.dtype != dt:
            raise RuntimeError("torchannel_amd: mixed dtypes")
    load()
    return dt


LINE_BYTES = 128      # cache-line size the padded layout aligns every column's row 1 to


def padded_pitch(nz, dtype) -> int:
    """Column stride (elements) of the padded layout: nz + 2 rounded up to whole 128-byte lines."""
    a = LINE_BYTES // torch.empty((), dtype=dtype).element_size()
    return -(-(nz + 2) // a) * a


def empty_field(shape, pitch, dtype, device, zero=True):
    """A ghosted array of `shape` (planes, columns, rows).  pitch 0: a dense tensor.  pitch > 0:
    a strided view whose columns sit `pitch` elements apart, with row 1 of every column on a
    128-byte line; it indexes like the dense array, and the storage behind the view extends at
    least one full pitch past every column (flat passes of the library cover
    planes * columns * pitch elements)."""
    n0, n1, n2 = (int(x) for x in shape)
    if not pitch:
        return (torch.zeros if zero else torch.empty)((n0, n1, n2), dtype=dtype, device=device)
    if pitch < n2:
        raise ValueError(f"pitch {pitch} is shorter than the column length {n2}")
    esz = torch.empty((), dtype=dtype).element_size()
    a = LINE_BYTES // esz
    if (pitch * esz) % LINE_BYTES:
        raise ValueError(f"pitch {pitch} is not a whole number of {LINE_BYTES}-byte lines")
    buf = (torch.zeros if zero else torch.empty)(n0 * n1 * pitch + 2 * a, dtype=dtype, device=device)
    off = (-(buf.data_ptr() // esz + 1)) % a             # element off + 1 on a line
    return buf.as_strided((n0, n1, n2), (n1 * pitch, pitch, 1), buf.storage_offset() + off)


def field_pitch(t) -> int:
    """0 for a contiguous (dense) array, else its column stride; raises for other layouts."""
    if t.is_contiguous():
        return 0
    if t.dim() == 3:
        P = t.stride(1)
        if t.stride() == (t.shape[1] * P, P, 1) and P >= t.shape[2]:
            return P
    raise RuntimeError("torchannel_amd: field arrays must be dense or have columns at one stride "
                       "(the padded layout)")


def field_like(t, zero=False):
    """A new array with t's shape, dtype, device and layout (dense or padded)."""
    P = field_pitch(t)
    if P == 0:
        return torch.zeros_like(t) if zero else torch.empty_like(t)
    return empty_field(t.shape, P, t.dtype, t.device, zero=zero)


def clone_field(t):
    out = field_like(t)
    out.copy_(t)
    return out


def clone_field_as(src, like):
    """A copy of src in the layout of `like` (same shape)."""
    out = field_like(like)
    out.copy_(src)
    return out


def require_fields(*tensors):
    """(dtype, pitch) of ghosted field arrays passed to one call: all dense (pitch 0), or all with
    one column stride P (a contiguous array whose column length is P counts as that layout)."""
    ts = [t for t in tensors if t is not None]
    dt = require_device(*ts, dense=False)
    pitches = {field_pitch(t) for t in ts}
    if pitches <= {0}:
        return dt, 0
    P = max(pitches)
    for t in ts:
        if field_pitch(t) not in (0, P) or (field_pitch(t) == 0 and t.shape[2] != P):
            raise RuntimeError("torchannel_amd: the field arrays of one call must share one layout")
    return dt, P


def dtype_code(dt):
    return TC_F64 if dt == torch.float64 else TC_F32


def top_code(top_wall_bc_type: str) -> int:
    if top_wall_bc_type not in ("dirichlet", "neumann"):
        raise ValueError(f"Invalid top wall BC type: {top_wall_bc_type}. Must be 'dirichlet' or 'neumann'")
    return TC_TOP_NEUMANN if top_wall_bc_type == "neumann" else TC_TOP_DIRICHLET


class Grid:
    """Owns a tc_grid (extents, spacings, metric tables on the device)."""

    def __init__(self, nx, ny, nz, dx, dy, nu, top, dtype, dz_f_host, dz_c_host, device, pitch=0):
        lib = load()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dtype = dtype
        self.device = device
        self.pitch = int(pitch)
        f = (ctypes.c_double * nz)(*dz_f_host)
        c = (ctypes.c_double * (nz + 1))(*dz_c_host)
        h = ctypes.c_void_p()
        with torch.cuda.device(device):
            check(lib.tc_grid_create(nx, ny, nz, float(dx), float(dy), float(nu), top, dtype_code(dtype),
                                     f, c, ctypes.byref(h)))
        self.handle = h
        if self.pitch:
            check(lib.tc_grid_set_pitch(h, self.pitch))
        self._poisson = None

    def set_tuning(self, key, value):
        """Set a launch-shape knob (TUNE_*); results do not depend on it."""
        check(load().tc_grid_set_tuning(self.handle, int(key), int(value)))

    def get_tuning(self, key):
        out = ctypes.c_int()
        check(load().tc_grid_get_tuning(self.handle, int(key), ctypes.byref(out)))
        return out.value

    def poisson(self):
        if self._poisson is None:
            h = ctypes.c_void_p()
            with torch.cuda.device(self.device):
                check(load().tc_poisson_create(self.handle, ctypes.byref(h)))
            self._poisson = h
        return self._poisson

    def new_poisson(self):
        """A Poisson plan of this grid with its own workspace, owned by the caller
        (destroy_poisson); channels stepped together each need one."""
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(load().tc_poisson_create(self.handle, ctypes.byref(h)))
        return h

    @staticmethod
    def destroy_poisson(h):
        load().tc_poisson_destroy(h)

    def __del__(self):
        try:
            lib = load()
            if self._poisson is not None:
                lib.tc_poisson_destroy(self._poisson)
            lib.tc_grid_destroy(self.handle)
        except Exception:
            pass


_grid_cache: dict = {}


def _metric_host(t, n):
    if t is None:
        return [1.0] * n
    return [float(x) for x in t.detach().to("cpu", torch.float64).reshape(-1)[:n].tolist()]


def device_key(device) -> str:
    """'cuda' and 'cuda:<current>' name the same device: one spelling for cache keys."""
    d = torch.device(device)
    if d.type == "cuda" and d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return str(d)


def get_grid(nx, ny, nz, dx, dy, nu, top, dtype, device, dz_f=None, dz_c=None, pitch=0) -> Grid:
    """Cached Grid for these arguments (pitch: the column stride of the arrays its calls take, 0
    for dense arrays).  The metric tensors are identified by object (weak reference) and version
    counter, so an in-place change or a new tensor rebuilds the grid."""
    key = (int(nx), int(ny), int(nz), float(dx), float(dy), float(nu), int(top), dtype, device_key(device),
           id(dz_f), getattr(dz_f, "_version", -1), id(dz_c), getattr(dz_c, "_version", -1), int(pitch))
    ent = _grid_cache.get(key)
    if ent is not None:
        g, rf, rc = ent
        if (dz_f is None or rf() is dz_f) and (dz_c is None or rc() is dz_c):
            return g
    g = Grid(int(nx), int(ny), int(nz), dx, dy, nu, top, dtype, _metric_host(dz_f, nz),
             _metric_host(dz_c, nz + 1), device, pitch=pitch)
    rf = weakref.ref(dz_f) if dz_f is not None else (lambda: None)
    rc = weakref.ref(dz_c) if dz_c is not None else (lambda: None)
    if len(_grid_cache) > 64:
        _grid_cache.clear()
    _grid_cache[key] = (g, rf, rc)
    return g
