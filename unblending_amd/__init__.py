"""unblending_amd -- MI355X (gfx950) backend of Unblending's two per-pixel passes.

Python mirror of the reference's packed-buffer C++ API
(unblending/include/unblending/pixels.hpp:34-44) over the C interface of
libunblending.so (include/unblending.h, include/unblending_hip.h):

    compute_color_unmixing(image, layer_infos, has_opaque_background, target_concurrency=0)
    perform_matte_refinement(image, layers, layer_infos, has_opaque_background,
                             force_smooth_background, target_concurrency=0)

Both return packed [layer][row][column][RGBA] float64 arrays.  As in the
reference (unblending/src/gpu_dispatch.cpp:57-68) the backend comes from the
UNBLENDING_BACKEND environment variable: unset/"auto"/"hip" run on the GPU,
"cpu" and "cuda" are rejected (this library ships only the HIP backend),
anything else raises ValueError.  Errors map as the reference's exceptions
do: invalid arguments -> ValueError, allocation -> MemoryError, device
failures -> DeviceError.

The native library is required: importing this package without a built
libunblending.so raises ImportError (run __graft_entry__.build()).
"""
from __future__ import annotations

import ctypes
import enum
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libunblending.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("libunblending.so is not built; run __graft_entry__.build() (hipcc, gfx950)")

# torch ships its own HIP runtime with the same SONAME (libamdhip64.so.7) as
# /opt/rocm's.  Loading torch first makes libunblending.so bind to that one
# runtime; the other order would put two runtimes in the process.
try:
    import torch  # noqa: F401
except ImportError:  # the library also works without torch (host-pointer API)
    torch = None

_lib = ctypes.CDLL(LIB_PATH)

# ---- status codes and enums (include/unblending.h) ----
OK, INVALID_ARGUMENT, BUFFER_TOO_SMALL, DEVICE_ERROR, OUT_OF_MEMORY, SOLVER_ERROR = range(6)
CUDA_ERROR = DEVICE_ERROR
BACKEND_AUTO, BACKEND_CPU, BACKEND_CUDA, BACKEND_HIP = range(4)


class BlendMode(enum.IntEnum):
    Normal = 0
    Multiply = 1
    Screen = 2
    Overlay = 3
    Darken = 4
    Lighten = 5
    ColorDodge = 6
    ColorBurn = 7
    HardLight = 8
    SoftLight = 9
    Difference = 10
    Exclusion = 11
    LinearDodge = 12


class DeviceError(RuntimeError):
    """The device could not run the pass (UNBLENDING_DEVICE_ERROR)."""


class SolverError(RuntimeError):
    """Any other failure (UNBLENDING_SOLVER_ERROR)."""


class BufferTooSmall(ValueError):
    """Output capacity below 4 * width * height * layers (UNBLENDING_BUFFER_TOO_SMALL)."""


@dataclass(frozen=True)
class CompOp:
    """Porter-Duff factors (unblending/include/unblending/comp_op.hpp:9-30)."""
    X: int = 1
    Y: int = 1
    Z: int = 1

    @staticmethod
    def SourceOver():
        return CompOp(1, 1, 1)

    @staticmethod
    def Plus():
        return CompOp(2, 1, 1)

    def is_source_over(self):
        return (self.X, self.Y, self.Z) == (1, 1, 1)

    def is_plus(self):
        return (self.X, self.Y, self.Z) == (2, 1, 1)


@dataclass
class GaussianColorModel:
    """Mean and precision (inverse covariance) of a layer's colours
    (unblending/include/unblending/color_model.hpp:19-59)."""
    mu: np.ndarray
    sigma_inv: np.ndarray

    def __post_init__(self):
        self.mu = np.asarray(self.mu, dtype=np.float64).reshape(3)
        self.sigma_inv = np.asarray(self.sigma_inv, dtype=np.float64).reshape(3, 3)

    @staticmethod
    def from_sigma(mu, sigma):
        return GaussianColorModel(mu, np.linalg.inv(np.asarray(sigma, dtype=np.float64)))

    def calculate_distance(self, color):
        d = np.asarray(color, dtype=np.float64) - self.mu
        return float(d @ self.sigma_inv @ d)


@dataclass
class LayerInfo:
    """unblending/include/unblending/layer_info.hpp:10-15."""
    comp_op: CompOp
    blend_mode: BlendMode
    color_model: GaussianColorModel


# ---- C structs ----
class RgbView(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_double)), ("size", ctypes.c_size_t),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("row_stride", ctypes.c_size_t)]


class CLayer(ctypes.Structure):
    _fields_ = [("comp_x", ctypes.c_int32), ("comp_y", ctypes.c_int32), ("comp_z", ctypes.c_int32),
                ("blend_mode", ctypes.c_int32), ("mean", ctypes.c_double * 3),
                ("precision", ctypes.c_double * 9)]


class Options(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("backend", ctypes.c_int32),
                ("opaque_background", ctypes.c_int32), ("smooth_background", ctypes.c_int32),
                ("target_concurrency", ctypes.c_int32)]


class HipParams(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("filter_radius", ctypes.c_int32),
                ("filter_epsilon", ctypes.c_double), ("alpha_fixed", ctypes.c_int32),
                ("solver", ctypes.c_int32), ("timing", ctypes.c_int32), ("warm_start", ctypes.c_int32),
                ("mlp", ctypes.c_void_p), ("chunk_size", ctypes.c_int32), ("filter_block", ctypes.c_int32),
                ("filter_rows", ctypes.c_int32), ("warm_outer_cap", ctypes.c_int32),
                ("lbfgs_memory", ctypes.c_int32), ("gray_layers", ctypes.c_int32), ("sparsity", ctypes.c_int32)]


class HipImage(ctypes.Structure):
    """unblending_hip_image: one image of a batch (device pointers)."""
    _fields_ = [("rgb", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("layers", ctypes.c_void_p), ("layer_count", ctypes.c_size_t), ("first_pass", ctypes.c_void_p),
                ("alpha_planes", ctypes.c_void_p), ("output", ctypes.c_void_p), ("filter_radius", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("mlp", ctypes.c_void_p)]


class HipStats(ctypes.Structure):
    _fields_ = [("evals", ctypes.c_uint64), ("iterations", ctypes.c_uint64), ("outer", ctypes.c_uint64),
                ("pixels", ctypes.c_uint64), ("solve_ms", ctypes.c_double), ("filter_ms", ctypes.c_double),
                ("filter_coeff_ms", ctypes.c_double), ("filter_out_ms", ctypes.c_double),
                ("filter_radius", ctypes.c_int32), ("filter_channels", ctypes.c_int32),
                ("filter_launches", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("fallback_pixels", ctypes.c_uint64), ("predict_ms", ctypes.c_double),
                ("lane_slots", ctypes.c_uint64), ("history_pairs", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_lib.unblending_abi_version.restype = ctypes.c_uint32
_lib.unblending_unmix.restype = ctypes.c_int32
_lib.unblending_unmix.argtypes = [ctypes.POINTER(RgbView), ctypes.POINTER(CLayer), ctypes.c_size_t,
                                  ctypes.POINTER(Options), _dp, ctypes.c_size_t]
_lib.unblending_refine.restype = ctypes.c_int32
_lib.unblending_refine.argtypes = [ctypes.POINTER(RgbView), ctypes.POINTER(CLayer), ctypes.c_size_t, _dp,
                                   ctypes.c_size_t, ctypes.POINTER(Options), _dp, ctypes.c_size_t]
_lib.unblending_hip_unmix_ex.restype = ctypes.c_int32
_lib.unblending_hip_unmix_ex.argtypes = [ctypes.POINTER(RgbView), ctypes.POINTER(CLayer), ctypes.c_size_t,
                                         ctypes.POINTER(Options), ctypes.POINTER(HipParams), _dp,
                                         ctypes.c_size_t, ctypes.POINTER(HipStats)]
_lib.unblending_hip_refine_ex.restype = ctypes.c_int32
_lib.unblending_hip_refine_ex.argtypes = [ctypes.POINTER(RgbView), ctypes.POINTER(CLayer), ctypes.c_size_t,
                                          _dp, ctypes.c_size_t, ctypes.POINTER(Options),
                                          ctypes.POINTER(HipParams), _dp, ctypes.c_size_t,
                                          ctypes.POINTER(HipStats)]
_lib.unblending_hip_workspace_size.restype = ctypes.c_size_t
_lib.unblending_hip_workspace_size.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_size_t,
                                               ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
_lib.unblending_hip_workspace_size_ex.restype = ctypes.c_size_t
_lib.unblending_hip_workspace_size_ex.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_size_t,
                                                  ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(HipParams)]
_lib.unblending_hip_unmix_device.restype = ctypes.c_int32
_lib.unblending_hip_unmix_device.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(CLayer),
                                             ctypes.c_size_t, ctypes.POINTER(Options),
                                             ctypes.POINTER(HipParams), _vp, _vp, _vp, ctypes.c_size_t, _vp,
                                             ctypes.POINTER(HipStats)]
_lib.unblending_hip_refine_device.restype = ctypes.c_int32
_lib.unblending_hip_refine_device.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(CLayer),
                                              ctypes.c_size_t, _vp, _vp, ctypes.POINTER(Options),
                                              ctypes.POINTER(HipParams), _vp, _vp, ctypes.c_size_t, _vp,
                                              ctypes.POINTER(HipStats)]
_lib.unblending_hip_filter_workspace_size.restype = ctypes.c_size_t
_lib.unblending_hip_filter_workspace_size.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
_lib.unblending_hip_guided_filter_device.restype = ctypes.c_int32
_lib.unblending_hip_guided_filter_device.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32,
                                                     ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int32,
                                                     ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                                     _vp, ctypes.c_size_t, _vp, ctypes.POINTER(HipStats)]
_lib.unblending_hip_guided_filter_device_noalloc.restype = ctypes.c_int32
_lib.unblending_hip_guided_filter_device_noalloc.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32,
                                                             ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                             ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                                             ctypes.c_int32, ctypes.c_int32, _vp,
                                                             ctypes.POINTER(HipStats)]
_lib.uhip_solve_planes.restype = ctypes.c_int32
_lib.uhip_solve_planes.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(CLayer), ctypes.c_size_t,
                                   _dp, _dp, ctypes.c_int32, ctypes.c_int32, _dp]
_lib.unblending_hip_mlp_model_size.restype = ctypes.c_size_t
_lib.unblending_hip_mlp_model_size.argtypes = []
_lib.unblending_hip_mlp_param_count.restype = ctypes.c_int32
_lib.unblending_hip_mlp_param_count.argtypes = []
_lib.unblending_hip_mlp_set_params.restype = ctypes.c_int32
_lib.unblending_hip_mlp_set_params.argtypes = [_vp, ctypes.POINTER(ctypes.c_float), _vp]
_lib.unblending_hip_mlp_get_params.restype = ctypes.c_int32
_lib.unblending_hip_mlp_get_params.argtypes = [_vp, ctypes.POINTER(ctypes.c_float), _vp]
_lib.unblending_hip_mlp_train_workspace_size.restype = ctypes.c_size_t
_lib.unblending_hip_mlp_train_workspace_size.argtypes = [ctypes.c_int32]
_lib.unblending_hip_mlp_train.restype = ctypes.c_int32
_lib.unblending_hip_mlp_train.argtypes = [_vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp,
                                          ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_float, ctypes.c_uint64, _vp, ctypes.c_int32, _vp,
                                          ctypes.c_size_t, _vp, ctypes.POINTER(ctypes.c_float),
                                          ctypes.POINTER(ctypes.c_float)]
_lib.unblending_hip_mlp_predict_workspace_size.restype = ctypes.c_size_t
_lib.unblending_hip_mlp_predict_workspace_size.argtypes = []
_lib.unblending_hip_mlp_predict.restype = ctypes.c_int32
_lib.unblending_hip_mlp_predict.argtypes = [_vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _vp,
                                            ctypes.c_size_t, _vp]
_lib.unblending_hip_batch_workspace_size.restype = ctypes.c_size_t
_lib.unblending_hip_batch_workspace_size.argtypes = [ctypes.POINTER(HipImage), ctypes.c_size_t, ctypes.c_int32,
                                                     ctypes.c_int32]
_lib.unblending_hip_batch_workspace_size_ex.restype = ctypes.c_size_t
_lib.unblending_hip_batch_workspace_size_ex.argtypes = [ctypes.POINTER(HipImage), ctypes.c_size_t, ctypes.c_int32,
                                                        ctypes.c_int32, ctypes.POINTER(HipParams)]
for _f in (_lib.unblending_hip_unmix_batch_device, _lib.unblending_hip_refine_batch_device):
    _f.restype = ctypes.c_int32
    _f.argtypes = [ctypes.POINTER(HipImage), ctypes.c_size_t, ctypes.POINTER(Options), ctypes.POINTER(HipParams),
                   _vp, ctypes.c_size_t, _vp, ctypes.POINTER(HipStats)]
_lib.unblending_hip_last_error.restype = ctypes.c_char_p
_lib.unblending_hip_release_buffers.restype = ctypes.c_int32
_lib.unblending_hip_release_buffers.argtypes = []


_lib.unblending_hip_composite_device.restype = ctypes.c_int32
_lib.unblending_hip_composite_device.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_size_t,
                                                 ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                                 _vp, _vp]
_lib.unblending_hip_composite.restype = ctypes.c_int32
_lib.unblending_hip_composite.argtypes = [_dp, ctypes.c_size_t, ctypes.c_int32, ctypes.c_int32, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), _dp,
                                          ctypes.c_size_t]


def lib():
    return _lib


def abi_version() -> int:
    return int(_lib.unblending_abi_version())


def last_error() -> str:
    return _lib.unblending_hip_last_error().decode()


def release_device_buffers():
    """Frees the device buffers the host-pointer calls keep on the current
    device between calls (unblending_hip_release_buffers)."""
    raise_for_status(_lib.unblending_hip_release_buffers())


def layer_array(layer_infos) -> ctypes.Array:
    """LayerInfo list -> unblending_layer[]."""
    arr = (CLayer * len(layer_infos))()
    for i, info in enumerate(layer_infos):
        op = info.comp_op
        arr[i].comp_x, arr[i].comp_y, arr[i].comp_z = int(op.X), int(op.Y), int(op.Z)
        arr[i].blend_mode = int(info.blend_mode)
        m = info.color_model
        for c in range(3):
            arr[i].mean[c] = float(m.mu[c])
        P = np.asarray(m.sigma_inv, dtype=np.float64).reshape(9)
        for k in range(9):
            arr[i].precision[k] = float(P[k])
    return arr


def options(backend=BACKEND_AUTO, opaque=True, smooth=False, target_concurrency=0) -> Options:
    o = Options()
    o.struct_size = ctypes.sizeof(Options)
    o.backend = int(backend)
    o.opaque_background = int(bool(opaque))
    o.smooth_background = int(bool(smooth))
    o.target_concurrency = int(target_concurrency)
    return o


def hip_params(radius=-1, eps=0.0, alpha_fixed=False, solver=0, timing=False, warm_start=None, chunk_size=0,
               filter_block=0, filter_rows=0, warm_outer_cap=0, lbfgs_memory=0, gray_layers=0,
               sparsity=False) -> HipParams:
    """warm_start: a trained ``mlp.WarmStartModel`` (first pass only), or None.
    chunk_size / filter_block / filter_rows: launch shapes (0: defaults; see
    ``unblending_amd.autotune``).  lbfgs_memory: (s, y) pairs of the inner
    minimiser; 0 = NLopt's default for LD_LBFGS (every pair of an inner
    solve), 1 = the one-pair fast mode.  gray_layers: bit mask of layers whose
    colour is constrained to gray (the reference's SECOND_LAYER_GRAY: 0b10);
    sparsity: the first pass's sparsity term (the reference's SPARCITY)."""
    p = HipParams()
    p.struct_size = ctypes.sizeof(HipParams)
    p.filter_radius = int(radius)
    p.filter_epsilon = float(eps)
    p.alpha_fixed = int(bool(alpha_fixed))
    p.solver = int(solver)
    p.timing = int(bool(timing))
    p.warm_start = int(warm_start is not None)
    p.mlp = warm_start.data_ptr() if warm_start is not None else None
    p.chunk_size = int(chunk_size)
    p.filter_block = int(filter_block)
    p.filter_rows = int(filter_rows)
    p.warm_outer_cap = int(warm_outer_cap)
    p.lbfgs_memory = int(lbfgs_memory)
    p.gray_layers = int(gray_layers)
    p.sparsity = int(bool(sparsity))
    return p


def environment_backend() -> int:
    """gpu_dispatch.cpp:57-68, with the 'hip' value."""
    value = os.environ.get("UNBLENDING_BACKEND", "")
    table = {"": BACKEND_AUTO, "auto": BACKEND_AUTO, "cpu": BACKEND_CPU, "cuda": BACKEND_CUDA,
             "hip": BACKEND_HIP}
    if value not in table:
        raise ValueError("UNBLENDING_BACKEND must be auto, cpu, cuda, or hip")
    return table[value]


def raise_for_status(status: int):
    if status == OK:
        return
    msg = last_error()
    if status == INVALID_ARGUMENT:
        raise ValueError(msg or "invalid argument")
    if status == BUFFER_TOO_SMALL:
        raise BufferTooSmall(msg or "buffer too small")
    if status == DEVICE_ERROR:
        raise DeviceError(msg or "device error")
    if status == OUT_OF_MEMORY:
        raise MemoryError(msg or "out of memory")
    raise SolverError(msg or "solver error (status %d)" % status)


def _view(image: np.ndarray):
    a = np.asarray(image, dtype=np.float64)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("image must be (height, width, 3)")
    a = np.ascontiguousarray(a)
    v = RgbView(a.ctypes.data_as(_dp), a.size, a.shape[1], a.shape[0], 0)
    return v, a


def compute_color_unmixing(image, layer_infos, has_opaque_background, target_concurrency=0,
                           *, params: HipParams | None = None, stats: HipStats | None = None):
    """First pass (pixels.hpp:34-37).  Returns (L, h, w, 4) packed RGBA."""
    v, keep = _view(image)
    L = len(layer_infos)
    out = np.empty((L, v.height, v.width, 4))
    opt = options(environment_backend(), has_opaque_background, False, target_concurrency)
    st = _lib.unblending_hip_unmix_ex(ctypes.byref(v), layer_array(layer_infos), L, ctypes.byref(opt),
                                      ctypes.byref(params) if params else None, out.ctypes.data_as(_dp),
                                      out.size, ctypes.byref(stats) if stats else None)
    raise_for_status(st)
    return out


def perform_matte_refinement(image, layers, layer_infos, has_opaque_background, force_smooth_background,
                             target_concurrency=0, *, params: HipParams | None = None,
                             stats: HipStats | None = None):
    """Second pass (pixels.hpp:40-44); layers is the packed first pass."""
    v, keep = _view(image)
    L = len(layer_infos)
    first = np.ascontiguousarray(np.asarray(layers, dtype=np.float64))
    out = np.empty((L, v.height, v.width, 4))
    opt = options(environment_backend(), has_opaque_background, force_smooth_background, target_concurrency)
    st = _lib.unblending_hip_refine_ex(ctypes.byref(v), layer_array(layer_infos), L, first.ctypes.data_as(_dp),
                                       first.size, ctypes.byref(opt), ctypes.byref(params) if params else None,
                                       out.ctypes.data_as(_dp), out.size, ctypes.byref(stats) if stats else None)
    raise_for_status(st)
    return out


def _composite_tables(comp_ops, modes):
    ops = (ctypes.c_int32 * (3 * len(comp_ops)))(*[v for op in comp_ops for v in (op.X, op.Y, op.Z)])
    md = (ctypes.c_int32 * len(modes))(*[int(m) for m in modes])
    return ops, md


def composite_layers(packed, comp_ops, modes=None):
    """Image-level recomposite (unblending/src/unblending.cpp:638-664) on the
    device: packed (L, h, w, 4) straight-alpha layers, bottom first, folded
    with each layer's comp-op and blend mode; returns (h, w, 4) RGBA.
    comp_ops may also be the list of LayerInfo (modes then taken from it)."""
    if modes is None:
        infos = comp_ops
        comp_ops = [i.comp_op for i in infos]
        modes = [i.blend_mode for i in infos]
    packed = np.ascontiguousarray(packed, dtype=np.float64)
    L, h, w, _ = packed.shape
    if len(comp_ops) != L or len(modes) != L:
        raise ValueError("one comp-op and one blend mode per layer")
    out = np.empty((h, w, 4), dtype=np.float64)
    ops, md = _composite_tables(comp_ops, modes)
    raise_for_status(_lib.unblending_hip_composite(
        packed.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), packed.size, w, h, L, ops, md,
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size))
    return out


from . import device  # noqa: E402  (device-resident API over torch tensors)
