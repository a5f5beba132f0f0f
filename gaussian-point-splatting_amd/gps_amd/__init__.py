"""ctypes binding of libgps_amd.so, the MI355X Gaussian point splatting hot path.

The C ABI is declared in include/gps.h.  This module mirrors the reference's
render-pass surface for Python callers (tests, benchmark, serving loop):

    Scene(...)                 GPUGaussianScene::upload_from_cpu
                               (src/core/gaussians/gaussian_types.cuh:442-503,568-608)
    Pass(settings, scene)      GaussianPointSplatting(const RenderSettings&)
                               (src/core/rendering/passes/gaussian_point_splatting.cu:777-816)
    Pass.render(frame, ...)    RenderPass::render(d_image, d_depth, settings)
                               (src/core/rendering/render_pass.cuh:17-24)
    resolve_taa / resolve_progressive
                               RenderServer::renderFrame post-processing
                               (src/core/rendering/render_server.cu:122-157)

Device memory is owned by the library (DeviceArray) so no second HIP runtime
(e.g. torch's bundled one) is needed in the process.  The shared library must
have been built (``make -C gaussian-point-splatting_amd``); importing this
module without it raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgps_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C {os.path.dirname(_HERE)}` "
        "(the MI355X path has no fallback)")

_lib = C.CDLL(LIB_PATH)

# ----------------------------------------------------------------- constants
GPS_OK = 0
FORMAT_FP32 = 0
FORMAT_HALF_FP8 = 1
FORMAT_QUANTIZED = 2

DBG_WEIGHTS = 0
DBG_MASK_CURRENT = 1
DBG_MASK_PREVIOUS = 2
DBG_VALID_IDS = 3
DBG_VALID_OFFSETS = 4
DBG_INDICES = 5
DBG_FRAMEBUFFER = 6
DBG_RECORDS = 7
DBG_HIZ = 8
DBG_CULL_MASK = 9
DBG_REJECTS = 10
COUNT_ATOMICS = 1
COUNT_REJECTS = 2
REJECT_REASONS = ("tested", "prev_valid", "opacity", "projection", "off_screen", "small", "hiz", "zero_count",
                  "valid")
_DBG_DTYPE = {DBG_FRAMEBUFFER: np.uint64, DBG_HIZ: np.float32}

ACCUM_PROGRESSIVE = 0
ACCUM_TAA = 1
ACCUM_NONE = 2


# ----------------------------------------------------------------- structs
class SceneDesc(C.Structure):
    _fields_ = [("num_gaussians", C.c_uint64), ("sh_degree", C.c_int32), ("format", C.c_int32),
                ("means", C.c_void_p), ("rotations", C.c_void_p), ("scales", C.c_void_p),
                ("opacities", C.c_void_p), ("shs", C.c_void_p)]


class QuantizedBounds(C.Structure):
    """gps_quantized_bounds: the scene box of FORMAT_QUANTIZED."""
    _fields_ = [("min_pos", C.c_float * 3), ("max_pos", C.c_float * 3), ("min_log_scale", C.c_float * 3),
                ("max_log_scale", C.c_float * 3), ("min_sh", C.c_float * 45), ("max_sh", C.c_float * 45)]

    def as_array(self) -> np.ndarray:
        return np.frombuffer(bytes(self), np.float32).copy()


class RenderSettingsC(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("render_pass_count", C.c_int32),
                ("supersampling_factor", C.c_int32), ("max_number_of_points", C.c_int64),
                ("fully_disable_hierarchical_culling", C.c_int32), ("gaussians_per_block", C.c_int32),
                ("flexible_frame_size", C.c_int32)]


class FrameParams(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("frustum_planes", C.c_float * 24),
                ("fovy_degrees", C.c_float), ("near_plane", C.c_float), ("far_plane", C.c_float),
                ("background", C.c_float * 3), ("width", C.c_int32), ("height", C.c_int32),
                ("supersampling_factor", C.c_int32), ("frame", C.c_int32),
                ("enable_occlusion_culling", C.c_int32), ("enable_hierarchical_culling", C.c_int32),
                ("cull_small_gaussians", C.c_int32), ("use_unbiased_2d_splatting", C.c_int32),
                ("reduce_point_count", C.c_int32), ("occlusion_culling_min_reduce_count", C.c_int32),
                ("samples_per_pixel", C.c_int32)]

    def copy(self) -> "FrameParams":
        f = FrameParams()
        C.memmove(C.addressof(f), C.addressof(self), C.sizeof(FrameParams))
        return f


class Stats(C.Structure):
    _fields_ = [("num_gaussians", C.c_uint64), ("valid_gaussians", C.c_uint64 * 2),
                ("total_weight", C.c_uint64 * 2), ("num_points", C.c_uint64 * 2), ("samples", C.c_uint64),
                ("culled_blocks", C.c_uint64 * 2), ("passes", C.c_int32), ("budget_truncated", C.c_int32),
                ("K", C.c_int32), ("gaussians_per_block", C.c_int32), ("atomics", C.c_uint64),
                ("num_blocks", C.c_uint64), ("splat_binned", C.c_int32), ("splat_fallback", C.c_int32),
                ("bin_entries", C.c_uint32), ("bin_jobs", C.c_uint32), ("bin_big", C.c_uint32),
                ("reserved", C.c_uint32), ("culled_groups", C.c_uint64 * 2)]

    def as_dict(self) -> dict:
        return {"num_gaussians": int(self.num_gaussians), "valid_gaussians": list(self.valid_gaussians),
                "total_weight": list(self.total_weight), "num_points": list(self.num_points),
                "samples": int(self.samples), "culled_blocks": list(self.culled_blocks),
                "passes": int(self.passes), "budget_truncated": int(self.budget_truncated), "K": int(self.K),
                "gpb": int(self.gaussians_per_block), "num_blocks": int(self.num_blocks),
                "atomics": int(self.atomics), "splat_binned": int(self.splat_binned),
                "splat_fallback": int(self.splat_fallback), "bin_entries": int(self.bin_entries),
                "bin_jobs": int(self.bin_jobs), "bin_big": int(self.bin_big),
                "culled_groups": list(self.culled_groups)}


class SynthParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("num_gaussians", C.c_uint64), ("kind", C.c_int32),
                ("sh_degree", C.c_int32), ("clusters", C.c_int32), ("reserved", C.c_int32),
                ("half_extent", C.c_double), ("cluster_sigma", C.c_double), ("log_scale_mu", C.c_double),
                ("log_scale_sigma", C.c_double), ("opacity_sigma", C.c_double), ("dc_sigma", C.c_double),
                ("rest_sigma", C.c_double), ("offset", C.c_double * 3)]


_vp = C.c_void_p
_fp = C.POINTER(C.c_float)


def _sig(name, res, *args):
    fn = getattr(_lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_sig("gps_abi_version", C.c_int32)
_sig("gps_status_str", C.c_char_p, C.c_int32)
_sig("gps_last_error", C.c_char_p)
_sig("gps_scene_create", C.c_int32, C.POINTER(SceneDesc), C.POINTER(_vp))
_sig("gps_scene_create_empty", C.c_int32, C.c_uint64, C.c_int32, C.c_int32, C.POINTER(_vp))
_sig("gps_scene_upload_range", C.c_int32, _vp, C.c_uint64, C.POINTER(SceneDesc))
_sig("gps_scene_destroy", None, _vp)
_sig("gps_scene_device_bytes", C.c_uint64, _vp)
_sig("gps_pass_create", C.c_int32, C.POINTER(RenderSettingsC), _vp, C.POINTER(_vp))
_sig("gps_pass_destroy", None, _vp)
_sig("gps_pass_render", C.c_int32, _vp, C.POINTER(FrameParams), _vp, _vp, _vp)
_sig("gps_pass_stats", C.c_int32, _vp, C.POINTER(Stats), _vp)
_sig("gps_pass_debug_size", C.c_uint64, _vp, C.c_int32)
_sig("gps_pass_debug_copy", C.c_int32, _vp, C.c_int32, _vp, C.c_uint64, _vp)
_sig("gps_resolve_taa", C.c_int32, _vp, _vp, _vp, _vp, _fp, C.c_float, C.c_int32, C.c_int32, C.c_int32, _vp, _vp,
     _vp)
_sig("gps_resolve_progressive", C.c_int32, _vp, _vp, C.c_int32, C.c_int32, C.c_int32, _vp, _vp)
_sig("gps_look_at", None, _fp, _fp, _fp, _fp)
_sig("gps_perspective", None, C.c_float, C.c_float, C.c_float, C.c_float, _fp)
_sig("gps_frustum_planes", None, C.c_float, _fp, C.c_float, C.c_float, C.c_float, _fp)
_sig("gps_frame_params_init", None, C.POINTER(FrameParams), _fp, _fp, _fp, C.c_float, C.c_int32, C.c_int32,
     C.c_int32, C.c_float, C.c_float)
_sig("gps_taa_reprojection", None, _fp, _fp, _fp, _fp)
_sig("gps_mat4_mul", None, _fp, _fp, _fp)
_sig("gps_synth_generate", C.c_int32, C.POINTER(SynthParams), C.c_uint64, C.c_uint64, _vp, _vp, _vp, _vp, _vp)
_sig("gps_synth_generate_indices", C.c_int32, C.POINTER(SynthParams), _vp, C.c_uint64, _vp, _vp, _vp, _vp, _vp)
_sig("gps_morton_order", C.c_int32, _vp, C.c_uint64, _vp)
_sig("gps_bvh_order", C.c_int32, _vp, _vp, _vp, C.c_uint64, _vp)
_sig("gps_encode_half_fp8", C.c_int32, C.POINTER(SceneDesc), _vp, _vp, _vp, _vp, _vp, _vp)
_sig("gps_quantized_bounds_compute", C.c_int32, C.POINTER(SceneDesc), C.POINTER(QuantizedBounds))
_sig("gps_encode_quantized", C.c_int32, C.POINTER(SceneDesc), C.POINTER(QuantizedBounds), _vp, _vp, _vp, _vp)
_sig("gps_scene_set_quantized_bounds", C.c_int32, _vp, C.POINTER(QuantizedBounds))
_sig("gps_float_to_half", C.c_uint16, C.c_float)
_sig("gps_float_to_fp8e4m3", C.c_uint8, C.c_float)
_sig("gps_device_count", C.c_int32, C.POINTER(C.c_int32))
_sig("gps_set_device", C.c_int32, C.c_int32)
_sig("gps_device_malloc", C.c_int32, C.POINTER(_vp), C.c_uint64)
_sig("gps_device_free", C.c_int32, _vp)
_sig("gps_memcpy_h2d", C.c_int32, _vp, _vp, C.c_uint64, _vp)
_sig("gps_memcpy_d2h", C.c_int32, _vp, _vp, C.c_uint64, _vp)
_sig("gps_memset_d8", C.c_int32, _vp, C.c_int32, C.c_uint64, _vp)
_sig("gps_stream_create", C.c_int32, C.POINTER(_vp))
_sig("gps_stream_destroy", C.c_int32, _vp)
_sig("gps_stream_synchronize", C.c_int32, _vp)
_sig("gps_device_free_memory", C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
_sig("gps_event_create", C.c_int32, C.POINTER(_vp))
_sig("gps_event_destroy", C.c_int32, _vp)
_sig("gps_event_record", C.c_int32, _vp, _vp)
_sig("gps_event_elapsed_ms", C.c_int32, _vp, _vp, C.POINTER(C.c_float))

_sig("gps_pass_set_timing", C.c_int32, _vp, C.c_int32)
_sig("gps_pass_set_splat_mode", C.c_int32, _vp, C.c_int32)
_sig("gps_pass_set_counters", C.c_int32, _vp, C.c_int32)
_sig("gps_pass_set_group_culling", C.c_int32, _vp, C.c_int32)
_sig("gps_pass_set_ordered_splat", C.c_int32, _vp, C.c_int32)
_sig("gps_pass_timing_summary", C.c_int32, _vp, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
     C.POINTER(C.c_uint64))
_sig("gps_selftest_math", C.c_int32, C.c_int32, _vp, _vp, C.c_uint64, _vp)
_sig("gps_selftest_rng", C.c_int32, C.c_uint32, C.c_uint32, C.c_int32, _vp, _vp)
_sig("gps_selftest_poisson", C.c_int32, C.c_uint32, _vp, _vp, _vp)
_sig("gps_selftest_copy", C.c_int32, _vp, _vp, C.c_uint64, _vp)
_sig("gps_selftest_atomics", C.c_int32, _vp, C.c_uint64, C.c_uint32, C.c_uint32, _vp)

lib = _lib


class GpsError(RuntimeError):
    def __init__(self, status: int, what: str):
        msg = _lib.gps_status_str(status).decode()
        last = _lib.gps_last_error().decode()
        super().__init__(f"{what}: {msg}" + (f" ({last})" if last else ""))
        self.status = status


def check(status: int, what: str) -> None:
    if status != GPS_OK:
        raise GpsError(status, what)


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f(a) -> "C.Array":
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


# ----------------------------------------------------------------- device
def device_count() -> int:
    n = C.c_int32(0)
    st = _lib.gps_device_count(C.byref(n))
    return int(n.value) if st == GPS_OK else 0


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        h = C.c_void_p()
        check(_lib.gps_stream_create(C.byref(h)), "stream create")
        self.handle = h

    def synchronize(self) -> None:
        check(_lib.gps_stream_synchronize(self.handle), "stream synchronize")

    def close(self) -> None:
        if self.handle:
            _lib.gps_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        h = C.c_void_p()
        check(_lib.gps_event_create(C.byref(h)), "event create")
        self.handle = h

    def record(self, stream: Optional[Stream] = None) -> None:
        check(_lib.gps_event_record(self.handle, stream.handle if stream else None), "event record")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0)
        check(_lib.gps_event_elapsed_ms(self.handle, end.handle, C.byref(ms)), "event elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                _lib.gps_event_destroy(self.handle)
        except Exception:
            pass


class DeviceArray:
    """A device allocation owned by libgps_amd with a numpy shape/dtype."""

    def __init__(self, shape, dtype=np.float32, zero: bool = False):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(_lib.gps_device_malloc(C.byref(p), max(self.nbytes, 1)), "device malloc")
        self.ptr = p
        if zero:
            self.zero()

    @property
    def addr(self):
        return self.ptr

    def zero(self, stream: Optional[Stream] = None) -> None:
        check(_lib.gps_memset_d8(self.ptr, 0, self.nbytes, stream.handle if stream else None), "memset")
        if stream is None:
            check(_lib.gps_stream_synchronize(None), "sync")

    def upload(self, a: np.ndarray, stream: Optional[Stream] = None) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        check(_lib.gps_memcpy_h2d(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes,
                                  stream.handle if stream else None), "h2d")

    def download(self, stream: Optional[Stream] = None) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(_lib.gps_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes,
                                  stream.handle if stream else None), "d2h")
        return out

    def free(self) -> None:
        if self.ptr:
            _lib.gps_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ----------------------------------------------------------------- scenes
@dataclass
class SceneArrays:
    """Host scene in the CPUGaussian layout (gaussian_types.cuh:89-100)."""
    means: np.ndarray          # (N, 3) float32
    rotations: np.ndarray      # (N, 4) float32, (w, x, y, z)
    scales: np.ndarray         # (N, 3) float32, linear
    opacities: np.ndarray      # (N,)   float32, post-sigmoid
    shs: np.ndarray            # (N, 3*(deg+1)^2) float32
    sh_degree: int = 0

    @property
    def n(self) -> int:
        return int(self.means.shape[0])

    def desc(self, fmt: int = FORMAT_FP32) -> SceneDesc:
        for name in ("means", "rotations", "scales", "opacities", "shs"):
            a = getattr(self, name)
            assert a.dtype == np.float32 and a.flags.c_contiguous, name
        d = SceneDesc()
        d.num_gaussians = self.n
        d.sh_degree = self.sh_degree
        d.format = fmt
        d.means = _ptr(self.means)
        d.rotations = _ptr(self.rotations)
        d.scales = _ptr(self.scales)
        d.opacities = _ptr(self.opacities)
        d.shs = _ptr(self.shs)
        return d

    def morton_order(self) -> np.ndarray:
        """perm[new] = old of the reference's --sort-morton-order."""
        perm = np.empty(self.n, np.uint32)
        check(_lib.gps_morton_order(_ptr(self.means), self.n, _ptr(perm)), "morton order")
        return perm

    def morton_sorted(self) -> "SceneArrays":
        return self.subset(self.morton_order())

    def bvh_order(self) -> np.ndarray:
        """perm[new] = old of the reference's default load-time reorder
        (GaussianBVH::build_bvh_and_reorder, gaussian_bvh.cu:146-169); runs on
        the GPU."""
        perm = np.empty(self.n, np.uint32)
        check(_lib.gps_bvh_order(_ptr(self.means), _ptr(self.rotations), _ptr(self.scales), self.n, _ptr(perm)),
              "bvh order")
        return perm

    def bvh_sorted(self) -> "SceneArrays":
        return self.subset(self.bvh_order())

    def subset(self, idx) -> "SceneArrays":
        return SceneArrays(np.ascontiguousarray(self.means[idx]), np.ascontiguousarray(self.rotations[idx]),
                           np.ascontiguousarray(self.scales[idx]), np.ascontiguousarray(self.opacities[idx]),
                           np.ascontiguousarray(self.shs[idx]), self.sh_degree)


def synth_params(seed: int, n: int, kind: str = "uniform", sh_degree: int = 0, clusters: int = 1,
                 half_extent: float = 10.0, cluster_sigma: float = 2.0, log_scale_mu: float = -4.0,
                 log_scale_sigma: float = 0.5, opacity_sigma: float = 1.0, dc_sigma: float = 0.3,
                 rest_sigma: float = 0.05, offset=(0.0, 0.0, 0.0)) -> SynthParams:
    p = SynthParams()
    p.seed = seed
    p.num_gaussians = n
    p.kind = 0 if kind == "uniform" else 1
    p.sh_degree = sh_degree
    p.clusters = clusters
    p.half_extent = half_extent
    p.cluster_sigma = cluster_sigma
    p.log_scale_mu = log_scale_mu
    p.log_scale_sigma = log_scale_sigma
    p.opacity_sigma = opacity_sigma
    p.dc_sigma = dc_sigma
    p.rest_sigma = rest_sigma
    for i in range(3):
        p.offset[i] = float(offset[i])
    return p


def synth_scene(p: SynthParams, first: int = 0, count: Optional[int] = None) -> SceneArrays:
    """Generate Gaussians [first, first+count) of a synthetic scene."""
    n = int(p.num_gaussians - first) if count is None else int(count)
    nc = (p.sh_degree + 1) ** 2
    means = np.empty((n, 3), np.float32)
    rots = np.empty((n, 4), np.float32)
    scales = np.empty((n, 3), np.float32)
    ops = np.empty((n,), np.float32)
    shs = np.empty((n, 3 * nc), np.float32)
    check(_lib.gps_synth_generate(C.byref(p), first, n, _ptr(means), _ptr(rots), _ptr(scales), _ptr(ops),
                                  _ptr(shs)), "synth")
    return SceneArrays(means, rots, scales, ops, shs, int(p.sh_degree))


def quantized_bounds(s: SceneArrays) -> QuantizedBounds:
    """Scene box of the reference's compressed format (GaussianSceneBounds)."""
    b = QuantizedBounds()
    d = s.desc(FORMAT_QUANTIZED)
    check(_lib.gps_quantized_bounds_compute(C.byref(d), C.byref(b)), "quantized bounds")
    return b


def merge_quantized_bounds(a: QuantizedBounds, b: QuantizedBounds) -> QuantizedBounds:
    """The box of two parts of one scene (componentwise min of the minima, max of the maxima)."""
    out = QuantizedBounds()
    for name in ("pos", "log_scale", "sh"):
        lo, hi = f"min_{name}", f"max_{name}"
        n = len(getattr(a, lo))
        for i in range(n):
            getattr(out, lo)[i] = min(getattr(a, lo)[i], getattr(b, lo)[i])
            getattr(out, hi)[i] = max(getattr(a, hi)[i], getattr(b, hi)[i])
    return out


def encode_quantized(s: SceneArrays, bounds: Optional[QuantizedBounds] = None) -> dict:
    """Host encoding of FORMAT_QUANTIZED (the bytes the GPU scene holds)."""
    n = s.n
    b = bounds if bounds is not None else quantized_bounds(s)
    out = {"geometry": np.empty((n, 4), np.uint32), "opacity": np.empty((n,), np.uint8),
           "sh_dc": np.empty((n, 3), np.float32), "sh_codes": np.empty((n, 45), np.uint8), "bounds": b}
    d = s.desc(FORMAT_QUANTIZED)
    check(_lib.gps_encode_quantized(C.byref(d), C.byref(b), *[_ptr(out[k]) for k in
                                                             ("geometry", "opacity", "sh_dc", "sh_codes")]),
          "encode quantized")
    return out


def encode_half_fp8(s: SceneArrays) -> dict:
    """Host encoding of the compact format (the bytes the GPU scene holds)."""
    n = s.n
    out = {"rotations": np.empty((n, 4), np.uint16), "scales": np.empty((n, 3), np.uint16),
           "opacities": np.empty((n,), np.uint16), "sh_dc": np.empty((n, 3), np.uint16),
           "sh_rest": np.empty((n, 45), np.uint8), "sh_scale": np.empty((n,), np.float32)}
    d = s.desc(FORMAT_HALF_FP8)
    check(_lib.gps_encode_half_fp8(C.byref(d), *[_ptr(out[k]) for k in
                                                 ("rotations", "scales", "opacities", "sh_dc", "sh_rest",
                                                  "sh_scale")]), "encode")
    return out


class Scene:
    """A scene resident in HBM."""

    def __init__(self, arrays: Optional[SceneArrays], fmt: int = FORMAT_FP32):
        self.handle = None
        if arrays is None:
            return
        h = C.c_void_p()
        d = arrays.desc(fmt)
        check(_lib.gps_scene_create(C.byref(d), C.byref(h)), "scene create")
        self.handle = h
        self.n = arrays.n
        self.format = fmt
        self.sh_degree = arrays.sh_degree

    @classmethod
    def empty(cls, n: int, sh_degree: int, fmt: int = FORMAT_FP32) -> "Scene":
        """An allocated scene of n Gaussians, filled by upload_range."""
        s = cls(None, fmt)
        h = C.c_void_p()
        check(_lib.gps_scene_create_empty(n, sh_degree, fmt, C.byref(h)), "scene create")
        s.handle = h
        s.n = n
        s.format = fmt
        s.sh_degree = sh_degree
        return s

    def set_quantized_bounds(self, bounds: QuantizedBounds) -> None:
        """FORMAT_QUANTIZED: the scene box, needed before upload_range."""
        check(_lib.gps_scene_set_quantized_bounds(self.handle, C.byref(bounds)), "quantized bounds")

    def upload_range(self, first: int, arrays: SceneArrays) -> None:
        d = arrays.desc(self.format)
        check(_lib.gps_scene_upload_range(self.handle, first, C.byref(d)), "scene upload")

    @property
    def device_bytes(self) -> int:
        return int(_lib.gps_scene_device_bytes(self.handle))

    def close(self) -> None:
        if self.handle:
            _lib.gps_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class RenderSettings:
    """The create-time RenderSettings fields (src/render_settings.cuh:36-88)."""
    width: int = 1920
    height: int = 1080
    render_pass_count: int = 1
    supersampling_factor: int = 2
    max_number_of_points: int = 250_000_000
    fully_disable_hierarchical_culling: bool = False
    gaussians_per_block: int = 0
    flexible_frame_size: bool = False

    def c(self) -> RenderSettingsC:
        s = RenderSettingsC()
        s.width, s.height = self.width, self.height
        s.render_pass_count = self.render_pass_count
        s.supersampling_factor = self.supersampling_factor
        s.max_number_of_points = self.max_number_of_points
        s.fully_disable_hierarchical_culling = int(self.fully_disable_hierarchical_culling)
        s.gaussians_per_block = self.gaussians_per_block
        s.flexible_frame_size = int(self.flexible_frame_size)
        return s


class Pass:
    """GaussianPointSplatting render pass over a resident scene."""

    def __init__(self, settings: RenderSettings, scene: Scene):
        h = C.c_void_p()
        self._settings_c = settings.c()
        check(_lib.gps_pass_create(C.byref(self._settings_c), scene.handle, C.byref(h)), "pass create")
        self.handle = h
        self.settings = settings
        self.scene = scene

    def render(self, frame: FrameParams, color: DeviceArray, depth: DeviceArray,
               stream: Optional[Stream] = None) -> None:
        check(_lib.gps_pass_render(self.handle, C.byref(frame), color.ptr, depth.ptr,
                                   stream.handle if stream else None), "pass render")

    def stats(self, stream: Optional[Stream] = None) -> dict:
        s = Stats()
        check(_lib.gps_pass_stats(self.handle, C.byref(s), stream.handle if stream else None), "stats")
        return s.as_dict()

    def debug(self, which: int, stream: Optional[Stream] = None) -> np.ndarray:
        n = int(_lib.gps_pass_debug_size(self.handle, which))
        dt = _DBG_DTYPE.get(which, np.uint32)
        out = np.empty((n,), dt)
        check(_lib.gps_pass_debug_copy(self.handle, which, _ptr(out), n, stream.handle if stream else None),
              f"debug copy {which}")
        return out

    STAGES = ("clear", "cull", "preprocess", "distribute", "splat", "combine", "hiz")

    SPLAT_AUTO = 0
    SPLAT_POINT_PARALLEL = 1
    SPLAT_BINNED = 2

    def set_splat_mode(self, mode: int) -> None:
        """SPLAT_AUTO, SPLAT_POINT_PARALLEL or SPLAT_BINNED."""
        check(_lib.gps_pass_set_splat_mode(self.handle, mode), "splat mode")

    def set_counters(self, count_atomics: bool, count_rejects: bool = False) -> None:
        flags = (COUNT_ATOMICS if count_atomics else 0) | (COUNT_REJECTS if count_rejects else 0)
        check(_lib.gps_pass_set_counters(self.handle, flags), "counters")

    def rejects(self, stream: Optional[Stream] = None) -> list:
        """Per render pass of the last frame, {reason: Gaussians} (counters on
        with set_counters(..., count_rejects=True))."""
        r = self.debug(DBG_REJECTS, stream)[:2 * len(REJECT_REASONS)].reshape(2, len(REJECT_REASONS))
        return [dict(zip(REJECT_REASONS, map(int, row))) for row in r]

    def group_violations(self, stream: Optional[Stream] = None) -> list:
        """Per render pass of the last frame (counters on with count_rejects):
        members of groups the group occlusion test skipped that survived every
        per-Gaussian test.  0 when the test is sound."""
        n = 2 * len(REJECT_REASONS)
        return [int(v) for v in self.debug(DBG_REJECTS, stream)[n:n + 2]]

    def ordered_skipped(self, stream: Optional[Stream] = None) -> list:
        """Per render pass of the last frame (counters on with count_rejects):
        points the ordered splat skipped behind nearer tiles."""
        n = 2 * len(REJECT_REASONS) + 2
        return [int(v) for v in self.debug(DBG_REJECTS, stream)[n:n + 2]]

    ORDERED_AUTO = 0
    ORDERED_OFF = 1
    ORDERED_ALWAYS = 2

    def set_ordered_splat(self, mode: int) -> None:
        """ORDERED_AUTO (dense passes), ORDERED_OFF or ORDERED_ALWAYS; the
        framebuffer is the same bit for bit in every mode."""
        check(_lib.gps_pass_set_ordered_splat(self.handle, mode), "ordered splat")

    def set_group_culling(self, enable: bool) -> None:
        """The group occlusion test (on by default; outputs are the same bits either way)."""
        check(_lib.gps_pass_set_group_culling(self.handle, 1 if enable else 0), "group culling")

    def set_timing(self, capacity: int) -> None:
        check(_lib.gps_pass_set_timing(self.handle, capacity), "set timing")

    def timing_summary(self) -> dict:
        ms = (C.c_double * 7)()
        n = (C.c_uint64 * 7)()
        frames = C.c_uint64(0)
        check(_lib.gps_pass_timing_summary(self.handle, ms, n, C.byref(frames)), "timing summary")
        return {"frames": int(frames.value),
                "stages": {k: {"ms": float(ms[i]), "launches": int(n[i])} for i, k in enumerate(self.STAGES)}}

    def close(self) -> None:
        if self.handle:
            _lib.gps_pass_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- camera
def look_at(eye, center, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    out = np.empty(16, np.float32)
    e, ep = _f(eye)
    c, cp = _f(center)
    u, up_ = _f(up)
    _lib.gps_look_at(ep, cp, up_, out.ctypes.data_as(_fp))
    return out


def perspective(fovy_deg, aspect, near, far) -> np.ndarray:
    out = np.empty(16, np.float32)
    _lib.gps_perspective(fovy_deg, aspect, near, far, out.ctypes.data_as(_fp))
    return out


def mat4_mul(a, b) -> np.ndarray:
    out = np.empty(16, np.float32)
    a_, ap = _f(a)
    b_, bp = _f(b)
    _lib.gps_mat4_mul(ap, bp, out.ctypes.data_as(_fp))
    return out


def taa_reprojection(prev_vp, view, proj) -> np.ndarray:
    out = np.empty(16, np.float32)
    a, ap = _f(prev_vp)
    v, vp = _f(view)
    p, pp = _f(proj)
    _lib.gps_taa_reprojection(ap, vp, pp, out.ctypes.data_as(_fp))
    return out


def frame_params(eye, center, up=(0.0, 1.0, 0.0), fovy: float = 45.0, width: int = 1280, height: int = 720,
                 ss: int = 1, near: float = 0.2, far: float = 1e4, frame: int = 1, occlusion: bool = False,
                 hierarchical: bool = False, unbiased: bool = True, reduce_point_count: bool = True,
                 cull_small: bool = False, background=(0.0, 0.0, 0.0), min_reduce_count: int = 0,
                 spp: int = 0) -> FrameParams:
    """Frame inputs for a look-at camera (view = glm::lookAt, proj = glm::perspective)."""
    f = FrameParams()
    e, ep = _f(eye)
    c, cp = _f(center)
    u, up_ = _f(up)
    _lib.gps_frame_params_init(C.byref(f), ep, cp, up_, fovy, width, height, ss, near, far)
    f.frame = frame
    f.enable_occlusion_culling = int(occlusion)
    f.enable_hierarchical_culling = int(hierarchical)
    f.use_unbiased_2d_splatting = int(unbiased)
    f.reduce_point_count = int(reduce_point_count)
    f.cull_small_gaussians = int(cull_small)
    f.occlusion_culling_min_reduce_count = min_reduce_count
    f.samples_per_pixel = spp
    for i in range(3):
        f.background[i] = background[i]
    return f


# ----------------------------------------------------------------- resolve
def resolve_taa(cur_color: DeviceArray, cur_depth: DeviceArray, prev_color: DeviceArray, prev_depth: DeviceArray,
                reprojection, blend: float, width: int, height: int, frame_number: int, out_color: DeviceArray,
                out_rgb8: DeviceArray, stream: Optional[Stream] = None) -> None:
    m, mp = _f(reprojection)
    check(_lib.gps_resolve_taa(cur_color.ptr, cur_depth.ptr, prev_color.ptr, prev_depth.ptr, mp, blend, width,
                               height, frame_number, out_color.ptr, out_rgb8.ptr,
                               stream.handle if stream else None), "taa")


def resolve_progressive(color: DeviceArray, accum: DeviceArray, width: int, height: int, num_frames: int,
                        out_rgb8: DeviceArray, stream: Optional[Stream] = None) -> None:
    check(_lib.gps_resolve_progressive(color.ptr, accum.ptr, width, height, num_frames, out_rgb8.ptr,
                                       stream.handle if stream else None), "progressive")


# ----------------------------------------------------------------- serving
class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("max_width", C.c_int32), ("max_height", C.c_int32), ("max_spp", C.c_int32),
                ("max_supersampling", C.c_int32), ("tick_us", C.c_int32), ("device", C.c_int32),
                ("hbm_budget_bytes", C.c_uint64)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_int32), ("reserved", C.c_int32), ("frame", FrameParams), ("out_rgb8", C.c_void_p),
                ("out_color", C.c_void_p), ("arrival_s", C.c_double)]


class RequestResult(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("status", C.c_int32), ("slot", C.c_int32), ("arrival_s", C.c_double),
                ("admit_s", C.c_double), ("done_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("passes_created", C.c_uint64),
                ("passes_evicted", C.c_uint64), ("max_queue", C.c_uint64), ("max_batch", C.c_uint64),
                ("ticks_with_admissions", C.c_uint64), ("hbm_budget_bytes", C.c_uint64),
                ("hbm_used_bytes", C.c_uint64), ("clock_s", C.c_double), ("slot_busy_s", C.c_double)]


_sig("gps_pass_device_bytes", C.c_uint64, _vp)
_sig("gps_pass_estimate_bytes", C.c_uint64, C.POINTER(RenderSettingsC), _vp)
_sig("gps_server_create", C.c_int32, C.POINTER(ServerConfig), C.POINTER(_vp))
_sig("gps_server_add_scene", C.c_int32, _vp, _vp, C.POINTER(C.c_int32))
_sig("gps_server_prepare", C.c_int32, _vp)
_sig("gps_server_start", C.c_int32, _vp)
_sig("gps_server_clock", C.c_double, _vp)
_sig("gps_server_submit", C.c_int32, _vp, C.POINTER(Request), C.POINTER(C.c_uint64))
_sig("gps_server_poll", C.c_int32, _vp, C.POINTER(RequestResult), C.c_int32, C.c_int32)
_sig("gps_server_drain", C.c_int32, _vp, C.c_int32)
_sig("gps_server_get_stats", C.c_int32, _vp, C.POINTER(ServerStats))
_sig("gps_server_last_error", C.c_char_p, _vp)
_sig("gps_server_destroy", None, _vp)


class Server:
    """Serving loop over resident scenes (gps_server_*): independent frames of
    mixed size / spp, admitted by a native scheduler thread into HIP-stream
    slots as they arrive and memory allows."""

    def __init__(self, slots: int = 4, max_width: int = 1920, max_height: int = 1080, max_spp: int = 4,
                 max_supersampling: int = 2, tick_us: int = 200, hbm_budget_bytes: int = 0):
        cfg = ServerConfig(slots, max_width, max_height, max_spp, max_supersampling, tick_us, 0, hbm_budget_bytes)
        h = C.c_void_p()
        check(_lib.gps_server_create(C.byref(cfg), C.byref(h)), "server create")
        self.handle = h
        self.max_width, self.max_height = max_width, max_height
        self.scenes = []
        self._keep = {}

    def add_scene(self, scene: Scene) -> int:
        i = C.c_int32(0)
        check(_lib.gps_server_add_scene(self.handle, scene.handle, C.byref(i)), "add scene")
        self.scenes.append(scene)
        return int(i.value)

    def prepare(self) -> None:
        """Create every slot's pass for every scene now (within the budget)."""
        check(_lib.gps_server_prepare(self.handle), "server prepare")

    def start(self) -> None:
        check(_lib.gps_server_start(self.handle), "server start")

    def clock(self) -> float:
        return float(_lib.gps_server_clock(self.handle))

    def submit(self, scene: int, frame: FrameParams, arrival_s: float = -1.0, out_rgb8: Optional[np.ndarray] = None,
               out_color: Optional[np.ndarray] = None) -> int:
        r = Request()
        r.scene = scene
        r.frame = frame
        r.arrival_s = arrival_s
        n = frame.width * frame.height
        if out_rgb8 is not None:
            assert out_rgb8.dtype == np.uint8 and out_rgb8.size == 3 * n and out_rgb8.flags.c_contiguous
            r.out_rgb8 = out_rgb8.ctypes.data
        if out_color is not None:
            assert out_color.dtype == np.float32 and out_color.size == 3 * n and out_color.flags.c_contiguous
            r.out_color = out_color.ctypes.data
        t = C.c_uint64(0)
        check(_lib.gps_server_submit(self.handle, C.byref(r), C.byref(t)), "submit")
        self._keep[int(t.value)] = (out_rgb8, out_color)  # alive until the result is polled
        return int(t.value)

    def poll(self, max_results: int = 1024, wait_ms: int = 0) -> list:
        buf = (RequestResult * max_results)()
        n = _lib.gps_server_poll(self.handle, buf, max_results, wait_ms)
        if n < 0:
            raise GpsError(1, "poll")
        out = []
        for i in range(n):
            r = buf[i]
            self._keep.pop(int(r.ticket), None)
            out.append({"ticket": int(r.ticket), "status": int(r.status), "slot": int(r.slot),
                        "arrival_s": r.arrival_s, "admit_s": r.admit_s, "done_s": r.done_s})
        return out

    def drain(self, timeout_ms: int = 600_000) -> None:
        check(_lib.gps_server_drain(self.handle, timeout_ms), "drain")

    def stats(self) -> dict:
        s = ServerStats()
        check(_lib.gps_server_get_stats(self.handle, C.byref(s)), "server stats")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self) -> None:
        if self.handle:
            _lib.gps_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- batches
class BatchSettings(C.Structure):
    _fields_ = [("max_views", C.c_int32), ("max_width", C.c_int32), ("max_height", C.c_int32),
                ("max_spp", C.c_int32), ("max_supersampling", C.c_int32),
                ("fully_disable_hierarchical_culling", C.c_int32), ("gaussians_per_block", C.c_int32),
                ("reserved", C.c_int32), ("max_number_of_points", C.c_int64)]


class BatchTuning(C.Structure):
    _fields_ = [("splat_threads", C.c_int32), ("splat_points_per_thread", C.c_int32), ("use_graph", C.c_int32),
                ("variants", C.c_int32), ("variant_threads", C.c_int32 * 8),
                ("variant_points_per_thread", C.c_int32 * 8), ("variant_ms", C.c_float * 8),
                ("preprocess_views_per_workgroup", C.c_int32), ("pre_variants", C.c_int32),
                ("pre_variant_views", C.c_int32 * 4), ("pre_variant_ms", C.c_float * 4)]


_sig("gps_batch_create", C.c_int32, C.POINTER(BatchSettings), _vp, C.POINTER(_vp))
_sig("gps_batch_render", C.c_int32, _vp, C.c_int32, C.POINTER(FrameParams), C.POINTER(_vp), C.POINTER(_vp), _vp)
_sig("gps_batch_autotune", C.c_int32, _vp, C.c_int32, C.POINTER(FrameParams), C.POINTER(_vp), C.POINTER(_vp), _vp,
     C.c_int32, C.POINTER(BatchTuning))
_sig("gps_batch_set_tuning", C.c_int32, _vp, C.POINTER(BatchTuning))
_sig("gps_batch_set_splat_mode", C.c_int32, _vp, C.c_int32)
_sig("gps_batch_view_stats", C.c_int32, _vp, C.c_int32, C.POINTER(Stats), _vp)
_sig("gps_batch_device_bytes", C.c_uint64, _vp)
_sig("gps_batch_destroy", None, _vp)


class Batch:
    """Many views of one scene per launch sequence (gps_batch_*)."""

    def __init__(self, scene: Scene, max_views: int, max_width: int, max_height: int, max_spp: int = 4,
                 max_supersampling: int = 2, gaussians_per_block: int = 0, hierarchical: bool = True,
                 max_number_of_points: int = 250_000_000):
        st = BatchSettings(max_views, max_width, max_height, max_spp, max_supersampling, int(not hierarchical),
                           gaussians_per_block, 0, max_number_of_points)
        h = C.c_void_p()
        check(_lib.gps_batch_create(C.byref(st), scene.handle, C.byref(h)), "batch create")
        self.handle = h
        self.scene = scene

    @staticmethod
    def _arrays(frames, colors, depths):
        n = len(frames)
        fa = (FrameParams * n)(*frames)
        ca = (_vp * n)(*[c.ptr for c in colors])
        da = (_vp * n)(*[d.ptr for d in depths])
        return n, fa, ca, da

    def render(self, frames, colors, depths, stream: Optional[Stream] = None) -> None:
        n, fa, ca, da = self._arrays(frames, colors, depths)
        check(_lib.gps_batch_render(self.handle, n, fa, ca, da, stream.handle if stream else None), "batch render")

    def autotune(self, frames, colors, depths, stream: Optional[Stream] = None, repeats: int = 3) -> dict:
        n, fa, ca, da = self._arrays(frames, colors, depths)
        t = BatchTuning()
        check(_lib.gps_batch_autotune(self.handle, n, fa, ca, da, stream.handle if stream else None, repeats,
                                      C.byref(t)), "batch autotune")
        return {"splat_threads": t.splat_threads, "splat_points_per_thread": t.splat_points_per_thread,
                "preprocess_views_per_workgroup": t.preprocess_views_per_workgroup,
                "variants": [{"threads": t.variant_threads[i], "points_per_thread": t.variant_points_per_thread[i],
                              "ms": round(float(t.variant_ms[i]), 4)} for i in range(t.variants)],
                "preprocess_variants": [{"views_per_workgroup": t.pre_variant_views[i],
                                         "ms": round(float(t.pre_variant_ms[i]), 4)} for i in range(t.pre_variants)]}

    def set_tuning(self, threads: int, points_per_thread: int, graph: bool, views_per_workgroup: int = 0) -> None:
        """Splat variant (threads x points per thread), graph replay, and the
        preprocess split (views per workgroup, 0 = all)."""
        t = BatchTuning()
        t.splat_threads, t.splat_points_per_thread, t.use_graph = threads, points_per_thread, int(graph)
        t.preprocess_views_per_workgroup = views_per_workgroup
        check(_lib.gps_batch_set_tuning(self.handle, C.byref(t)), "batch tuning")

    def set_splat_mode(self, mode: int) -> None:
        """Pass.SPLAT_AUTO / SPLAT_POINT_PARALLEL / SPLAT_BINNED, per view."""
        check(_lib.gps_batch_set_splat_mode(self.handle, mode), "batch splat mode")

    def view_stats(self, view: int, stream: Optional[Stream] = None) -> dict:
        s = Stats()
        check(_lib.gps_batch_view_stats(self.handle, view, C.byref(s), stream.handle if stream else None), "stats")
        return s.as_dict()

    @property
    def device_bytes(self) -> int:
        return int(_lib.gps_batch_device_bytes(self.handle))

    def close(self) -> None:
        if self.handle:
            _lib.gps_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- camera paths
class CameraState(C.Structure):
    """gps_camera_state: one keyframe of a camera-path JSON file."""
    _fields_ = [("id", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("position", C.c_float * 3),
                ("rotation", C.c_float * 9), ("fy", C.c_float), ("fx", C.c_float), ("img_name", C.c_char * 64)]


_sig("gps_camera_path_load", C.c_int32, C.c_char_p, C.POINTER(CameraState), C.c_int32, C.POINTER(C.c_int32))
_sig("gps_camera_path_save", C.c_int32, C.c_char_p, C.POINTER(CameraState), C.c_int32)
_sig("gps_camera_fovy", C.c_float, C.POINTER(CameraState))
_sig("gps_camera_resolution", C.c_int32, C.POINTER(CameraState), C.c_int32, C.POINTER(C.c_int32),
     C.POINTER(C.c_int32))
_sig("gps_frame_params_from_camera", C.c_int32, C.POINTER(CameraState), C.c_int32, C.c_int32, C.c_float, C.c_int32,
     C.c_float, C.c_float, C.POINTER(FrameParams))
_sig("gps_camera_path_frame", C.c_int32, C.POINTER(CameraState), C.c_int32, C.c_float, C.c_float, C.c_int32,
     C.c_int32, C.c_int32, C.c_float, C.c_float, C.POINTER(FrameParams), C.POINTER(C.c_int32))


def camera_path_load(path: str) -> list:
    """Cameras of a camera-path JSON file, sorted by id."""
    n = C.c_int32()
    check(_lib.gps_camera_path_load(path.encode(), None, 0, C.byref(n)), f"camera path {path}")
    arr = (CameraState * max(1, n.value))()
    check(_lib.gps_camera_path_load(path.encode(), arr, n.value, C.byref(n)), f"camera path {path}")
    return list(arr)[:n.value]


def camera_path_save(path: str, cams) -> None:
    arr = (CameraState * max(1, len(cams)))(*cams)
    check(_lib.gps_camera_path_save(path.encode(), arr, len(cams)), f"camera path {path}")


def camera_fovy(cam: CameraState) -> float:
    return float(_lib.gps_camera_fovy(C.byref(cam)))


def camera_resolution(cam: CameraState, width: int) -> tuple:
    w, h = C.c_int32(), C.c_int32()
    check(_lib.gps_camera_resolution(C.byref(cam), width, C.byref(w), C.byref(h)), "camera resolution")
    return w.value, h.value


def frame_from_camera(cam: CameraState, width: int, height: int, fovy: float = 0.0, ss: int = 1,
                      near: float = 0.2, far: float = 1e4) -> FrameParams:
    """Frame of a camera keyframe as the viewer builds it (fovy <= 0: the camera's own)."""
    f = FrameParams()
    check(_lib.gps_frame_params_from_camera(C.byref(cam), width, height, fovy, ss, near, far, C.byref(f)), "camera")
    return f


def camera_path_frame(cams, time_s: float, width: int, height: int, fovy: float = 0.0, ss: int = 1,
                      near: float = 0.2, far: float = 1e4) -> tuple:
    """(frame, finished) of camera-path playback at time_s (keyframes 1/24 s apart)."""
    arr = (CameraState * len(cams))(*cams)
    f = FrameParams()
    fin = C.c_int32()
    check(_lib.gps_camera_path_frame(arr, len(cams), time_s, fovy, width, height, ss, near, far, C.byref(f),
                                     C.byref(fin)), "camera path frame")
    return f, bool(fin.value)


# ----------------------------------------------------------------- PLY files
class PlyInfo(C.Structure):
    _fields_ = [("num_gaussians", C.c_uint64), ("record_bytes", C.c_uint64), ("has_rest", C.c_int32),
                ("bvh_sorted", C.c_int32)]


_sig("gps_ply_info_read", C.c_int32, C.c_char_p, C.POINTER(PlyInfo))
_sig("gps_ply_load", C.c_int32, C.c_char_p, C.c_int32, C.c_uint64, _vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_uint64))
_sig("gps_ply_write", C.c_int32, C.c_char_p, C.POINTER(SceneDesc), C.c_int32)


def ply_info(path: str) -> dict:
    i = PlyInfo()
    check(_lib.gps_ply_info_read(os.fsencode(path), C.byref(i)), f"ply header {path}")
    return {"num_gaussians": int(i.num_gaussians), "record_bytes": int(i.record_bytes),
            "has_rest": bool(i.has_rest), "bvh_sorted": bool(i.bvh_sorted)}


def load_ply(path: str, sort_morton_order: bool = False) -> SceneArrays:
    """A 3DGS PLY file as SceneArrays (SH degree 3), with the reference
    loader's activations and optional load-time Morton order."""
    n = ply_info(path)["num_gaussians"]
    a = SceneArrays(np.empty((n, 3), np.float32), np.empty((n, 4), np.float32), np.empty((n, 3), np.float32),
                    np.empty((n,), np.float32), np.empty((n, 48), np.float32), 3)
    got = C.c_uint64(0)
    check(_lib.gps_ply_load(os.fsencode(path), int(sort_morton_order), n, _ptr(a.means), _ptr(a.rotations),
                            _ptr(a.scales), _ptr(a.opacities), _ptr(a.shs), C.byref(got)), f"ply load {path}")
    return a


def write_ply(path: str, s: SceneArrays, mark_sorted: bool = False) -> None:
    d = s.desc(FORMAT_FP32)
    check(_lib.gps_ply_write(os.fsencode(path), C.byref(d), int(mark_sorted)), f"ply write {path}")
