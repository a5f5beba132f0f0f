"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE: only tests/, bench.py's CPU-baseline leg and the smoke
check import this module, and only as the checker.  The oracle is a plain C++
transcription of the reference's per-frame hot path (see gps_oracle.cpp).
It shares the POD structs of include/gps.h with the product so both can be
driven with identical inputs.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build() -> None:
    subprocess.run(["make", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


if not os.path.exists(LIB_PATH):
    build()

_lib = C.CDLL(LIB_PATH)

# The product binding defines the shared structs; import it for the layouts
# only (no GPU call happens here).
sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "gaussian-point-splatting_amd"))
import gps_amd  # noqa: E402
from gps_amd import FrameParams, RenderSettingsC, SceneDesc, Stats  # noqa: E402,F401


class HalfFp8Scene(C.Structure):
    _fields_ = [("num_gaussians", C.c_uint64), ("sh_degree", C.c_int32), ("reserved", C.c_int32),
                ("means", C.c_void_p), ("rotations", C.c_void_p), ("scales", C.c_void_p),
                ("opacities", C.c_void_p), ("sh_dc", C.c_void_p), ("sh_rest", C.c_void_p),
                ("sh_scale", C.c_void_p)]


class QuantizedScene(C.Structure):
    _fields_ = [("num_gaussians", C.c_uint64), ("bounds", gps_amd.QuantizedBounds), ("geometry", C.c_void_p),
                ("opacity", C.c_void_p), ("sh_dc", C.c_void_p), ("sh_codes", C.c_void_p)]


_vp = C.c_void_p


def _sig(name, res, *args):
    fn = getattr(_lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_sig("orc_pass_create", _vp, C.POINTER(RenderSettingsC), C.POINTER(SceneDesc), C.c_int32)
_sig("orc_pass_create_half_fp8", _vp, C.POINTER(RenderSettingsC), C.POINTER(HalfFp8Scene), C.c_int32)
_sig("orc_pass_create_quantized", _vp, C.POINTER(RenderSettingsC), C.POINTER(QuantizedScene), C.c_int32)
_sig("orc_quantized_decode", None, C.POINTER(QuantizedScene), C.c_uint64, _vp)
_sig("orc_pass_destroy", None, _vp)
_sig("orc_pass_set_threads", None, _vp, C.c_int32)
_sig("orc_pass_render", C.c_int32, _vp, C.POINTER(FrameParams), _vp, _vp)
_sig("orc_pass_stats", C.c_int32, _vp, C.POINTER(Stats))
_sig("orc_pass_debug_size", C.c_uint64, _vp, C.c_int32)
_sig("orc_pass_debug_copy", C.c_int32, _vp, C.c_int32, _vp, C.c_uint64)
_sig("orc_pass_stage_ms", None, _vp, C.POINTER(C.c_double))
_sig("orc_pass_biased_events", None, _vp, _vp)
_sig("orc_pass_rejects", None, _vp, _vp)
_sig("orc_biased_sample", C.c_uint64, _vp, C.c_uint64, C.c_int32, C.c_int32)
_sig("orc_pass_cull", C.c_int32, _vp, C.POINTER(FrameParams))
_sig("orc_pass_set_state", C.c_int32, _vp, _vp, C.c_uint64, _vp, C.c_uint64)
_sig("orc_project", C.c_int32, _vp, C.POINTER(FrameParams), C.c_int32, _vp)
_sig("orc_project_ids", C.c_int32, _vp, C.POINTER(FrameParams), _vp, C.c_uint64, _vp)
_sig("orc_resolve_taa", None, _vp, _vp, _vp, _vp, C.POINTER(C.c_float), C.c_float, C.c_int32, C.c_int32, C.c_int32,
     _vp, _vp)
_sig("orc_resolve_progressive", None, _vp, _vp, C.c_int32, C.c_int32, C.c_int32, _vp)
_sig("orc_rng_probe", None, C.c_uint32, C.c_uint32, C.c_int32, _vp)
_sig("orc_dilog", C.c_float, C.c_float)
_sig("orc_inv_dilog", C.c_float, C.c_float)
_sig("orc_corrected_box_muller", None, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float))
_sig("orc_poisson", C.c_uint32, C.c_uint32, C.c_uint32, C.c_float)
_sig("orc_stochastic_round", C.c_int32, C.c_float, C.c_float)
_sig("orc_pos_to_memory", C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32)
_sig("orc_depth_bits", C.c_uint64, C.c_float, C.c_float, C.c_float)

_sig("orc_detmath", None, C.c_int32, _vp, _vp, C.c_uint64)
_sig("orc_sh_eval", None, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32, C.POINTER(C.c_double))
_sig("orc_sh_eval_compact", None, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32, C.POINTER(C.c_double))

lib = _lib


def biased_sample(word: int, value: int, precheck: bool = True, add_if_equal: bool = True):
    """Apply one biased-splat sample to a framebuffer word.  precheck: the
    splat's pre-check + atomicAddColor (gaussian_point_splatting.cu:130-138);
    otherwise atomicAddColor alone (post_processing_kernels.cuh:13-42).
    Returns (new word, atomicAddColor's result or 1)."""
    w = np.array([word], np.uint64)
    r = _lib.orc_biased_sample(_ptr(w), int(value), int(precheck), int(add_if_equal))
    return int(w[0]), int(r)


def detmath(kind: int, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    _lib.orc_detmath(kind, _ptr(x), _ptr(out), x.size)
    return out


def sh_eval(direction, coeffs48, fp64: bool, compact: bool = False) -> np.ndarray:
    """computeColorFromSHFloat (or, compact, computeColorFromCompactSH) at a view direction."""
    d = np.ascontiguousarray(direction, np.float32)
    c = np.ascontiguousarray(coeffs48, np.float32)
    out = (C.c_double * 3)()
    fn = _lib.orc_sh_eval_compact if compact else _lib.orc_sh_eval
    fn(d.ctypes.data_as(C.POINTER(C.c_float)), c.ctypes.data_as(C.POINTER(C.c_float)), int(fp64), out)
    return np.array(list(out))


def quantized_scene(enc: dict) -> QuantizedScene:
    """The oracle's view of gps_amd.encode_quantized output (arrays stay owned by `enc`)."""
    q = QuantizedScene()
    q.num_gaussians = enc["geometry"].shape[0]
    q.bounds = enc["bounds"]
    for k in ("geometry", "opacity", "sh_dc", "sh_codes"):
        setattr(q, k, _ptr(enc[k]))
    return q


def quantized_decode(enc: dict) -> np.ndarray:
    """(N, 59): decoded mean, rotation (stored order), scale, opacity, 16 x rgb SH coefficients."""
    q = quantized_scene(enc)
    n = enc["geometry"].shape[0]
    out = np.empty((n, 59), np.float32)
    for g in range(n):
        _lib.orc_quantized_decode(C.byref(q), g, _ptr(out[g]))
    return out

_DTYPE = {gps_amd.DBG_FRAMEBUFFER: np.uint64, gps_amd.DBG_HIZ: np.float32}

_sig("orc_bvh_fragments", C.c_int, _vp, _vp, _vp, C.c_uint64, _vp)
_sig("orc_bvh_leaves", C.c_int64, _vp, C.c_uint64, _vp, _vp)
_sig("orc_bvh_order", C.c_int, _vp, _vp, _vp, C.c_uint64, _vp)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def bvh_fragments(means, rotations, scales) -> np.ndarray:
    """tinybvh input of the load-time reorder: 3 vec4 vertices per Gaussian."""
    means, rotations, scales = _f32(means), _f32(rotations), _f32(scales)
    n = len(means)
    v = np.empty((n, 12), np.float32)
    _lib.orc_bvh_fragments(_ptr(means), _ptr(rotations), _ptr(scales), n, _ptr(v))
    return v


def bvh_leaves(verts: np.ndarray):
    """(first, count) of the leaves of the tinybvh build, left-first order."""
    verts = _f32(verts)
    n = len(verts)
    first = np.empty(n, np.uint32)
    count = np.empty(n, np.uint32)
    k = _lib.orc_bvh_leaves(_ptr(verts), n, _ptr(first), _ptr(count))
    assert k >= 0
    return first[:k], count[:k]


def bvh_order(means, rotations, scales) -> np.ndarray:
    """perm[new] = old of the reference's BVH reorder."""
    means, rotations, scales = _f32(means), _f32(rotations), _f32(scales)
    n = len(means)
    perm = np.empty(n, np.uint32)
    assert _lib.orc_bvh_order(_ptr(means), _ptr(rotations), _ptr(scales), n, _ptr(perm)) == 0
    return perm


REF_TINYBVH_PATH = os.path.join(_HERE, "_ref", "libtinybvh_ref.so")


def tinybvh_ref():
    """The reference's own tinybvh built into _ref/ (None where the reference
    tree was absent at build time)."""
    if not os.path.exists(REF_TINYBVH_PATH):
        return None
    lib = C.CDLL(REF_TINYBVH_PATH)
    lib.ref_tinybvh_leaf_order.restype = C.c_int
    lib.ref_tinybvh_leaf_order.argtypes = [_vp, C.c_uint32, _vp]
    return lib


def tinybvh_ref_order(lib, verts: np.ndarray) -> np.ndarray:
    """perm[new] = old from tinybvh's leaves and the reference's reorder:
    new[leafOrder[m]] = old[m] (gaussian_bvh.cu:47-84)."""
    verts = _f32(verts)
    n = len(verts)
    lo = np.empty(n, np.uint32)
    assert lib.ref_tinybvh_leaf_order(_ptr(verts), n, _ptr(lo)) == 0
    perm = np.empty(n, np.uint32)
    perm[lo] = np.arange(n, dtype=np.uint32)
    return perm


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class OraclePass:
    """CPU transcription of GaussianPointSplatting over a host scene."""

    def __init__(self, settings: "gps_amd.RenderSettings", scene: "gps_amd.SceneArrays", fmt: int = 0,
                 threads: int = 1):
        self._s = settings.c()
        self._keep = scene
        if fmt == gps_amd.FORMAT_FP32:
            d = scene.desc(fmt)
            self.handle = _lib.orc_pass_create(C.byref(self._s), C.byref(d), threads)
        elif fmt == gps_amd.FORMAT_QUANTIZED:
            self._enc = gps_amd.encode_quantized(scene)
            q = quantized_scene(self._enc)
            self.handle = _lib.orc_pass_create_quantized(C.byref(self._s), C.byref(q), threads)
        else:
            enc = gps_amd.encode_half_fp8(scene)
            self._enc = enc
            h = HalfFp8Scene()
            h.num_gaussians = scene.n
            h.sh_degree = scene.sh_degree
            h.means = _ptr(scene.means)
            for k in ("rotations", "scales", "opacities", "sh_dc", "sh_rest", "sh_scale"):
                setattr(h, k, _ptr(enc[k]))
            self.handle = _lib.orc_pass_create_half_fp8(C.byref(self._s), C.byref(h), threads)
        if not self.handle:
            raise RuntimeError("oracle pass create failed")
        self.width = settings.width
        self.height = settings.height

    @classmethod
    def from_half_fp8(cls, settings: "gps_amd.RenderSettings", means: np.ndarray, enc: dict, sh_degree: int,
                      threads: int = 1) -> "OraclePass":
        """A pass over a scene already encoded in the compact format (the
        arrays gps_amd.encode_half_fp8 returns, concatenated over chunks), so
        a large scene need not be held in fp32 on the host."""
        self = cls.__new__(cls)
        self._s = settings.c()
        self._keep = (means, enc)
        self._enc = enc
        h = HalfFp8Scene()
        h.num_gaussians = means.shape[0]
        h.sh_degree = sh_degree
        h.means = _ptr(means)
        for k in ("rotations", "scales", "opacities", "sh_dc", "sh_rest", "sh_scale"):
            setattr(h, k, _ptr(enc[k]))
        self.handle = _lib.orc_pass_create_half_fp8(C.byref(self._s), C.byref(h), threads)
        if not self.handle:
            raise RuntimeError("oracle pass create failed")
        self.width = settings.width
        self.height = settings.height
        return self

    def set_threads(self, n: int) -> None:
        _lib.orc_pass_set_threads(self.handle, n)

    def render(self, frame: FrameParams):
        color = np.zeros((self.height, self.width, 3), np.float32)
        depth = np.zeros((self.height, self.width), np.float32)
        st = _lib.orc_pass_render(self.handle, C.byref(frame), _ptr(color), _ptr(depth))
        if st != 0:
            raise RuntimeError(f"oracle render failed: {st}")
        return color, depth

    def stats(self) -> dict:
        s = Stats()
        _lib.orc_pass_stats(self.handle, C.byref(s))
        return s.as_dict()

    def biased_events(self) -> dict:
        """Order-sensitive events of the last render's biased splat: samples of
        colour 0 past the pre-check, colour sums carried out of bit 35."""
        out = np.zeros(2, np.uint64)
        _lib.orc_pass_biased_events(self.handle, _ptr(out))
        return {"zero_colour": int(out[0]), "wrapped": int(out[1])}

    def rejects(self) -> list:
        """Per render pass of the last frame, {reason: Gaussians} in the
        reasons of gps_amd.REJECT_REASONS."""
        out = np.zeros(18, np.uint64)
        _lib.orc_pass_rejects(self.handle, _ptr(out))
        return [dict(zip(gps_amd.REJECT_REASONS, map(int, out[9 * s:9 * s + 9]))) for s in range(2)]

    def stage_ms(self) -> list:
        out = (C.c_double * 8)()
        _lib.orc_pass_stage_ms(self.handle, out)
        return list(out)

    def debug(self, which: int) -> np.ndarray:
        n = int(_lib.orc_pass_debug_size(self.handle, which))
        out = np.empty((n,), _DTYPE.get(which, np.uint32))
        st = _lib.orc_pass_debug_copy(self.handle, which, _ptr(out), n)
        if st != 0:
            raise RuntimeError(f"oracle debug copy {which} failed: {st}")
        return out

    def set_state(self, hiz: np.ndarray, mask_other: np.ndarray) -> None:
        """Carried frame state (Hi-Z chain, the valid mask the next first pass
        leaves in place), as read from another pass after the frame before."""
        hiz = np.ascontiguousarray(hiz, np.float32)
        mask_other = np.ascontiguousarray(mask_other, np.uint32)
        if _lib.orc_pass_set_state(self.handle, _ptr(hiz), hiz.size, _ptr(mask_other), mask_other.size) != 0:
            raise RuntimeError("oracle set_state: sizes do not match the pass")

    def cull(self, frame: FrameParams) -> np.ndarray:
        """Block-cull mask for `frame` against the Hi-Z chain held now."""
        if _lib.orc_pass_cull(self.handle, C.byref(frame)) != 0:
            raise RuntimeError("oracle cull failed (block culling disabled?)")
        return self.debug(gps_amd.DBG_CULL_MASK)

    def project(self, frame: FrameParams, fp64: bool) -> np.ndarray:
        n = self._keep.n
        out = np.zeros((n, 16), np.float64)
        _lib.orc_project(self.handle, C.byref(frame), int(fp64), _ptr(out))
        return out

    def project_ids(self, frame: FrameParams, ids: np.ndarray) -> np.ndarray:
        """project(frame, fp64=False) rows of the Gaussians `ids` only."""
        ids = np.ascontiguousarray(ids, np.uint32)
        out = np.zeros((ids.size, 16), np.float64)
        _lib.orc_project_ids(self.handle, C.byref(frame), _ptr(ids), ids.size, _ptr(out))
        return out

    def close(self) -> None:
        if self.handle:
            _lib.orc_pass_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def resolve_taa(cur, cur_depth, prev, prev_depth, reprojection, blend, width, height, frame_number):
    out = np.empty_like(cur)
    rgb = np.empty(cur.shape, np.uint8)
    m = np.ascontiguousarray(reprojection, np.float32)
    _lib.orc_resolve_taa(_ptr(cur), _ptr(cur_depth), _ptr(prev), _ptr(prev_depth),
                         m.ctypes.data_as(C.POINTER(C.c_float)), blend, width, height, frame_number, _ptr(out),
                         _ptr(rgb))
    return out, rgb


def resolve_progressive(color, accum, width, height, num_frames):
    color = np.ascontiguousarray(color, np.float32).copy()
    accum = np.ascontiguousarray(accum, np.float32).copy()
    rgb = np.empty(color.shape, np.uint8)
    _lib.orc_resolve_progressive(_ptr(color), _ptr(accum), width, height, num_frames, _ptr(rgb))
    return color, accum, rgb
