"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the benchmark's CPU-baseline leg,
never by the product path.  See hdsr_oracle.cpp for what each output restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

OUT = {
    "cache": 0, "depth": 1, "large_list": 2, "tile_counts": 3, "tile_offsets": 4, "entries_canon": 5,
    "small_tiles": 6, "large_tiles": 7, "large_bases": 8, "chunks": 9, "entries": 10, "keys": 11, "rgba": 12,
    "out_depth": 13, "gb_normal": 14, "gb_material": 15, "stats": 16, "rgba64": 17, "depth64": 18,
    "raw_inversions": 19, "diffrec": 20, "tile_loaded": 21,
}
STATS = ["refs", "small_tiles", "large_tiles", "chunks", "large_splats", "visible", "cap_refs_tripped",
         "cap_large_splats_tripped", "cap_chunks_tripped", "ref_cap"]
FLAG_GBUFFER, FLAG_FP64, FLAG_NO_COMPOSITE, FLAG_FP64_NOQUANT, FLAG_DIFF = 1, 2, 4, 8, 16
FLAG_F16 = 32  # composite with binary16 transmittance / colour (the f16 shader variant)
FLAG_REF_CAPS = 64  # the reference's WebGPU caps enforced (65,535 large splats, maxTotalTileRefs, 1,024 chunks)


class OrcSplats(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("logscale", C.c_void_p), ("rot", C.c_void_p), ("opacity", C.c_void_p),
                ("sorted", C.c_void_p), ("material", C.c_void_p), ("normals", C.c_void_p),
                ("normals_len", C.c_uint32), ("shdata", C.c_void_p), ("n", C.c_uint32), ("sh_degree", C.c_uint32)]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    lib.orc_run.restype = C.c_void_p
    lib.orc_run.argtypes = [C.POINTER(OrcSplats), C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]
    lib.orc_size.restype = C.c_uint64
    lib.orc_size.argtypes = [C.c_void_p, C.c_int]
    lib.orc_get.restype = C.c_int
    lib.orc_get.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64]
    lib.orc_free.argtypes = [C.c_void_p]
    lib.orc_free.restype = None
    for name in ("orc_expf", "orc_logf"):
        getattr(lib, name).restype = C.c_float
        getattr(lib, name).argtypes = [C.c_float]
    for name in ("orc_expf_n", "orc_logf_n", "orc_expf_gauss_n"):
        getattr(lib, name).restype = None
        getattr(lib, name).argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
    lib.orc_f16_bits_n.restype = None
    lib.orc_f16_bits_n.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
    lib.orc_q20_folded_n.restype = None
    lib.orc_q20_folded_n.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
    lib.orc_f16_bits.restype = C.c_uint32
    lib.orc_f16_bits.argtypes = [C.c_float]
    lib.orc_f16_to_f32.restype = C.c_float
    lib.orc_f16_to_f32.argtypes = [C.c_uint32]
    lib.orc_f16_round.restype = None
    lib.orc_f16_round.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
    lib.orc_mat4_inverse.restype = None
    lib.orc_mat4_inverse.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
    lib.orc_screen_space_normals.restype = None
    lib.orc_screen_space_normals.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                             C.c_uint32, C.c_float, C.c_float, C.c_void_p]
    lib.orc_bilateral_normals.restype = None
    lib.orc_bilateral_normals.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                          C.c_float, C.c_int, C.c_void_p]
    lib.orc_deferred_lighting.restype = None
    lib.orc_deferred_lighting.argtypes = [C.c_void_p] * 8 + [C.c_uint32, C.c_uint32, C.c_void_p]
    for name, nargs in (("orc_ggx_d", 2), ("orc_smith_g", 3), ("orc_fresnel", 2)):
        getattr(lib, name).restype = C.c_float
        getattr(lib, name).argtypes = [C.c_float] * nargs
    lib.orc_u32sat.restype = C.c_uint32
    lib.orc_u32sat.argtypes = [C.c_float]
    lib.orc_morton2.restype = C.c_uint32
    lib.orc_morton2.argtypes = [C.c_uint32, C.c_uint32]
    _lib = lib
    return lib


def morton2(x: int, y: int) -> int:
    """mortonEncode2D (the Morton slot of tile (x, y))."""
    return int(load().orc_morton2(int(x), int(y)))


def mat4_inverse(m16):
    """mat4Inverse (src/splatRenderer.ts:148-168): (fp32 inverse, True), or (None, False) when
    |det| < 1e-10."""
    m = np.ascontiguousarray(m16, np.float32)
    out = np.zeros(16, np.float32)
    ok = C.c_int(0)
    load().orc_mat4_inverse(m.ctypes.data, out.ctypes.data, C.byref(ok))
    return (out, True) if ok.value else (None, False)


def screen_space_normals(depth, src_words, inv16, cam3, mode=0, grazing_start=0.45, grazing_end=0.80):
    """gpu_screen_space_normals.wgsl over an (H, W) depth image and normal words."""
    d = np.ascontiguousarray(depth, np.float32)
    s = np.ascontiguousarray(src_words, np.uint32)
    H, W = d.shape
    inv = np.ascontiguousarray(inv16, np.float32)
    cam = np.ascontiguousarray(cam3, np.float32)
    out = np.zeros((H, W), np.uint32)
    load().orc_screen_space_normals(d.ctypes.data, s.ctypes.data, W, H, inv.ctypes.data, cam.ctypes.data, mode,
                                    grazing_start, grazing_end, out.ctypes.data)
    return out


def bilateral_normals(words, depth, spatial_sigma=3.0, depth_sigma=0.05, normal_sigma=1.5, radius=3):
    """bilateral_normal_filter.wgsl over (H, W) normal words and depth."""
    w = np.ascontiguousarray(words, np.uint32)
    d = np.ascontiguousarray(depth, np.float32)
    H, W = d.shape
    out = np.zeros((H, W), np.uint32)
    load().orc_bilateral_normals(w.ctypes.data, d.ctypes.data, W, H, spatial_sigma, depth_sigma, normal_sigma, radius,
                                 out.ctypes.data)
    return out


def deferred_lighting(rgba, depth, normal_words, material_words, params44, luts, ao=None, bloom=None):
    """gpu_deferred_lighting.wgsl (no environment map) over an (H, W) G-buffer: rgba (H, W, 4)
    fp32, depth (H, W), normal words (H, W), material words (H, W, 4), the 44-float Params block,
    the 3 x 256 curve tables; ao (H, W, 4) and bloom (ceil(H/2), ceil(W/2), 4) optional.
    Returns (H, W, 4) binary16 bits (uint16)."""
    c = np.ascontiguousarray(rgba, np.float32)
    d = np.ascontiguousarray(depth, np.float32)
    H, W = d.shape
    n = np.ascontiguousarray(normal_words, np.uint32)
    m = np.ascontiguousarray(material_words, np.uint32)
    P = np.ascontiguousarray(params44, np.float32)
    lt = np.ascontiguousarray(luts, np.float32)
    assert c.size == H * W * 4 and n.size == H * W and m.size == H * W * 4 and P.size == 44 and lt.size == 768
    a = None if ao is None else np.ascontiguousarray(ao, np.float32)
    b = None if bloom is None else np.ascontiguousarray(bloom, np.float32)
    if a is not None:
        assert a.size == H * W * 4
    if b is not None:
        assert b.size == ((H + 1) // 2) * ((W + 1) // 2) * 4
    out = np.zeros((H, W, 4), np.uint16)
    load().orc_deferred_lighting(c.ctypes.data, d.ctypes.data, n.ctypes.data, m.ctypes.data,
                                 None if a is None else a.ctypes.data, None if b is None else b.ctypes.data,
                                 lt.ctypes.data, P.ctypes.data, W, H, out.ctypes.data)
    return out


def oct_words_to_normals(words) -> np.ndarray:
    """Unpack oct-encoded normal words to (..., 3) float64 unit vectors (for checks)."""
    w = np.asarray(words, np.uint32)
    e = np.stack([(w & 0xFFFF).astype(np.uint16).view(np.float16), (w >> 16).astype(np.uint16).view(np.float16)],
                 -1).astype(np.float64)
    n = np.concatenate([e, (1 - np.abs(e[..., :1]) - np.abs(e[..., 1:]))], -1)
    neg = n[..., 2] < 0
    fx = (1 - np.abs(n[..., 1])) * np.where(n[..., 0] >= 0, 1, -1)
    fy = (1 - np.abs(n[..., 0])) * np.where(n[..., 1] >= 0, 1, -1)
    n[..., 0] = np.where(neg, fx, n[..., 0])
    n[..., 1] = np.where(neg, fy, n[..., 1])
    return n / np.linalg.norm(n, axis=-1, keepdims=True)


def normals_to_oct_words(n) -> np.ndarray:
    """octEncode + pack2x16float (round to nearest even) of (..., 3) unit vectors."""
    n = np.asarray(n, np.float64)
    s = np.abs(n).sum(-1, keepdims=True)
    p = n[..., :2] / s
    neg = n[..., 2] < 0
    q0 = (1 - np.abs(p[..., 1])) * np.where(p[..., 0] >= 0, 1, -1)
    q1 = (1 - np.abs(p[..., 0])) * np.where(p[..., 1] >= 0, 1, -1)
    p = np.stack([np.where(neg, q0, p[..., 0]), np.where(neg, q1, p[..., 1])], -1).astype(np.float32)
    h = p.astype(np.float16).view(np.uint16).astype(np.uint32)
    return h[..., 0] | (h[..., 1] << 16)


def f16_round(x) -> np.ndarray:
    """The oracle's binary16 rounding (f16r) of float64 values, as float64."""
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    load().orc_f16_round(x.ctypes.data, out.ctypes.data, x.size)
    return out


_DTYPES = {"rgba": np.float32, "out_depth": np.float32, "rgba64": np.float64, "depth64": np.float64,
           "diffrec": np.float32}


class OracleFrame:
    """All outputs of one oracle frame as numpy arrays."""

    def __init__(self, scene_arrays: dict, uniforms: np.ndarray, width: int, height: int, flags: int = 0):
        lib = load()
        self._arrays = {k: (v if v is None or k == "sh_degree" else np.ascontiguousarray(v))
                        for k, v in scene_arrays.items()}
        a = self._arrays
        nrm = a.get("normals")
        s = OrcSplats(a["positions"].ctypes.data, a["scales"].ctypes.data, a["rotations"].ctypes.data,
                      a["opacities"].ctypes.data, a["sorted"].ctypes.data, a["materials"].ctypes.data,
                      None if nrm is None else nrm.ctypes.data, 0 if nrm is None else int(nrm.size),
                      a["sh_data"].ctypes.data, int(a["positions"].shape[0]), int(a["sh_degree"]))
        u = np.ascontiguousarray(uniforms, dtype=np.float32)
        h = lib.orc_run(C.byref(s), u.ctypes.data, width, height, flags)
        if not h:
            raise RuntimeError("oracle rejected the input")
        self.out = {}
        try:
            for name, which in OUT.items():
                n = lib.orc_size(h, which)
                dt = _DTYPES.get(name, np.uint32)
                arr = np.empty(n // np.dtype(dt).itemsize, dt)
                if n:
                    lib.orc_get(h, which, arr.ctypes.data, n)
                self.out[name] = arr
        finally:
            lib.orc_free(h)
        self.stats = dict(zip(STATS, self.out["stats"].tolist()))
        self.width, self.height = width, height

    def __getitem__(self, k):
        return self.out[k]


def scene_arrays(scene, sorted_indices) -> dict:
    return {
        "positions": np.ascontiguousarray(scene.positions, np.float32),
        "scales": np.ascontiguousarray(scene.scales, np.float32),
        "rotations": np.ascontiguousarray(scene.rotations, np.float32),
        "opacities": np.ascontiguousarray(scene.opacities, np.float32),
        "materials": np.ascontiguousarray(scene.materials, np.uint32),
        "normals": None if scene.normals is None else np.ascontiguousarray(scene.normals, np.float32),
        "sh_data": np.ascontiguousarray(scene.sh_data(), np.float32),
        "sorted": np.ascontiguousarray(sorted_indices, np.uint32),
        "sh_degree": scene.sh_degree,
    }


def run(scene, sorted_indices, uniforms, width, height, flags=0) -> OracleFrame:
    return OracleFrame(scene_arrays(scene, sorted_indices), uniforms, width, height, flags)


def time_frame(scene, sorted_indices, uniforms, width, height, row_step: int) -> dict:
    """Times one single-threaded oracle frame, compositing every row_step-th tile row; returns
    seconds per phase and the whole-frame estimate (composite time scaled by row_step)."""
    lib = load()
    lib.orc_run_timed.restype = C.c_void_p
    lib.orc_run_timed.argtypes = [C.POINTER(OrcSplats), C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.POINTER(C.c_double)]
    a = scene_arrays(scene, sorted_indices)
    s = OrcSplats(a["positions"].ctypes.data, a["scales"].ctypes.data, a["rotations"].ctypes.data,
                  a["opacities"].ctypes.data, a["sorted"].ctypes.data, a["materials"].ctypes.data, None, 0,
                  a["sh_data"].ctypes.data, int(a["positions"].shape[0]), int(a["sh_degree"]))
    u = np.ascontiguousarray(uniforms, dtype=np.float32)
    t = (C.c_double * 3)()
    h = lib.orc_run_timed(C.byref(s), u.ctypes.data, width, height, row_step, t)
    if not h:
        raise RuntimeError("oracle rejected the input")
    lib.orc_free(h)
    proj, binsort, comp = t[0], t[1], t[2]
    return {"project_s": proj, "bin_sort_s": binsort, "composite_sample_s": comp, "row_step": row_step,
            "frame_s_est": proj + binsort + comp * row_step, "measured_s": proj + binsort + comp}


# ---- the GPU sortedIndices producer (f32 keys) ----------------------------------------------

def view_depth_keys_f32(positions, view16) -> np.ndarray:
    """createViewDepthSortInput (src/gpuSortPrototype.ts:120-140, 308-320): key of splat i is
    -(view-space z), z = ((V[2] x + V[6] y) + V[10] z) + V[14] in float64, stored as float32."""
    p = np.asarray(positions, dtype=np.float32).reshape(-1, 3).astype(np.float64)
    v = np.asarray(view16, dtype=np.float32).astype(np.float64)
    z = ((v[2] * p[:, 0] + v[6] * p[:, 1]) + v[10] * p[:, 2]) + v[14]
    return (-z).astype(np.float32)


def gpu_sort_order(keys) -> np.ndarray:
    """The order of the bitonic network of gpuSortPrototype.ts:4-55 in descending mode (the
    renderer's sortedIndices), every stage vectorised over its pairs: keys padded to a power of
    two with (-inf, 0xFFFFFFFF); stage (k, j) compare-exchanges (i, i + j) for i with bit j clear,
    ascending when i & k != 0, with should_swap's tie rule (smaller index first in both directions,
    so exact key ties come out in a network-dependent order).  Returns the first n indices."""
    keys = np.asarray(keys, dtype=np.float32)
    n = len(keys)
    P = 1
    while P < max(1, n):
        P *= 2
    k = np.full(P, -np.inf, np.float32)
    idx = np.full(P, 0xFFFFFFFF, np.uint64)
    k[:n] = keys
    idx[:n] = np.arange(n)
    t = np.arange(P // 2, dtype=np.int64)
    kk = 2
    while kk <= P:
        j = kk // 2
        while j >= 1:
            i = (t // j) * 2 * j + t % j
            q = i + j
            lk, rk, li, ri = k[i], k[q], idx[i], idx[q]
            asc = (i & kk) != 0
            tie = (lk == rk) & (li > ri)
            swap = np.where(asc, (lk > rk) | tie, (lk < rk) | tie)
            k[i], k[q] = np.where(swap, rk, lk), np.where(swap, lk, rk)
            idx[i], idx[q] = np.where(swap, ri, li), np.where(swap, li, ri)
            j //= 2
        kk *= 2
    return idx[:n].astype(np.uint32)


def bitonic_sort_restated(keys) -> np.ndarray:
    """simulateGpuSortPrototype (src/gpuSortPrototype.ts:142-163) pair by pair in the loop order of
    the reference, for small inputs (pins the vectorised gpu_sort_order)."""
    n = len(keys)
    padded = 1
    while padded < max(1, n):
        padded *= 2
    k = np.full(padded, -np.inf, np.float32)
    idx = np.full(padded, 0xFFFFFFFF, np.uint64)
    k[:n] = keys
    idx[:n] = np.arange(n)
    kk = 2
    while kk <= padded:
        j = kk // 2
        while j >= 1:
            for i in range(padded):
                partner = i ^ j
                if partner <= i or partner >= padded:
                    continue
                ascending = not ((i & kk) == 0)
                lk, rk, li, ri = k[i], k[partner], idx[i], idx[partner]
                if ascending:
                    swap = lk > rk or (lk == rk and li > ri)
                else:
                    swap = lk < rk or (lk == rk and li > ri)
                if swap:
                    k[i], k[partner] = rk, lk
                    idx[i], idx[partner] = ri, li
            j //= 2
        kk *= 2
    return idx[idx != 0xFFFFFFFF].astype(np.uint32)
