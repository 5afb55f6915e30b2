"""Camera conventions of the reference, for building frame uniforms.

- View matrix: column-major, right-handed, camera looks down -z (src/camera.ts:335-349).
- Projection: WebGPU [0, 1] depth, column-major (src/camera.ts:353-363); for pinhole cameras
  given by fx, fy the diagonal is P[0] = 2 fx / W, P[5] = 2 fy / H.
- viewProj = FLIP * P * V with FLIP = diag(1, -1, 1, 1) (src/realSmokeScene.ts:27-32,239-241).
- The 256-byte frame uniform block of writeTileSplatFrameUniforms
  (src/gpuTileSplatCompositor.ts:730-774).
"""
from __future__ import annotations

import math

import numpy as np

NEAR = 0.01
FAR = 100.0
FLIP = np.diag([1.0, -1.0, 1.0, 1.0])


def _normalize(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Column-major (16,) float32 view matrix with rows right / up / back (viewMatrixFromPose)."""
    eye = np.asarray(eye, dtype=np.float64)
    fz = _normalize(eye - np.asarray(target, dtype=np.float64))
    right = _normalize(np.cross(np.asarray(up, dtype=np.float64), fz))
    true_up = np.cross(fz, right)
    m = np.array([
        right[0], true_up[0], fz[0], 0.0,
        right[1], true_up[1], fz[1], 0.0,
        right[2], true_up[2], fz[2], 0.0,
        -right.dot(eye), -true_up.dot(eye), -fz.dot(eye), 1.0,
    ])
    return m.astype(np.float32)


def pinhole_projection(fx: float, fy: float, width: int, height: int, near=NEAR, far=FAR) -> np.ndarray:
    """Column-major (16,) float32 perspective matrix with focal lengths in pixels."""
    range_inv = 1.0 / (near - far)
    m = np.zeros(16)
    m[0] = 2.0 * fx / width
    m[5] = 2.0 * fy / height
    m[10] = far * range_inv
    m[11] = -1.0
    m[14] = near * far * range_inv
    return m.astype(np.float32)


def col_major(m16) -> np.ndarray:
    """(16,) column-major -> (4, 4) matrix M with M[row, col]."""
    return np.asarray(m16, dtype=np.float64).reshape(4, 4).T


def to_col_major(M) -> np.ndarray:
    return np.asarray(M, dtype=np.float64).T.reshape(16)


def view_projection(proj16, view16) -> np.ndarray:
    vp = FLIP @ col_major(proj16) @ col_major(view16)
    return to_col_major(vp).astype(np.float32)


def orbit_eye(radius: float, azimuth: float, elevation: float, target=(0.0, 0.0, 0.0)) -> np.ndarray:
    t = np.asarray(target, dtype=np.float64)
    return t + radius * np.array([
        math.cos(elevation) * math.sin(azimuth),
        math.sin(elevation),
        math.cos(elevation) * math.cos(azimuth),
    ])


class Camera:
    """A pinhole camera with everything the frame uniforms need."""

    def __init__(self, eye, target, fx, fy, width, height, up=(0.0, 1.0, 0.0)):
        self.eye = np.asarray(eye, dtype=np.float64)
        self.target = np.asarray(target, dtype=np.float64)
        self.fx, self.fy = float(fx), float(fy)
        self.width, self.height = int(width), int(height)
        self.view = look_at(self.eye, self.target, up)
        self.proj = pinhole_projection(fx, fy, width, height)
        self.view_proj = view_projection(self.proj, self.view)
        self.camera_pos = self.eye.astype(np.float32)

    def __repr__(self):
        return f"Camera(eye={self.eye.tolist()}, {self.width}x{self.height}, fx={self.fx})"


def write_frame_uniforms(view_proj, plan, splat_scale=1.0, sh_degree=0, camera_pos=None, view_matrix=None,
                         proj_matrix=None, normal_matrix=None) -> np.ndarray:
    """writeTileSplatFrameUniforms (gpuTileSplatCompositor.ts:730-774): a (64,) float32 array whose
    bytes are the 256-byte uniform block."""
    t = np.zeros(64, dtype=np.float32)
    u = t.view(np.uint32)
    t[0:16] = np.asarray(view_proj, dtype=np.float32)
    t[16] = plan["viewportWidth"]
    t[17] = plan["viewportHeight"]
    t[18] = plan["tileSizePx"]
    t[19] = 0.0
    u[20] = plan["tileColumns"]
    u[21] = plan["tileRows"]
    u[22] = plan["splatCount"]
    u[23] = plan["maxTotalTileRefs"]
    t[24] = splat_scale
    u[25] = sh_degree
    if camera_pos is not None:
        t[28:31] = np.asarray(camera_pos, dtype=np.float32)
    if view_matrix is not None:
        t[32:48] = np.asarray(view_matrix, dtype=np.float32)
    if proj_matrix is not None:
        p = np.asarray(proj_matrix, dtype=np.float32)
        t[48] = np.float32(float(plan["viewportWidth"]) * float(p[0]) * 0.5)
        t[49] = np.float32(float(plan["viewportHeight"]) * float(p[5]) * 0.5)
    nm = normal_matrix if normal_matrix is not None else [1, 0, 0, 0, 1, 0, 0, 0, 1]
    t[52:55] = nm[0:3]
    t[56:59] = nm[3:6]
    t[60:63] = nm[6:9]
    return t


def frame_uniforms_for(camera: Camera, plan, sh_degree: int, splat_scale=1.0) -> np.ndarray:
    return write_frame_uniforms(camera.view_proj, plan, splat_scale, sh_degree, camera.camera_pos, camera.view,
                                camera.proj)
