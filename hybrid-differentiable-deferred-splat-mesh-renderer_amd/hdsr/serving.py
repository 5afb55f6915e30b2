"""Serving form: a continuously batching request server over the C ABI (hdsr_server_*).

    srv = Server(hbm_budget_bytes=200 << 30, max_views_per_batch=64, slots=2, use_graphs=True)
    srv.add_scene(7, scene)                       # hdsr.scenes.Scene, host arrays kept alive here
    t = srv.submit(7, camera)                     # camera: hdsr.camera.Camera
    srv.wait(t); img = srv.read_rgba(t); srv.release(t)

The server owns device memory for resident scenes (LRU-evicted under the budget) and for the
outputs of in-flight requests; a worker thread in the library batches queued requests of any
scene / resolution mix into one launch per pipeline stage.  The plain per-request drop-in is the
same server with max_views_per_batch=1 and one slot.
"""
from __future__ import annotations

import ctypes as C
import threading
import time

import numpy as np

from . import _lib, compositor
from .camera import frame_uniforms_for


class ServerError(_lib.HdsrError):
    pass


class Server:
    def __init__(self, hbm_budget_bytes=0, max_views_per_batch=64, slots=2, use_graphs=True,
                 max_splats_per_batch=0):
        self._lib = _lib.load()
        cfg = _lib.ServerConfig(int(hbm_budget_bytes), int(max_splats_per_batch), int(max_views_per_batch),
                                int(slots), 1 if use_graphs else 0, 0)
        h = C.c_void_p()
        rc = self._lib.hdsr_server_create(C.byref(cfg), C.byref(h))
        if rc:
            raise ServerError(rc, "hdsr_server_create failed")
        self.handle = h
        self._keep = {}
        self._meta = {}

    def _check(self, rc):
        if rc:
            m = self._lib.hdsr_server_last_error(self.handle)
            raise ServerError(rc, m.decode() if m else "")

    def close(self):
        if self.handle:
            self._lib.hdsr_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, scene_id: int, scene, sorted_indices=None):
        arrays = [np.ascontiguousarray(scene.positions, np.float32), np.ascontiguousarray(scene.scales, np.float32),
                  np.ascontiguousarray(scene.rotations, np.float32),
                  np.ascontiguousarray(scene.opacities, np.float32),
                  np.ascontiguousarray(scene.materials, np.uint32), np.ascontiguousarray(scene.sh_data(), np.float32)]
        srt = None if sorted_indices is None else np.ascontiguousarray(sorted_indices, np.uint32)
        s = _lib.Splats(arrays[0].ctypes.data, arrays[1].ctypes.data, arrays[2].ctypes.data, arrays[3].ctypes.data,
                        arrays[4].ctypes.data, None, 0, arrays[5].ctypes.data,
                        None if srt is None else srt.ctypes.data, scene.count, scene.sh_degree, _lib.HDSR_MEM_HOST,
                        arrays[5].size)
        self._check(self._lib.hdsr_server_add_scene(self.handle, scene_id, C.byref(s)))
        self._keep[scene_id] = (arrays, srt)
        self._meta[scene_id] = (scene.count, scene.sh_degree)

    def uniforms(self, scene_id: int, cam) -> np.ndarray:
        n, deg = self._meta[scene_id]
        plan = compositor.planTileSplatCompositor({"viewportWidth": cam.width, "viewportHeight": cam.height,
                                                   "tileSizePx": 16, "splatCount": n})
        return frame_uniforms_for(cam, plan, deg)

    def submit_uniforms(self, scene_id: int, width: int, height: int, uniforms: np.ndarray) -> int:
        u = np.ascontiguousarray(uniforms, np.float32)
        req = _lib.Request(scene_id, width, height, 0, u.ctypes.data)
        t = C.c_uint64()
        self._check(self._lib.hdsr_server_submit(self.handle, C.byref(req), C.byref(t)))
        return t.value

    def submit(self, scene_id: int, cam) -> int:
        return self.submit_uniforms(scene_id, cam.width, cam.height, self.uniforms(scene_id, cam))

    def wait(self, ticket: int, timeout_s: float = 0.0):
        self._check(self._lib.hdsr_server_wait(self.handle, ticket, timeout_s))

    def result(self, ticket: int) -> dict:
        r = _lib.ResultInfo()
        self._check(self._lib.hdsr_server_result(self.handle, ticket, C.byref(r)))
        return {"done": bool(r.done), "width": r.width, "height": r.height, "batch_views": r.batch_views,
                "latency_ms": r.latency_ms}

    def read_rgba(self, ticket: int) -> np.ndarray:
        info = self.result(ticket)
        out = np.empty((info["height"], info["width"], 4), np.float32)
        self._check(self._lib.hdsr_server_read(self.handle, ticket, _lib.OUT_RGBA, out.ctypes.data, out.nbytes))
        return out

    def read_depth(self, ticket: int) -> np.ndarray:
        info = self.result(ticket)
        out = np.empty((info["height"], info["width"]), np.float32)
        self._check(self._lib.hdsr_server_read(self.handle, ticket, _lib.OUT_DEPTH, out.ctypes.data, out.nbytes))
        return out

    def release(self, ticket: int):
        self._check(self._lib.hdsr_server_release(self.handle, ticket))

    def hold(self, on: bool = True):
        """While held, no new batch is formed; a burst submitted meanwhile shares launches
        (wait and drain release the hold)."""
        self._check(self._lib.hdsr_server_hold(self.handle, 1 if on else 0))

    def drain(self):
        self._check(self._lib.hdsr_server_drain(self.handle))

    def stats(self) -> dict:
        s = _lib.ServerStats()
        self._check(self._lib.hdsr_server_get_stats(self.handle, C.byref(s)))
        return {name: getattr(s, name) for name, _ in _lib.ServerStats._fields_}

    def reset_stats(self):
        self._check(self._lib.hdsr_server_reset_stats(self.handle))


def poisson_run(server: Server, requests, rate_hz: float, duration_s: float, seed=0, max_requests=None) -> dict:
    """Submits requests (a list of (scene_id, w, h, uniforms)) cyclically as a Poisson stream of
    the given rate for duration_s, releasing results as they complete; returns the server stats
    (p50/p99 latency, frames/s over the busy interval)."""
    rng = np.random.default_rng(seed)
    server.drain()
    server.reset_stats()
    tickets = []
    lock = threading.Lock()
    done_flag = [False]

    def releaser():
        i = 0
        while True:
            with lock:
                if i < len(tickets):
                    t = tickets[i]
                elif done_flag[0]:
                    return
                else:
                    t = None
            if t is None:
                time.sleep(0.0005)
                continue
            server.wait(t)
            server.release(t)
            i += 1

    th = threading.Thread(target=releaser, daemon=True)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    k = 0
    while True:
        now = time.perf_counter()
        if now - t0 >= duration_s or (max_requests and k >= max_requests):
            break
        if now < t_next:
            if t_next - now > 0.0005:
                time.sleep(min(t_next - now - 0.0003, 0.01))
            continue
        sid, w, h, u = requests[k % len(requests)]
        t = server.submit_uniforms(sid, w, h, u)
        with lock:
            tickets.append(t)
        k += 1
        t_next += rng.exponential(1.0 / rate_hz)
    with lock:
        done_flag[0] = True
    th.join()
    server.drain()
    st = server.stats()
    st["submitted"] = k
    st["offered_hz"] = rate_hz
    st["fps"] = st["frames"] / st["busy_s"] if st["busy_s"] > 0 else 0.0
    return st
