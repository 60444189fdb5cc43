"""Serving loop (libavbd_hip.so's avbd_server_*): step/rollout requests for many worlds, admitted as they arrive, with
world state resident in HBM under a byte budget and least recently used idle worlds evicted to host memory and restored
exactly. Running requests overlap on the device, one HIP stream per world, each turn of the loop launching the next
step of every running request. Requests of one world run in submission order, so its result does not depend on what
else is being served.

    server = avbd.serve.Server(hbmBudget=64 << 30, maxActive=32)
    w = server.addWorld(scene)
    r = server.submit(w, steps=60)
    server.wait(r); bodies = server.readBodies(w)
"""
import ctypes

import numpy as np

from . import _native
from ._native import check, ptr
from .scene import AvbdScene3D
from .world import _scene_desc


class Server:
    def __init__(self, hbmBudget=200 << 30, maxActive=32, lookahead=2, device=0, colorCap=16, colorRounds=16,
                 broadphase="auto", smallWorldBodies=0, smallWorldBlock=1024, largeWorldBodies=0, largeWorldShare=4):
        lib = _native.load()
        o = _native.ServerOptions()
        o.hbm_budget_bytes = int(hbmBudget)
        o.max_active = int(maxActive)
        o.lookahead = int(lookahead)
        o.device = int(device)
        o.color_cap = int(colorCap)
        o.color_rounds = int(colorRounds)
        o.use_graph = 1
        if broadphase not in ("auto", "grid", "hploc"):
            raise ValueError(f"Unknown broadphase: {broadphase}")
        o.broadphase = {"auto": 0, "grid": 1, "hploc": 2}[broadphase]
        if smallWorldBlock not in (128, 256, 512, 1024):
            raise ValueError(f"smallWorldBlock must be 128, 256, 512 or 1024, not {smallWorldBlock}")
        o.small_world_bodies = int(smallWorldBodies)
        o.small_world_block = int(smallWorldBlock)
        o.large_world_bodies = int(largeWorldBodies)
        o.large_world_share = int(largeWorldShare)
        h = ctypes.c_void_p()
        check(lib.avbd_server_create(ctypes.byref(o), ctypes.byref(h)))
        self._lib, self._h, self.destroyed = lib, h, False
        self._bodies = {}

    def _alive(self):
        if self.destroyed:
            raise RuntimeError("The server has been destroyed")

    def addWorld(self, scene):
        """Register a built scene; returns its world id."""
        self._alive()
        if not isinstance(scene, AvbdScene3D):
            raise TypeError("An AVBD scene builder is required")
        scene.assertEditable()
        keep = []
        desc = _scene_desc(scene, keep)
        wid = ctypes.c_uint32()
        check(self._lib.avbd_server_add_world(self._h, ctypes.byref(desc), ctypes.byref(wid)))
        scene.created = True
        self._bodies[wid.value] = scene.body_count
        return wid.value

    def submit(self, world, steps=1):
        self._alive()
        rid = ctypes.c_uint64()
        check(self._lib.avbd_server_submit(self._h, int(world), int(steps), ctypes.byref(rid)))
        return rid.value

    def wait(self, request):
        """Block until the request finished; returns its latency in ms (submission to last step done)."""
        self._alive()
        ms = ctypes.c_double()
        check(self._lib.avbd_server_wait(self._h, int(request), ctypes.byref(ms)))
        return ms.value

    def done(self, request):
        self._alive()
        d = ctypes.c_int()
        check(self._lib.avbd_server_poll(self._h, int(request), ctypes.byref(d)))
        return bool(d.value)

    def readBodies(self, world):
        self._alive()
        out = np.zeros((self._bodies[int(world)], 40), np.float32)
        check(self._lib.avbd_server_read_bodies(self._h, int(world), ptr(out, ctypes.c_float)))
        return out

    def stats(self):
        self._alive()
        s = _native.ServerStats()
        check(self._lib.avbd_server_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def resetStats(self):
        self._alive()
        check(self._lib.avbd_server_reset_stats(self._h))

    def destroy(self):
        if self.destroyed:
            return
        check(self._lib.avbd_server_destroy(self._h))
        self.destroyed = True

    def __del__(self):
        try:
            if not self.destroyed and _native._lib is not None:
                self._lib.avbd_server_destroy(self._h)
                self.destroyed = True
        except Exception:
            pass
