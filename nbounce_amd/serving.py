"""Serving over libnbounce: resident scenes, continuous batching of render requests
(nb_server_* in include/nbounce.h), and the seeded mixed-request workload."""
import ctypes
import math

import numpy as np

from . import layout
from ._lib import check, lib, ptr


class Server:
    def __init__(self, ctx, hbm_budget_bytes=200 << 30, max_paths=1 << 23, max_bounces=12):
        h = ctypes.c_void_p()
        check(lib.nb_server_create(ctx.handle, hbm_budget_bytes, max_paths, max_bounces, ctypes.byref(h)))
        self.handle, self.ctx = h, ctx
        self._scenes = []

    def add_scene(self, scene_buffers, envmap):
        idx = ctypes.c_uint32()
        check(lib.nb_server_add_scene(self.handle, scene_buffers.handle, envmap.handle, ctypes.byref(idx)))
        self._scenes.append((scene_buffers, envmap))
        return idx.value

    def set_cache(self, scene, cache):
        """Cached requests for `scene` use (and train) this RadianceCache."""
        check(lib.nb_server_set_cache(self.handle, scene, cache.handle))
        self._scenes.append(cache)

    def submit(self, camera_data, scene, width, height, spp, bounces=8, contribution_factor=4.0, arrival_s=-1.0,
               cached=False, train_paths=0, lr=1e-3):
        r = np.zeros(1, dtype=layout.REQUEST)
        r["camera"] = camera_data
        r["scene"], r["width"], r["height"], r["spp"] = scene, width, height, spp
        r["bounces"], r["contribution_factor"], r["arrival_s"] = bounces, contribution_factor, arrival_s
        r["mode"], r["train_paths"], r["lr"] = (1 if cached else 0), train_paths, lr
        rid = ctypes.c_uint64()
        check(lib.nb_server_submit(self.handle, ptr(r), ctypes.byref(rid)))
        return rid.value

    def step(self):
        a, q = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib.nb_server_step(self.handle, ctypes.byref(a), ctypes.byref(q)))
        return a.value, q.value

    def poll(self, cap=1024):
        out = np.zeros(cap, dtype=layout.COMPLETION)
        n = ctypes.c_uint32()
        check(lib.nb_server_poll(self.handle, ptr(out), cap, ctypes.byref(n)))
        return out[:n.value]

    def read(self, rid, width, height):
        img = np.zeros((height, width, 4), dtype=np.float32)
        check(lib.nb_server_read(self.handle, rid, ptr(img)))
        return img

    def release(self, rid):
        check(lib.nb_server_release(self.handle, rid))

    def now(self):
        t = ctypes.c_double()
        check(lib.nb_server_now(self.handle, ctypes.byref(t)))
        return t.value

    def stats(self):
        s = np.zeros(1, dtype=layout.SERVER_STATS)
        check(lib.nb_server_stats_get(self.handle, ptr(s)))
        return {k: s[k][0].item() for k in layout.SERVER_STATS.names if k != "pad"}

    def close(self):
        if getattr(self, "handle", None):
            lib.nb_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


RESOLUTIONS = [(256, 256), (512, 512), (1024, 768), (1920, 1080)]
RES_P = [0.4, 0.3, 0.2, 0.1]
SPPS = [1, 4, 16, 64]


def mixed_requests(n, qps, n_scenes, seed=0, extent=200.0):
    """Poisson arrivals at `qps`; resolution from RESOLUTIONS with RES_P, spp uniform over SPPS,
    scene uniform, camera: eye uniform in the scene box, looking at a uniform point of the box.
    Returns a list of dicts (arrival offsets in seconds from the start)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / qps, size=n))
    reqs = []
    for k in range(n):
        w, h = RESOLUTIONS[rng.choice(4, p=RES_P)]
        spp = int(rng.choice(SPPS))
        scene = int(rng.integers(0, n_scenes))
        eye = rng.uniform(0.0, extent, size=3)
        target = rng.uniform(0.0, extent, size=3)
        while np.linalg.norm(target - eye) < 1.0:
            target = rng.uniform(0.0, extent, size=3)
        reqs.append(dict(arrival=float(t[k]), width=w, height=h, spp=spp, scene=scene, eye=eye, target=target))
    return reqs


def run_open_loop(server, requests, camera_data_fn, bounces=8, contribution_factor=4.0, keep_images=False,
                  train_paths=0, lr=1e-3):
    """Drives the server in real time: submits each request when its arrival time has passed,
    steps while there is work, and collects completions.  Returns (latencies, wall seconds,
    images by request index if keep_images)."""
    import time
    n = len(requests)
    t0 = server.now()
    ids, i, done = {}, 0, 0
    lat = np.zeros(n)
    images = {}
    while done < n:
        now = server.now() - t0
        while i < n and requests[i]["arrival"] <= now:
            r = requests[i]
            rid = server.submit(camera_data_fn(r), r["scene"], r["width"], r["height"], r["spp"], bounces,
                                contribution_factor, arrival_s=t0 + r["arrival"], cached=r.get("cached", False),
                                train_paths=train_paths, lr=lr)
            ids[rid] = i
            i += 1
        active, queued = server.step()
        for c in server.poll():
            k = ids[int(c["id"])]
            lat[k] = c["done_s"] - c["arrival_s"]
            if keep_images:
                r = requests[k]
                images[k] = server.read(int(c["id"]), r["width"], r["height"])
            server.release(int(c["id"]))
            done += 1
        if active == 0 and queued == 0 and i < n:
            wait = requests[i]["arrival"] - (server.now() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.05))
    return lat, server.now() - t0, images
