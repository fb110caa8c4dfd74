"""ctypes binding of libfrep.so (the C ABI declared in include/frep.h).

This module only marshals arguments; all work runs in the HIP library.  If
libfrep.so is missing or cannot be loaded, every entry point raises
:class:`FrepError` -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libfrep.so")

FREP_OK, FREP_INVALID_ARG, FREP_BAD_TAPE, FREP_OUT_OF_MEMORY, FREP_DEVICE, FREP_UNSUPPORTED = range(6)
STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "BAD_TAPE", 3: "OUT_OF_MEMORY", 4: "DEVICE",
                5: "UNSUPPORTED"}
FORMAT_REF, FORMAT_WIDE = 1, 2
MAX_LEVELS = 8

EXPORTED = [
    "frep_init", "frep_shutdown", "frep_last_error", "frep_set_timing", "frep_get_timing",
    "frep_device_info", "frep_tape_upload", "frep_tape_free", "frep_tree_create",
    "frep_voxelize", "frep_voxelize_batch", "frep_tree_info_get", "frep_tree_export", "frep_tree_free",
    "frep_tree_save", "frep_tree_load",
    "frep_camera_look", "frep_render", "frep_render_noalloc", "frep_target_create",
    "frep_target_free", "frep_render_layer", "frep_target_resolve", "frep_eval_points", "frep_eval_intervals",
    "frep_eval_gradients", "frep_math", "frep_root_schedule", "frep_compile",
    "frep_compile_last_error",
    "frep_server_create", "frep_server_destroy", "frep_server_last_error", "frep_server_scene",
    "frep_server_submit", "frep_server_wait", "frep_server_poll", "frep_server_stats_get",
    "frep_server_reset_stats",
]


class FrepError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Grid(C.Structure):
    _fields_ = [("n_levels", C.c_uint32), ("dims", C.c_uint32 * MAX_LEVELS),
                ("low", C.c_float * 3), ("world_size", C.c_float), ("tape_time", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("forward", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("screen_world", C.c_float * 2),
                ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("ortho", C.c_uint32), ("jitter", C.c_uint32), ("seed", C.c_uint32),
                ("background", C.c_float * 4), ("light_count", C.c_uint32),
                ("light_dir", (C.c_float * 4) * 4), ("light_color", (C.c_float * 4) * 4),
                ("ambient", C.c_float * 4)]


class TreeInfo(C.Structure):
    _fields_ = [("n_levels", C.c_uint32), ("dims", C.c_uint32 * MAX_LEVELS),
                ("node_count", C.c_uint32 * MAX_LEVELS),
                ("tape_len_max", C.c_uint32 * MAX_LEVELS),
                ("slot_count", C.c_uint32), ("retries", C.c_uint32), ("pad", C.c_uint32),
                ("tape_len_sum", C.c_uint64 * MAX_LEVELS),
                ("op_evals", C.c_uint64 * MAX_LEVELS),
                ("work_flops", C.c_uint64 * MAX_LEVELS),
                ("tape_pool_words", C.c_uint64), ("device_bytes", C.c_uint64)]


class Timing(C.Structure):
    _fields_ = [("n_levels", C.c_uint32), ("level_ms", C.c_float * MAX_LEVELS),
                ("init_ms", C.c_float), ("render_ms", C.c_float), ("batch_ms", C.c_float)]


class ServerConfig(C.Structure):
    _fields_ = [("device_budget_bytes", C.c_uint64), ("max_batch", C.c_uint32),
                ("max_inflight", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("flags", C.c_uint32), ("camera", Camera),
                ("tape_time", C.c_float),
                ("rgba", C.POINTER(C.c_uint8)), ("hit_ids", C.POINTER(C.c_uint64)),
                ("normals", C.POINTER(C.c_float))]


REQUEST_AT_TIME = 1


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("voxelized", C.c_uint32), ("batch_size", C.c_uint32),
                ("latency_ms", C.c_float)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("batches", C.c_uint64), ("voxelizations", C.c_uint64),
                ("evictions", C.c_uint64), ("resident_bytes", C.c_uint64),
                ("budget_bytes", C.c_uint64), ("resident_scenes", C.c_uint32),
                ("queued", C.c_uint32), ("p50_ms", C.c_float), ("p99_ms", C.c_float),
                ("mean_ms", C.c_float), ("max_ms", C.c_float), ("mean_batch", C.c_float),
                ("elapsed_s", C.c_float)]


_lib = None


def load():
    """Load libfrep.so and declare the prototypes of every exported symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FrepError(FREP_DEVICE, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, u32p, f32p = C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_float)
    u8p, u64p = C.POINTER(C.c_uint8), C.POINTER(C.c_uint64)
    proto = {
        "frep_init": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
        "frep_shutdown": (None, [vp]),
        "frep_last_error": (C.c_char_p, [vp]),
        "frep_set_timing": (C.c_int, [vp, C.c_int]),
        "frep_get_timing": (C.c_int, [vp, C.POINTER(Timing)]),
        "frep_device_info": (C.c_int, [vp, u32p, u32p]),
        "frep_tape_upload": (C.c_int, [vp, u32p, C.c_uint32, f32p, C.c_uint32, C.c_uint32,
                                       C.c_uint32, u32p]),
        "frep_tape_free": (C.c_int, [vp, C.c_uint32]),
        "frep_tree_create": (C.c_int, [vp, u32p]),
        "frep_voxelize": (C.c_int, [vp, C.c_uint32, C.POINTER(Grid), C.c_uint32]),
        "frep_voxelize_batch": (C.c_int, [vp, C.c_uint32, u32p, C.POINTER(Grid), u32p]),
        "frep_tree_info_get": (C.c_int, [vp, C.c_uint32, C.POINTER(TreeInfo)]),
        "frep_tree_export": (C.c_int, [vp, C.c_uint32, u32p, C.c_uint64, u64p]),
        "frep_tree_free": (C.c_int, [vp, C.c_uint32]),
        "frep_tree_save": (C.c_int, [vp, C.c_uint32, u32p, C.c_uint64, u64p]),
        "frep_tree_load": (C.c_int, [vp, u32p, C.c_uint64, u32p]),
        "frep_camera_look": (C.c_int, [C.POINTER(Camera), f32p, C.c_float]),
        "frep_render": (C.c_int, [vp, C.c_uint32, C.POINTER(Camera), u8p, u64p, f32p]),
        "frep_render_noalloc": (C.c_int, [vp, C.c_uint32, C.POINTER(Camera), u8p, u64p, f32p]),
        "frep_target_create": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32, u32p]),
        "frep_target_free": (C.c_int, [vp, C.c_uint32]),
        "frep_render_layer": (C.c_int, [vp, C.c_uint32, C.POINTER(Camera), C.c_uint32,
                                        C.c_uint32]),
        "frep_target_resolve": (C.c_int, [vp, C.c_uint32, u8p]),
        "frep_eval_points": (C.c_int, [vp, C.c_uint32, f32p, C.c_uint32, C.c_float, f32p]),
        "frep_eval_intervals": (C.c_int, [vp, C.c_uint32, f32p, f32p, C.c_uint32, C.c_float,
                                          f32p, u8p]),
        "frep_eval_gradients": (C.c_int, [vp, C.c_uint32, f32p, C.c_uint32, C.c_float, f32p]),
        "frep_math": (C.c_int, [vp, C.c_uint32, f32p, C.c_uint32, f32p]),
        "frep_root_schedule": (C.c_int, [u32p, C.c_uint32, C.c_uint32, C.c_uint32, u32p, u32p,
                                         C.c_uint32]),
        "frep_compile": (C.c_int, [vp, C.c_uint32, C.c_uint32, u32p, C.c_uint32, f32p,
                                   C.c_uint32, vp]),
        "frep_compile_last_error": (C.c_char_p, []),
        "frep_server_create": (C.c_int, [vp, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
        "frep_server_destroy": (C.c_int, [vp]),
        "frep_server_last_error": (C.c_char_p, [vp]),
        "frep_server_scene": (C.c_int, [vp, C.c_uint32, C.POINTER(Grid), u32p]),
        "frep_server_submit": (C.c_int, [vp, C.POINTER(Request), u64p]),
        "frep_server_wait": (C.c_int, [vp, C.c_uint64, C.POINTER(Result)]),
        "frep_server_poll": (C.c_int, [vp, C.c_uint64, C.POINTER(Result), C.POINTER(C.c_int)]),
        "frep_server_stats_get": (C.c_int, [vp, C.POINTER(ServerStats)]),
        "frep_server_reset_stats": (C.c_int, [vp]),
    }
    for name, (res, args) in proto.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def make_grid(grid: dict) -> Grid:
    g = Grid()
    g.n_levels = len(grid["dims"])
    for i, d in enumerate(grid["dims"]):
        g.dims[i] = d
    for i in range(3):
        g.low[i] = grid["low"][i]
    g.world_size = grid["world_size"]
    g.tape_time = grid.get("tape_time", 0.0)
    return g


def make_camera(basis: dict, width, height, spp=1, ortho=False, jitter=False, seed=0,
                background=(0.0, 0.0, 0.0, 1.0), lights=None,
                ambient=(0.2, 0.2, 0.2, 1.0)) -> Camera:
    cam = Camera()
    for i in range(3):
        cam.pos[i] = basis["pos"][i]
        cam.forward[i] = basis["forward"][i]
        cam.right[i] = basis["right"][i]
        cam.up[i] = basis["up"][i]
    cam.screen_world[0], cam.screen_world[1] = basis["screen_world"]
    cam.width, cam.height, cam.spp = width, height, spp
    cam.ortho, cam.jitter, cam.seed = int(ortho), int(jitter), seed
    for i in range(4):
        cam.background[i] = background[i]
        cam.ambient[i] = ambient[i]
    lights = lights if lights is not None else default_lights()
    cam.light_count = len(lights)
    for l, (d, col) in enumerate(lights):
        for i in range(4):
            cam.light_dir[l][i] = d[i]
            cam.light_color[l][i] = col[i]
    return cam


def default_lights():
    """The reference's two directional lights (vulkan/src/engine/raytracer.cpp:187-192)."""
    def nrm(v):
        v = np.array(v, dtype=np.float32)
        ln = np.float32(np.sqrt(np.float32(np.float32(v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])))
        return [float(np.float32(c / ln)) for c in v]
    return [(nrm([-1, -1, -0.2, 0]), [1, 0, 0, 0]), (nrm([1, -1, -0.2, 0]), [0, 0, 2, 0])]


class Context:
    """One device context (frep_init .. frep_shutdown)."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = C.c_void_p()
        st = self.L.frep_init(device, C.byref(h))
        if st != FREP_OK:
            raise FrepError(st, f"frep_init(device={device}) failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.frep_shutdown(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, st):
        if st != FREP_OK:
            raise FrepError(st, self.L.frep_last_error(self.h).decode())

    # ---- context ----
    def set_timing(self, on: bool):
        self._check(self.L.frep_set_timing(self.h, 1 if on else 0))

    def timing(self) -> Timing:
        t = Timing()
        self._check(self.L.frep_get_timing(self.h, C.byref(t)))
        return t

    def device_info(self):
        cu, grid = C.c_uint32(), C.c_uint32()
        self._check(self.L.frep_device_info(self.h, C.byref(cu), C.byref(grid)))
        return cu.value, grid.value

    # ---- tapes ----
    def upload(self, tape, fmt=None) -> int:
        w = np.ascontiguousarray(tape.words, dtype=np.uint32)
        c = np.ascontiguousarray(tape.consts, dtype=np.float32)
        if c.size == 0:
            c = np.zeros(1, dtype=np.float32)
            nc = 0
        else:
            nc = len(tape.consts)
        if fmt is None:
            fmt = FORMAT_REF if nc <= 256 else FORMAT_WIDE
        out = C.c_uint32()
        self._check(self.L.frep_tape_upload(self.h, _p(w, C.c_uint32), len(w), _p(c, C.c_float),
                                            nc, tape.slot_count, fmt, C.byref(out)))
        return out.value

    def upload_raw(self, words, consts, slot_count, fmt=FORMAT_REF) -> int:
        """Upload raw arrays (no validation on the Python side)."""
        w = np.ascontiguousarray(words, dtype=np.uint32)
        c = np.ascontiguousarray(consts, dtype=np.float32)
        nc = len(c)
        if c.size == 0:
            c = np.zeros(1, dtype=np.float32)
        out = C.c_uint32()
        self._check(self.L.frep_tape_upload(self.h, _p(w, C.c_uint32), len(w), _p(c, C.c_float),
                                            nc, slot_count, fmt, C.byref(out)))
        return out.value

    def tape_free(self, t: int):
        self._check(self.L.frep_tape_free(self.h, t))

    # ---- trees ----
    def tree_create(self) -> int:
        out = C.c_uint32()
        self._check(self.L.frep_tree_create(self.h, C.byref(out)))
        return out.value

    def voxelize(self, tape: int, grid: dict, tree: int | None = None) -> int:
        if tree is None:
            tree = self.tree_create()
        g = make_grid(grid)
        self._check(self.L.frep_voxelize(self.h, tape, C.byref(g), tree))
        return tree

    def voxelize_batch(self, tapes, grids, trees=None):
        """Voxelize many (tape, grid) pairs with one launch per level; returns the trees."""
        n = len(tapes)
        if trees is None:
            trees = [self.tree_create() for _ in range(n)]
        th = np.ascontiguousarray(tapes, dtype=np.uint32)
        tr = np.ascontiguousarray(trees, dtype=np.uint32)
        gs = (Grid * max(n, 1))()
        for i, g in enumerate(grids):
            gs[i] = make_grid(g)
        self._check(self.L.frep_voxelize_batch(self.h, n, _p(th, C.c_uint32), gs,
                                               _p(tr, C.c_uint32)))
        return list(trees)

    def tree_info(self, tree: int) -> TreeInfo:
        info = TreeInfo()
        self._check(self.L.frep_tree_info_get(self.h, tree, C.byref(info)))
        return info

    def save(self, tree: int, path=None) -> np.ndarray:
        """frep_tree_save: the tree file as u32 words (also written to path)."""
        n = C.c_uint64()
        self._check(self.L.frep_tree_save(self.h, tree, None, 0, C.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint32)
        self._check(self.L.frep_tree_save(self.h, tree, _p(buf, C.c_uint32), n.value, C.byref(n)))
        if path is not None:
            buf.astype("<u4").tofile(path)
        return buf

    def load(self, src) -> int:
        """frep_tree_load from u32 words or a file path; returns the new tree."""
        buf = np.fromfile(src, dtype="<u4") if isinstance(src, (str, bytes)) or \
            hasattr(src, "__fspath__") else np.ascontiguousarray(src, dtype=np.uint32)
        out = C.c_uint32()
        self._check(self.L.frep_tree_load(self.h, _p(buf, C.c_uint32), len(buf), C.byref(out)))
        return out.value

    def export(self, tree: int) -> np.ndarray:
        n = C.c_uint64()
        self._check(self.L.frep_tree_export(self.h, tree, None, 0, C.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint32)
        self._check(self.L.frep_tree_export(self.h, tree, _p(buf, C.c_uint32), n.value,
                                            C.byref(n)))
        return buf

    def tree_free(self, tree: int):
        self._check(self.L.frep_tree_free(self.h, tree))

    # ---- render ----
    def render(self, tree: int, cam: Camera, want=("rgba", "hits", "normals"), noalloc=False):
        """frep_render, or with noalloc=True frep_render_noalloc (fixed staging)."""
        spp = max(1, cam.spp)
        npx = cam.width * cam.height
        rgba = np.zeros((cam.height, cam.width, 4), dtype=np.uint8) if "rgba" in want else None
        hits = np.zeros(npx * spp, dtype=np.uint64) if "hits" in want else None
        nrm = np.zeros((npx * spp, 3), dtype=np.float32) if "normals" in want else None
        fn = self.L.frep_render_noalloc if noalloc else self.L.frep_render
        self._check(fn(
            self.h, tree, C.byref(cam),
            _p(rgba, C.c_uint8) if rgba is not None else None,
            _p(hits, C.c_uint64) if hits is not None else None,
            _p(nrm, C.c_float) if nrm is not None else None))
        return rgba, hits, nrm

    # ---- layered render target ----
    def target_create(self, width: int, height: int, layers: int) -> int:
        out = C.c_uint32()
        self._check(self.L.frep_target_create(self.h, width, height, layers, C.byref(out)))
        self._targets = getattr(self, "_targets", {})
        self._targets[out.value] = (width, height)
        return out.value

    def target_free(self, target: int):
        self._check(self.L.frep_target_free(self.h, target))

    def render_layer(self, tree: int, cam: Camera, target: int, layer: int):
        self._check(self.L.frep_render_layer(self.h, tree, C.byref(cam), target, layer))

    def target_resolve(self, target: int) -> np.ndarray:
        w, h = self._targets[target]
        out = np.zeros((h, w, 4), dtype=np.uint8)
        self._check(self.L.frep_target_resolve(self.h, target, _p(out, C.c_uint8)))
        return out

    # ---- batch evaluation ----
    def eval_points(self, tape: int, xyz, t=0.0) -> np.ndarray:
        p = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        out = np.zeros(len(p), dtype=np.float32)
        self._check(self.L.frep_eval_points(self.h, tape, _p(p, C.c_float), len(p), t,
                                            _p(out, C.c_float)))
        return out

    def eval_intervals(self, tape: int, lo, hi, t=0.0, n_mm=0):
        lo = np.ascontiguousarray(lo, dtype=np.float32).reshape(-1, 3)
        hi = np.ascontiguousarray(hi, dtype=np.float32).reshape(-1, 3)
        out = np.zeros((len(lo), 2), dtype=np.float32)
        ch = np.zeros((len(lo), max(n_mm, 1)), dtype=np.uint8)
        self._check(self.L.frep_eval_intervals(self.h, tape, _p(lo, C.c_float), _p(hi, C.c_float),
                                               len(lo), t, _p(out, C.c_float),
                                               _p(ch, C.c_uint8) if n_mm else None))
        return out, ch[:, :n_mm]

    def eval_gradients(self, tape: int, xyz, t=0.0) -> np.ndarray:
        p = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        out = np.zeros((len(p), 4), dtype=np.float32)
        self._check(self.L.frep_eval_gradients(self.h, tape, _p(p, C.c_float), len(p), t,
                                               _p(out, C.c_float)))
        return out

    def math(self, fn: int, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros_like(x)
        self._check(self.L.frep_math(self.h, fn, _p(x, C.c_float), len(x), _p(out, C.c_float)))
        return out


ROOT_SCHEDULE_FIELDS = ("scheduled", "passes", "consts", "c0", "m0", "e0", "res", "slots",
                        "table_bytes", "cell_bytes", "cells_per_block", "words")


def root_schedule(words, lanes_per_cell=64, lds_per_block=160 * 1024):
    """The level-0 schedule frep_tape_upload builds for a tape (host only):
    (info dict, schedule words), or (info, None) when the tape has none."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    info = np.zeros(12, dtype=np.uint32)
    L = load()
    st = L.frep_root_schedule(_p(w, C.c_uint32), len(w), lanes_per_cell, lds_per_block,
                              _p(info, C.c_uint32), None, 0)
    if st != FREP_OK:
        raise FrepError(st, "frep_root_schedule failed")
    d = {k: int(v) for k, v in zip(ROOT_SCHEDULE_FIELDS, info)}
    if not d["scheduled"]:
        return d, None
    out = np.zeros(max(1, d["words"]), dtype=np.uint32)
    st = L.frep_root_schedule(_p(w, C.c_uint32), len(w), lanes_per_cell, lds_per_block,
                              _p(info, C.c_uint32), _p(out, C.c_uint32), len(out))
    if st != FREP_OK:
        raise FrepError(st, "frep_root_schedule failed")
    return d, out[:d["words"]]


def camera_look(cam: Camera, initial_up, fov_deg) -> Camera:
    up = (C.c_float * 3)(*initial_up)
    st = load().frep_camera_look(C.byref(cam), up, fov_deg)
    if st != FREP_OK:
        raise FrepError(st, "frep_camera_look failed")
    return cam


class Server:
    """Request server on a context (frep_server_*): scenes stay resident,
    requests that arrive together are batched, batches overlap on the GPU."""

    def __init__(self, ctx: Context, budget_bytes=0, max_batch=0, max_inflight=0):
        self.ctx = ctx
        self.L = ctx.L
        cfg = ServerConfig(budget_bytes, max_batch, max_inflight)
        h = C.c_void_p()
        st = self.L.frep_server_create(ctx.h, C.byref(cfg), C.byref(h))
        if st != FREP_OK:
            raise FrepError(st, ctx.L.frep_last_error(ctx.h).decode())
        self.h = h
        self._out = {}

    def _check(self, st):
        if st != FREP_OK:
            raise FrepError(st, self.L.frep_server_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.frep_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def scene(self, tape: int, grid: dict) -> int:
        g = make_grid(grid)
        out = C.c_uint32()
        self._check(self.L.frep_server_scene(self.h, tape, C.byref(g), C.byref(out)))
        return out.value

    def submit(self, scene: int, cam: Camera, want=("rgba",), tape_time=None) -> int:
        """Queues a render of `scene`; with tape_time, of its tree built at
        that time (FREP_REQUEST_AT_TIME)."""
        spp = max(1, cam.spp)
        npx = cam.width * cam.height
        rgba = np.zeros((cam.height, cam.width, 4), dtype=np.uint8) if "rgba" in want else None
        hits = np.zeros(npx * spp, dtype=np.uint64) if "hits" in want else None
        nrm = np.zeros((npx * spp, 3), dtype=np.float32) if "normals" in want else None
        req = Request()
        req.scene = scene
        req.camera = cam
        if tape_time is not None:
            req.flags = REQUEST_AT_TIME
            req.tape_time = tape_time
        req.rgba = _p(rgba, C.c_uint8) if rgba is not None else None
        req.hit_ids = _p(hits, C.c_uint64) if hits is not None else None
        req.normals = _p(nrm, C.c_float) if nrm is not None else None
        t = C.c_uint64()
        self._check(self.L.frep_server_submit(self.h, C.byref(req), C.byref(t)))
        self._out[t.value] = (rgba, hits, nrm)
        return t.value

    def wait(self, ticket: int):
        """Returns (Result, rgba, hits, normals) of a submitted request."""
        r = Result()
        self._check(self.L.frep_server_wait(self.h, ticket, C.byref(r)))
        rgba, hits, nrm = self._out.pop(ticket)
        return r, rgba, hits, nrm

    def poll(self, ticket: int):
        r, done = Result(), C.c_int()
        self._check(self.L.frep_server_poll(self.h, ticket, C.byref(r), C.byref(done)))
        if not done.value:
            return None
        rgba, hits, nrm = self._out.pop(ticket)
        return r, rgba, hits, nrm

    def stats(self) -> ServerStats:
        s = ServerStats()
        self._check(self.L.frep_server_stats_get(self.h, C.byref(s)))
        return s

    def reset_stats(self):
        self._check(self.L.frep_server_reset_stats(self.h))
