"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class Grid(C.Structure):
    _fields_ = [("n_levels", C.c_uint32), ("dims", C.c_uint32 * 8), ("low", C.c_float * 3),
                ("world_size", C.c_float), ("tape_time", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("forward", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("screen_world", C.c_float * 2),
                ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("ortho", C.c_uint32), ("jitter", C.c_uint32), ("seed", C.c_uint32),
                ("background", C.c_float * 4), ("light_count", C.c_uint32),
                ("light_dir", (C.c_float * 4) * 4), ("light_color", (C.c_float * 4) * 4),
                ("ambient", C.c_float * 4)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with its Makefile (g++, seconds)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u32p, f32p, u8p = C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_uint8)
        L.orc_sinf.restype = C.c_float
        L.orc_sinf.argtypes = [C.c_float]
        L.orc_cosf.restype = C.c_float
        L.orc_cosf.argtypes = [C.c_float]
        L.orc_expf.restype = C.c_float
        L.orc_expf.argtypes = [C.c_float]
        L.orc_tape_validate.restype = C.c_int
        L.orc_tape_validate.argtypes = [u32p, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_eval_point.restype = C.c_float
        L.orc_eval_point.argtypes = [u32p, f32p, C.c_float, C.c_float, C.c_float, C.c_float]
        L.orc_eval_interval.restype = C.c_uint32
        L.orc_eval_interval.argtypes = [u32p, f32p, f32p, f32p, C.c_float, f32p, u8p]
        L.orc_shorten.restype = C.c_uint32
        L.orc_shorten.argtypes = [u32p, u8p, u32p, C.POINTER(C.c_int)]
        L.orc_eval_gradient.restype = None
        L.orc_eval_gradient.argtypes = [u32p, f32p, C.c_float, C.c_float, C.c_float, C.c_float,
                                        f32p]
        L.orc_voxelize.restype = C.c_void_p
        L.orc_voxelize.argtypes = [u32p, f32p, C.POINTER(Grid)]
        L.orc_tree_free.restype = None
        L.orc_tree_free.argtypes = [C.c_void_p]
        L.orc_tree_node_count.restype = C.c_uint32
        L.orc_tree_node_count.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_tree_canonical.restype = C.c_uint64
        L.orc_tree_canonical.argtypes = [C.c_void_p, u32p, C.c_uint64]
        L.orc_tree_flatten.restype = None
        L.orc_tree_flatten.argtypes = [C.c_void_p, u8p]
        L.orc_tree_stats.restype = None
        L.orc_tree_stats.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.orc_dense_occupancy.restype = None
        L.orc_dense_occupancy.argtypes = [u32p, f32p, C.c_uint32, C.POINTER(Grid), C.c_uint32,
                                          C.c_uint32, C.c_int, C.c_int, u8p]
        L.orc_render.restype = None
        L.orc_render.argtypes = [C.c_void_p, C.POINTER(Camera), C.POINTER(C.c_uint64), f32p, u8p]
        L.orc_render_rows.restype = None
        L.orc_render_rows.argtypes = [C.c_void_p, C.POINTER(Camera), C.c_uint32, C.c_uint32,
                                      C.POINTER(C.c_uint64), f32p, u8p]
        L.orc_check_blob.restype = C.c_int64
        L.orc_check_blob.argtypes = [u32p, f32p, C.POINTER(Grid), u32p, C.c_uint64, C.c_uint32,
                                     C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]
        L.orc_tree_from_blob.restype = C.c_void_p
        L.orc_tree_from_blob.argtypes = [u32p, C.c_uint64, f32p, C.c_uint32, C.POINTER(Grid)]
        _lib = L
    return _lib


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint32))


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.size == 0:
        a = np.zeros(1, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


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


def sinf(x):
    return lib().orc_sinf(x)


def cosf(x):
    return lib().orc_cosf(x)


def expf(x):
    return lib().orc_expf(x)


def validate(words, n_consts, slot_count) -> int:
    w, wp = _u32(words)
    return lib().orc_tape_validate(wp, len(w), n_consts, slot_count)


def eval_point(tape, x, y, z, t=0.0) -> float:
    w, wp = _u32(tape.words)
    c, cp = _f32(tape.consts)
    return lib().orc_eval_point(wp, cp, x, y, z, t)


def eval_interval(tape, lo, hi, t=0.0):
    """Returns ((low, high), choices array)."""
    w, wp = _u32(tape.words)
    c, cp = _f32(tape.consts)
    lo_, lop = _f32(lo)
    hi_, hip = _f32(hi)
    out = np.zeros(2, dtype=np.float32)
    ch = np.zeros(max(1, tape.n), dtype=np.uint8)
    n = lib().orc_eval_interval(wp, cp, lop, hip, t, out.ctypes.data_as(C.POINTER(C.c_float)),
                                ch.ctypes.data_as(C.POINTER(C.c_uint8)))
    return (float(out[0]), float(out[1])), ch[:n].copy()


def shorten(tape_words, choices):
    """Returns (new words, reused flag)."""
    w, wp = _u32(tape_words)
    ch = np.ascontiguousarray(choices, dtype=np.uint8)
    if ch.size == 0:
        ch = np.zeros(1, dtype=np.uint8)
    out = np.zeros(int(w[0]) + 1, dtype=np.uint32)
    reused = C.c_int(0)
    k = lib().orc_shorten(wp, ch.ctypes.data_as(C.POINTER(C.c_uint8)),
                          out.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(reused))
    return out[:k + 1].copy(), bool(reused.value)


def eval_gradient(tape, x, y, z, t=0.0):
    w, wp = _u32(tape.words)
    c, cp = _f32(tape.consts)
    out = np.zeros(4, dtype=np.float32)
    lib().orc_eval_gradient(wp, cp, x, y, z, t, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


class Tree:
    """Owning wrapper of an oracle tree."""

    def __init__(self, tape, grid: dict, blob=None):
        """The oracle's voxelization of `tape` over `grid`, or with `blob` the
        tree a canonical blob describes (for the oracle's raytracer)."""
        self._w, wp = _u32(tape.words)
        self._c, cp = _f32(tape.consts)
        self.grid = grid
        self._g = make_grid(grid)
        if blob is None:
            self.h = lib().orc_voxelize(wp, cp, C.byref(self._g))
        else:
            b, bp = _u32(blob)
            self.h = lib().orc_tree_from_blob(bp, b.size, cp, len(tape.consts),
                                              C.byref(self._g))
            if not self.h:
                raise ValueError("malformed tree blob")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    def counts(self):
        return [lib().orc_tree_node_count(self.h, L) for L in range(len(self.grid["dims"]))]

    def canonical(self) -> np.ndarray:
        n = lib().orc_tree_canonical(self.h, None, 0)
        buf = np.zeros(n, dtype=np.uint32)
        lib().orc_tree_canonical(self.h, buf.ctypes.data_as(C.POINTER(C.c_uint32)), n)
        return buf

    def flatten(self) -> np.ndarray:
        F = int(np.prod(self.grid["dims"]))
        bits = np.zeros(F ** 3, dtype=np.uint8)
        lib().orc_tree_flatten(self.h, bits.ctypes.data_as(C.POINTER(C.c_uint8)))
        return bits.reshape(F, F, F)

    def stats(self) -> np.ndarray:
        nl = len(self.grid["dims"])
        out = np.zeros(nl * 4, dtype=np.float64)
        lib().orc_tree_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(nl, 4)

    def render(self, cam: Camera):
        n = cam.width * cam.height * max(1, cam.spp)
        hits = np.zeros(n, dtype=np.uint64)
        normals = np.zeros(n * 3, dtype=np.float32)
        rgba = np.zeros(cam.width * cam.height * 4, dtype=np.uint8)
        lib().orc_render(self.h, C.byref(cam), hits.ctypes.data_as(C.POINTER(C.c_uint64)),
                         normals.ctypes.data_as(C.POINTER(C.c_float)),
                         rgba.ctypes.data_as(C.POINTER(C.c_uint8)))
        return hits, normals.reshape(-1, 3), rgba.reshape(cam.height, cam.width, 4)


    def render_rows(self, cam: Camera, row_begin: int, row_step: int):
        """render() over rows row_begin, row_begin + row_step, ... only;
        returns full-frame arrays with just those rows written."""
        n = cam.width * cam.height * max(1, cam.spp)
        hits = np.zeros(n, dtype=np.uint64)
        normals = np.zeros(n * 3, dtype=np.float32)
        rgba = np.zeros(cam.width * cam.height * 4, dtype=np.uint8)
        lib().orc_render_rows(self.h, C.byref(cam), row_begin, row_step,
                              hits.ctypes.data_as(C.POINTER(C.c_uint64)),
                              normals.ctypes.data_as(C.POINTER(C.c_float)),
                              rgba.ctypes.data_as(C.POINTER(C.c_uint8)))
        return hits, normals.reshape(-1, 3), rgba.reshape(cam.height, cam.width, 4)


REPORT_KINDS = {1: "malformed", 2: "node count", 3: "coords", 4: "tape", 5: "leaf mask",
                6: "interior mask", 7: "child ranks"}


def check_blob(tape, grid: dict, blob, n_full: int, sample_every: int = 1,
               dense_every: int = 1, threads: int = 1):
    """orc_check_blob: levels 0..n_full-1 of a canonical blob compared in
    full with the oracle's, then a sampled descent (every sample_every-th
    node per level, found in the blob by coords), with every dense_every-th
    sampled last-level node also checked against dense sampling of the full
    tape.  Returns (mismatches, report dict)."""
    w, wp = _u32(tape.words)
    c, cp = _f32(tape.consts)
    g = make_grid(grid)
    b, bp = _u32(blob)
    rep = (C.c_uint64 * 24)()
    bad = lib().orc_check_blob(wp, cp, C.byref(g), bp, b.size, n_full, sample_every,
                               dense_every, threads, rep)
    nl = len(grid["dims"])
    r = {"full_records": rep[0], "sampled_records": rep[1], "dense_voxels": rep[2],
         "dense_mismatches": rep[3], "bad_records": rep[7],
         "counts": [int(rep[8 + L]) for L in range(nl)]}
    if rep[6]:
        r["first"] = {"level": int(rep[4]), "record": int(rep[5]),
                      "kind": REPORT_KINDS.get(int(rep[6]), int(rep[6]))}
    return int(bad), r


def dense_occupancy(tape, grid: dict, x_begin=0, x_end=None, faithful=False, threads=1):
    F = int(np.prod(grid["dims"]))
    if x_end is None:
        x_end = F
    w, wp = _u32(tape.words)
    c, cp = _f32(tape.consts)
    g = make_grid(grid)
    out = np.zeros((x_end - x_begin) * F * F, dtype=np.uint8)
    lib().orc_dense_occupancy(wp, cp, tape.slot_count, C.byref(g), x_begin, x_end,
                              1 if faithful else 0, threads,
                              out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out.reshape(x_end - x_begin, F, F)


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
    if lights is None:
        lights = default_lights()
    cam.light_count = len(lights)
    for l, (d, col) in enumerate(lights):
        for i in range(4):
            cam.light_dir[l][i] = d[i]
            cam.light_color[l][i] = col[i]
    return cam


def default_lights():
    """The reference's two directional lights (raytracer.cpp:187-192)."""
    def nrm(v):
        v = np.array(v, dtype=np.float32)
        ln = np.float32(np.sqrt(np.float32(np.float32(v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])))
        return [float(np.float32(c / ln)) for c in v]
    return [(nrm([-1, -1, -0.2, 0]), [1, 0, 0, 0]), (nrm([1, -1, -0.2, 0]), [0, 0, 2, 0])]
