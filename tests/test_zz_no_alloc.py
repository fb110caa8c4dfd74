"""frep_render_noalloc (fixed device staging, no allocation) against
frep_render (library-grown staging): identical outputs on two calls in a
row, the second with a smaller frame, through the C ABI; requests beyond
the fixed staging fail with FREP_INVALID_ARG before any device work."""
import ctypes as C

import numpy as np
import pytest

from frep import lib, shapes

pytestmark = pytest.mark.gpu

STAGING = 64 << 20


def _same_bits(a, b):
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.uint32) == b.view(np.uint32))))


def _camera(w, h, spp, jitter, pos):
    basis = shapes.camera_basis(pos, tuple(-c for c in pos), (0, 1, 0), 45.0, w, h)
    return lib.make_camera(basis, w, h, spp=spp, jitter=jitter, seed=77)


def test_render_noalloc_matches_render():
    ctx = lib.Context(0)
    try:
        th = ctx.upload(shapes.random_csg(24, 2))
        tree = ctx.voxelize(th, shapes.make_grid([16, 8, 8]))
        # first call 960x540 at 2 jittered spp (rgba 2.1 MB, hit ids 8.3 MB,
        # normals 12.4 MB), second a smaller 320x200 frame at 1 spp
        for w, h, spp, jit, pos in [(960, 540, 2, True, (1.7, 1.2, 2.1)),
                                    (320, 200, 1, False, (-2.0, 0.8, 1.9))]:
            cam = _camera(w, h, spp, jit, pos)
            want = ctx.render(tree, cam)
            got = ctx.render(tree, cam, noalloc=True)
            assert (want[1] != np.uint64(2 ** 64 - 1)).sum() > 1000
            assert np.array_equal(got[0], want[0])
            assert np.array_equal(got[1], want[1])
            assert _same_bits(got[2], want[2])
            # outputs requested one at a time land at the start of the staging
            rgba_only, _, _ = ctx.render(tree, cam, want=("rgba",), noalloc=True)
            assert np.array_equal(rgba_only, want[0])
        ctx.tree_free(tree)
        ctx.tape_free(th)
    finally:
        ctx.close()


def test_render_noalloc_rejects_outputs_beyond_its_staging():
    ctx = lib.Context(0)
    try:
        th = ctx.upload(shapes.spheres_union(8, 1))
        tree = ctx.voxelize(th, shapes.make_grid([8, 8, 4]))
        # 1920x1080 at 4 spp with hit ids and normals needs about 174 MB
        cam = _camera(1920, 1080, 4, True, (0.0, 0.5, 3.0))
        hits = np.zeros(1920 * 1080 * 4, dtype=np.uint64)
        nrm = np.zeros((1920 * 1080 * 4, 3), dtype=np.float32)
        st = ctx.L.frep_render_noalloc(ctx.h, tree, C.byref(cam), None,
                                       hits.ctypes.data_as(C.POINTER(C.c_uint64)),
                                       nrm.ctypes.data_as(C.POINTER(C.c_float)))
        assert st == 1   # FREP_INVALID_ARG
        assert b"staging" in ctx.L.frep_last_error(ctx.h)
        # the rgba alone (8.3 MB) fits and matches frep_render
        want, _, _ = ctx.render(tree, cam, want=("rgba",))
        got, _, _ = ctx.render(tree, cam, want=("rgba",), noalloc=True)
        assert np.array_equal(got, want)
        ctx.tree_free(tree)
        ctx.tape_free(th)
    finally:
        ctx.close()
