"""GPU parity on the reference's seven demo shapes (vulkan/src/shapes.cpp:5-117),
compiled by frep_compile and voxelized over the reference's world box
(low -20, size 40: vulkan/src/application.cpp:21-22): canonical trees equal
to the oracle's at several tape times, and the frame seen from the
reference's default camera (position (0, 0, 40) looking down -z, 45 degrees:
application.cpp:50-54) with hit ids and pixels equal to the oracle's."""
import numpy as np
import pytest

from frep import csg, lib, shapes

pytestmark = pytest.mark.gpu

MISS = np.uint64(0xFFFFFFFFFFFFFFFF)
TIMED = {"screw", "elastic_cube", "morph", "organic_balls"}


def grid(dims, t):
    return shapes.make_grid(dims, low=csg.DEMO_LOW, world_size=csg.DEMO_WORLD, tape_time=t)


@pytest.mark.parametrize("name", sorted(csg.DEMO_SHAPES))
def test_demo_shape_trees_match_oracle(ctx, oracle, name):
    tape = csg.compile(csg.DEMO_SHAPES[name]())
    print(f"{name}: {tape.n} instructions, {tape.slot_count} slots", flush=True)
    th = ctx.upload(tape, fmt=tape.meta["format"])
    tree = ctx.tree_create()
    times = (0.0, 1.7, 4.25) if name in TIMED else (0.0,)
    for dims in ([8, 8, 4], [16, 4, 4]):
        for t in times:
            g = grid(dims, t)
            ctx.voxelize(th, g, tree)
            got = ctx.export(tree)
            ref = oracle.Tree(tape, g).canonical()
            assert np.array_equal(got, ref), (name, dims, t)
    ctx.tree_free(tree)
    ctx.tape_free(th)


@pytest.mark.parametrize("name,dims,pos", [
    ("organic_balls", [8, 8, 4], (0.0, 0.0, 40.0)),
    ("menger_sponge3", [8, 8, 4], (0.0, 0.0, 40.0)),
    # the tower is about one unit wide: a finer grid and a closer camera
    ("twisted_tower", [16, 8, 4], (0.0, 2.5, 9.0)),
])
def test_demo_shape_render_matches_oracle(ctx, oracle, name, dims, pos):
    tape = csg.compile(csg.DEMO_SHAPES[name]())
    th = ctx.upload(tape, fmt=tape.meta["format"])
    g = grid(dims, 0.6)
    tree = ctx.voxelize(th, g)
    w, h = 160, 120
    basis = shapes.camera_basis(pos, (0, 0, -1), (0, 1, 0), 45.0, w, h)
    rgba, hits, nrm = ctx.render(tree, lib.make_camera(basis, w, h))
    ohits, onrm, orgba = oracle.Tree(tape, g).render(oracle.make_camera(basis, w, h))
    assert (hits != MISS).sum() > 100
    assert np.array_equal(hits, ohits)
    assert np.array_equal(rgba, orgba)
    ctx.tree_free(tree)
    ctx.tape_free(th)
