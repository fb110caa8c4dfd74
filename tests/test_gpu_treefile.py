"""Tree files (frep_tree_save / frep_tree_load): a saved tree loads into a
new tree -- in the same context, in another context from a file on disk --
that exports, renders and re-voxelizes exactly like the original and like
the oracle; a tree the CPU oracle built loads and renders like the oracle's
own render; malformed files are rejected before any device work."""
import numpy as np
import pytest

from frep import csg, lib, shapes
from frep import tree as T
from frep.tape import Builder, compile_tape

pytestmark = pytest.mark.gpu

MISS = np.uint64(0xFFFFFFFFFFFFFFFF)


def trig_tape():
    b = Builder()
    x, y, z, t = b.x(), b.y(), b.z(), b.t()
    g = b.add(b.mul(b.sin(b.mul(b.const(5.0), x)), b.cos(b.mul(b.const(4.0), y))),
              b.div(b.exp(b.mul(b.const(-1.0), b.square(z))), b.add(b.const(1.5), t)))
    s = shapes.sphere_sdf(b, (0.1, -0.2, 0.05), 0.6)
    return compile_tape(b, b.max(b.min(b.sub(g, b.const(0.3)), s), b.neg(b.add(s, b.const(0.4)))))


CASES = {
    "spheres8": (lambda: shapes.spheres_union(8, 1), [8, 8, 4], 0.0),
    "trig": (trig_tape, [8, 8, 4], 0.5),
    "menger": (lambda: csg.compile(csg.MengerSponge(3)), [8, 8, 4], 0.0),
}


def _frame(w=128, h=96, pos=(1.6, 1.1, 2.4)):
    basis = shapes.camera_basis(pos, tuple(-c / np.linalg.norm(pos) for c in pos), (0, 1, 0),
                                45.0, w, h)
    return basis, w, h


def _grid(name):
    fn, dims, t = CASES[name]
    if name == "menger":
        return shapes.make_grid(dims, low=csg.DEMO_LOW, world_size=csg.DEMO_WORLD, tape_time=t)
    return shapes.make_grid(dims, tape_time=t)


@pytest.mark.parametrize("name", sorted(CASES))
def test_saved_tree_loads_identically(ctx, oracle, tmp_path, name):
    tape = CASES[name][0]()
    g = _grid(name)
    th = ctx.upload(tape, fmt=2 if len(tape.consts) > 256 else 1)
    tree = ctx.voxelize(th, g)
    words = ctx.save(tree, tmp_path / "tree.frp")
    grid2, consts2, slots2, blob2 = T.parse_file(words)
    assert grid2["dims"] == g["dims"] and slots2 == tape.slot_count
    assert np.array_equal(consts2.view(np.uint32), np.asarray(tape.consts, np.float32).view(np.uint32))
    ref = oracle.Tree(tape, g).canonical()
    assert np.array_equal(blob2, ref)
    # same context
    t2 = ctx.load(words)
    assert np.array_equal(ctx.export(t2), ref)
    wpos = (25.0, 14.0, 30.0) if name == "menger" else (1.6, 1.1, 2.4)
    basis, w, h = _frame(pos=wpos)
    cam = lib.make_camera(basis, w, h, spp=2, jitter=True, seed=5)
    a = ctx.render(tree, cam)
    b = ctx.render(t2, cam)
    assert (a[1] != MISS).sum() > 100
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    i1, i2 = ctx.tree_info(tree), ctx.tree_info(t2)
    assert list(i1.node_count) == list(i2.node_count)
    assert list(i1.tape_len_sum) == list(i2.tape_len_sum)
    # another context, from the file on disk
    c2 = lib.Context(0)
    try:
        t3 = c2.load(tmp_path / "tree.frp")
        assert np.array_equal(c2.export(t3), ref)
        c = c2.render(t3, cam)
        for x, y in zip(a, c):
            assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
        # a loaded tree's buffers take a new build
        th3 = c2.upload(shapes.spheres_union(5, 9))
        g3 = shapes.make_grid([8, 4, 4])
        c2.voxelize(th3, g3, t3)
        assert np.array_equal(c2.export(t3), oracle.Tree(shapes.spheres_union(5, 9), g3).canonical())
    finally:
        c2.close()
    ctx.tree_free(t2)
    ctx.tree_free(tree)
    ctx.tape_free(th)


def test_oracle_built_tree_loads_and_renders_like_the_oracle(ctx, oracle):
    tape = trig_tape()
    g = shapes.make_grid([8, 4, 4], tape_time=0.5)
    ot = oracle.Tree(tape, g)
    words = T.file_words(g, tape.consts, tape.slot_count, ot.canonical())
    t = ctx.load(words)
    basis, w, h = _frame()
    kw = dict(spp=1, jitter=False)
    rgba, hits, nrm = ctx.render(t, lib.make_camera(basis, w, h, **kw))
    ohits, onrm, orgba = ot.render(oracle.make_camera(basis, w, h, **kw))
    assert (hits != MISS).sum() > 100
    assert np.array_equal(hits, ohits)
    assert np.array_equal(rgba, orgba)
    ctx.tree_free(t)


def test_malformed_tree_files_are_rejected(ctx, oracle):
    tape = shapes.spheres_union(8, 1)
    g = shapes.make_grid([8, 4])
    blob = oracle.Tree(tape, g).canonical()
    good = T.file_words(g, tape.consts, tape.slot_count, blob)
    t = ctx.load(good)
    ctx.tree_free(t)
    hdr = T.FILE_HEADER + len(tape.consts)
    nl = 2
    # first level-1 node's position in the blob: after level 0's one node
    dim0 = 8
    cells0, W0 = dim0 ** 3, dim0 ** 3 // 32
    root_tape_len = int(blob[3 + 2 * nl + 3])
    p_root = hdr + 3 + 2 * nl
    p_masks = p_root + 3 + 1 + root_tape_len
    p_children = p_masks + 2 * W0
    child = blob[3 + 2 * nl + 3 + 1 + root_tape_len + 2 * W0:][:cells0]
    first = int(np.nonzero(child != 0xFFFFFFFF)[0][0])
    cases = {}
    w = good.copy(); w[0] ^= 1; cases["magic"] = w
    w = good.copy(); w[p_children + first] = 100000; cases["child link out of range"] = w
    w = good.copy(); w[p_masks + W0 + (first >> 5)] ^= 1 << (first & 31); cases["interior bit"] = w
    w = good.copy(); w[p_root + 4] = (int(w[p_root + 4]) & 0xFFFFFF00) | 13; cases["bad op"] = w
    w = good.copy(); w[p_root] = 1; cases["root coordinates"] = w
    cases["truncated"] = good[:-5].copy()
    cases["trailing"] = np.concatenate([good, np.zeros(3, np.uint32)])
    w = good.copy(); w[3] = 7; cases["dims"] = w
    for name, words in cases.items():
        with pytest.raises(lib.FrepError) as e:
            ctx.load(words)
        assert e.value.status in (1, 2, 5), name   # INVALID_ARG, BAD_TAPE, UNSUPPORTED (grid)
