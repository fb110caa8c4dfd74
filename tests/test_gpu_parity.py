"""GPU parity: libfrep.so (HIP, through the C ABI) against the CPU oracle.

Bar: bit-exact for trees (canonical blob: structure, masks, child links and
every shortened tape), hit ids, densities, interval bounds and min/max
choices; normals and gradients within 2 ulp (FP32), which the shared
operation order makes exact in practice.
"""
import numpy as np
import pytest

from frep import lib, shapes
from frep.tape import Builder, compile_tape

pytestmark = pytest.mark.gpu

MISS = np.uint64(0xFFFFFFFFFFFFFFFF)


def same_bits(a, b):
    """Bitwise float equality, all NaNs equal."""
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.uint32) == b.view(np.uint32))))


def ulp_diff(a, b):
    a = np.asarray(a, dtype=np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, dtype=np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


def trig_tape():
    """A tape touching every op: sin, cos, exp, div, neg, sqrt, min, max, t."""
    b = Builder()
    x, y, z, t = b.x(), b.y(), b.z(), b.t()
    g = b.add(b.mul(b.sin(b.mul(b.const(6.0), x)), b.cos(b.mul(b.const(5.0), y))),
              b.div(b.exp(b.mul(b.const(-1.5), b.square(z))), b.add(b.const(1.5), t)))
    s = shapes.sphere_sdf(b, (0.1, -0.2, 0.05), 0.6)
    f = b.max(b.min(b.sub(g, b.const(0.3)), s), b.neg(b.add(s, b.const(0.4))))
    f = b.min(f, b.sub(b.sqrt(b.add(b.square(x), b.const(0.01))), b.const(0.05)))
    return compile_tape(b, f)


TAPES = {
    "spheres8": lambda: shapes.spheres_union(8, 1),
    "csg24": lambda: shapes.random_csg(24, 2),
    "trig": trig_tape,
}


@pytest.mark.parametrize("fn", [0, 1, 2])
def test_math_bit_exact(ctx, oracle, fn):
    rng = np.random.default_rng(10 + fn)
    if fn == 2:
        x = np.concatenate([rng.uniform(-110, 95, 20000), rng.uniform(-2, 2, 20000)])
    else:
        x = np.concatenate([rng.uniform(-50, 50, 20000), rng.uniform(-7000, 7000, 5000),
                            rng.uniform(-1e6, 1e6, 1000)])
    special = [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-30, -1e-30, 3e38, -3e38, 88.72, -87.3,
               -103.9, -104.0, 1.5707964, 3.1415927, 6.2831855]
    x = np.concatenate([x, special]).astype(np.float32)
    got = ctx.math(fn, x)
    f = [oracle.sinf, oracle.cosf, oracle.expf][fn]
    ref = np.array([f(float(v)) for v in x], dtype=np.float32)
    assert same_bits(got, ref)


@pytest.mark.parametrize("name", sorted(TAPES))
def test_eval_points_bit_exact(ctx, oracle, name):
    tape = TAPES[name]()
    th = ctx.upload(tape)
    rng = np.random.default_rng(1)
    p = rng.uniform(-1.2, 1.2, (4096, 3)).astype(np.float32)
    for t in (0.0, 0.75):
        got = ctx.eval_points(th, p, t)
        ref = np.array([oracle.eval_point(tape, *q, t) for q in p], dtype=np.float32)
        assert same_bits(got, ref), name


@pytest.mark.parametrize("dims", [[8, 8], [16, 8]])
def test_cpp_point_slot_file_10_vs_oracle(ctx, oracle, dims):
    """The gyroid blend (sin/cos, 9 slots) at an 8-wide last level runs the
    C++ point interpreter, 8 voxels per lane, on its 10-slot file: its tree
    equals the oracle's."""
    tape = shapes.gyroid_blend()
    assert tape.slot_count == 9
    grid = shapes.make_grid(dims)
    th = ctx.upload(tape)
    tree = ctx.voxelize(th, grid)
    assert np.array_equal(ctx.export(tree), oracle.Tree(tape, grid).canonical())
    ctx.tree_free(tree)


def _word(op, o, a, b):
    return op | o << 8 | a << 16 | b << 24


def _fusion_tapes():
    """Raw tapes (8 slots, constant pool [0.75]) for every CONST + ADD/SUB/MUL
    pattern the point interpreter fuses, their non-fusable neighbours, and a
    CONST on either side of a 64-word window edge.  Result in slot 0."""
    from frep.tape import tape_from_words
    const = _word(0, 4, 0, 0)                      # slot 4 = 0.75
    cases = {}
    for op, name in ((3, "add"), (4, "sub"), (5, "mul")):
        cases[name + "_in_place_on_a"] = [_word(12, 5, 0, 0), const, _word(op, 5, 5, 4),
                                          _word(12, 0, 5, 5)]
        cases[name + "_in_place_on_b"] = [_word(12, 5, 1, 1), const, _word(op, 5, 4, 5),
                                          _word(12, 0, 5, 5)]
        cases[name + "_on_const_a"] = [const, _word(op, 4, 4, 1), _word(12, 0, 4, 4)]
        cases[name + "_on_const_b"] = [const, _word(op, 4, 2, 4), _word(12, 0, 4, 4)]
        cases[name + "_both_const"] = [const, _word(op, 4, 4, 4), _word(op, 0, 0, 4)]
        cases[name + "_new_slot"] = [const, _word(op, 5, 4, 1), _word(op, 0, 5, 4)]
    # 62 words that stay live through shortening (slot 6 = y, 2y, y, ...,
    # read at the end), so a shortened copy keeps the window edge too
    fill = [_word(12, 6, 1, 1)] + [_word(3 + (i & 1), 6, 6, 1) for i in range(61)]
    tail = [_word(5, 4, 4, 1), _word(3, 5, 4, 6), _word(3, 0, 5, 0)]
    cases["const_at_word_62"] = fill + [const] + tail
    cases["const_at_word_63"] = fill + [_word(3, 6, 6, 1), const] + tail
    cases["const_last"] = [_word(12, 5, 0, 0), _word(0, 0, 0, 0)]
    return {k: tape_from_words(np.array([len(w)] + w, dtype=np.uint32),
                               np.array([0.75], dtype=np.float32), 8)
            for k, w in cases.items()}


def test_fused_const_handlers_bit_exact(ctx, oracle):
    """Point evaluation (assembly interpreter, 8-slot file) of tapes built to
    reach each fused CONST handler and the cases next to them, against the
    oracle, bit for bit."""
    rng = np.random.default_rng(5)
    p = rng.uniform(-1.5, 1.5, (4096, 3)).astype(np.float32)
    for name, tape in _fusion_tapes().items():
        th = ctx.upload(tape)
        got = ctx.eval_points(th, p, 0.25)
        ref = np.array([oracle.eval_point(tape, *q, 0.25) for q in p], dtype=np.float32)
        assert same_bits(got, ref), name
        ctx.tape_free(th)


def test_asm_sqrt_special_inputs_bit_exact(ctx, oracle):
    """SQRT on the assembly interpreter over normal inputs, and over inputs
    mixed with tiny, denormal, zero, negative, infinite and NaN ones (the
    scaling below 2^-96 and the class fix-ups): equal to the oracle (IEEE
    sqrt) bit for bit."""
    from frep.tape import tape_from_words
    words = [_word(11, 0, 0, 0)]                   # slot 0 = sqrt(x)
    tape = tape_from_words(np.array([1] + words, dtype=np.uint32),
                           np.array([0.0], dtype=np.float32), 8)
    rng = np.random.default_rng(9)
    normal = rng.uniform(2.0 ** -90, 4.0, 1024).astype(np.float32)
    special = np.array([0.0, -0.0, 1e-30, 2.0 ** -97, 2.0 ** -96, 1e-40, 1e-45, -1.0,
                        np.inf, -np.inf, np.nan, 3.0e38], dtype=np.float32)
    mixed = normal.copy()
    mixed[::85][:len(special)] = special
    th = ctx.upload(tape)
    for xs in (normal, mixed):
        p = np.zeros((len(xs), 3), dtype=np.float32)
        p[:, 0] = xs
        got = ctx.eval_points(th, p, 0.0)
        ref = np.array([oracle.eval_point(tape, *q, 0.0) for q in p], dtype=np.float32)
        assert same_bits(got, ref)
    ctx.tape_free(th)


@pytest.mark.parametrize("dims", [[4, 4, 8], [4, 8, 4]])
def test_fused_const_trees_match(ctx, oracle, dims):
    """The same tapes voxelized into three-level trees whose last level runs
    the assembly interpreter on the shortened copies the level before writes
    (8-wide nodes: 8 voxels per lane; 4-wide nodes: the one-voxel form): the
    trees equal the oracle's."""
    grid = shapes.make_grid(dims)
    for name, tape in _fusion_tapes().items():
        th = ctx.upload(tape)
        tr = ctx.voxelize(th, grid)
        got = ctx.export(tr)
        ctx.tree_free(tr)
        ctx.tape_free(th)
        assert np.array_equal(got, oracle.Tree(tape, grid).canonical()), name


@pytest.mark.parametrize("name", sorted(TAPES))
def test_eval_intervals_bit_exact(ctx, oracle, name):
    tape = TAPES[name]()
    th = ctx.upload(tape)
    n_mm = sum(1 for op, *_ in tape.instrs() if op in (7, 8))
    rng = np.random.default_rng(2)
    lo = rng.uniform(-1.2, 1.0, (2048, 3)).astype(np.float32)
    size = rng.choice([0.0, 1e-3, 0.05, 0.3, 1.0], size=(2048, 1)).astype(np.float32)
    hi = (lo + size).astype(np.float32)
    got, ch = ctx.eval_intervals(th, lo, hi, 0.25, n_mm=n_mm)
    for k in range(len(lo)):
        (rlo, rhi), rch = oracle.eval_interval(tape, lo[k], hi[k], 0.25)
        assert same_bits(got[k], [rlo, rhi]), (name, k)
        assert np.array_equal(ch[k], rch), (name, k)


@pytest.mark.parametrize("name", sorted(TAPES))
def test_eval_gradients(ctx, oracle, name):
    tape = TAPES[name]()
    th = ctx.upload(tape)
    rng = np.random.default_rng(3)
    p = rng.uniform(-1, 1, (2048, 3)).astype(np.float32)
    got = ctx.eval_gradients(th, p, 0.5)
    ref = np.array([oracle.eval_gradient(tape, *q, 0.5) for q in p], dtype=np.float32)
    finite = np.isfinite(ref) & np.isfinite(got)
    assert np.array_equal(np.isfinite(ref), np.isfinite(got))
    assert ulp_diff(got[finite], ref[finite]).max(initial=0) <= 2


CASES = [
    ("spheres8", [8, 8, 4]),       # 256^3, the plain small workload
    ("spheres8", [4, 4, 4, 4]),    # 4 levels
    ("csg24", [8, 8, 4]),
    ("csg24", [16, 8, 8]),         # 1024^3, the headline workload
    ("csg24", [12, 8]),            # non power-of-two branching
    ("csg24", [8, 16]),            # 16-wide last level: 8 chunks per node
    ("trig", [8, 8, 4]),
    ("trig", [16, 4]),
]


@pytest.mark.parametrize("name,dims", CASES)
def test_voxelize_canonical_bit_exact(ctx, oracle, name, dims):
    tape = TAPES[name]()
    grid = shapes.make_grid(dims, tape_time=0.5 if name == "trig" else 0.0)
    th = ctx.upload(tape)
    tree = ctx.voxelize(th, grid)
    got = ctx.export(tree)
    ref = oracle.Tree(tape, grid)
    want = ref.canonical()
    info = ctx.tree_info(tree)
    assert list(info.node_count[:len(dims)]) == ref.counts()
    assert got.shape == want.shape and np.array_equal(got, want)
    ctx.tree_free(tree)


@pytest.mark.parametrize("n_prims,seed,slots", [(8, 3, 9), (20, 2, 11), (48, 5, 13)])
def test_point_slot_files_vs_oracle(ctx, oracle, n_prims, seed, slots):
    """The last level's assembly interpreter sizes its register file by the
    tape's slot count (10 slots for 9..10, 16 for 11..16): trees equal the
    oracle's on either side of the boundary."""
    tape = shapes.random_csg(n_prims, seed)
    assert tape.slot_count == slots
    grid = shapes.make_grid([8, 8, 4])
    tree = ctx.voxelize(ctx.upload(tape), grid)
    assert np.array_equal(ctx.export(tree), oracle.Tree(tape, grid).canonical())
    ctx.tree_free(tree)


def test_repeated_voxelize_replays_equal_direct(oracle):
    """With FREP_GRAPH=1, repeating a voxelization with unchanged arguments
    replays it from a graph (captured on the first repeat); a call with new
    arguments (tape time) runs directly.  Every call's tree equals the
    oracle's, and the replays equal the direct launches of a default
    context."""
    import os
    tape = TAPES["trig"]()
    direct = lib.Context(0)
    os.environ["FREP_GRAPH"] = "1"
    try:
        fast = lib.Context(0)
    finally:
        del os.environ["FREP_GRAPH"]
    try:
        th = fast.upload(tape)
        tree = fast.tree_create()
        other = fast.tree_create()
        dh = direct.upload(tape)
        seq = [0.5, 0.5, 0.5, 0.25, 0.25, 0.5, 0.5]
        for k, tt in enumerate(seq):
            grid = shapes.make_grid([8, 8, 4], tape_time=tt)
            fast.voxelize(th, grid, tree)
            fast.voxelize(th, shapes.make_grid([16, 4], tape_time=tt), other)  # interleaved tree
            got = fast.export(tree)
            want = direct.export(direct.voxelize(dh, grid))
            assert np.array_equal(got, want), k
            if k in (0, 3):
                assert np.array_equal(got, oracle.Tree(tape, grid).canonical()), k
            info = fast.tree_info(tree)
            assert list(info.node_count[:3]) == oracle.Tree(tape, grid).counts(), k
    finally:
        fast.close()
        direct.close()


def _slot0_unwritten_tape():
    """Slot 0 is never written, so the result is x itself: the root's
    shortened tapes keep nothing ([0])."""
    from frep.tape import Tape
    words = np.array([2, 0 | (4 << 8), 3 | (5 << 8) | (1 << 16) | (4 << 24)], dtype=np.uint32)
    return Tape(words=words, consts=np.array([0.25], dtype=np.float32), slot_count=6)


ROOT_CASES = CASES + [("edge_x", [8, 4]), ("spheres8", [4, 8]), ("union400", [8, 4, 4]),
                      ("union700", [8, 4, 4]), ("gyroid", [16, 4])]


def test_root_pass_kernel_equals_per_cell_kernel(oracle):
    """Level 0 evaluated in dependency passes over the tape's instructions
    (one wavefront per root cell; FREP_ROOT_OPS=2 makes its use mandatory)
    gives the same tree as the per-cell kernel (FREP_ROOT_OPS=0) and as the
    oracle, across root sizes 64 to 4096 cells, a non power-of-two root,
    every op type at the root (trig), a result that is an input, unions of
    6k and 10k instructions (constant table, reused value slots, 8-cell
    blocks for their LDS, the larger one above 64 KiB per block) and the
    gyroid blend (a long dependency chain)."""
    import os
    ctxs = {}
    for mode in ("0", "2"):
        os.environ["FREP_ROOT_OPS"] = mode
        try:
            ctxs[mode] = lib.Context(0)
        finally:
            del os.environ["FREP_ROOT_OPS"]
    tapes = dict(TAPES, edge_x=_slot0_unwritten_tape,
                 union400=lambda: shapes.spheres_union(400, 1035),   # 6k instructions
                 union700=lambda: shapes.spheres_union(700, 1036),   # 10k instructions
                 gyroid=shapes.gyroid_blend)                          # sin/cos chain, 266 passes
    try:
        for name, dims in ROOT_CASES:
            tape = tapes[name]()
            grid = shapes.make_grid(dims, tape_time=0.5 if name == "trig" else 0.0)
            got = {}
            for mode, c in ctxs.items():
                th = c.upload(tape)
                tree = c.voxelize(th, grid)
                got[mode] = c.export(tree)
                c.tree_free(tree)
            want = oracle.Tree(tape, grid).canonical()
            assert np.array_equal(got["2"], want), (name, dims)
            assert np.array_equal(got["0"], want), (name, dims)
    finally:
        for c in ctxs.values():
            c.close()


def test_wide_format_tape_vs_oracle(ctx, oracle):
    """A tape with more than 256 constants (format 2: 16-bit constant index),
    as in the batched mix: tree, densities and a render equal the oracle's."""
    tape = shapes.spheres_union(120, 1017)
    assert len(tape.consts) > 256
    grid = shapes.make_grid([8, 4, 4])
    th = ctx.upload(tape)
    tree = ctx.voxelize(th, grid)
    ref = oracle.Tree(tape, grid)
    assert np.array_equal(ctx.export(tree), ref.canonical())
    rng = np.random.default_rng(17)
    p = rng.uniform(-1, 1, (1000, 3)).astype(np.float32)
    want = np.array([oracle.eval_point(tape, *q, 0.0) for q in p], dtype=np.float32)
    assert same_bits(ctx.eval_points(th, p), want)
    basis = shapes.camera_basis((0, 0, 3), (0, 0, -1), (0, 1, 0), 45.0, 96, 96)
    cam = lib.make_camera(basis, 96, 96)
    _, hits, _ = ctx.render(tree, cam)
    ohits, _, _ = ref.render(oracle.make_camera(basis, 96, 96))
    assert np.array_equal(hits, ohits)
    ctx.tree_free(tree)


@pytest.mark.parametrize("name,dims", [("csg24", [16, 8, 8]), ("trig", [8, 8, 4])])
def test_tree_info_work_matches_exported_tapes(ctx, name, dims):
    """The device-reduced work statistics of frep_tree_info equal the same
    sums taken over the exported tree's tapes on the host."""
    from frep import tree as ftree
    tape = TAPES[name]()
    tr = ctx.voxelize(ctx.upload(tape), shapes.make_grid(dims, tape_time=0.5))
    info = ctx.tree_info(tr)
    levels = ftree.parse(ctx.export(tr))
    work = ftree.level_flops(levels)
    for L, lv in enumerate(levels):
        lens = [int(t[0]) for t in lv.tapes]
        assert info.tape_len_sum[L] == sum(lens)
        assert info.tape_len_max[L] == max(lens)
        assert info.op_evals[L] == work[L]["op_evals"]
        assert info.work_flops[L] == work[L]["flops"]
    ctx.tree_free(tr)


def test_voxelize_deterministic_and_reusable(ctx, oracle):
    tape = shapes.random_csg(24, 2)
    grid = shapes.make_grid([16, 8, 8])
    th = ctx.upload(tape)
    tree = ctx.tree_create()
    ctx.voxelize(th, grid, tree)
    first = ctx.export(tree)
    # a different shape into the same tree, then the first again
    t1 = ctx.upload(shapes.spheres_union(8, 1))
    ctx.voxelize(t1, shapes.make_grid([8, 8, 4]), tree)
    ctx.voxelize(th, grid, tree)
    assert np.array_equal(ctx.export(tree), first)
    ctx.tree_free(tree)


def test_builds_without_init_stay_exact(ctx, oracle):
    """A plain build leaves the tree's counters as the init kernel would for
    its tape, and the next plain build of that tape into the same buffers
    skips the init launch.  Every build below, whichever path preceded it
    (repeat, batched build of another tape into the same tree, regrowth,
    another tape, a grid with more nodes), equals the oracle's tree, and a
    render of a tree built that way equals one of a fresh tree."""
    a = shapes.random_csg(24, 2)
    b = shapes.spheres_union(8, 1)
    ga = shapes.make_grid([16, 8, 8])
    gb = shapes.make_grid([8, 8, 4])
    want_a = oracle.Tree(a, ga).canonical()
    want_b = oracle.Tree(b, gb).canonical()
    ha, hb = ctx.upload(a), ctx.upload(b)
    tree = ctx.tree_create()
    for k in range(3):
        ctx.voxelize(ha, ga, tree)
        assert np.array_equal(ctx.export(tree), want_a), k
        assert list(ctx.tree_info(tree).node_count[:3]) == oracle.Tree(a, ga).counts(), k
    ctx.voxelize_batch([hb], [gb], [tree])
    assert np.array_equal(ctx.export(tree), want_b)
    for k in range(2):
        ctx.voxelize(ha, ga, tree)
        assert np.array_equal(ctx.export(tree), want_a), ("after batch", k)
    g2 = shapes.make_grid([16, 16, 8])
    ctx.voxelize(ha, g2, tree)
    assert np.array_equal(ctx.export(tree), oracle.Tree(a, g2).canonical())
    ctx.voxelize(hb, gb, tree)
    ctx.voxelize(hb, gb, tree)
    assert np.array_equal(ctx.export(tree), want_b)
    for k in range(2):
        ctx.voxelize(ha, ga, tree)
        assert np.array_equal(ctx.export(tree), want_a), ("after b", k)
    basis = shapes.camera_basis((1.8, 1.2, 2.4), (-0.55, -0.35, -0.75), (0, 1, 0), 45.0, 96, 64,
                                ortho=False, extent=(2.4, 2.4))
    cam = lib.make_camera(basis, 96, 64, spp=1, ortho=False, jitter=False, seed=7)
    rgba, hits, nrm = ctx.render(tree, cam)
    fresh = ctx.voxelize(ha, ga)
    frgba, fhits, fnrm = ctx.render(fresh, cam)
    assert (hits != MISS).sum() > 0
    assert np.array_equal(rgba, frgba)
    assert np.array_equal(hits, fhits)
    assert same_bits(nrm, fnrm)
    ctx.tree_free(fresh)
    ctx.tree_free(tree)


def _pulsing_spheres():
    """Two spheres whose radii grow with t: an assembly-interpreter tape
    (CONST ADD SUB MUL SQRT MIN) that reads the input t."""
    b = Builder()
    r = b.add(b.const(0.3), b.mul(b.const(0.25), b.t()))
    d0 = b.sub(b.sqrt(b.add(b.add(b.square(b.x()), b.square(b.y())), b.square(b.z()))), r)
    xs = b.sub(b.x(), b.const(0.45))
    d1 = b.sub(b.sqrt(b.add(b.add(b.square(xs), b.square(b.y())), b.square(b.z()))),
               b.mul(r, b.const(0.7)))
    return compile_tape(b, b.min(d0, d1))


def test_time_dependent_tape_on_assembly_path(ctx, oracle):
    """Slot 3 is filled with t for tapes that read it (and left alone for
    those that write it first): a time-dependent tape on the assembly point
    interpreter equals the oracle at several times, and the times differ."""
    tape = _pulsing_spheres()
    th = ctx.upload(tape)
    tree = ctx.tree_create()
    seen = []
    for tt in (0.0, 1.0, 0.5, 1.0):
        for dims in ([8, 8, 4], [16, 8, 8]):
            grid = shapes.make_grid(dims, tape_time=tt)
            ctx.voxelize(th, grid, tree)
            got = ctx.export(tree)
            assert np.array_equal(got, oracle.Tree(tape, grid).canonical()), (tt, dims)
            if dims == [16, 8, 8]:
                seen.append(got)
    assert not np.array_equal(seen[0], seen[1])
    assert np.array_equal(seen[1], seen[3])
    ctx.tree_free(tree)


def test_edge_shapes(ctx, oracle):
    b = Builder()
    empty = compile_tape(b, b.add(b.const(1.0), b.mul(b.x(), b.const(0.0))))
    b2 = Builder()
    full = compile_tape(b2, b2.sub(b2.mul(b2.x(), b2.const(0.0)), b2.const(1.0)))
    for tape in (empty, full):
        grid = shapes.make_grid([8, 4])
        th = ctx.upload(tape)
        tree = ctx.voxelize(th, grid)
        assert np.array_equal(ctx.export(tree), oracle.Tree(tape, grid).canonical())


def test_bad_tapes_rejected(ctx):
    good = shapes.spheres_union(2, 3)
    w = [int(v) for v in good.words]

    def patched(i, value):
        v = list(w)
        v[i] = value
        return np.array(v, dtype=np.uint32)

    bad_op = patched(1, (w[1] & 0xFFFFFF00) | 13)                  # op 13 does not exist
    bad_slot = patched(2, (w[2] & 0xFFFF00FF) | (200 << 8))        # out slot >= slot_count
    bad_const = patched(1, (w[1] & 0xFFFF) | (250 << 16))          # pool index >= n_consts
    for words in (bad_op, bad_slot, bad_const):
        with pytest.raises(lib.FrepError) as e:
            ctx.upload_raw(words, good.consts, good.slot_count)
        assert e.value.status == lib.FREP_BAD_TAPE
    with pytest.raises(lib.FrepError) as e:
        ctx.upload_raw(w[:5], good.consts, good.slot_count)
    assert e.value.status == lib.FREP_BAD_TAPE
    th = ctx.upload(good)
    with pytest.raises(lib.FrepError) as e:
        ctx.voxelize(th, shapes.make_grid([6, 4]))
    assert e.value.status == lib.FREP_UNSUPPORTED


RENDERS = [
    ("spheres8", [8, 8, 4], dict(pos=(0, 0, 3), fwd=(0, 0, -1), up=(0, 1, 0), w=128, h=96,
                                 ortho=False, spp=1, jitter=False)),
    ("spheres8", [8, 8, 4], dict(pos=(0, 0, 3), fwd=(0, 0, -1), up=(0, 1, 0), w=96, h=96,
                                 ortho=True, spp=1, jitter=False)),
    ("csg24", [16, 8, 8], dict(pos=(1.8, 1.2, 2.4), fwd=(-0.55, -0.35, -0.75), up=(0, 1, 0),
                               w=160, h=90, ortho=False, spp=2, jitter=True)),
    ("trig", [8, 8, 4], dict(pos=(0.3, 2.5, 0.2), fwd=(-0.1, -1, -0.05), up=(0, 0, 1), w=96,
                             h=80, ortho=False, spp=1, jitter=True)),
]


@pytest.mark.parametrize("name,dims,cfg", RENDERS)
def test_render_bit_exact(ctx, oracle, name, dims, cfg):
    tape = TAPES[name]()
    grid = shapes.make_grid(dims, tape_time=0.5 if name == "trig" else 0.0)
    th = ctx.upload(tape)
    tree = ctx.voxelize(th, grid)
    basis = shapes.camera_basis(cfg["pos"], cfg["fwd"], cfg["up"], 45.0, cfg["w"], cfg["h"],
                                ortho=cfg["ortho"], extent=(2.4, 2.4))
    kw = dict(spp=cfg["spp"], ortho=cfg["ortho"], jitter=cfg["jitter"], seed=1234)
    rgba, hits, nrm = ctx.render(tree, lib.make_camera(basis, cfg["w"], cfg["h"], **kw))
    ohits, onrm, orgba = oracle.Tree(tape, grid).render(
        oracle.make_camera(basis, cfg["w"], cfg["h"], **kw))
    assert (hits != MISS).sum() > 0
    assert np.array_equal(hits, ohits)
    hit = hits != MISS
    assert ulp_diff(nrm[hit], onrm[hit]).max(initial=0) <= 2
    assert np.array_equal(rgba, orgba)
    ctx.tree_free(tree)


def test_voxelize_batch_heterogeneous_vs_oracle(ctx, oracle):
    """One batch mixing tapes (slot files 8 and 16, trig ops, asm and C++
    interpreters), level counts and branchings; every tree equals the
    oracle's."""
    jobs = [
        (TAPES["spheres8"](), shapes.make_grid([8, 8, 4])),
        (TAPES["csg24"](), shapes.make_grid([16, 8, 8])),
        (TAPES["trig"](), shapes.make_grid([8, 8, 4], tape_time=0.5)),
        (shapes.spheres_union(3, 11), shapes.make_grid([4, 4, 4, 4])),
        (TAPES["csg24"](), shapes.make_grid([12, 8])),
        (shapes.spheres_union(20, 12), shapes.make_grid([8, 4, 4])),
        (TAPES["trig"](), shapes.make_grid([16, 4])),
        (shapes.spheres_union(5, 13), shapes.make_grid([8])),
    ]
    ths = [ctx.upload(t) for t, _ in jobs]
    trees = ctx.voxelize_batch(ths, [g for _, g in jobs])
    for (tape, grid), tr in zip(jobs, trees):
        ref = oracle.Tree(tape, grid)
        assert list(ctx.tree_info(tr).node_count[:len(grid["dims"])]) == ref.counts()
        assert np.array_equal(ctx.export(tr), ref.canonical())
        ctx.tree_free(tr)


def test_voxelize_batch_root_levels_vs_oracle(oracle):
    """Level 0 of every multi-level tree of a batch by the batched root
    kernel (FREP_ROOT_OPS=2 makes it mandatory): one launch over trees that
    differ in tape, root size and LDS layout, where a 10k-instruction union
    sets 8-cell blocks for the whole launch.  Every tree equals the
    oracle's."""
    import os
    os.environ["FREP_ROOT_OPS"] = "2"
    try:
        c = lib.Context(0)
    finally:
        del os.environ["FREP_ROOT_OPS"]
    jobs = [
        (shapes.spheres_union(700, 1036), shapes.make_grid([8, 4, 4])),
        (TAPES["csg24"](), shapes.make_grid([16, 8, 8])),
        (TAPES["trig"](), shapes.make_grid([8, 8, 4], tape_time=0.5)),
        (shapes.spheres_union(3, 11), shapes.make_grid([4, 4, 4, 4])),
        (TAPES["csg24"](), shapes.make_grid([12, 8])),
        (shapes.gyroid_blend(), shapes.make_grid([16, 4])),
        (shapes.spheres_union(5, 13), shapes.make_grid([8])),
    ]
    try:
        ths = [c.upload(t) for t, _ in jobs]
        trees = c.voxelize_batch(ths, [g for _, g in jobs])
        for (tape, grid), tr in zip(jobs, trees):
            ref = oracle.Tree(tape, grid)
            assert list(c.tree_info(tr).node_count[:len(grid["dims"])]) == ref.counts()
            assert np.array_equal(c.export(tr), ref.canonical())
    finally:
        c.close()


@pytest.mark.parametrize("jobs", [
    [("csg24", [16, 8, 8]), ("spheres8", [8, 8, 8]), ("csg24", [8, 8, 8])],   # <= 10 slots
    [(20, 2, [8, 8, 8]), (48, 5, [8, 8, 8]), ("spheres8", [8, 8])],            # 11..13 slots
    [("spheres8", [8, 4, 8]), (20, 2, [16, 8])],                                 # <= 8 and 11
])
def test_voxelize_batch_assembly_point_kernels_vs_oracle(ctx, oracle, jobs):
    """Batches whose every last level takes the assembly point interpreter
    run the assembly-only batched kernel with the 8-, 10- or 16-slot file;
    every tree equals the oracle's."""
    tapes = []
    for j in jobs:
        tape = TAPES[j[0]]() if isinstance(j[0], str) else shapes.random_csg(j[0], j[1])
        tapes.append((tape, shapes.make_grid(j[-1])))
    ths = [ctx.upload(t) for t, _ in tapes]
    trees = ctx.voxelize_batch(ths, [g for _, g in tapes])
    for (tape, grid), tr in zip(tapes, trees):
        ref = oracle.Tree(tape, grid)
        assert np.array_equal(ctx.export(tr), ref.canonical())
        ctx.tree_free(tr)


def test_voxelize_batch_tuned_and_graph_replays(ctx, oracle):
    """The same batch run repeatedly: the first calls time the autotuner's
    candidates (grid sizes x one or two cells per lane at the interval
    levels), later ones replay a captured graph.  Every call, with other
    work on the library's stream in between, gives the oracle's trees."""
    jobs = [
        (TAPES["csg24"](), shapes.make_grid([16, 8, 8])),
        (shapes.spheres_union(20, 12), shapes.make_grid([8, 4, 4])),
        (TAPES["trig"](), shapes.make_grid([8, 8, 4], tape_time=0.5)),
        (shapes.spheres_union(40, 14), shapes.make_grid([8, 8, 8])),
    ]
    want = [oracle.Tree(t, g).canonical() for t, g in jobs]
    ths = [ctx.upload(t) for t, _ in jobs]
    trees = [ctx.tree_create() for _ in jobs]
    other = ctx.upload(shapes.spheres_union(8, 1))
    for call in range(11):
        ctx.voxelize_batch(ths, [g for _, g in jobs], trees)
        for tr, w in zip(trees, want):
            assert np.array_equal(ctx.export(tr), w), call
        # eager work between replays
        tr = ctx.voxelize(other, shapes.make_grid([8, 8, 4]))
        ctx.tree_free(tr)
    for tr in trees:
        ctx.tree_free(tr)


def test_voxelize_batch_mix_equals_plain(ctx):
    """The 64-shape batched mix (4..400 spheres, 128^3..1024^3): each tree
    built in the batch equals the same shape voxelized alone."""
    mix = shapes.batch_mix(64)
    ths = [ctx.upload(t) for t, _, _ in mix]
    trees = ctx.voxelize_batch(ths, [g for _, g, _ in mix])
    alone = ctx.tree_create()
    for (tape, grid, _), th, tr in zip(mix, ths, trees):
        ctx.voxelize(th, grid, alone)
        assert np.array_equal(ctx.export(tr), ctx.export(alone)), tape.n
    # the same batch again into the same trees (buffers reused) is unchanged
    first = ctx.export(trees[5])
    ctx.voxelize_batch(ths, [g for _, g, _ in mix], trees)
    assert np.array_equal(ctx.export(trees[5]), first)
    for tr in trees + [alone]:
        ctx.tree_free(tr)
    for th in ths:
        ctx.tape_free(th)


def test_voxelize_batch_rejects_duplicate_tree(ctx):
    tape = shapes.spheres_union(2, 3)
    th = ctx.upload(tape)
    tr = ctx.tree_create()
    with pytest.raises(lib.FrepError) as e:
        ctx.voxelize_batch([th, th], [shapes.make_grid([8, 4])] * 2, [tr, tr])
    assert e.value.status == lib.FREP_INVALID_ARG
    ctx.tree_free(tr)


def _plain_outputs(ctx, tape, grid, cam):
    th = ctx.upload(tape)
    tr = ctx.voxelize(th, grid)
    out = ctx.render(tr, cam)
    ctx.tree_free(tr)
    return out


def _served_case(n_req=36, size=(96, 64), cyclic=False):
    scenes = [
        (TAPES["spheres8"](), shapes.make_grid([8, 8, 4])),
        (TAPES["csg24"](), shapes.make_grid([16, 8, 8])),
        (TAPES["trig"](), shapes.make_grid([8, 8, 4], tape_time=0.5)),
        (shapes.random_csg(24, 3), shapes.make_grid([16, 8, 8])),
        (shapes.random_csg(24, 4), shapes.make_grid([8, 8, 8])),
        (shapes.spheres_union(30, 21), shapes.make_grid([8, 4, 4])),
    ]
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(n_req):
        k = i % len(scenes) if cyclic else int(rng.integers(0, len(scenes)))
        b = shapes.orbit_basis(float(rng.uniform(0, 2 * np.pi)), float(rng.uniform(-1.0, 1.0)),
                               3.0, size[0], size[1])
        cam = lib.make_camera(b, size[0], size[1], spp=1 + (i % 2), jitter=True, seed=100 + i)
        reqs.append((k, cam))
    return scenes, reqs


def _check_served(ctx, srv, scenes, reqs, handles=None):
    if handles is None:
        handles = [srv.scene(ctx.upload(t), g) for t, g in scenes]
    tickets = [srv.submit(handles[k], cam, want=("rgba", "hits", "normals")) for k, cam in reqs]
    results = [srv.wait(t) for t in tickets]
    for (k, cam), (res, rgba, hits, nrm) in zip(reqs, results):
        assert res.status == lib.FREP_OK
        prgba, phits, pnrm = _plain_outputs(ctx, scenes[k][0], scenes[k][1], cam)
        assert np.array_equal(hits, phits)
        assert same_bits(nrm, pnrm)
        assert np.array_equal(rgba, prgba)
    return results, handles


def test_server_results_equal_plain_path(ctx):
    """Requests submitted together are batched (trees built by the batched
    kernels, renders on the server's streams); every output equals what
    frep_voxelize + frep_render give for the same scene and camera."""
    scenes, reqs = _served_case()
    srv = lib.Server(ctx, max_batch=16, max_inflight=2)
    try:
        results, handles = _check_served(ctx, srv, scenes, reqs)
        st = srv.stats()
        assert st.completed == len(reqs) and st.failed == 0
        assert max(r[0].batch_size for r in results) > 1
        assert st.voxelizations == len(scenes)       # each scene built once, then resident
        assert st.p99_ms >= st.p50_ms > 0
        # a second round renders resident trees only
        srv.reset_stats()
        _check_served(ctx, srv, scenes, reqs[:12], handles)
        assert srv.stats().voxelizations == 0
    finally:
        srv.close()


def test_server_evicts_under_a_small_budget(ctx):
    """A budget that holds one 1024^3 CSG tree (about 33 MB) but never two:
    with the scenes requested in turn, trees are evicted and rebuilt, and
    every output is still the plain path's."""
    scenes, reqs = _served_case(n_req=24, size=(64, 48), cyclic=True)
    srv = lib.Server(ctx, budget_bytes=56 << 20, max_batch=4, max_inflight=2)
    try:
        _check_served(ctx, srv, scenes, reqs)
        st = srv.stats()
        assert st.failed == 0 and st.evictions > 0
    finally:
        srv.close()


def test_asm_and_cpp_interpreters_agree(oracle):
    """The hand-scheduled point interpreter (8 points per lane) and the C++
    interpreter give identical trees and point values (FREP_NO_ASM=1 forces
    the C++ path in a second context)."""
    import os
    tape = shapes.random_csg(24, 2)
    grid = shapes.make_grid([16, 8, 8])
    os.environ["FREP_NO_ASM"] = "1"
    try:
        plain = lib.Context(0)
    finally:
        del os.environ["FREP_NO_ASM"]
    fast = lib.Context(0)
    try:
        outs = []
        for c in (fast, plain):
            th = c.upload(tape)
            outs.append(c.export(c.voxelize(th, grid)))
        assert np.array_equal(outs[0], outs[1])
        rng = np.random.default_rng(9)
        p = rng.uniform(-1.1, 1.1, (5000, 3)).astype(np.float32)
        a = fast.eval_points(fast.upload(tape), p)
        b = plain.eval_points(plain.upload(tape), p)
        assert same_bits(a, b)
    finally:
        fast.close()
        plain.close()


@pytest.mark.parametrize("name,dims,t", [("csg24", [16, 8, 8], 0.0), ("spheres8", [8, 8, 4], 0.0),
                                          ("trig", [8, 8, 4], 0.5), ("gyroid", [8, 8, 4], 0.0),
                                          ("csg24", [8, 8, 8, 4], 0.0)])
def test_interval_one_and_two_cells_per_lane_agree(ctx, oracle, name, dims, t):
    """Interval levels evaluate two cells per lane wherever a node holds a
    multiple of 128 cells and the slot file is at most 16 (one choice row
    per 64-cell group); FREP_INTERVAL_K1=1 forces one cell per lane.  Both
    give the oracle's tree, in 3- and 4-level trees, with sin/cos/exp/div
    and time-dependent tapes."""
    import os
    tape = shapes.gyroid_blend() if name == "gyroid" else TAPES[name]()
    grid = shapes.make_grid(dims, tape_time=t)
    th = ctx.upload(tape)
    outs = []
    for k1 in ("1", None):
        if k1:
            os.environ["FREP_INTERVAL_K1"] = k1
        try:
            tr = ctx.voxelize(th, grid)
        finally:
            os.environ.pop("FREP_INTERVAL_K1", None)
        outs.append(ctx.export(tr))
        ctx.tree_free(tr)
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[1], oracle.Tree(tape, grid).canonical())
    ctx.tape_free(th)


@pytest.mark.parametrize("name,dims,t", [("csg24", [16, 8, 8], 0.0), ("trig", [8, 8, 4], 0.5),
                                          ("csg24", [8, 8, 8, 4], 0.0), ("spheres8", [4, 4, 4, 4], 0.0),
                                          ("spheres8", [8, 4, 8], 0.0)])
def test_shortening_log_and_second_walk_agree(ctx, oracle, name, dims, t):
    """Shortened tapes are written from the wavefront's LDS log of the words
    its cells keep; a chunk whose log outgrows its LDS walks the parent tape
    a second time instead.  FREP_LOG_CAP (read at each launch) shrinks the
    log so that every chunk, or every chunk with more than 3 kept words,
    takes the second walk, with one and two cells per lane: the trees equal
    the oracle's every time."""
    import os
    tape = TAPES[name]()
    grid = shapes.make_grid(dims, tape_time=t)
    th = ctx.upload(tape)
    want = oracle.Tree(tape, grid).canonical()
    for cap, k1 in ((None, None), ("1", None), ("1", "1"), ("3", None), ("3", "1")):
        for key, v in (("FREP_LOG_CAP", cap), ("FREP_INTERVAL_K1", k1)):
            if v:
                os.environ[key] = v
        try:
            tr = ctx.voxelize(th, grid)
        finally:
            os.environ.pop("FREP_LOG_CAP", None)
            os.environ.pop("FREP_INTERVAL_K1", None)
        got = ctx.export(tr)
        ctx.tree_free(tr)
        assert np.array_equal(got, want), (cap, k1)
    ctx.tape_free(th)


@pytest.mark.parametrize("name,dims", [("csg24", [16, 8, 8]), ("spheres8", [8, 8, 4]),
                                       ("trig", [8, 8, 4]), ("csg24", [8, 8, 8, 4])])
def test_tape_pool_holds_only_written_tapes(ctx, name, dims):
    """Each level reserves exactly the words of the tapes it writes: one
    [k, k words] tape per class of cells of a wavefront that keep the same
    words, none for a child that keeps its whole parent tape.  So the words
    in use past the root tape lie between one copy of each distinct
    shortened tape per level and one tape per shortened child, and far
    below a region of the parent's length per child."""
    from frep import tree as T
    tape = TAPES[name]()
    grid = shapes.make_grid(dims)
    th = ctx.upload(tape)
    tr = ctx.voxelize(th, grid)
    levels = T.parse(ctx.export(tr))
    used = int(ctx.tree_info(tr).tape_pool_words)
    ctx.tree_free(tr)
    ctx.tape_free(th)
    root = 16 + ((tape.n + 1 + 7) & ~7)          # root tape at word 16, 8-word aligned
    lo = hi = per_child_region = 0
    for L in range(1, len(levels)):
        par = levels[L - 1]
        plen = np.zeros(len(levels[L].tapes), dtype=np.int64)
        for j in range(len(par.tapes)):
            for c in par.child[j]:
                if c != 0xFFFFFFFF:
                    plen[c] = len(par.tapes[j]) - 1
        distinct = set()
        for j, tp in enumerate(levels[L].tapes):
            k = len(tp) - 1
            per_child_region += plen[j] + 1
            if k < plen[j]:
                hi += k + 1
                distinct.add(tp.tobytes())
        lo += sum(len(d) // 4 for d in distinct)
    assert lo <= used - root <= hi, (lo, used - root, hi)
    assert hi == 0 or used - root < per_child_region / 2


def test_cpp_engine_classes_match_the_c_abi(ctx, tmp_path):
    """The C++ engine classes (include/frep_engine.hpp), driven like the
    reference's Application::Draw by cpp/engine_demo, give the same tree,
    image and hit ids as the C ABI called directly."""
    import os
    import struct
    import subprocess
    exe = os.path.join(os.path.dirname(lib.LIB_PATH), "engine_demo")
    if not os.path.exists(exe):
        pytest.fail("engine_demo is not built (run __graft_entry__.build())")
    tape = TAPES["csg24"]()
    dims, w, h = [16, 8, 8], 160, 120
    path = tmp_path / "tape.bin"
    with open(path, "wb") as f:
        words = np.asarray(tape.words, dtype=np.uint32)
        consts = np.asarray(tape.consts, dtype=np.float32)
        f.write(struct.pack("<I", len(words)) + words.tobytes())
        f.write(struct.pack("<I", len(consts)) + consts.tobytes())
        f.write(struct.pack("<I", tape.slot_count))
    prefix = str(tmp_path / "out")
    r = subprocess.run([exe, str(path), ",".join(map(str, dims)), str(w), str(h), prefix],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got_tree = np.fromfile(prefix + ".tree", dtype=np.uint32)
    got_rgba = np.fromfile(prefix + ".rgba", dtype=np.uint8).reshape(h, w, 4)
    got_hits = np.fromfile(prefix + ".hits", dtype=np.uint64)
    th = ctx.upload(tape)
    tree = ctx.voxelize(th, shapes.make_grid(dims))
    cam = lib.Camera()
    cam.pos[:] = [0.0, 0.0, 3.0]
    cam.forward[:] = [0.0, 0.0, -1.0]
    cam.width, cam.height, cam.spp = w, h, 1
    cam.background[:] = [0.0, 0.0, 0.0, 1.0]
    cam.ambient[:] = [0.2, 0.2, 0.2, 1.0]
    lights = lib.default_lights()
    cam.light_count = len(lights)
    for l, (d, col) in enumerate(lights):
        cam.light_dir[l][:] = d
        cam.light_color[l][:] = col
    lib.camera_look(cam, [0.0, 1.0, 0.0], 45.0)
    rgba, hits, _ = ctx.render(tree, cam, want=("rgba", "hits"))
    assert np.array_equal(got_tree, ctx.export(tree))
    assert np.array_equal(got_hits, hits)
    assert np.array_equal(got_rgba, rgba)
    # RenderTarget + layered Raytracer: frames 0..3 jittered with seeds 0..3
    got_present = np.fromfile(prefix + ".present", dtype=np.uint8).reshape(h, w, 4)
    tg = ctx.target_create(w, h, 4)
    cam.jitter = 1
    for frame in range(4):
        cam.seed = frame
        ctx.render_layer(tree, cam, tg, frame)
    assert np.array_equal(got_present, ctx.target_resolve(tg))
    ctx.target_free(tg)
    ctx.tree_free(tree)
