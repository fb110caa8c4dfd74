"""Pins the CPU oracle (oracle/frep_oracle.cpp) independently of the GPU path.

The reference ships no tests, fixtures or golden vectors, so each oracle
function is pinned by what the mathematics fixes: libm in double precision,
closed forms, Moore interval arithmetic, brute force on small inputs, and
invariants that a dropped term, wrong sign or swapped operand would break.
"""
import math

import numpy as np
import pytest

from frep import shapes
from frep.tape import Builder, compile_tape, encode, tape_from_words

F32 = np.float32


def ulps(a, b):
    a = np.asarray(a, dtype=np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, dtype=np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


def single_op_tape(op, a=0, b=1, out=4, consts=()):
    """Tape computing one op on slots a, b, then COPY to slot 0."""
    words = [2, encode(op, out, a, b), encode(12, 0, out, 0)]
    return tape_from_words(words, list(consts) or [0.0], 5)


# --------------------------------------------------------------- transcendentals
@pytest.mark.parametrize("name,ref,lo,hi,bound", [
    ("sinf", math.sin, -100.0, 100.0, 2),
    ("cosf", math.cos, -100.0, 100.0, 2),
    ("sinf", math.sin, -3000.0, 3000.0, 2),
    ("expf", math.exp, -87.0, 88.5, 2),
])
def test_transcendentals_vs_libm(oracle, name, ref, lo, hi, bound):
    rng = np.random.default_rng(7)
    xs = rng.uniform(lo, hi, 20000).astype(np.float32)
    f = getattr(oracle, name)
    got = np.array([f(float(x)) for x in xs], dtype=np.float32)
    want = np.array([ref(float(x)) for x in xs], dtype=np.float32)
    # absolute error for sin/cos near zeros: compare in ulps of max(|want|, 2^-24)
    err = ulps(got, want)
    if name != "expf":
        small = np.abs(want) < 1e-3
        assert np.all(np.abs(got[small] - want[small]) < 2e-7)
        err = err[~small]
    assert err.max() <= bound


def test_transcendentals_special(oracle):
    assert oracle.sinf(0.0) == 0.0 and oracle.cosf(0.0) == 1.0 and oracle.expf(0.0) == 1.0
    assert math.isnan(oracle.sinf(float("inf"))) and math.isnan(oracle.cosf(float("nan")))
    assert oracle.expf(100.0) == float("inf") and oracle.expf(-200.0) == 0.0
    assert math.isnan(oracle.expf(float("nan")))
    # odd / even symmetry and the Pythagorean identity
    for x in np.linspace(-20, 20, 401, dtype=np.float32):
        x = float(x)
        assert oracle.sinf(-x) == -oracle.sinf(x)
        assert oracle.cosf(-x) == oracle.cosf(x)
        s, c = oracle.sinf(x), oracle.cosf(x)
        assert abs(s * s + c * c - 1.0) < 4e-7
    # denormal results of exp are produced, not flushed
    assert 0.0 < oracle.expf(-100.0) < 1.2e-38


# --------------------------------------------------------------- point evaluation
def test_point_matches_dag_evaluation(oracle):
    fns = {"sin": oracle.sinf, "cos": oracle.cosf, "exp": oracle.expf}
    rng = np.random.default_rng(3)
    for tape in (shapes.spheres_union(8, 1), shapes.random_csg(24, 2),
                 shapes.gyroid_blend(20.0, 4, 5)):
        b, root = tape.meta["builder"], tape.meta["root"]
        for p in rng.uniform(-1.2, 1.2, (200, 3)).astype(np.float32):
            got = oracle.eval_point(tape, *p)
            want = b.eval(root, *p, fns=fns)
            assert np.float32(got).view(np.uint32) == np.float32(want).view(np.uint32)


def test_sphere_closed_form(oracle):
    b = Builder()
    tape = compile_tape(b, shapes.sphere_sdf(b, (0.25, -0.5, 0.125), 0.5))
    assert oracle.eval_point(tape, 0.25, -0.5, 0.125) == -0.5      # centre: -r exactly
    assert oracle.eval_point(tape, 0.75, -0.5, 0.125) == 0.0       # on the surface
    assert oracle.eval_point(tape, 0.25, 0.5, 0.125) == 0.5        # distance 1 - r


def test_single_ops_against_numpy(oracle):
    a, b = np.float32(1.75), np.float32(-0.3)
    cases = {3: a + b, 4: a - b, 5: a * b, 6: a / b, 7: min(a, b), 8: max(a, b),
             10: -a, 11: np.sqrt(a)}
    for op, want in cases.items():
        t = single_op_tape(op)
        assert oracle.eval_point(t, float(a), float(b), 0.0) == np.float32(want), op
    # operand order matters for SUB / DIV (a swapped operand must fail here)
    assert oracle.eval_point(single_op_tape(4, 1, 0), float(a), float(b), 0.0) == np.float32(b - a)
    # glm min/max: NaN in the second operand returns the first
    nan = float("nan")
    assert oracle.eval_point(single_op_tape(7), 1.0, nan, 0.0) == 1.0
    assert math.isnan(oracle.eval_point(single_op_tape(7), nan, 1.0, 0.0))
    # t is slot 3, constants come from the pool with a 16-bit index
    t = tape_from_words([2, encode(0, 4, 0x34, 0x12), encode(3, 0, 4, 3)],
                        [0.0] * 0x1234 + [2.5], 5)
    assert oracle.eval_point(t, 0.0, 0.0, 0.0, 0.5) == 3.0


# --------------------------------------------------------------- interval evaluation
def test_interval_rules_by_hand(oracle):
    lo, hi = [-1.0, 2.0, 0.0], [3.0, 5.0, 0.0]
    # x in [-1, 3], y in [2, 5]
    exp = {3: (1.0, 8.0), 4: (-6.0, 1.0), 5: (-5.0, 15.0), 7: (-1.0, 3.0), 8: (2.0, 5.0),
           10: (-3.0, 1.0), 11: (0.0, math.sqrt(3.0))}
    for op, (l, h) in exp.items():
        (gl, gh), _ = oracle.eval_interval(single_op_tape(op), lo, hi)
        assert (gl, gh) == (np.float32(l), np.float32(h)), op
    # division: [2,5] / [-1,3] contains 0 -> everything; [-1,3] / [2,5] = [-1/2, 3/2]
    (gl, gh), _ = oracle.eval_interval(single_op_tape(6), lo, hi)
    assert gl == -0.5 and gh == 1.5
    (gl, gh), _ = oracle.eval_interval(single_op_tape(6, 1, 0), lo, hi)
    assert gl == -math.inf and gh == math.inf
    # min/max choices: disjoint -> single side, overlap -> both
    _, ch = oracle.eval_interval(single_op_tape(7), [0.0, 2.0, 0.0], [1.0, 3.0, 0.0])
    assert list(ch) == [1]
    _, ch = oracle.eval_interval(single_op_tape(7), [4.0, 2.0, 0.0], [5.0, 3.0, 0.0])
    assert list(ch) == [2]
    _, ch = oracle.eval_interval(single_op_tape(8), [0.0, 2.0, 0.0], [1.0, 3.0, 0.0])
    assert list(ch) == [2]
    _, ch = oracle.eval_interval(single_op_tape(8), [0.0, 1.0, 0.0], [1.0, 3.0, 0.0])
    assert list(ch) == [3]
    # sin over a full period is [-1, 1]; cos near 0 keeps 1 as upper bound
    (gl, gh), _ = oracle.eval_interval(single_op_tape(1), [0.0, 0, 0], [7.0, 0, 0])
    assert (gl, gh) == (-1.0, 1.0)
    (gl, gh), _ = oracle.eval_interval(single_op_tape(2), [-0.5, 0, 0], [0.25, 0, 0])
    assert gh == 1.0 and gl == np.float32(oracle.cosf(-0.5))


@pytest.mark.parametrize("maker", [lambda: shapes.spheres_union(8, 1),
                                   lambda: shapes.random_csg(24, 2),
                                   lambda: shapes.gyroid_blend(20.0, 4, 5)])
def test_interval_conservative_vs_dense_sampling(oracle, maker):
    tape = maker()
    rng = np.random.default_rng(11)
    for _ in range(60):
        lo = rng.uniform(-1.1, 0.9, 3).astype(np.float32)
        size = np.float32(rng.choice([0.01, 0.05, 0.2]))
        hi = (lo + size).astype(np.float32)
        (ilo, ihi), _ = oracle.eval_interval(tape, lo, hi)
        g = np.linspace(0, 1, 6, dtype=np.float32)
        for u in g:
            for v in g:
                for w in g:
                    p = lo + np.array([u, v, w], dtype=np.float32) * size
                    p = np.minimum(p, hi)
                    d = oracle.eval_point(tape, *p)
                    if tape.meta["kind"] == "gyroid_blend":
                        # transcendentals / divide: conservative up to a few ulp
                        assert ilo - 1e-5 <= d <= ihi + 1e-5
                    else:
                        assert ilo <= d <= ihi


# --------------------------------------------------------------- shortening
def test_shorten_kat_min_disjoint_becomes_copy(oracle):
    # slot4 = 1; slot5 = y + slot4; slot4 = min(x, slot5); slot0 = slot4
    words = [4, encode(0, 4, 0, 0), encode(3, 5, 1, 4), encode(7, 4, 0, 5), encode(12, 0, 4, 0)]
    t = tape_from_words(words, [1.0], 6)
    # x in [0,1], y+1 in [3,3.5]: FIRST -> the y branch and the constant die
    (_, _), ch = oracle.eval_interval(t, [0.0, 2.0, 0.0], [1.0, 2.5, 0.0])
    assert list(ch) == [1]
    new, reused = oracle.shorten(t.words, ch)
    assert not reused
    assert [int(w) for w in new] == [2, encode(12, 4, 0, 5), encode(12, 0, 4, 0)]
    # min(y+1, x) with x below: SECOND -> COPY with A := B (slot x)
    words2 = [4, encode(0, 4, 0, 0), encode(3, 5, 1, 4), encode(7, 4, 5, 0), encode(12, 0, 4, 0)]
    t2 = tape_from_words(words2, [1.0], 6)
    (_, _), ch = oracle.eval_interval(t2, [0.0, 2.0, 0.0], [1.0, 2.5, 0.0])
    assert list(ch) == [2]
    new, reused = oracle.shorten(t2.words, ch)
    assert not reused
    assert [int(w) for w in new] == [2, encode(12, 4, 0, 0), encode(12, 0, 4, 0)]
    # overlapping intervals: BOTH, nothing removed -> the parent tape is reused
    (_, _), ch = oracle.eval_interval(t, [0.0, 0.0, 0.0], [3.0, 2.5, 0.0])
    assert list(ch) == [3]
    new, reused = oracle.shorten(t.words, ch)
    assert reused and np.array_equal(new, t.words)


def test_shorten_drops_dead_branch(oracle):
    # min(sphere_a, sphere_b) over a box touching only sphere a
    b = Builder()
    # distinct constants, so no sub-expression is shared between the spheres
    sa = shapes.sphere_sdf(b, (-0.5, 0.05, -0.03), 0.2)
    sb = shapes.sphere_sdf(b, (0.5, -0.02, 0.04), 0.25)
    t = compile_tape(b, b.min(sa, sb))
    lo, hi = [-0.7, -0.1, -0.1], [-0.4, 0.1, 0.1]
    (_, _), ch = oracle.eval_interval(t, lo, hi)
    new, reused = oracle.shorten(t.words, ch)
    assert not reused and new[0] == 15       # sphere a (14 ops) + COPY
    # kept set == backward closure computed in Python
    assert new[0] < t.n


def _closure(words, choices):
    n = int(words[0])
    active, keep, mm = {0}, [], len(choices) - 1
    for i in range(n - 1, -1, -1):
        w = int(words[1 + i])
        op, o, a, b = w & 0xFF, (w >> 8) & 0xFF, (w >> 16) & 0xFF, w >> 24
        is_mm = op in (7, 8)
        if o in active:
            keep.append(i)
            active.discard(o)
            if op == 0:
                sel = 0
            elif op in (3, 4, 5, 6):
                sel = 3
            elif is_mm:
                sel = int(choices[mm])
            else:
                sel = 1
            if sel & 1:
                active.add(a)
            if sel & 2:
                active.add(b)
        if is_mm:
            mm -= 1
    return sorted(keep)


@pytest.mark.parametrize("maker", [lambda: shapes.random_csg(24, 2),
                                   lambda: shapes.spheres_union(8, 1),
                                   lambda: shapes.gyroid_blend(20.0, 4, 5)])
def test_shorten_invariants(oracle, maker):
    tape = maker()
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(80):
        lo = rng.uniform(-1.0, 0.9, 3).astype(np.float32)
        size = np.float32(rng.choice([0.03, 0.1, 0.25]))
        hi = (lo + size).astype(np.float32)
        (ilo, ihi), ch = oracle.eval_interval(tape, lo, hi)
        new, reused = oracle.shorten(tape.words, ch)
        k = int(new[0])
        assert k <= tape.n
        keep = _closure(tape.words, ch)
        assert len(keep) == k
        short = tape_from_words(new, tape.consts, tape.slot_count)
        if not reused:
            # kept instructions in order; MIN/MAX with one side rewritten as COPY
            for j, i in enumerate(keep):
                ow, nw = int(tape.words[1 + i]), int(new[1 + j])
                assert (ow >> 8) & 0xFF == (nw >> 8) & 0xFF
        for p in rng.uniform(0, 1, (40, 3)).astype(np.float32):
            q = (lo + p * size).astype(np.float32)
            a = oracle.eval_point(tape, *q)
            b = oracle.eval_point(short, *q)
            assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32)
        checked += 1
    assert checked == 80


# --------------------------------------------------------------- gradient
def test_gradient_closed_forms(oracle):
    b = Builder()
    c = (0.1, -0.2, 0.3)
    tape = compile_tape(b, shapes.sphere_sdf(b, c, 0.4))
    rng = np.random.default_rng(2)
    for p in rng.uniform(-1, 1, (100, 3)).astype(np.float32):
        g = oracle.eval_gradient(tape, *p)
        d = p.astype(np.float64) - np.array(c)
        n = d / np.linalg.norm(d)
        assert np.allclose(g[:3], n, atol=2e-6)
        assert g[3] == np.float32(oracle.eval_point(tape, *p))
    b2 = Builder()
    nrm = (0.6, 0.0, 0.8)
    tape2 = compile_tape(b2, shapes.halfplane(b2, nrm, 0.1))
    g = oracle.eval_gradient(tape2, 0.3, 0.4, -0.2)
    assert tuple(g[:3]) == nrm


def test_gradient_finite_differences(oracle):
    # smooth field using every differentiable op (no min/max kinks)
    b = Builder()
    x, y, z = b.x(), b.y(), b.z()
    f = b.add(b.mul(b.sin(b.mul(b.const(6.0), x)), b.cos(b.mul(b.const(5.0), y))),
              b.div(b.exp(b.mul(b.const(-1.5), b.square(z))), b.add(b.const(1.5), x)))
    f = b.sub(b.add(f, b.sqrt(b.add(b.square(y), b.const(0.01)))), b.neg(b.mul(x, z)))
    tape = compile_tape(b, f)
    rng = np.random.default_rng(4)
    h = 1e-3
    for p in rng.uniform(-0.7, 0.7, (50, 3)):
        g = oracle.eval_gradient(tape, *p.astype(np.float32))
        fd = []
        for k in range(3):
            e = np.zeros(3); e[k] = h
            fp = oracle.eval_point(tape, *(p + e).astype(np.float32))
            fm = oracle.eval_point(tape, *(p - e).astype(np.float32))
            fd.append((fp - fm) / (2 * h))
        assert np.allclose(g[:3], fd, atol=0.05 * (1 + np.abs(fd).max()))


# --------------------------------------------------------------- tree build
@pytest.mark.parametrize("maker,dims", [(lambda: shapes.spheres_union(8, 1), [8, 8, 4]),
                                        (lambda: shapes.random_csg(24, 2), [8, 8, 4]),
                                        (lambda: shapes.spheres_union(8, 1), [4, 4, 4, 4]),
                                        (lambda: shapes.random_csg(24, 2), [16, 8])])
def test_tree_occupancy_equals_dense_sampling(oracle, maker, dims):
    tape = maker()
    grid = shapes.make_grid(dims)
    tree = oracle.Tree(tape, grid)
    flat = tree.flatten()
    dense = oracle.dense_occupancy(tape, grid, threads=8)
    assert flat.sum() > 0
    assert np.array_equal(flat, dense)


def test_tree_canonical_blob_parses_and_matches_flatten(oracle):
    from frep import tree as ftree
    tape = shapes.spheres_union(8, 1)
    grid = shapes.make_grid([8, 4, 4])
    t = oracle.Tree(tape, grid)
    levels = ftree.parse(t.canonical())
    assert [len(l.tapes) for l in levels] == t.counts()
    assert np.array_equal(ftree.occupancy(levels), t.flatten())
    # every child tape is a shortening of its parent's: never longer
    for L in range(len(levels) - 1):
        for j, ch in enumerate(levels[L].child):
            for c in ch[ch != 0xFFFFFFFF]:
                assert levels[L + 1].tapes[c][0] <= levels[L].tapes[j][0]


def _record_fields(blob):
    """Word offsets of each record's fields: {(level, j): (coords, tape,
    leaf, interior, child)} (interior/child None at the last level)."""
    nl = int(blob[2])
    dims = [int(v) for v in blob[3:3 + nl]]
    counts = [int(v) for v in blob[3 + nl:3 + 2 * nl]]
    pos, out = 3 + 2 * nl, {}
    for L in range(nl):
        cells = dims[L] ** 3
        last = L == nl - 1
        for j in range(counts[L]):
            k = int(blob[pos + 3])
            leaf = pos + 4 + k
            out[(L, j)] = (pos, pos + 3, leaf, None if last else leaf + cells // 32,
                           None if last else leaf + cells // 16)
            pos = leaf + cells // 32 + (0 if last else cells // 32 + cells)
    return out


@pytest.mark.parametrize("n_full,sample,dense", [(0, 1, 1), (1, 3, 2), (2, 1, 1), (2, 4, 3),
                                                 (3, 1, 1)])
def test_check_blob_accepts_the_oracles_tree(oracle, n_full, sample, dense):
    """The sampled checker that validates 4096^3 trees passes the oracle's
    own canonical blob at every depth of full evaluation and sampling rate,
    and reports the oracle's node counts for the levels it evaluated."""
    tape, grid = shapes.random_csg(24, 2), shapes.make_grid([8, 8, 4])
    t = oracle.Tree(tape, grid)
    bad, rep = oracle.check_blob(tape, grid, t.canonical(), n_full, sample, dense, threads=4)
    assert bad == 0 and rep["bad_records"] == 0 and "first" not in rep
    known = min(n_full, 2) + 1
    assert rep["counts"][:known] == t.counts()[:known]
    if n_full < 3:
        assert rep["sampled_records"] > 0 and rep["dense_voxels"] > 0


def test_check_blob_finds_planted_mistakes(oracle):
    """One flipped bit in any field of any level is reported, with its kind
    and record, whether that level is checked in full or by sampling."""
    tape, grid = shapes.random_csg(24, 2), shapes.make_grid([8, 8, 4])
    blob = oracle.Tree(tape, grid).canonical()
    fields = _record_fields(blob)
    rng = np.random.default_rng(3)
    kinds = {0: "coords", 1: "tape", 2: "leaf mask", 3: "interior mask", 4: "child ranks"}
    for n_full in (1, 3):
        for L in (1, 2):
            n = int(blob[3 + 3 + L])
            for f, kind in kinds.items():
                if fields[(L, 0)][f] is None:
                    continue
                j = int(rng.integers(0, n))
                off = fields[(L, j)][f]
                if f == 1:
                    off += 1 + int(rng.integers(0, int(blob[off])))      # an instruction word
                elif f == 4:
                    off += int(np.flatnonzero(blob[off:off + grid["dims"][L] ** 3]
                                              != 0xFFFFFFFF)[0])     # an existing child
                elif f == 2 or f == 3:
                    off += int(rng.integers(0, grid["dims"][L] ** 3 // 32))
                b2 = blob.copy()
                b2[off] ^= np.uint32(1 << (0 if f in (0, 4) else int(rng.integers(0, 24))))
                bad, rep = oracle.check_blob(tape, grid, b2, n_full, 1, 1, threads=4)
                assert bad > 0, (n_full, L, kind)
                if f == 0:          # also seen from the parent's child entry
                    continue
                first = rep["first"]
                assert first["level"] == L, (n_full, L, kind, first)
                if not (f == 4 and n_full == 1):
                    assert first["kind"] == kind and first["record"] == j, (n_full, L, kind, first)
    # a changed node count, and a truncated blob
    b2 = blob.copy()
    b2[3 + 3 + 1] += 1
    assert oracle.check_blob(tape, grid, b2, 1)[0] != 0
    assert oracle.check_blob(tape, grid, blob[:-1], 1)[0] == -1


def test_tree_from_blob_renders_rows_like_the_tree(oracle):
    """A tree read back from its canonical blob is the same tree, and the
    row-subset render equals the full render on those rows."""
    tape, grid = shapes.random_csg(24, 2), shapes.make_grid([8, 8, 4])
    t = oracle.Tree(tape, grid)
    blob = t.canonical()
    t2 = oracle.Tree(tape, grid, blob=blob)
    assert np.array_equal(t2.canonical(), blob)
    cam = oracle.make_camera(shapes.orbit_basis(0.7, 0.3, 3.0, 160, 96), 160, 96, spp=3,
                             jitter=True, seed=9)
    h, n, rgba = t.render(cam)
    h2, n2, rgba2 = t2.render_rows(cam, 2, 5)
    rows = np.arange(2, 96, 5)
    assert np.array_equal(h.reshape(96, 160, 3)[rows], h2.reshape(96, 160, 3)[rows])
    assert np.array_equal(rgba[rows], rgba2[rows])
    assert (h.reshape(96, 160, 3)[rows] != np.uint64(2 ** 64 - 1)).sum() > 100
    with pytest.raises(ValueError):
        oracle.Tree(tape, grid, blob=blob[:-3])


def test_node_counts_brute_force(oracle):
    # level-1 node count == number of root cells whose interval is ambiguous
    tape = shapes.spheres_union(8, 1)
    grid = shapes.make_grid([8, 8])
    t = oracle.Tree(tape, grid)
    cs = np.float32(np.float32(2.0) / np.float32(8.0))
    amb = 0
    for x in range(8):
        for y in range(8):
            for z in range(8):
                lo = np.float32(-1.0) + np.array([x, y, z], dtype=np.float32) * cs
                (ilo, ihi), _ = oracle.eval_interval(tape, lo, (lo + cs).astype(np.float32))
                amb += not (ilo > 0 or ihi < 0)
    assert t.counts() == [1, amb]


# --------------------------------------------------------------- render
# The render is pinned in two independent parts:
#   * the DDA, by an exact float64 voxel walk (Amanatides-Woo) over the
#     oracle tree's leaf cells: every ray whose first occupied cell is
#     unambiguous must report exactly that cell, at the level whose leaf
#     bit covers it;
#   * the shading, by half-space shapes n.p - d, whose gradient is the
#     constant n: every hit sample then has the closed-form colour
#     ambient + sum_l max(dot(-L_l, n/|n|), 0) * colour_l, stored as
#     round(255 * clamp(v, 0, 1)) (vulkan/shaders/raytracer/main.comp:115-139,
#     the lights of vulkan/src/engine/raytracer.cpp:187-192).
MISS = np.uint64(0xFFFFFFFFFFFFFFFF)
DDA_EPS = 1e-4          # utils.comp:19


def _leaf_levels(blob):
    """Finest-grid array: 0 for empty, L + 1 where a leaf bit of level L
    covers the fine cell (from the canonical blob, not the oracle)."""
    from frep import tree as T
    levels = T.parse(blob)
    dims = [lv.dim for lv in levels]
    F = int(np.prod(dims))
    out = np.zeros((F, F, F), dtype=np.int8)

    def rec(L, j):
        lv = levels[L]
        dim = lv.dim
        span = int(np.prod(dims[L + 1:])) if L + 1 < len(dims) else 1
        for idx in range(dim ** 3):
            x, y, z = idx // (dim * dim), (idx // dim) % dim, idx % dim
            g = [int(lv.coords[j][0]) * dim + x, int(lv.coords[j][1]) * dim + y,
                 int(lv.coords[j][2]) * dim + z]
            if (int(lv.leaf[j][idx >> 5]) >> (idx & 31)) & 1:
                out[g[0] * span:(g[0] + 1) * span, g[1] * span:(g[1] + 1) * span,
                    g[2] * span:(g[2] + 1) * span] = L + 1
            elif lv.interior is not None and (int(lv.interior[j][idx >> 5]) >> (idx & 31)) & 1:
                rec(L + 1, int(lv.child[j][idx]))

    rec(0, 0)
    return out, dims


def _exact_first_hit(occ, low, size, o, d, margin=5e-4):
    """First occupied fine cell along o + t d by an exact walk in float64.

    Returns ('hit', cell), ('miss', None) or ('ambiguous', None).  The DDA
    under test samples 1e-4 to 2e-4 past each boundary crossing and works
    inside the grid box inset by 1e-4, so a cell whose usable segment is
    shorter than `margin`, or that the ray only grazes along an edge or a
    corner, may or may not be seen: a ray meeting such an occupied cell
    before its first clearly traversed occupied cell is ambiguous."""
    F = occ.shape[0]
    cs = size / F
    o = np.asarray(o, dtype=np.float64)
    d = np.asarray(d, dtype=np.float64)
    with np.errstate(divide="ignore"):
        inv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), np.inf)
    lo_b, hi_b = low + DDA_EPS, low + size - DDA_EPS
    t0 = np.where(d != 0.0, (lo_b - o) * inv, -np.inf)
    t1 = np.where(d != 0.0, (hi_b - o) * inv, np.inf)
    for a in range(3):
        if d[a] == 0.0 and not (lo_b <= o[a] <= hi_b):
            return "miss", None
    tmin, tmax = np.minimum(t0, t1), np.maximum(t0, t1)
    t_enter, t_leave = float(np.max(tmin)), float(np.min(tmax))
    if not (t_enter <= t_leave and t_leave >= 0.0):
        return "miss", None
    win_lo, win_hi = max(t_enter, 0.0) + 3e-4, t_leave - 3e-4
    t = max(t_enter, 0.0)
    p = o + t * d
    cell = np.clip(np.floor((p - low) / cs).astype(np.int64), 0, F - 1)
    step = np.where(d > 0, 1, np.where(d < 0, -1, 0))
    ambiguous = False
    while True:
        nxt = np.full(3, np.inf)
        for a in range(3):
            if step[a] != 0:
                bound = low + (cell[a] + (1 if step[a] > 0 else 0)) * cs
                nxt[a] = (bound - o[a]) / d[a]
        t_exit = float(np.min(nxt))
        seg = min(t_exit, win_hi) - max(t, win_lo)
        if occ[tuple(cell)]:
            if seg >= margin and not ambiguous:
                return "hit", tuple(int(c) for c in cell)
            if seg > -margin:
                ambiguous = True
                if seg >= margin:
                    return "ambiguous", None
        near = [a for a in range(3) if nxt[a] - t_exit < margin]
        if len(near) > 1:
            # grazing an edge or corner: the cells beside the path are touched too
            for mask in range(1, 1 << len(near)):
                if mask == (1 << len(near)) - 1:
                    continue
                c2 = cell.copy()
                for i, a in enumerate(near):
                    if mask >> i & 1:
                        c2[a] += step[a]
                if np.all(c2 >= 0) and np.all(c2 < F) and occ[tuple(c2)] and \
                        t_exit < win_hi + margin:
                    ambiguous = True
        a = int(np.argmin(nxt))
        cell[a] += step[a]
        t = t_exit
        if t >= t_leave + margin or cell[a] < 0 or cell[a] >= F:
            return ("ambiguous", None) if ambiguous else ("miss", None)


def _rays(basis, W, H, ortho):
    """Pixel rays without jitter (main.comp:161-184), in float64."""
    fwd, right, up = (np.array(basis[k], dtype=np.float64) for k in ("forward", "right", "up"))
    pos = np.array(basis["pos"], dtype=np.float64)
    sw = basis["screen_world"]
    for y in range(H):
        for x in range(W):
            dx = 2.0 * (x / W) - 1.0
            dy = 2.0 * (y / H) - 1.0
            if ortho:
                yield x, y, pos + right * (dx * sw[0] * 0.5) + up * (dy * sw[1] * 0.5), fwd
            else:
                d = fwd + dx * sw[0] * right + dy * sw[1] * up
                yield x, y, pos, d / np.linalg.norm(d)


def _check_dda_exact(oracle, tape, dims, basis, W, H, ortho, min_decided=0.9):
    grid = shapes.make_grid(dims)
    tr = oracle.Tree(tape, grid)
    lvl, dims = _leaf_levels(tr.canonical())
    assert np.array_equal(lvl > 0, tr.flatten().astype(bool))
    occ = lvl > 0
    F = occ.shape[0]
    cam = oracle.make_camera(basis, W, H, ortho=ortho)
    hits, _, _ = tr.render(cam)
    decided = hit_count = 0
    for x, y, o, d in _rays(basis, W, H, ortho):
        kind, cell = _exact_first_hit(occ, -1.0, 2.0, o, d)
        h = hits[y * W + x]
        if kind == "ambiguous":
            continue
        decided += 1
        if kind == "miss":
            assert h == MISS, (x, y)
            continue
        L = int(lvl[cell]) - 1
        span = int(np.prod(dims[L + 1:])) if L + 1 < len(dims) else 1
        gc = [c // span for c in cell]
        want = (L << 60) | (gc[0] << 40) | (gc[1] << 20) | gc[2]
        assert int(h) == want, (x, y, hex(int(h)), hex(want))
        hit_count += 1
    assert decided >= min_decided * W * H
    assert hit_count > 0.1 * W * H
    return decided, hit_count


def test_dda_exact_axis_rays_through_cell_centres(oracle):
    # orthographic rays along -z, -x and +y through the centres of the
    # 16^3 fine cells: every pixel's first occupied cell is unique
    tape = shapes.spheres_union(8, 1)
    F = 16
    h = 1.0 / F
    for pos, fwd, up0 in [((h, h, 3.0), (0, 0, -1), (0, 1, 0)),
                          ((3.0, h, h), (-1, 0, 0), (0, 1, 0)),
                          ((h, -3.0, h), (0, 1, 0), (0, 0, 1))]:
        basis = shapes.camera_basis(pos, fwd, up0, 45.0, F, F, ortho=True, extent=(2.0, 2.0))
        decided, hits = _check_dda_exact(oracle, tape, [4, 4], basis, F, F, True,
                                         min_decided=1.0)
        assert decided == F * F


@pytest.mark.parametrize("dims,pos,target", [
    ([4, 4], (2.3, 1.7, 2.9), (0.05, -0.03, 0.02)),
    ([8], (-2.1, 2.6, -1.9), (0.1, 0.0, -0.1)),
    ([4, 4, 2 * 2], (0.4, -2.8, 1.3), (0.0, 0.1, 0.0)),
])
def test_dda_exact_perspective_rays(oracle, dims, pos, target):
    tape = shapes.random_csg(6, 3) if len(dims) == 3 else shapes.spheres_union(8, 1)
    fwd = np.array(target) - np.array(pos)
    fwd = fwd / np.linalg.norm(fwd)
    W, H = 24, 20
    basis = shapes.camera_basis(pos, tuple(fwd), (0, 1, 0), 30.0, W, H)
    decided, hits = _check_dda_exact(oracle, tape, dims, basis, W, H, False)
    print(f"decided {decided} of {W * H} rays, {hits} hits")


def _halfspace(n, d):
    from frep.tape import compile_tape
    b = Builder()
    return compile_tape(b, shapes.halfplane(b, n, d))


def _expected_colour(n, lights, ambient):
    n = np.array([float(F32(v)) for v in n], dtype=np.float64)
    n = n / np.linalg.norm(n)
    v = np.array(ambient, dtype=np.float64)
    for ld, col in lights:
        it = max(float(np.dot(-np.array(ld[:3], dtype=np.float64), n)), 0.0)
        v = v + it * np.array(col, dtype=np.float64)
    v = np.clip(v, 0.0, 1.0) * 255.0
    # keep clear of the rounding boundary so float32 evaluation cannot flip it
    frac = v - np.floor(v)
    assert np.all(np.abs(frac - 0.5) > 0.02), v
    return np.floor(v + 0.5).astype(np.uint8)


def _unit(v):
    v = np.array(v, dtype=np.float64)
    return list(v / np.linalg.norm(v)) + [0.0]


REF_LIGHTS = [(_unit([-1, -1, -0.2]), [1, 0, 0, 0]), (_unit([1, -1, -0.2]), [0, 0, 2, 0])]


@pytest.mark.parametrize("n,d,lights,ambient", [
    # the reference's two lights and ambient; plane facing the camera:
    # R = 0.2 + 0.2/|(-1,-1,-0.2)| -> 87, G = 51, B = 0.2 + 2 * 0.14 -> 122
    ((0.0, 0.0, 1.0), 0.3, None, (0.2, 0.2, 0.2, 1.0)),
    # oblique normal: one light's intensity clamps at zero
    ((0.6, 0.3, 0.742), 0.1, None, (0.2, 0.2, 0.2, 1.0)),
    ((-0.5, 0.25, 0.83), -0.2, None, (0.2, 0.2, 0.2, 1.0)),
    # four lights, non-default ambient, one channel saturating
    ((0.2, -0.3, 0.93), 0.0,
     [(_unit([0, 0, -1]), [0.3, 0.1, 0.9, 0]), (_unit([0.3, 0.2, -1]), [0.05, 0.4, 0.3, 0.1]),
      (_unit([-0.7, 0.1, 0.4]), [1, 1, 1, 1]), (_unit([0.1, 0.9, -0.4]), [0.1, 0.05, 0.2, 0])],
     (0.13, 0.07, 0.3, 0.6)),
])
def test_shading_closed_form_half_spaces(oracle, n, d, lights, ambient):
    tape = _halfspace(n, d)
    lights = REF_LIGHTS if lights is None else lights
    want = _expected_colour(n, lights, ambient)
    W, H = 32, 24
    grid = shapes.make_grid([4, 4])
    tr = oracle.Tree(tape, grid)
    for ortho in (False, True):
        basis = shapes.camera_basis((0.15, 0.1, 3.0), (0, 0, -1), (0, 1, 0), 45.0, W, H,
                                    ortho=ortho, extent=(1.6, 1.2))
        bg = (0.25, 0.5, 0.75, 1.0)
        cam = oracle.make_camera(basis, W, H, ortho=ortho, lights=lights, ambient=ambient,
                                 background=bg)
        hits, normals, rgba = tr.render(cam)
        rgba = rgba.reshape(-1, 4)
        hit = hits != MISS
        assert hit.sum() > W * H // 8
        assert np.all(rgba[hit] == want), (rgba[hit][:3], want)
        nn = np.array(n, dtype=np.float64) / np.linalg.norm(np.array(n, dtype=np.float64))
        assert np.allclose(normals[hit], nn, atol=2e-7)
        # misses carry the background colour
        bgb = np.floor(np.array(bg) * 255.0 + 0.5).astype(np.uint8)
        assert np.all(rgba[~hit] == bgb)


def test_shading_spp_average(oracle):
    # jittered samples: a pixel whose k of 4 samples hit the half space shows
    # round(255 * (k c + (4 - k) bg) / 4) per channel
    n, d = (0.0, 0.0, 1.0), 0.3
    tape = _halfspace(n, d)
    grid = shapes.make_grid([4, 4])
    tr = oracle.Tree(tape, grid)
    W, H = 40, 30
    basis = shapes.camera_basis((0.0, 0.0, 3.0), (0, 0, -1), (0, 1, 0), 60.0, W, H)
    bg = (0.0, 0.0, 0.0, 1.0)
    cam = oracle.make_camera(basis, W, H, spp=4, jitter=True, seed=99, background=bg)
    hits, _, rgba = tr.render(cam)
    k = (hits.reshape(H * W, 4) != MISS).sum(axis=1)
    assert (k == 0).any() and (k == 4).any() and ((k > 0) & (k < 4)).any()
    nrm = np.array(n) / np.linalg.norm(n)
    c = np.array([0.2, 0.2, 0.2, 1.0])
    for ld, col in REF_LIGHTS:
        c = c + max(float(np.dot(-np.array(ld[:3]), nrm)), 0.0) * np.array(col)
    want = np.floor(np.clip((k[:, None] * c + (4 - k[:, None]) * np.array(bg)) / 4.0, 0, 1)
                    * 255.0 + 0.5).astype(np.uint8)
    assert np.array_equal(rgba.reshape(-1, 4), want)


def test_render_miss_pixels_show_background(oracle):
    tape = shapes.spheres_union(8, 1)
    grid = shapes.make_grid([4, 4, 4])
    t = oracle.Tree(tape, grid)
    W = H = 24
    basis = shapes.camera_basis((0.2, 0.1, 3.0), (0, 0, -1), (0, 1, 0), 45.0, W, H)
    cam = oracle.make_camera(basis, W, H)
    hits, normals, rgba = t.render(cam)
    assert ((hits != MISS).sum()) > 20
    assert np.all(rgba.reshape(-1, 4)[hits == MISS][:, :3] == 0)
    assert np.all(np.isnan(normals[hits == MISS]))


@pytest.mark.parametrize("axis", ["x", "y", "z", "t"])
def test_bare_axis_root_evaluates_to_that_axis(oracle, axis):
    # a tape whose result is an input must still leave it in slot 0
    b = Builder()
    root = b.axis(axis)
    tp = compile_tape(b, root)
    for p in [(1.0, 2.0, 3.0, 4.0), (-0.5, 0.25, -7.0, 0.125)]:
        want = b.eval(root, *p)
        assert oracle.eval_point(tp, *p) == want == p["xyzt".index(axis)]
