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
def _brute_first_hit(occ, low, cs_fine, o, d, tmax=10.0, steps=20000):
    """March the ray in tiny steps over the finest occupancy grid."""
    F = occ.shape[0]
    for i in range(steps):
        t = tmax * i / steps
        p = (o + t * d - low) / cs_fine
        c = np.floor(p).astype(int)
        if np.all(c >= 0) and np.all(c < F) and occ[c[0], c[1], c[2]]:
            return t
    return None


def test_render_hits_against_brute_force(oracle):
    tape = shapes.spheres_union(8, 1)
    grid = shapes.make_grid([4, 4, 4])
    t = oracle.Tree(tape, grid)
    occ = t.flatten()
    W = H = 24
    basis = shapes.camera_basis((0.2, 0.1, 3.0), (0, 0, -1), (0, 1, 0), 45.0, W, H)
    cam = oracle.make_camera(basis, W, H)
    hits, normals, rgba = t.render(cam)
    miss = np.uint64(0xFFFFFFFFFFFFFFFF)
    F = 64
    agree = 0
    for y in range(H):
        for x in range(W):
            dx = np.float32(2.0) * (np.float32(x) / np.float32(W)) - np.float32(1.0)
            dy = np.float32(2.0) * (np.float32(y) / np.float32(H)) - np.float32(1.0)
            d = (np.array(basis["forward"]) + dx * basis["screen_world"][0] * np.array(basis["right"])
                 + dy * basis["screen_world"][1] * np.array(basis["up"]))
            d = d / np.linalg.norm(d)
            o = np.array(basis["pos"], dtype=np.float64)
            tb = _brute_first_hit(occ, -1.0, 2.0 / F, o, d)
            h = hits[y * W + x]
            if tb is None:
                agree += h == miss
                continue
            if h == miss:
                continue
            level = int(h) >> 60
            gx, gy, gz = (int(h) >> 40) & 0xFFFFF, (int(h) >> 20) & 0xFFFFF, int(h) & 0xFFFFF
            span = F // (4 ** (level + 1))
            p = (o + tb * d + 1.0) / (2.0 / F)
            c = np.floor(p).astype(int) // span
            agree += bool(np.all(np.abs(c - np.array([gx, gy, gz])) <= 1))
    assert agree >= 0.97 * W * H
    assert ((hits != miss).sum()) > 20
    # miss pixels carry the background colour, hit pixels are lit
    assert np.all(rgba.reshape(-1, 4)[hits == miss][:, :3] == 0)
