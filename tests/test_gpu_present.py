"""Layered render target and presentation (frep_target_*, frep_render_layer):
each layer equals the oracle's 1-spp render of its frame; the resolved image
equals the mean of the layers as vulkan/shaders/renderer/triangle.frag:14-23
computes it (rgb/255 summed in layer order, divided by the layer count,
alpha 1), on the oracle's layers and in closed form for constant layers;
the frame is written as a PNG file that reads back unchanged."""
import numpy as np
import pytest

from frep import image, lib, shapes
from frep.tape import Builder, compile_tape

pytestmark = pytest.mark.gpu

F32 = np.float32


def resolve_like_frag(layers):
    """The fragment shader's average of rgba8 layers, in float32."""
    acc = np.zeros(layers[0].shape[:2] + (3,), dtype=F32)
    for q in layers:
        acc = acc + q[:, :, :3].astype(F32) / F32(255.0)
    v = np.clip(acc / F32(len(layers)), F32(0), F32(1))
    rgb = (v * F32(255.0) + F32(0.5)).astype(np.uint8)
    return np.concatenate([rgb, np.full(rgb.shape[:2] + (1,), 255, np.uint8)], axis=2)


def test_layers_and_resolve_match_the_oracle(ctx, oracle, tmp_path):
    tape = shapes.spheres_union(8, 1)
    g = shapes.make_grid([8, 8, 4])
    th = ctx.upload(tape)
    tree = ctx.voxelize(th, g)
    otree = oracle.Tree(tape, g)
    w, h, L = 96, 72, 5
    basis = shapes.camera_basis((0.4, 0.3, 3.0), (0, 0, -1), (0, 1, 0), 45.0, w, h)
    tg = ctx.target_create(w, h, L)
    olayers = []
    for frame in range(L):
        # one jittered sample per frame, a new seed each frame
        kw = dict(spp=1, jitter=True, seed=100 + frame)
        ctx.render_layer(tree, lib.make_camera(basis, w, h, **kw), tg, frame % L)
        _, _, orgba = otree.render(oracle.make_camera(basis, w, h, **kw))
        olayers.append(orgba)
        # the layer alone, resolved through a one-layer target, is the render
        one = ctx.target_create(w, h, 1)
        ctx.render_layer(tree, lib.make_camera(basis, w, h, **kw), one, 0)
        got = ctx.target_resolve(one)
        assert np.array_equal(got[:, :, :3], orgba[:, :, :3]) and np.all(got[:, :, 3] == 255)
        ctx.target_free(one)
    got = ctx.target_resolve(tg)
    want = resolve_like_frag(olayers)
    assert np.array_equal(got, want)
    # jitter makes the layers differ at edges: the mean is not any one layer
    assert not np.array_equal(got[:, :, :3], olayers[0][:, :, :3])
    p = tmp_path / "frame.png"
    image.write_png(p, got)
    assert np.array_equal(image.read_png(p), got)
    ctx.target_free(tg)
    ctx.tree_free(tree)


def test_resolve_closed_form(ctx):
    # a half space z < 0.3 fills the view: every sample of layer 0 has the
    # closed-form colour (87, 51, 122) under the reference's lights; the two
    # other layers stay zero, so the mean is round(c / 3)
    b = Builder()
    tape = compile_tape(b, shapes.halfplane(b, (0.0, 0.0, 1.0), 0.3))
    th = ctx.upload(tape)
    tree = ctx.voxelize(th, shapes.make_grid([4, 4]))
    w, h = 32, 24
    basis = shapes.camera_basis((0.0, 0.0, 3.0), (0, 0, -1), (0, 1, 0), 45.0, w, h, ortho=True,
                                extent=(1.0, 0.75))
    tg = ctx.target_create(w, h, 3)
    ctx.render_layer(tree, lib.make_camera(basis, w, h, ortho=True), tg, 0)
    got = ctx.target_resolve(tg)
    c = np.array([87, 51, 122], dtype=np.float64)
    want = np.floor(c / 255.0 / 3.0 * 255.0 + 0.5).astype(np.uint8)   # 29, 17, 41
    assert np.all(got[:, :, :3] == want) and np.all(got[:, :, 3] == 255)
    ctx.target_free(tg)


def test_render_layer_argument_checks(ctx):
    th = ctx.upload(shapes.spheres_union(4, 2))
    tree = ctx.voxelize(th, shapes.make_grid([8, 4]))
    basis = shapes.camera_basis((0, 0, 3.0), (0, 0, -1), (0, 1, 0), 45.0, 16, 16)
    tg = ctx.target_create(16, 16, 2)
    for cam, layer in [(lib.make_camera(basis, 16, 16, spp=2), 0),
                       (lib.make_camera(basis, 16, 12), 0),
                       (lib.make_camera(basis, 16, 16), 2)]:
        with pytest.raises(lib.FrepError):
            ctx.render_layer(tree, cam, tg, layer)
    with pytest.raises(lib.FrepError):
        ctx.target_create(0, 16, 1)
    ctx.target_free(tg)
    with pytest.raises(lib.FrepError):
        ctx.target_free(tg)
