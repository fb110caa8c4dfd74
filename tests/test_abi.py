"""CPU-side checks of the C ABI: the library loads, exports every entry point
declared in include/frep.h, and the host-only helpers behave."""
import ctypes as C
import os
import re

import numpy as np

from frep import lib, shapes
from frep.tape import OP_NAMES, encode

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "frep.h")).read()
    return sorted(set(re.findall(r"^\s*(?:frep_status|void|const char\*)\s+(frep_\w+)\s*\(",
                                 text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "frep_tape_upload" in names and "frep_voxelize" in names and "frep_render" in names
    assert sorted(lib.EXPORTED) == names


def test_library_exports_every_declared_symbol():
    L = lib.load()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_struct_layouts_match_header_sizes():
    # sizes implied by the header's field lists (all 4-byte fields, one u64 pair)
    assert C.sizeof(lib.Grid) == 4 * (1 + 8 + 3 + 1 + 1)
    assert C.sizeof(lib.Camera) == 4 * (3 * 4 + 2 + 6 + 4 + 1 + 16 + 16 + 4)
    assert C.sizeof(lib.TreeInfo) == 4 * (1 + 8 * 3 + 3) + 8 * (8 * 3 + 2)
    assert C.sizeof(lib.Timing) == 4 * (1 + 8 + 3)
    assert C.sizeof(lib.ServerConfig) == 16
    assert C.sizeof(lib.Request) == 8 + C.sizeof(lib.Camera) + 4 + 3 * 8
    assert C.sizeof(lib.Result) == 16
    assert C.sizeof(lib.ServerStats) == 8 * 8 + 4 * 2 + 4 * 6


def test_camera_look_host_helper():
    cam = lib.Camera()
    cam.forward[:] = [0.0, 0.0, -1.0]
    cam.width, cam.height = 640, 480
    lib.camera_look(cam, [0.0, 1.0, 0.0], 45.0)
    assert list(cam.right) == [1.0, -0.0, 0.0] or list(cam.right) == [1.0, 0.0, 0.0]
    assert list(cam.up) == [0.0, 1.0, 0.0]
    assert abs(cam.screen_world[0] - 2 * np.tan(np.pi / 8)) < 1e-6
    assert abs(cam.screen_world[1] - cam.screen_world[0] * 0.75) < 1e-6
    # same basis as the Python helper used by the tests
    b = shapes.camera_basis((0, 0, 0), (0, 0, -1), (0, 1, 0), 45.0, 640, 480)
    assert np.allclose(b["right"], list(cam.right)) and np.allclose(b["up"], list(cam.up))


def test_tape_compiler_reference_rules():
    """Slot allocation follows tape.cpp:151-228: inputs released before the
    output takes the lowest free slot; result in slot 0."""
    t = shapes.spheres_union(2, 3)
    ins = list(t.instrs())
    assert ins[-1][1] == 0
    # t is never used, so slot 3 starts empty and the first CONST takes it
    # (GetFreeSlot returns the lowest slot without a value, tape.cpp:81-92)
    assert ins[0][0] == 0 and ins[0][1] == 3
    # x - cx reads x (slot 0) and cx (slot 3); cx dies here, so its slot is
    # released before the output is placed and the SUB writes slot 3 again
    assert OP_NAMES[ins[1][0]] == "SUB" and ins[1][2] == 0 and ins[1][3] == 3
    assert ins[1][1] == 3
    assert encode(3, 1, 2, 3) == 3 | 1 << 8 | 2 << 16 | 3 << 24
    assert t.n == 2 * 14 + 1


def test_tree_file_helpers_round_trip(oracle):
    from frep import tree as T
    tape = shapes.spheres_union(8, 1)
    g = shapes.make_grid([8, 4], tape_time=0.25)
    blob = oracle.Tree(tape, g).canonical()
    w = T.file_words(g, tape.consts, tape.slot_count, blob)
    assert int(w[0]) == 0x46505246 and int(w[2]) == 2
    g2, c2, s2, b2 = T.parse_file(w)
    assert g2["dims"] == [8, 4] and g2["tape_time"] == 0.25 and g2["world_size"] == 2.0
    assert s2 == tape.slot_count and np.array_equal(b2, blob)
    assert np.array_equal(c2, np.asarray(tape.consts, np.float32))
    levels = T.parse(b2)
    assert [len(lv.tapes) for lv in levels] == oracle.Tree(tape, g).counts()


def test_png_round_trip(tmp_path):
    from frep import image
    a = np.random.default_rng(3).integers(0, 256, (9, 13, 4)).astype(np.uint8)
    image.write_png(tmp_path / "a.png", a)
    assert np.array_equal(image.read_png(tmp_path / "a.png"), a)
    image.write_ppm(tmp_path / "a.ppm", a)
    data = open(tmp_path / "a.ppm", "rb").read()
    assert data.startswith(b"P6 13 9 255\n") and len(data) == len(b"P6 13 9 255\n") + 13 * 9 * 3
