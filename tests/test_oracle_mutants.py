"""The render pins of test_oracle.py must catch plausible mistakes.

Each case plants one mistake in a copy of the oracle's C++ source (on the
CPU, in a temporary directory), builds it, and checks that the exact DDA
walk or the closed-form shading tests fail against the mutant.  The
mistakes are the ones a restatement of vulkan/shaders/raytracer/main.comp
could plausibly make: a flipped light direction, a wrong light colour, a
dropped ambient term, truncating instead of rounding, the INDEX3D bit order,
the DDA's step sign and the child-node offset on descent.
"""
import importlib.util
import os
import subprocess

import numpy as np
import pytest

import test_oracle as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "frep_oracle.cpp")

MUTANTS = {
    "light_sign": ("const F3 ld = F3{-cam->light_dir[l][0], -cam->light_dir[l][1],\n"
                   "                                             -cam->light_dir[l][2]};",
                   "const F3 ld = F3{cam->light_dir[l][0], cam->light_dir[l][1],\n"
                   "                                             cam->light_dir[l][2]};"),
    "light_colour": ("it * cam->light_color[l][c];", "it * cam->light_color[l][(c + 1) & 3];"),
    "no_ambient": ("col[c] = cam->ambient[c] + diffuse[c];", "col[c] = diffuse[c];"),
    "truncate": ("(uint8_t)(v * 255.0f + 0.5f)", "(uint8_t)(v * 255.0f)"),
    "index_order": ("const uint32_t idx = (uint32_t)cell[2] + dim * (uint32_t)cell[1] +\n"
                    "                                 dim * dim * (uint32_t)cell[0];",
                    "const uint32_t idx = (uint32_t)cell[0] + dim * (uint32_t)cell[1] +\n"
                    "                                 dim * dim * (uint32_t)cell[2];"),
    "step_sign": ("iv[0] >= 0.0f ? 1.0f : 0.0f, iv[1]", "iv[0] >= 0.0f ? 0.0f : 1.0f, iv[1]"),
    "descend_offset": ("F3{(float)cell[0] * cs, (float)cell[1] * cs,",
                       "F3{(float)cell[0] * cs, (float)cell[0] * cs,"),
}


def _load_mutant(tmp_path, name):
    src = open(SRC).read()
    old, new = MUTANTS[name]
    assert src.count(old) == 1, f"mutation site of {name} not found exactly once"
    d = tmp_path / name
    d.mkdir()
    (d / "frep_oracle.cpp").write_text(src.replace(old, new))
    (d / "frep_oracle.h").write_text(open(os.path.join(ROOT, "oracle", "frep_oracle.h")).read())
    so = d / "liboracle.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-shared", "-o",
                    str(so), str(d / "frep_oracle.cpp"), "-lpthread"], check=True)
    spec = importlib.util.spec_from_file_location(f"oracle_{name}",
                                                  os.path.join(ROOT, "oracle", "oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.LIB_PATH = str(so)
    mod._lib = None
    return mod


def _render_checks(O):
    yield "dda_axis", lambda: T.test_dda_exact_axis_rays_through_cell_centres(O)
    yield "dda_persp", lambda: T.test_dda_exact_perspective_rays(
        O, [4, 4], (2.3, 1.7, 2.9), (0.05, -0.03, 0.02))
    yield "dda_persp3", lambda: T.test_dda_exact_perspective_rays(
        O, [4, 4, 4], (0.4, -2.8, 1.3), (0.0, 0.1, 0.0))
    yield "shade_facing", lambda: T.test_shading_closed_form_half_spaces(
        O, (0.0, 0.0, 1.0), 0.3, None, (0.2, 0.2, 0.2, 1.0))
    yield "shade_oblique", lambda: T.test_shading_closed_form_half_spaces(
        O, (0.6, 0.3, 0.742), 0.1, None, (0.2, 0.2, 0.2, 1.0))


def test_unmutated_oracle_passes_the_render_checks(oracle):
    for name, check in _render_checks(oracle):
        check()


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_render_checks_catch_the_mutant(tmp_path, name):
    O = _load_mutant(tmp_path, name)
    caught = []
    with np.errstate(all="ignore"):
        for check_name, check in _render_checks(O):
            try:
                check()
            except AssertionError:
                caught.append(check_name)
    assert caught, f"no render check fails on the {name} mutant"
