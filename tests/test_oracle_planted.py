"""The closed-form pins of the oracle's shading composition catch the
mistakes they are meant to catch.  Each case copies oracle/mlbvh_oracle.c,
plants one plausible mistake in the copy (a dropped term, a wrong sign, a
transposed operand, a wrong constant), builds it with the oracle's own flags
into a temporary library, and runs the named closed-form tests of
tests/test_oracle_sampling.py against it: each must fail.  The unplanted
oracle passes the same tests in the ordinary suite.

Composition restated (main_gpu.cu): normal image :84-109, AO :111-177, ONB
:39-45, hemisphere direction :152-153 and :231-235, path tracing :179-271."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from oracle import pyoracle
from tests import test_oracle_sampling as pins

ORACLE_SRC = os.path.join(os.path.dirname(pyoracle.__file__), "mlbvh_oracle.c")
ORACLE_HDR = os.path.join(os.path.dirname(pyoracle.__file__), "mlbvh_oracle.h")
SORT_OBJ = os.path.join(os.path.dirname(pyoracle.__file__), "_build", "sort_comparison.o")
CFLAGS = ["-std=gnu11", "-O2", "-fPIC", "-fno-semantic-interposition", "-fopenmp", "-ffp-contract=off",
          "-fno-fast-math", "-w"]

# name: (text in the oracle, planted text, [(test, kwargs), ...] of which at least one must fail)
PLANTS = {
    "ao normal not turned toward the ray": (
        "            if (0.0f < vdot(n, rd)) n = vneg(n);\n            orc_v3 xaxis, zaxis;",
        "            orc_v3 xaxis, zaxis;",
        [(pins.test_ao_closed_box_is_fully_occluded, {"n_sample": 4, "winding": "as_built"}),
         (pins.test_ao_closed_box_is_fully_occluded, {"n_sample": 4, "winding": "reversed"})]),
    "ao origin lifted to the wrong side": (
        "vmuls(n, 0.001f)", "vmuls(n, -0.001f)",
        [(pins.test_ao_lone_quad_is_unoccluded, {"up": True}),
         (pins.test_ao_between_parallel_planes_closed_form, {})]),
    "ao origin not lifted": (
        "vmuls(n, 0.001f)", "vmuls(n, 0.0f)",
        [(pins.test_ao_between_parallel_planes_closed_form, {})]),
    "ao accumulates 1/(N+1)": (
        "if (occluded) ao += 1.0f / (float)nSample;", "if (occluded) ao += 1.0f / (float)(nSample + 1);",
        [(pins.test_ao_closed_box_is_fully_occluded, {"n_sample": 16, "winding": "as_built"})]),
    "ao gamma 1/2": (
        "float brightness = orc_pow_pos(1.0f - ao, 1.0f / 2.2f);",
        "float brightness = orc_pow_pos(1.0f - ao, 1.0f / 2.0f);",
        [(pins.test_ao_between_parallel_planes_closed_form, {})]),
    "onb y axis z component sign": (
        "*yaxis = V(b, sign + z.y * z.y * a, -z.y);", "*yaxis = V(b, sign + z.y * z.y * a, z.y);",
        [(pins.test_onb_is_orthonormal_and_right_handed, {})]),
    "onb without the sign of z": (
        "const float sign = copysignf(1.0f, z.z);", "const float sign = 1.0f;",
        [(pins.test_onb_is_orthonormal_and_right_handed, {})]),
    "direction with normal and tangent components transposed": (
        "return vadd(vadd(vmuls(xaxis, dl.x), vmuls(zaxis, dl.z)), vmuls(n, dl.y));",
        "return vadd(vadd(vmuls(xaxis, dl.x), vmuls(zaxis, dl.y)), vmuls(n, dl.z));",
        [(pins.test_cosine_directions_around_a_normal, {})]),
    "normal image of the flipped normal": (
        "                orc_v3 n = vnormalize(V(hit.ngx, hit.ngy, hit.ngz));\n                orc_v3 c = ",
        "                orc_v3 n = vnormalize(V(hit.ngx, hit.ngy, hit.ngz));\n"
        "                if (0.0f < vdot(n, rd)) n = vneg(n);\n                orc_v3 c = ",
        [(pins.test_normal_image_of_an_axis_aligned_quad, {"up": False})]),
    "path throughput without the albedo": (
        "tp = vmul(tp, at->reflectance);", "tp = vmul(tp, V(1.0f, 1.0f, 1.0f));",
        [(pins.test_pt_one_bounce_to_an_emissive_ceiling, {"up": True})]),
    "path origin lifted to the wrong side": (
        "vmuls(hn, 0.0001f)", "vmuls(hn, -0.0001f)",
        [(pins.test_pt_one_bounce_to_an_emissive_ceiling, {"up": True})]),
}


@pytest.fixture(scope="module")
def planted_libs(tmp_path_factory):
    pyoracle.build_lib()
    if not os.path.exists(SORT_OBJ):
        pytest.skip("oracle object files not built")
    src = open(ORACLE_SRC).read()
    root = tmp_path_factory.mktemp("planted")

    def build(item):
        i, (name, (old, new, _)) = item
        assert src.count(old) == 1, name  # the planted spot exists exactly once
        d = root / f"p{i}"
        d.mkdir()
        (d / "mlbvh_oracle.c").write_text(src.replace(old, new))
        shutil.copy(ORACLE_HDR, d / "mlbvh_oracle.h")
        out = d / "liboracle_planted.so"
        r = subprocess.run(["gcc", *CFLAGS, "-shared", "-o", str(out), str(d / "mlbvh_oracle.c"), SORT_OBJ, "-lm",
                            "-lstdc++"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        return name, str(out)

    with ThreadPoolExecutor(4) as ex:
        return dict(ex.map(build, enumerate(PLANTS.items())))


@pytest.mark.parametrize("name", list(PLANTS))
def test_a_planted_mistake_fails_a_closed_form_pin(planted_libs, name):
    _, _, checks = PLANTS[name]
    saved = pyoracle._lib
    pyoracle._lib = pyoracle.load(planted_libs[name])
    failed = []
    try:
        for fn, kwargs in checks:
            try:
                fn(pyoracle, **kwargs)
            except AssertionError:
                failed.append(fn.__name__)
    finally:
        pyoracle._lib = saved
    assert failed, f"no closed-form pin caught: {name}"
