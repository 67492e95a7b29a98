"""Regenerates the golden vectors in this folder from the reference's own
sobol.h and camera.h, compiled from their sources by `make -C oracle ref`
into oracle/_ref/ref_driver, and copies the reference demo's Cornell box
scene (bin/assets/cornelbox.obj, gzip-compressed, and cornelbox.mtl; CC BY
3.0, attribution in cornelbox/LICENSE.txt) into cornelbox/.  Run in a
checkout that has the reference sources; the outputs are committed so the
tests need no reference.

    python tests/golden/make_golden.py
"""
import gzip
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
REF_ASSETS = "/root/reference/bin/assets"


def copy_cornell_box():
    out = os.path.join(HERE, "cornelbox")
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(REF_ASSETS, "cornelbox.obj"), "rb") as fh:
        data = fh.read()
    with open(os.path.join(out, "cornelbox.obj.gz"), "wb") as fh:
        # mtime 0: the same bytes on every run
        with gzip.GzipFile(filename="cornelbox.obj", mode="wb", fileobj=fh, compresslevel=9, mtime=0) as gz:
            gz.write(data)
    with open(os.path.join(REF_ASSETS, "cornelbox.mtl"), "rb") as src, \
            open(os.path.join(out, "cornelbox.mtl"), "wb") as dst:
        dst.write(src.read())
    lic = ["Scene data of the MinimumLBVH demo (bin/assets/cornelbox.obj, cornelbox.mtl),",
           "stored here as a test fixture.  Its licence and attribution files follow.", ""]
    for name in ("cornelbox_copyright.txt", "Head_Infinite-Scan_License.txt"):
        with open(os.path.join(REF_ASSETS, "license", name), "r", encoding="utf-8", errors="replace") as fh:
            lic += [f"===== {name} =====", fh.read().rstrip(), ""]
    with open(os.path.join(out, "LICENSE.txt"), "w") as fh:
        fh.write("\n".join(lic) + "\n")


def run(mode, inputs, out_cols):
    with tempfile.TemporaryDirectory() as d:
        i = os.path.join(d, "in.bin")
        o = os.path.join(d, "out.bin")
        inputs.tofile(i)
        subprocess.run([DRIVER, mode, i, o], check=True)
        return np.fromfile(o, np.float32).reshape(-1, out_cols)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    rng = np.random.default_rng(20240601)
    # Sobol: sample indices used by the AO/PT kernels (0..16, spp counters),
    # pixel coordinates, seeds 0 and 12345 and bounce dimensions, plus random
    # 32-bit quadruples and the extremes.
    q = [[i, x, y, s] for i in range(17) for (x, y) in [(0, 0), (255, 255), (17, 200), (1919, 1079)]
         for s in (0, 12345, 3)]
    q += [[0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF], [0x80000000, 1, 2, 3]]
    q += rng.integers(0, 2**32, size=(400, 4), dtype=np.uint64).tolist()
    q = np.array(q, dtype=np.uint64).astype(np.uint32)
    xy = run("sobol", q, 2)
    np.savez(os.path.join(HERE, "sobol.npz"), inputs=q, outputs=xy)

    cams = []
    setups = [((0.5, 0.5, -1.5), (0.5, 0.5, -0.5), (0, 1, 0), 60.0, 256, 256),
              ((0.0, 0.3, -2.6), (0.0, 0.0, 0.0), (0, 1, 0), 60.0, 1920, 1080),
              ((0.0, 8.0, -26.0), (0.0, 1.0, 0.0), (0, 1, 0), 60.0, 1920, 1080),
              ((0.0, 1.0, 4.0), (0.0, 1.0, 0.0), (0, 1, 0), 45.0, 640, 480)]
    for eye, ctr, up, fov, w, h in setups:
        for (u, v) in [(0.0, 0.0), (0.5, 0.5), (1.0, 1.0), (0.25, 0.75)] + rng.random((20, 2)).tolist():
            cams.append(list(eye) + list(ctr) + list(up) + [math.radians(fov), w, h, u, v])
    cams = np.array(cams, np.float32)
    out = run("camera", cams, 15)
    np.savez(os.path.join(HERE, "camera.npz"), inputs=cams, outputs=out)
    print("wrote", len(q), "sobol and", len(cams), "camera vectors")
    copy_cornell_box()
    print("copied the Cornell box scene")


if __name__ == "__main__":
    sys.exit(main())
