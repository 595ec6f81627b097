"""The drop-in, executed: Atrc JSON through the C++ host.

* atrc_host_execute (RenderSession::execute, render_session.cpp:76-96, with
  the "gpu_pt" renderer created from pt's parameter names) gives bit for bit
  the image and AOVs that agx_render gives for the same descriptor and
  parameters, in Atrc's and in the BVH4 acceleration mode.
* atrc_cli -d scene.json (main.cpp:14-72) runs the same session and its
  save_to_img post-processor (save_to_img.cpp:41-67) writes the image flipped
  vertically, each channel as uint8(clamp(c^gamma, 0, 1) * 255) in PNG and
  as RGBE in HDR: the files equal what that rule gives on the session's own
  float image.
* Orientation: an emitter filling the upper half of the view (world +z, the
  camera's up) lands in the top rows of the saved PNG.
"""
import json
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

from atrc_amd import capi, scenes

from conftest import log

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "atrc_amd" / "atrc_cli"


def read_png(path: Path) -> np.ndarray:
    data = path.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    w = h = None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 2
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()          # filter type None on every row
    return rows[:, 1:].reshape(h, w, 3)


def read_hdr(path: Path) -> np.ndarray:
    data = path.read_bytes()
    head_end = data.index(b"\n\n") + 2
    line_end = data.index(b"\n", head_end)
    dims = data[head_end:line_end].split()
    assert dims[0] == b"-Y" and dims[2] == b"+X"
    h, w = int(dims[1]), int(dims[3])
    px = np.frombuffer(data[line_end + 1:], np.uint8)[:h * w * 4].reshape(h, w, 4).astype(np.float64)
    scale = np.where(px[..., 3:] > 0, np.ldexp(1.0, px[..., 3:].astype(np.int64) - 136), 0.0)
    return (px[..., :3] + 0.5) * scale * (px[..., 3:] > 0)


def expected_png(img: np.ndarray, gamma: float) -> np.ndarray:
    v = np.power(img[::-1].astype(np.float32), np.float32(gamma))
    return (np.clip(v, 0, 1) * np.float32(255)).astype(np.uint8)


@pytest.mark.parametrize("accel", ["atrc", "bvh4"])
def test_host_execute_equals_agx_render(gpu_ctx, tmp_path, accel):
    s, cam, tr = scenes.cornell_box()
    js = scenes.to_atrc_json(s, cam, tr, 64, 48, 4, tmp_path, renderer="gpu_pt", seed=42,
                             renderer_extra={"accel": accel})
    log(f"[drop-in] cornell through atrc_host_execute (accel {accel})")
    hs = capi.HostScene(js, str(tmp_path))
    got = hs.execute(0)
    desc = hs.desc()
    sess = hs.session(0)
    desc.accel = capi.ACCEL_ATRC if accel == "atrc" else capi.ACCEL_BVH4
    gpu_ctx.upload_scene(desc)
    p = scenes.render_params(sess.width, sess.height, 4, cam, seed=42, eps=sess.eps,
                             filter_radius=sess.filter_radius, **tr)
    p.camera = sess.camera
    ref = gpu_ctx.render(p)
    for k in ("image", "albedo", "normal", "denoise"):
        assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), k
    assert got["image"].mean() > 0.05
    hs.close()


def test_cli_saves_flipped_gamma_images(tmp_path):
    assert CLI.exists(), "atrc_cli is built by atrc_amd/build.py"
    s, cam, tr = scenes.cornell_box()
    png, hdr = tmp_path / "out" / "cornell.png", tmp_path / "out" / "cornell.hdr"
    pps = [{"type": "save_to_img", "filename": "${scene-directory}/out/cornell.png", "inv_gamma": 2.2},
           {"type": "save_to_img", "filename": "${scene-directory}/out/cornell.hdr"}]
    js = scenes.to_atrc_json(s, cam, tr, 40, 32, 4, tmp_path, renderer="gpu_pt", seed=7, post_processors=pps)
    (tmp_path / "scene.json").write_text(js)
    log("[drop-in] atrc_cli -d scene.json with save_to_img png + hdr")
    res = subprocess.run([str(CLI), "-d", str(tmp_path / "scene.json")], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    # the same session in process gives the float image the files were written from
    hs = capi.HostScene(js, str(tmp_path))
    img = hs.execute(0)["image"]
    hs.close()
    got = read_png(png)
    np.testing.assert_array_equal(got, expected_png(img, 1 / 2.2))
    h = read_hdr(hdr)
    ref = img[::-1].astype(np.float64)
    m = ref.max(axis=2, keepdims=True)
    assert np.all(np.abs(h - ref) <= m * 2.0 ** -7 + 1e-30)


def test_saved_png_keeps_up_at_the_top(tmp_path):
    sc = scenes.Scene(aggregate=capi.AGGREGATE_NATIVE)
    black = sc.add_material(scenes.Material(kind=capi.MAT_IDEAL_BLACK))
    # facing the camera (normal -y), covering world z > 0 of the plane y = 0
    sc.add(scenes.Entity(capi.GEOM_QUAD, black, emit=(0.5, 0.5, 0.5),
                         points=((-100.0, 0.0, 0.0), (100.0, 0.0, 0.0), (100.0, 0.0, 100.0), (-100.0, 0.0, 100.0))))
    cam = {"pos": (0.0, -5.0, 0.0), "dst": (0.0, 0.0, 0.0), "up": (0, 0, 1), "fov": 40.0}
    tr = {"min_depth": 5, "max_depth": 1, "cont_prob": 0.9}
    pps = [{"type": "save_to_img", "filename": "${scene-directory}/half.png", "gamma": 1.0}]
    js = scenes.to_atrc_json(sc, cam, tr, 16, 16, 4, tmp_path, renderer="gpu_pt", post_processors=pps)
    (tmp_path / "half.json").write_text(js)
    res = subprocess.run([str(CLI), "-d", str(tmp_path / "half.json")], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    img = read_png(tmp_path / "half.png")[..., 0]
    log(f"[drop-in] half-plane png rows: {img[:, 8].tolist()}")
    assert (img[:7] == 127).all() and (img[9:] == 0).all()
