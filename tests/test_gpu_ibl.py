"""The ibl environment light on the gfx950 path against the oracle
(tests/test_oracle_ibl.py pins it): path for path at 1 spp on spheres lit by
an image environment with a bright spot (importance-sampled NEE and MIS), and
a constant-ibl white furnace."""
import numpy as np
import pytest

from atrc_amd import capi, scenes

from conftest import log
from test_gpu_render_parity import _path_parity
from test_oracle_ibl import _bright_env

pytestmark = pytest.mark.gpu


def _ibl_scene(linear):
    s, cam, tr = scenes.procedural_spheres(n_spheres=10, subdiv=3, seed=8, extent=12.0)
    s.env = {"image": _bright_env(64, 32) * 0.5, "linear": linear, "wrap": ("repeat", "clamp")}
    cam = {"pos": (-8.0, -10.0, 14.0), "dst": (6.0, 6.0, 4.0), "up": (0, 0, 1), "fov": 45.0}
    return s, cam, tr


@pytest.mark.parametrize("linear", [True, False])
def test_ibl_path_parity(gpu_ctx, linear):
    s, cam, tr = _ibl_scene(linear)
    s.accel = capi.ACCEL_BVH4
    _path_parity(gpu_ctx, s, cam, tr, 96, 64, f"spheres under an image ibl (linear={linear})", min_frac=0.97)


def test_ibl_constant_furnace(gpu_ctx):
    s = scenes.Scene(aggregate=capi.AGGREGATE_NATIVE)
    m = s.add_material(scenes.Material(color=(1.0, 1.0, 1.0)))
    s.add(scenes.Entity(capi.GEOM_SPHERE, m, radius=1.0))
    s.env = {"texel": (1.0, 1.0, 1.0)}
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    cam = {"pos": (0.0, -4.0, 0.0), "dst": (0.0, 0.0, 0.0), "up": (0, 0, 1), "fov": 40.0}
    img = gpu_ctx.render(scenes.render_params(32, 32, 256, cam, seed=2, min_depth=5, max_depth=40, cont_prob=0.9),
                         aovs=False)["image"].astype(np.float64)
    log(f"[ibl] constant furnace mean {img.mean():.5f}")
    assert abs(img.mean() - 1) < 0.01
