"""An imported glTF scene (kittlespt_amd/gltf.py) through the GPU path: its
render equals the oracle's render of the same imported arrays."""
import numpy as np
import pytest

import oracle
from kittlespt_amd import abi, scenes
from kittlespt_amd.context import Context

from common import settings
from test_gltf import _write

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("binary", [False, True])
def test_imported_scene_renders_like_the_oracle(tmp_path, binary):
    from kittlespt_amd import gltf
    path, _, _ = _write(tmp_path, binary)
    sc, cams = gltf.load_gltf(str(path))
    W, H = 48, 40
    ctx = Context(0)
    ctx.upload(sc)
    orc = oracle.OracleScene(sc)
    st = settings(max_depth=4, integrator=abi.INTEGRATOR_WAVEFRONT)
    ctx.set_settings(st)
    ctx.resize(W, H)
    proj = scenes.perspective_fov_lh(cams[0].yfov, W, H, 0.1, 100.0)
    ctx.set_view(proj, cams[0].view)
    cam = ctx.camera()
    print(f"[gltf] binary={binary} render {W}x{H}", flush=True)
    ctx.render(1)
    img = ctx.estimate().reshape(-1, 3)
    o = orc.render_pixels(st, cam, W, H, np.arange(W * H, dtype=np.int32), 0, 1)
    o = o[0] if isinstance(o, tuple) else o
    assert np.array_equal(img.view(np.uint32), np.asarray(o, np.float32).reshape(-1, 3).view(np.uint32))
    ctx.close()
    orc.close()
