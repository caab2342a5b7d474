"""Field export (examples/dec_fluid.rs:143-164): the GPU's RGB8 pixels equal
the oracle's bit for bit (transfer, row flip, f32/f64, 3D slices), and the
PNG written from them decodes to the same pixels."""
import numpy as np
import pytest

import panopaea_amd as pa
from oracle import oracle as o
from test_oracle import _png_decode

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gpu_field_rgb8_matches_oracle(ctx, dtype, tmp_path):
    g = pa.Grid2d((48, 80), dtype)
    rng = np.random.default_rng(5)
    rho = (rng.standard_normal(48 * 80) * 1.5).astype(dtype)
    rho[:7] = [np.nan, np.inf, -np.inf, -2.0, 2.0, 0.0, 1e-30]
    f = g.new_simplex_2().upload(rho)
    img = pa.field_rgb8(f, -2.0, 2.0)
    want = o.field_rgb8(48, 80, rho, -2.0, 2.0)
    assert np.array_equal(img, want)
    path = tmp_path / "density.png"
    pa.png_export(str(path), img)
    assert np.array_equal(_png_decode(path.read_bytes()), want)


def test_gpu_field_rgb8_3d_slices(ctx):
    g = pa.Grid3d((6, 16, 24), np.float32)
    rho = np.random.default_rng(9).uniform(-3, 3, 6 * 16 * 24).astype(np.float32)
    f = g.new_simplex_2().upload(rho)
    for z in (0, 3, 5):
        plane = rho[z * 16 * 24:(z + 1) * 16 * 24]
        assert np.array_equal(pa.field_rgb8(f, -1.0, 1.5, z=z), o.field_rgb8(16, 24, plane, -1.0, 1.5))
    with pytest.raises(pa.PanoError):
        pa.field_rgb8(f, z=6)
