"""Height maps on disk and the reference's synthetic maps (texture_grid.cpp),
pinned by closed forms and by round trips through PNG files written with
every row filter, at 8 and 16 bits, grey to RGBA."""
import numpy as np
import pytest

from dmap_amd import tiles as T


@pytest.mark.parametrize("channels", [1, 2, 3, 4])
@pytest.mark.parametrize("depth", [8, 16])
def test_png_round_trip_all_filters(tmp_path, channels, depth):
    rng = np.random.default_rng(channels * 10 + depth)
    top = 255 if depth == 8 else 65535
    img = rng.integers(0, top + 1, size=(13, 11, channels))
    path = tmp_path / "m.png"
    T.write_png(path, img, depth=depth, filters=[0, 1, 2, 3, 4])
    vals, raw = T.read_png(path)
    assert np.array_equal(raw, img)
    nc = min(channels, 3)
    assert np.allclose(vals, img[:, :, :nc].sum(axis=2) / nc / top, rtol=0, atol=1e-15)
    assert vals.min() >= 0 and vals.max() <= 1


def test_png_scaling_known_values(tmp_path):
    # 8-bit grey 0, 51, 255 -> 0, 0.2, 1; 16-bit 65535 -> 1
    T.write_png(tmp_path / "a.png", np.array([[0, 51, 255]]), depth=8)
    assert np.allclose(T.read_png(tmp_path / "a.png")[0], [[0.0, 0.2, 1.0]])
    T.write_png(tmp_path / "b.png", np.array([[0, 32768, 65535]]), depth=16)
    assert np.allclose(T.read_png(tmp_path / "b.png")[0], [[0.0, 32768 / 65535, 1.0]])
    # RGB: the mean of the three channels; alpha ignored
    T.write_png(tmp_path / "c.png", np.array([[[255, 0, 0, 7]]]), depth=8)
    assert np.allclose(T.read_png(tmp_path / "c.png")[0], [[1 / 3]])


def test_tile_file_round_trip(tmp_path):
    tex = np.random.default_rng(1).random((32, 32)).astype(np.float32)
    T.save_tile(tmp_path / "t.dmt", tex, strength=0.05, offset=-0.025, repeat=False)
    got, s, off, rep = T.load_tile(tmp_path / "t.dmt")
    assert np.array_equal(np.asarray(got), tex) and s == 0.05 and off == -0.025 and rep is False
    assert np.array_equal(T.load_height_texture(tmp_path / "t.dmt"), tex.astype(np.float64))
    np.save(tmp_path / "t.npy", tex)
    assert np.array_equal(T.load_height_texture(tmp_path / "t.npy"), tex.astype(np.float64))
    with pytest.raises(ValueError):
        T.load_height_texture(tmp_path / "x.exr")


def test_synthetic_maps_closed_forms():
    cb = T.checkerboard(8, 4, 0.1, 0.9)
    assert cb[0, 0] == 0.1 and cb[0, 2] == 0.9 and cb[2, 0] == 0.9 and cb[2, 2] == 0.1 and cb[1, 1] == 0.1
    g = T.gradient_map(16, 0.0, 1.0)
    assert np.isclose(g.mean(), 0.5) and np.allclose(g, g.T) and g[0, 0] < g[-1, -1]
    assert np.isclose(g[0, 0], 0.5 / 16)
    s = T.gaussian_spot(17, 0.5, 0.5, 0.1, 2.0, 0.25)
    assert np.isclose(s[8, 8], 2.25) and s.min() > 0.25 and np.allclose(s, s.T)


def test_downsample_box():
    assert np.allclose(T.downsample_box(np.full((64, 64), 0.3), 16), 0.3)
    # a linear ramp's block means are the ramp at the block centres
    x = np.arange(64) + 0.5
    ramp = 2.0 * x[None, :] + 3.0 * x[:, None]
    d = T.downsample_box(ramp, 8)
    c = np.arange(8) * 8 + 4.0
    assert np.allclose(d, 2.0 * c[None, :] + 3.0 * c[:, None])
    # extra rows and columns beyond whole blocks are dropped
    assert T.downsample_box(np.arange(70 * 70, dtype=float).reshape(70, 70), 8).shape == (8, 8)
