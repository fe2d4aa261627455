"""Kernel #1 parity: the device pyramid against the oracle's double pyramid.

The device builds in double with the reference's operation order and
stores fp32 rounded outward, so per node: h_min is the oracle's value
rounded down and h_max rounded up (bit for bit); h0, gu, gv are the
oracle's values rounded to nearest (bit for bit); r, ru, rv enclose the
oracle's remainders plus the rounding of h0, gu, gv, and exceed that bound
by at most a few ulps.
"""
import numpy as np
import pytest

import oracle
from dmap_amd import workload as W

pytestmark = pytest.mark.gpu


def f_down(x):
    f = x.astype(np.float32)
    lo = f.astype(np.float64) > x
    f[lo] = np.nextafter(f[lo], np.float32(-np.inf))
    return f


def f_up(x):
    f = x.astype(np.float32)
    hi = f.astype(np.float64) < x
    f[hi] = np.nextafter(f[hi], np.float32(np.inf))
    return f


def check_tile(tile, texels, strength, offset, repeat, build):
    nodes = W.close_tile(texels, repeat).astype(np.float64)
    ref = oracle.build_pyramid(nodes, strength, build, offset)
    assert tile.n_levels == len(ref)
    np.testing.assert_array_equal(tile.nodes(), W.close_tile(texels, repeat))
    n = texels.shape[0]
    for k, lv in enumerate(ref):
        g = tile.level(k)
        half = 0.5 * 2.0 ** k / n
        np.testing.assert_array_equal(g["h_min"], f_down(lv["h_min"]), err_msg=f"h_min L{k}")
        np.testing.assert_array_equal(g["h_max"], f_up(lv["h_max"]), err_msg=f"h_max L{k}")
        for ch in ("h0", "gu", "gv"):
            np.testing.assert_array_equal(g[ch], lv[ch].astype(np.float32), err_msg=f"{ch} L{k}")
        egu = np.abs(lv["gu"] - g["gu"].astype(np.float64))
        egv = np.abs(lv["gv"] - g["gv"].astype(np.float64))
        need_r = lv["r"] + np.abs(lv["h0"] - g["h0"].astype(np.float64)) + (egu + egv) * half
        assert (g["r"].astype(np.float64) >= need_r).all(), f"r L{k}"
        assert (g["ru"].astype(np.float64) >= lv["ru"] + egu).all(), f"ru L{k}"
        assert (g["rv"].astype(np.float64) >= lv["rv"] + egv).all(), f"rv L{k}"
        for ch, need in (("r", need_r), ("ru", lv["ru"] + egu), ("rv", lv["rv"] + egv)):
            slack = g[ch].astype(np.float64) - need
            ulp = np.spacing(np.abs(need).astype(np.float32)).astype(np.float64)
            assert (slack <= 2 * ulp + 1e-30).all(), f"{ch} L{k} too wide"


@pytest.mark.parametrize("build", ["fold", "direct"])
@pytest.mark.parametrize("n,strength,offset,repeat", [(64, 0.7, -0.1, True), (256, 0.05, 0.0, True),
                                                      (32, 1.0, 0.25, False), (1, 0.3, 0.0, True),
                                                      (2, 0.3, 0.0, True), (8, 0.3, 0.0, False)])
def test_pyramid_matches_oracle(torch_gpu, build, n, strength, offset, repeat):
    import dmap_amd
    rng = np.random.default_rng(n * 7 + (build == "direct"))
    texels = rng.random((n, n)).astype(np.float32)
    tile = dmap_amd.DisplacementTile(texels, strength, offset, repeat, build)
    check_tile(tile, texels, strength, offset, repeat, build)


def test_pyramid_c2_map_fold(torch_gpu):
    """The full 2048^2 map of the incoherent workload, levels checked against the oracle."""
    import dmap_amd
    texels = W.c2_map(2048)
    tile = dmap_amd.DisplacementTile(texels, 0.1, 0.0, True, "fold")
    check_tile(tile, texels, 0.1, 0.0, True, "fold")
