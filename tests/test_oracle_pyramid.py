"""The oracle's pyramid (taylor_pyramid.cpp restated) pinned by the
reference's numpy golden vectors and by the invariants of
code/python/poc/tests/test_pyramid.py and validate_pyramid.cpp:121-184."""
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"
RNG = np.random.default_rng(11)


def bilinear(values, scale, X, Y):
    """h, h_u, h_v of the bilinear interpolant at (X, Y) in [0, 1]^2 (numpy, for the checks)."""
    V = np.asarray(values, dtype=np.float64) * scale
    n = V.shape[0] - 1
    fx, fy = X * n, Y * n
    i = np.minimum(fx.astype(np.int64), n - 1)
    j = np.minimum(fy.astype(np.int64), n - 1)
    s, t = fx - i, fy - j
    v00, v10, v01, v11 = V[j, i], V[j, i + 1], V[j + 1, i], V[j + 1, i + 1]
    h = (1 - s) * (1 - t) * v00 + s * (1 - t) * v10 + (1 - s) * t * v01 + s * t * v11
    hu = ((1 - t) * (v10 - v00) + t * (v11 - v01)) * n
    hv = ((1 - s) * (v01 - v00) + s * (v11 - v10)) * n
    return h, hu, hv


@pytest.mark.parametrize("case", ["random_65", "step_17", "random_17"])
@pytest.mark.parametrize("build", ["fold", "direct"])
def test_matches_reference_golden(case, build):
    """validate_pyramid.cpp:111-119 at tolerance 1e-12 (validate_pyramid.toml:4)."""
    z = np.load(GOLDEN / f"pyramid_{case}.npz")
    levels = oracle.build_pyramid(z["values"], z["scale"][0], build)
    assert len(levels) == (z["values"].shape[0] - 1).bit_length()
    worst = 0.0
    for k, lv in enumerate(levels):
        for ch in oracle.CHANNELS:
            ref = z[f"{build}/{ch}_L{k}"]
            worst = max(worst, np.abs(lv[ch] - ref).max() / max(np.abs(ref).max(), 1e-300))
    assert worst <= 1e-12


def containment(values, scale, build, m=5, tol=1e-10):
    """Dense h and gradient samples inside every ancestor node (test_pyramid.py:34-51)."""
    levels = oracle.build_pyramid(values, scale, build)
    n = values.shape[0] - 1
    x = (np.arange(n * m) + 0.5) / (n * m)
    X, Y = np.meshgrid(x, x, indexing="xy")
    h, hu, hv = bilinear(values, scale, X, Y)
    for k, lv in enumerate(levels):
        w = 2.0 ** k / n
        mk = lv["h0"].shape[0]
        ci = np.minimum((X / w).astype(np.int64), mk - 1)
        cj = np.minimum((Y / w).astype(np.int64), mk - 1)
        u0, v0 = (ci + 0.5) * w, (cj + 0.5) * w
        plane = lv["h0"][cj, ci] + lv["gu"][cj, ci] * (X - u0) + lv["gv"][cj, ci] * (Y - v0)
        assert (np.abs(h - plane) <= lv["r"][cj, ci] + tol).all(), k
        assert (np.abs(hu - lv["gu"][cj, ci]) <= lv["ru"][cj, ci] + tol).all(), k
        assert (np.abs(hv - lv["gv"][cj, ci]) <= lv["rv"][cj, ci] + tol).all(), k
        assert (h >= lv["h_min"][cj, ci] - tol).all() and (h <= lv["h_max"][cj, ci] + tol).all(), k


@pytest.mark.parametrize("build", ["fold", "direct"])
def test_containment_random(build):
    containment(RNG.random((33, 33)), 0.4, build)


@pytest.mark.parametrize("build", ["fold", "direct"])
def test_containment_step(build):
    v = np.zeros((17, 17))
    v[:, 9:] = 1.0
    containment(v, 1.0, build)


def test_leaf_is_exact():
    """The leaf remainder equals the supremum over the texel corners (test_pyramid.py:54-77)."""
    values, scale = RNG.random((9, 9)), 0.7
    lv = oracle.build_pyramid(values, scale)[0]
    n, w = 8, 1.0 / 8
    for i, j in [(0, 0), (3, 5), (7, 7), (2, 6)]:
        X = np.array([i, i + 1, i, i + 1]) * w
        Y = np.array([j, j, j + 1, j + 1]) * w
        h, _, _ = bilinear(values, scale, X, Y)
        plane = lv["h0"][j, i] + lv["gu"][j, i] * (X - (i + 0.5) * w) + lv["gv"][j, i] * (Y - (j + 0.5) * w)
        assert abs(np.abs(h - plane).max() - lv["r"][j, i]) < 1e-12


@pytest.mark.parametrize("build", ["fold", "direct"])
def test_minmax_channel_is_exact(build):
    """h_min/h_max equal the brute-force min-max pyramid bit for bit (test_pyramid.py:93-104)."""
    values, scale = RNG.random((33, 33)), 0.4
    levels = oracle.build_pyramid(values, scale, build)
    V = values * scale
    for k, lv in enumerate(levels):
        c = 2 ** k
        m = 32 // c
        for J in range(m):
            for I in range(m):
                block = V[J * c:(J + 1) * c + 1, I * c:(I + 1) * c + 1]
                assert lv["h_min"][J, I] == block.min() and lv["h_max"][J, I] == block.max()


def test_direct_vs_fold():
    """Same gradient hulls and range; direct r only smaller; |dh0| <= r_f - r_d (test_pyramid.py:107-123)."""
    values, scale = RNG.random((33, 33)), 0.4
    f = oracle.build_pyramid(values, scale, "fold")
    d = oracle.build_pyramid(values, scale, "direct")
    for k in range(len(f)):
        for key in ("gu", "gv", "ru", "rv", "h_min", "h_max"):
            mag = max(np.abs(f[k][key]).max(), 1.0)
            assert np.abs(d[k][key] - f[k][key]).max() <= 1e-12 * mag, (key, k)
        assert (d[k]["r"] <= f[k]["r"] + 1e-15).all()
        assert (np.abs(d[k]["h0"] - f[k]["h0"]) <= f[k]["r"] - d[k]["r"] + 1e-12).all()
    for key in oracle.CHANNELS:
        assert np.abs(d[0][key] - f[0][key]).max() <= 1e-12 * max(np.abs(f[0][key]).max(), 1.0)


@pytest.mark.parametrize("build", ["fold", "direct"])
def test_ramp_has_zero_remainders(build):
    """A plane is explained exactly at every level (test_pyramid.py:137-146)."""
    jj, ii = np.mgrid[0:17, 0:17]
    values = 0.3 * ii / 16 - 0.2 * jj / 16
    for lv in oracle.build_pyramid(values, 1.0, build):
        assert (lv["r"] < 1e-14).all()
        assert (lv["ru"] < 1e-13).all() and (lv["rv"] < 1e-13).all()
        assert np.allclose(lv["gu"], 0.3) and np.allclose(lv["gv"], -0.2)


def test_offset_moves_value_and_range_only():
    """taylor_pyramid.cpp:242-250: the offset shifts h0, h_min, h_max."""
    values = RNG.random((17, 17))
    a = oracle.build_pyramid(values, 0.5, "fold", 0.0)
    b = oracle.build_pyramid(values, 0.5, "fold", -0.25)
    for la, lb in zip(a, b):
        for ch in ("gu", "gv", "r", "ru", "rv"):
            assert (la[ch] == lb[ch]).all()
        for ch in ("h0", "h_min", "h_max"):
            assert np.allclose(lb[ch], la[ch] - 0.25, atol=1e-15)


def test_recovery_contains_stored_range():
    """validate_pyramid.cpp:153-163 (minmax_from_taylor)."""
    values = RNG.random((33, 33))
    for k, lv in enumerate(oracle.build_pyramid(values, 0.6)):
        s = 0.5 * 2.0 ** k / 32
        half = np.abs(lv["gu"]) * s + np.abs(lv["gv"]) * s + lv["r"]
        assert (lv["h0"] - half <= lv["h_min"] + 1e-12).all()
        assert (lv["h0"] + half >= lv["h_max"] - 1e-12).all()


def test_rejects_bad_grid():
    with pytest.raises(ValueError):
        oracle.build_pyramid(np.zeros((10, 10)), 1.0)


def test_close_tile():
    t = RNG.random((4, 4))
    r = oracle.close_tile(t, repeat=True)
    assert (r[:4, :4] == t).all() and (r[4, :4] == t[0]).all() and (r[:4, 4] == t[:, 0]).all() and r[4, 4] == t[0, 0]
    c = oracle.close_tile(t, repeat=False)
    assert (c[4, :4] == t[3]).all() and (c[:4, 4] == t[:, 3]).all()
