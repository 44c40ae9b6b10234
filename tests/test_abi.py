"""The C-ABI library loads and exports every symbol include/*.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(msbg_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_header_declares_entry_points():
    d = _declared()
    assert "msbg_apply_channel_pde_fast" in d and "msbg_splat_points" in d
    assert len(d) >= 25


def test_library_exports_every_declared_symbol():
    from msbg_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_scene_generator_is_deterministic():
    import numpy as np
    from msbg_amd import scene_spheres
    a = scene_spheres(1234, 5, 64, 448, 20, 60, 1000, 0.5, 0.5)
    b = scene_spheres(1234, 5, 64, 448, 20, 60, 1000, 0.5, 0.5)
    c = scene_spheres(1235, 5, 64, 448, 20, 60, 1000, 0.5, 0.5)
    assert a.shape == (5000, 3) and np.array_equal(a, b) and not np.array_equal(a, c)
    # sphere points lie at radius r +- a few sigma from their centre
    pts = scene_spheres(7, 1, 256, 256, 40, 40, 20000, 0.5, 0.0)
    r = np.linalg.norm(pts.astype(np.float64) - 256.0, axis=1)
    assert abs(r.mean() - 40) < 0.05 and 0.4 < r.std() < 0.6


def test_create_without_device_fails_cleanly(have_gpu):
    if have_gpu:
        pytest.skip("a device is present")
    import msbg_amd as M
    with pytest.raises(M.MsbgError) as e:
        M.MultiresSparseGrid.create("x", 64, 64, 64, 16)
    assert "EDEVICE" in str(e.value)


def test_scene_objects_perlin_displacement():
    import numpy as np
    from msbg_amd import scene_objects, scene_spheres
    a = scene_spheres(99, 3, 100, 400, 30, 60, 2000, 0.5, 0.5)
    b = scene_objects(99, 3, 100, 400, 30, 60, 2000, 0.5, 0.5, 0.0, 0.0)
    assert np.array_equal(a, b)
    # one sphere of radius 50, 30 % noise amplitude: radii stay inside the
    # displaced band and vary far more than the jitter alone
    p = scene_objects(5, 1, 256, 256, 50, 50, 20000, 0.0, 0.0, 0.3, 1.0 / 16)
    r = np.linalg.norm(p.astype(np.float64) - 256.0, axis=1)
    assert r.min() >= 50 * 0.7 - 1e-3 and r.max() <= 50 * 1.3 + 1e-3
    assert r.std() > 2.0
    # the same seed gives the same field
    assert np.array_equal(p, scene_objects(5, 1, 256, 256, 50, 50, 20000, 0.0, 0.0, 0.3, 1.0 / 16))
