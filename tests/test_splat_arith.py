"""The float identities the tiled splat kernel (msbg_amd/csrc/k_splat.hip)
relies on in place of the reference's expressions, checked on the CPU in
IEEE single precision:

- the distance code (uint16_t)roundf(x), x = f * 65535 (renderDensFromFloat_,
  sbg.h:272-315), as trunc(x + 0.49999997f) clamped to 65535, for every float
  x in [0, 65536], and 65535 or more for every squared distance just past
  distSqMax;
- the footprint face split t -> (t / nx, t % nx) as (t * magic) >> 16 with
  magic = trunc(65536 * rcp(nx)) + 1, for every rcp within one ulp;
- the tile halo H = floor(rScan - 0.5 + eps) + 1, which must hold every
  footprint box of msbg_demo.cpp:571-614 for particles of the cell."""
import numpy as np
import pytest

HALF_DOWN = np.float32(0.49999997)


def test_half_down_constant():
    assert HALF_DOWN == np.nextafter(np.float32(0.5), np.float32(0))


def test_code_rounding_every_float():
    hi = int(np.float32(65536.0).view(np.uint32)) + 1
    step = 1 << 25
    for s in range(0, hi, step):
        x = np.arange(s, min(s + step, hi), dtype=np.uint32).view(np.float32)
        got = np.minimum(np.trunc(x + HALF_DOWN), np.float32(65535.0))
        want = np.minimum(np.floor(x.astype(np.float64) + 0.5), 65535.0)  # roundf, x >= 0
        bad = np.nonzero(got.astype(np.float64) != want)[0]
        assert bad.size == 0, x[bad[:5]]


def test_code_rounding_plants_fail():
    # the plain floor(x + 0.5f) rounds 0.49999997 up: the identity needs the
    # smaller constant
    x = np.float32(0.49999997)
    assert np.trunc(x + np.float32(0.5)) == 1.0
    assert np.trunc(x + HALF_DOWN) == 0.0


@pytest.mark.parametrize("rscan", [1.0, 1.5, 2.0, 3.0, 4.0, 4.25])
def test_out_of_range_codes_are_empty(rscan):
    # d2 just past distSqMax, through both multiply forms, never codes below 65535
    D = np.float32(rscan) * np.float32(rscan)
    inv = np.float32(1.0) / D
    d2 = D
    for _ in range(64):
        d2 = np.nextafter(d2, np.float32(np.inf))
        x = (d2 * inv) * np.float32(65535.0)
        assert np.trunc(x + HALF_DOWN) >= 65535.0
        m, e = np.frexp(inv)
        if m == 0.5:
            x1 = d2 * (np.float32(65535.0) * inv)
            assert x1 == x


def test_one_multiply_matches_two_for_power_of_two_inverse():
    rng = np.random.default_rng(5)
    for inv in (np.float32(1 / 16), np.float32(1 / 4), np.float32(1.0), np.float32(1 / 64)):
        d2 = rng.uniform(0, 1.0 / inv, 200000).astype(np.float32)
        two = (d2 * inv) * np.float32(65535.0)
        one = d2 * (np.float32(65535.0) * inv)
        assert np.array_equal(one, two)


def test_face_split_magic_every_rcp():
    t = np.arange(64)
    for nx in range(1, 65):
        r = np.float32(1.0) / np.float32(nx)
        for rr in (np.nextafter(r, np.float32(0)), r, np.nextafter(r, np.float32(2))):
            magic = int(np.float32(65536.0) * rr) + 1
            assert np.array_equal((t * magic) >> 16, t // nx), nx


def _halo(S, rscan):
    m = S + rscan + 1.0
    eps = 2.0 ** (np.ceil(np.log2(m)) - 23)
    return max(int(np.floor(rscan - 0.5 + eps)) + 1, 1)


@pytest.mark.parametrize("rscan", [1.0, 1.5, 2.0, 2.5, 3.0, 3.75, 4.0, 4.25, 4.4])
def test_halo_holds_every_box(rscan):
    # particles inside cells and on both sides of their faces, in float32, with
    # the box formed as the kernel forms it: ceil((p - r) - 0.5), floor((p + r) - 0.5)
    S, cell = 4096, 16
    H = _halo(S, rscan)
    assert H <= 4
    r = np.float32(rscan)
    rng = np.random.default_rng(int(rscan * 100))
    faces = (rng.integers(1, S // cell, 3000) * cell).astype(np.float32)
    p = np.concatenate([
        rng.uniform(0, S, 20000).astype(np.float32),
        faces, np.nextafter(faces, np.float32(0)), np.nextafter(faces, np.float32(S)),
        faces + np.float32(0.5), faces - np.float32(0.5)])
    c = np.floor(p / cell).astype(np.int64) * cell  # the particle's cell (floor(p) / kCell)
    lo = np.ceil((p - r) - np.float32(0.5)).astype(np.int64)
    hi = np.floor((p + r) - np.float32(0.5)).astype(np.int64)
    assert np.all(lo >= c - H)
    assert np.all(hi <= c + cell - 1 + H)
    assert np.all(hi - lo + 1 <= 2 * H + 1)


def test_halo_plants_fail():
    # one voxel less of halo than the rule gives lets boxes out of the tile
    S, cell, r = 4096, 16, np.float32(4.0)
    p = np.float32(2000.0)
    c = 2000 - 2000 % cell
    lo = int(np.ceil((p - r) - np.float32(0.5)))
    assert lo < c - (_halo(S, 4.0) - 1)
