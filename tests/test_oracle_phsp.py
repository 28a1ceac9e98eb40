"""The oracle's phase-space source (an extension: one history per record,
with its own particle, energy, position and direction), pinned by what
reduces it to sources already pinned and by closed forms:

* records that restate the reference's pencil beam (SimDPMLike.cc:64-75: the
  entry point on the source plane, direction +z, no source draws) give the
  pencil run history by history, bit for bit -- in the reference slab and in a
  voxel phantom, for electrons and photons;
* a record outside a voxel phantom starts where its ray enters the phantom,
  1e-3 mm inside (ray-box intersection by hand); a ray that misses scores
  nothing and its energy escapes;
* energy conservation for mixed records (e-, e+, photons, random directions).
"""
import numpy as np
import pytest

import oracle_lib

EMC2 = 0.510991


def _records(n, ptype, e, x, y, z):
    r = np.zeros((n, 8))
    r[:, 0], r[:, 1], r[:, 2], r[:, 3], r[:, 4], r[:, 7] = ptype, e, x, y, z, 1.0
    return r


@pytest.mark.parametrize("ptype,e0", [(-1, 15.0), (0, 10.0)])
def test_pencil_records_equal_the_pencil_beam_in_the_slab(tables3, ptype, e0):
    n = 1500
    ref = oracle_lib.run_slab(tables3, n, e0=e0, ptype=ptype, rng=oracle_lib.RNG_PHILOX, seed=21, per_history=True)
    got = oracle_lib.run_slab(tables3, n, e0=e0, ptype=ptype, rng=oracle_lib.RNG_PHILOX, seed=21, per_history=True,
                              phsp=_records(n, ptype, e0, 0.0, 0.0, -0.5))
    np.testing.assert_array_equal(got["hist_edep"], ref["hist_edep"])
    np.testing.assert_array_equal(got["edep"], ref["edep"])


def test_pencil_records_equal_the_pencil_beam_in_a_voxel_phantom(tables0):
    mat, rho = np.zeros((20, 16, 16), np.uint8), np.ones((20, 16, 16), np.float32)
    n = 2000
    for ptype, e0, x0, y0 in ((-1, 9.0, 3.0, -2.0), (0, 6.0, -5.0, 4.0)):
        ref = oracle_lib.run_voxel(tables0, mat, rho, 2.0, n, ptype=ptype, e0=e0, x0=x0, y0=y0, seed=4,
                                   per_history=True)
        got = oracle_lib.run_voxel(tables0, mat, rho, 2.0, n, seed=4, per_history=True,
                                   phsp=_records(n, ptype, e0, x0, y0, -20.0))
        np.testing.assert_array_equal(got["hist_edep"], ref["hist_edep"])
        np.testing.assert_array_equal(got["edep"], ref["edep"])


def test_records_outside_move_to_the_phantom_surface(tables0):
    # 16 x 16 x 20 voxels of 2 mm: the box is [-16, 16] x [-16, 16] x [-20, 20] mm
    mat, rho = np.zeros((20, 16, 16), np.uint8), np.ones((20, 16, 16), np.float32)
    s = 1 / np.sqrt(2.0)
    rows = np.array([
        [0, 5.0, 3.0, 4.0, -60.0, 0.0, 0.0, 1.0],     # below the -z face, heading up: enters at z = -20
        [0, 5.0, -100.0, 1.0, 2.0, 1.0, 0.0, 0.0],    # left of the -x face: enters at x = -16
        [-1, 5.0, -30.0, 0.0, -40.0, s, 0.0, s],      # oblique: enters the -z face at x = -10
        [0, 5.0, 1.0, 2.0, 3.0, 0.0, 1.0, 0.0],       # inside: stays
        [0, 5.0, 50.0, 0.0, 0.0, 1.0, 0.0, 0.0],      # heading away: misses
        [0, 5.0, 0.0, 30.0, -60.0, 0.0, 0.0, 1.0],    # parallel to the box, outside in y: misses
    ])
    r = oracle_lib.run_voxel(tables0, mat, rho, 2.0, len(rows), seed=2, phsp=rows, primaries=True,
                             per_history=True)
    p = r["prim"]
    np.testing.assert_allclose(p[0, :3], [3.0, 4.0, -20.0 + 1e-3], atol=1e-12)
    np.testing.assert_allclose(p[1, :3], [-16.0 + 1e-3, 1.0, 2.0], atol=1e-12)
    np.testing.assert_allclose(p[2, :3], [-10.0 + 1e-3 * s, 0.0, -20.0 + 1e-3 * s], atol=1e-12)
    np.testing.assert_allclose(p[3, :3], [1.0, 2.0, 3.0], atol=0)
    np.testing.assert_allclose(p[:, 3:6], rows[:, 5:8], atol=0)
    assert r["hist_edep"][4] == 0.0 and r["hist_edep"][5] == 0.0
    assert r["e_escaped"] >= 10.0 - 1e-12


def test_mixed_records_conserve_energy(tables3):
    rng = np.random.default_rng(8)
    mat = rng.integers(0, 3, size=(14, 12, 10)).astype(np.uint8)
    rho = np.array([1.0, 4.54, 1.85], dtype=np.float32)[mat]
    n = 3000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rows = np.column_stack([rng.choice([-1, 0, 1], n), rng.uniform(1.0, 15.0, n),
                            rng.uniform(-12.0, 12.0, n), rng.uniform(-14.0, 14.0, n), rng.uniform(-16.0, 16.0, n), d])
    r = oracle_lib.run_voxel(tables3, mat, rho, 2.0, n, seed=6, phsp=rows)
    np.testing.assert_allclose(r["e_primary"], rows[:, 1].sum(), rtol=1e-12)
    # a positron record's kinetic energy is a primary's; its annihilation adds
    # 2 mc^2 like that of a pair's positron
    bal = (r["e_scored"] + r["e_escaped"]) - (r["e_primary"] + 2 * EMC2 * (r["n_annih"] - r["n_pair"]))
    assert abs(bal) < 1e-9 * r["e_primary"], bal
    assert r["n_annih"] > 0 and r["e_escaped"] > 0
