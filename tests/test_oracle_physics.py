"""Invariants of the oracle that the mathematics fixes (no reference needed).

* energy conservation per run: scored + escaped = sum of primary energies
  + 2 mc^2 (annihilations - pairs), for electrons and photons, slab and voxel;
* Philox-mode results do not depend on history order or chunking: a run of
  histories [0, N) equals the concatenation of runs [0, k) and [k, N);
* a planted mistake in the oracle (dropping the Compton electron) must break
  energy conservation -- the balance check is sensitive;
* the voxel phantom (an extension of the reference's Geom) reduces to the
  reference slab geometry when its voxels and materials replicate the slab's
  layers: the same Philox histories give the same depth dose.
"""
import numpy as np
import pytest

import oracle_lib

EMC2 = 0.510991


def balance(r, total_in):
    return (r["e_scored"] + r["e_escaped"]) - (total_in + 2 * EMC2 * (r["n_annih"] - r["n_pair"]))


@pytest.mark.parametrize("ptype,e0", [(-1, 15.0), (0, 12.0), (-1, 3.0), (0, 0.8)])
def test_slab_energy_conservation(tables3, ptype, e0):
    n = 3000
    r = oracle_lib.run_slab(tables3, n, e0=e0, ptype=ptype, rng=oracle_lib.RNG_PHILOX, seed=17)
    assert abs(balance(r, n * e0)) < 1e-9 * n * e0


def test_voxel_energy_conservation_and_tally(tables3):
    rng = np.random.default_rng(0)
    mat = rng.integers(0, 3, size=(12, 10, 8)).astype(np.uint8)
    rho = np.array([1.0, 4.54, 1.85], dtype=np.float32)[mat] * rng.uniform(0.9, 1.1, mat.shape).astype(np.float32)
    scale = float(np.max(rho / np.array([1.0, 4.54, 1.85])[mat]))
    n = 1500
    r = oracle_lib.run_voxel(tables3, mat, rho, 2.0, n, ptype=-1, e0=6.0, field=10.0, seed=5, majorant_scale=scale)
    assert abs(balance(r, n * 6.0)) < 1e-9 * n * 6.0
    assert abs(r["edep"].sum() - r["e_scored"]) < 1e-9 * r["e_scored"]


def test_history_order_independence(tables3):
    a = oracle_lib.run_slab(tables3, 400, rng=oracle_lib.RNG_PHILOX, seed=9, per_history=True)
    b1 = oracle_lib.run_slab(tables3, 150, rng=oracle_lib.RNG_PHILOX, seed=9, per_history=True)
    b2 = oracle_lib.run_slab(tables3, 250, rng=oracle_lib.RNG_PHILOX, seed=9, hist_begin=150, per_history=True)
    assert np.array_equal(a["hist_edep"], np.concatenate([b1["hist_edep"], b2["hist_edep"]]))


def test_batch_statistics(tables0):
    mat = np.zeros((20, 20, 20), np.uint8)
    rho = np.ones((20, 20, 20), np.float32)
    r = oracle_lib.run_voxel(tables0, mat, rho, 2.0, 4000, ptype=0, emin=0.5, emax=20.0, field=30.0, seed=1, nbatch=8)
    np.testing.assert_allclose(r["batch_sum"], r["edep"].ravel(), rtol=1e-12, atol=1e-12)
    assert np.all(r["batch_sum2"] * 8 + 1e-12 >= r["batch_sum"] ** 2 * (1 - 1e-12))


def test_planted_mistake_breaks_conservation(tables3):
    n = 2000
    oracle_lib.set_fault(1)
    try:
        r = oracle_lib.run_slab(tables3, n, e0=5.0, ptype=0, rng=oracle_lib.RNG_PHILOX, seed=17)
    finally:
        oracle_lib.set_fault(0)
    assert abs(balance(r, n * 5.0)) > 1e-3 * n * 5.0


def test_voxel_geometry_reproduces_the_reference_slab(tables3):
    # special case that reduces the voxel phantom (an extension) to the
    # reference geometry: 1 mm voxels whose faces coincide with the slab's box
    # faces, materials and nominal densities by depth index as Geom.cc:133-164.
    # Same Philox histories; the two coordinate systems differ by a shift, so
    # rounding can send a history another way: the depth doses agree closely
    # but not bit for bit.
    n, lbox, nxy, nz = 6000, 1.0, 201, 101
    slab = oracle_lib.run_slab(tables3, n, e0=15.0, ptype=-1, lbox=lbox, geom_index=3,
                               rng=oracle_lib.RNG_PHILOX, seed=29, per_history=True)
    mats = np.array([max(oracle_lib.slab_layer_material(iz, lbox, 3), 0) for iz in range(nz)], dtype=np.uint8)
    mat = np.broadcast_to(mats[:, None, None], (nz, nxy, nxy)).copy()
    rho = np.array([1.0, 4.54, 1.85], dtype=np.float32)[mat]
    vox = oracle_lib.run_voxel(tables3, mat, rho, lbox, n, ptype=-1, e0=15.0, seed=29, per_history=True)
    same = np.abs(vox["hist_edep"] - slab["hist_edep"]) <= 1e-9 * np.abs(slab["hist_edep"])
    assert np.mean(same) >= 0.9, np.mean(same)
    depth_vox = vox["edep"].reshape(nz, nxy, nxy).sum(axis=(1, 2))
    depth_slab = slab["edep"][:nz]
    m = depth_slab > 0.2 * depth_slab.max()
    np.testing.assert_allclose(depth_vox[m], depth_slab[m], rtol=0.01)
    np.testing.assert_allclose(vox["e_scored"], slab["e_scored"], rtol=1e-3)
    # and a planted mistake in the phantom (Ti layer given water's density) is seen
    rho_bad = rho.copy()
    rho_bad[mat == 1] = 1.0
    bad = oracle_lib.run_voxel(tables3, mat, rho_bad, lbox, n, ptype=-1, e0=15.0, seed=29)
    depth_bad = bad["edep"].reshape(nz, nxy, nxy).sum(axis=(1, 2))
    assert np.max(np.abs(depth_bad[m] - depth_slab[m]) / depth_slab[m]) > 0.1


def water(n):
    return np.zeros((n, n, n), np.uint8), np.ones((n, n, n), np.float32)


def test_gantry_source_conserves_energy_and_samples_the_spectrum(tables0):
    # a 30-degree field of photons with a triangular (0.1, 2, 18) MeV spectrum
    mat, rho = water(24)
    n = 3000
    r = oracle_lib.run_voxel(tables0, mat, rho, 2.0, n, ptype=0, emin=0.1, emax=18.0, emode=2.0, field=20.0,
                             angle=30.0, seed=7)
    assert abs(balance(r, r["e_primary"])) < 1e-9 * r["e_primary"]
    mean = r["e_primary"] / n      # triangular mean (a + b + c) / 3, sd 3.8 MeV
    assert abs(mean - (0.1 + 2.0 + 18.0) / 3) < 4 * 3.8 / np.sqrt(n)


def test_gantry_at_zero_degrees_is_the_plain_field(tables0):
    # same histories; the gantry start point lies 1e-3 mm inside the face
    mat, rho = water(20)
    kw = dict(ptype=0, emin=0.5, emax=20.0, field=24.0, seed=3, per_history=True)
    a = oracle_lib.run_voxel(tables0, mat, rho, 2.0, 2000, **kw)
    b = oracle_lib.run_voxel(tables0, mat, rho, 2.0, 2000, angle=0.0, **kw)
    same = np.abs(a["hist_edep"] - b["hist_edep"]) <= 1e-6 * np.maximum(np.abs(a["hist_edep"]), 1e-12)
    assert np.mean(same) > 0.95
    np.testing.assert_allclose(b["e_scored"], a["e_scored"], rtol=0.01)


@pytest.mark.parametrize("angle,axis", [(90.0, 2), (180.0, 0)])
def test_gantry_rotation_rotates_the_dose(tables0, angle, axis):
    # water cube: the depth dose along the beam is the same for every angle;
    # 90 deg runs along +x (array axis 2), 180 deg along -z (axis 0, reversed)
    mat, rho = water(30)
    kw = dict(ptype=0, e0=6.0, field=20.0, seed=11)
    n = 60000
    ref = oracle_lib.run_voxel(tables0, mat, rho, 2.0, n, **kw)["edep"].sum(axis=(1, 2))
    d = oracle_lib.run_voxel(tables0, mat, rho, 2.0, n, angle=angle, **kw)["edep"]
    prof = d.sum(axis=tuple(i for i in range(3) if i != axis))
    if angle == 180.0:
        prof = prof[::-1]
    # 6 mm depth bins (statistical spread ~1.5 %)
    ref3, prof3 = ref[:30].reshape(10, 3).sum(axis=1), prof[:30].reshape(10, 3).sum(axis=1)
    m = ref3 > 0.3 * ref3.max()
    np.testing.assert_allclose(prof3[m], ref3[m], rtol=0.05)
    # a wrong rotation sense (the 90-degree beam entering from +x) fails
    if angle == 90.0:
        np.testing.assert_array_less(0.2, np.max(np.abs(prof3[::-1][m] / ref3[m] - 1.0)))


def test_gantry_ray_missing_the_phantom_scores_nothing(tables0):
    mat, rho = water(10)
    r = oracle_lib.run_voxel(tables0, mat, rho, 2.0, 200, ptype=0, e0=5.0, x0=50.0, angle=45.0, seed=1)
    assert r["e_scored"] == 0.0 and abs(r["e_escaped"] - 1000.0) < 1e-9
