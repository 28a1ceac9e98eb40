"""The oracle's collimated gantry field (an extension: a point source at
distance sad upstream of the phantom centre, X and Y jaw pairs projected on
the isocentre plane, unscattered transmission exp(-n tau / cos theta) through
the n jaw pairs a particle is aimed past), pinned by closed forms and by what
reduces it to the parallel field already pinned:

* every primary travels on the line from the point source through its field
  point, the field points are uniform over the field, and it starts on the
  phantom surface where that line enters;
* given the field points, the number of primaries that pass the jaws in each
  region (outside X only, outside Y only, outside both) has the exact
  expectation sum exp(-n tau / cos theta); a planted mistake (no obliquity
  factor) fails it;
* jaws open over the whole field draw nothing and give the parallel field
  bit for bit; a far source gives the parallel field's histories;
* energy conservation with absorbed primaries counted as escaped.
"""
import math

import numpy as np
import pytest

import oracle_lib

N_SIDE, LBOX = 64, 4.0                       # 256 mm water cube
H = 0.5 * N_SIDE * LBOX
EMC2 = 0.510991


def _balance(r):
    # scored + escaped = primaries + rest mass released by annihilations - rest mass made by pairs
    return (r["e_scored"] + r["e_escaped"]) - (r["e_primary"] + 2 * EMC2 * (r["n_annih"] - r["n_pair"]))


def _water():
    return np.zeros((N_SIDE,) * 3, np.uint8), np.ones((N_SIDE,) * 3, np.float32)


def _plane_points(prim, ang, sad, x0, y0):
    """Field points (relative to the field centre) of the started primaries:
    where their ray crosses the plane through the centre perpendicular to the
    beam axis d = (sin a, 0, cos a), in the plane axes (cos a, 0, -sin a), (0, 1, 0)."""
    a = math.radians(ang)
    d = np.array([math.sin(a), 0.0, math.cos(a)])
    ax = np.array([math.cos(a), 0.0, -math.sin(a)])
    pos, r = prim[:, 0:3], prim[:, 3:6]
    t = -(pos @ d) / (r @ d)
    p = pos + t[:, None] * r
    return p @ ax - x0, p[:, 1] - y0


def test_primaries_leave_the_point_source_uniformly_over_the_field(tables0):
    mat, rho = _water()
    n, sad, ang, x0, y0, f = 4000, 300.0, 30.0, 10.0, -5.0, 200.0
    r = oracle_lib.run_voxel(tables0, mat, rho, LBOX, n, ptype=-1, e0=10.0, field=f, angle=ang, x0=x0, y0=y0,
                             sad=sad, seed=3, primaries=True, per_history=True)
    prim = r["prim"]
    a = math.radians(ang)
    d = np.array([math.sin(a), 0.0, math.cos(a)])
    src = -sad * d
    assert np.allclose(np.linalg.norm(prim[:, 3:6], axis=1), 1.0, atol=1e-12)
    # on the line from the source: (pos - s) parallel to r, pointing away from it
    v = prim[:, 0:3] - src
    cr = np.cross(v, prim[:, 3:6])
    assert np.abs(cr).max() < 1e-9 * np.linalg.norm(v, axis=1).max()
    assert np.all(np.einsum("ij,ij->i", v, prim[:, 3:6]) > 0.0)
    # on the phantom surface: one coordinate within 1e-3 mm of a face, all inside
    gap = H - np.abs(prim[:, 0:3])
    assert gap.min() > 0.0 and np.all(gap.min(axis=1) <= 1e-3 + 1e-9)
    # field points uniform over the field: mean, variance, Kolmogorov-Smirnov
    fu, fv = _plane_points(prim, ang, sad, x0, y0)
    for c in (fu, fv):
        assert np.abs(c).max() <= 0.5 * f + 1e-9
        assert abs(c.mean()) < 4.0 * f / math.sqrt(12.0 * n)
        assert abs(c.var() / (f * f / 12.0) - 1.0) < 0.06
        u = np.sort(c / f + 0.5)
        ks = np.max(np.maximum(np.arange(1, n + 1) / n - u, u - np.arange(n) / n))
        assert ks < 1.63 / math.sqrt(n)          # 1 % level
    assert r["n_jaw"] == 0


def _jaw_z(tables0, fault=oracle_lib.FAULT_NONE):
    """z scores of the passes through the jaws per region, against their
    exact expectation given each primary's field point."""
    mat, rho = _water()
    n, sad, f, tau = 60000, 300.0, 200.0, 1.0
    jaws = (-50.0, 40.0, -30.0, 35.0)
    with oracle_lib.fault(fault):
        r = oracle_lib.run_voxel(tables0, mat, rho, LBOX, n, ptype=-1, e0=8.0, field=f, angle=0.0, sad=sad,
                                 jaws=jaws, jaw_tau=tau, seed=11, primaries=True, per_history=True)
    prim = r["prim"]
    # every ray enters the phantom, and an 8 MeV electron that enters water
    # deposits energy: the absorbed primaries are those that score nothing;
    # they keep their field point (x0 + du, y0 + dv) as position
    started = r["hist_edep"] > 0.0
    du, dv = np.empty(n), np.empty(n)
    du[~started], dv[~started] = prim[~started, 0], prim[~started, 1]
    fu, fv = _plane_points(prim[started], 0.0, sad, 0.0, 0.0)
    du[started], dv[started] = fu, fv
    assert int((~started).sum()) == r["n_jaw"]
    nx = (du < jaws[0]) | (du >= jaws[1])
    ny = (dv < jaws[2]) | (dv >= jaws[3])
    nj = nx.astype(int) + ny.astype(int)
    cost = sad / np.sqrt(sad * sad + du * du + dv * dv)
    tr = np.exp(-nj * tau / cost)
    assert np.all(started[nj == 0])                 # the aperture is open
    z = []
    for sel in (nx & ~ny, ny & ~nx, nx & ny):
        exp_pass = tr[sel].sum()
        var = (tr[sel] * (1.0 - tr[sel])).sum()
        z.append((started[sel].sum() - exp_pass) / math.sqrt(var))
    return np.array(z), r


def test_jaw_transmission_has_its_exact_expectation(tables0):
    z, r = _jaw_z(tables0)
    assert np.sum(z * z) < 20.0, z          # chi-square, 3 degrees of freedom (p ~ 2e-4)
    assert abs(_balance(r)) < 1e-9 * r["e_primary"]


def test_jaw_check_catches_a_missing_obliquity(tables0):
    z, _ = _jaw_z(tables0, oracle_lib.FAULT_JAW_OBLIQUITY)
    assert np.sum(z * z) > 40.0, z


def test_open_jaws_and_far_source_reduce_to_the_parallel_field(tables0):
    mat, rho = _water()
    n = 1500
    kw = dict(ptype=0, e0=6.0, field=120.0, angle=40.0, x0=5.0, seed=5, per_history=True)
    par = oracle_lib.run_voxel(tables0, mat, rho, LBOX, n, **kw)
    openj = oracle_lib.run_voxel(tables0, mat, rho, LBOX, n, jaws=(-61.0, 61.0, -61.0, 61.0), jaw_tau=3.0, **kw)
    np.testing.assert_array_equal(openj["hist_edep"], par["hist_edep"])
    np.testing.assert_array_equal(openj["edep"], par["edep"])
    far = oracle_lib.run_voxel(tables0, mat, rho, LBOX, n, sad=1e9, **kw)
    same = np.isclose(far["hist_edep"], par["hist_edep"], rtol=1e-6, atol=1e-9)
    assert same.mean() > 0.95, same.mean()


@pytest.mark.parametrize("ptype", [0, -1])
def test_collimated_fields_conserve_energy(tables3, ptype):
    mat, rho = _water()
    mat[20:40, 10:50, 25:45] = 1             # a titanium block
    rho[20:40, 10:50, 25:45] = 4.54
    r = oracle_lib.run_voxel(tables3, mat, rho, LBOX, 3000, ptype=ptype, e0=12.0, field=180.0, angle=250.0,
                             sad=400.0, jaws=(-40.0, 60.0, -70.0, 20.0), jaw_tau=0.7, seed=9)
    assert r["n_jaw"] > 0
    assert abs(_balance(r)) < 1e-9 * r["e_primary"]
