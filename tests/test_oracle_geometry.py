"""The oracle's geometry and rotation against closed forms.

DistanceToBoundary (Geom.cc:20-117, and the voxel phantom of the same
logic): the distance to the next voxel face along a ray equals the nearest
intersection with the box's face planes, computed here in global
coordinates; the box is floor of the position over the voxel size; positions
outside give -1; a position within half the tolerance of the face ahead is
pushed by the tolerance into the next box. RotateToLabFrame
(SimDPMLike.cc:638-656): the result is a unit vector at angle acos(w) to the
frame axis, the identity for the +z axis, and vectors of opposite azimuth are
mirror images about the axis. The GPU kernels follow the oracle history by
history (tests/test_gpu_transport.py), so these pin both.
"""
import numpy as np

import oracle_lib as O


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def test_voxel_distance_is_the_ray_box_intersection():
    rng = np.random.default_rng(5)
    n, L = (7, 5, 9), 2.0
    org = -0.5 * np.array(n) * L
    dirs = _unit(rng, 500)
    for k in range(500):
        box = np.array([rng.integers(0, m) for m in n])
        frac = rng.uniform(0.01, 0.99, 3)   # well away from the faces (no push)
        r = org + (box + frac) * L
        v = dirs[k]
        d, b, r2 = O.distance(O.GEOM_VOXEL, L, r, v, n)
        lo, hi = org + box * L, org + (box + 1) * L
        t = [((hi[a] if v[a] > 0 else lo[a]) - r[a]) / v[a] for a in range(3) if v[a] != 0]
        assert b == tuple(box)
        assert np.allclose(r2, r, rtol=0, atol=0)
        assert abs(d - min(t)) <= 1e-12 * L, (d, min(t))
        # the point reached lies on the box surface
        p = r + d * v
        assert np.all(p >= lo - 1e-9) and np.all(p <= hi + 1e-9)
        assert np.min(np.minimum(np.abs(p - lo), np.abs(p - hi))) < 1e-9


def test_voxel_outside_and_tolerance_push():
    n, L = (4, 4, 4), 2.0
    org = -0.5 * np.array(n) * L
    for r in ([org[0] - 0.1, 0, 0], [0, 0, -org[2] + 0.1], [0, org[1] - 1e-9, 0]):
        assert O.distance(O.GEOM_VOXEL, L, r, [0, 0, 1], n)[0] == -1.0
    # 3e-5 mm below the face x = 0 between boxes 1 and 2, moving +x: pushed
    # by 1e-4 mm into box 2, whose far face is then L - 7e-5 away
    r = [-3e-5, 0.3, 0.3]
    d, b, r2 = O.distance(O.GEOM_VOXEL, L, r, [1, 0, 0], n)
    assert b[0] == 2 and abs(r2[0] - 7e-5) < 1e-15 and abs(d - (L - 7e-5)) < 1e-12
    # the same position moving -x is not pushed
    d, b, r2 = O.distance(O.GEOM_VOXEL, L, r, [-1, 0, 0], n)
    assert b[0] == 1 and r2[0] == r[0] and abs(d - (L - 3e-5)) < 1e-12
    # 6e-5 mm below the face (beyond half the tolerance): no push
    d, b, r2 = O.distance(O.GEOM_VOXEL, L, [-6e-5, 0.3, 0.3], [1, 0, 0], n)
    assert b[0] == 1 and abs(d - 6e-5) < 1e-15


def test_slab_boxes_extent_and_distance():
    # the reference slab: boxes of size L centred on i*L, |x|, |y| <= 100 mm,
    # -L <= z <= 100 mm
    rng = np.random.default_rng(9)
    L = 1.0
    for k, v in enumerate(_unit(rng, 300)):
        c = rng.integers(-80, 80, 3).astype(float)
        c[2] = abs(c[2])
        r = c * L + rng.uniform(-0.49, 0.49, 3) * L
        d, b, _ = O.distance(O.GEOM_SLAB, L, r, v)
        assert b == tuple(int(x) for x in c)
        lo, hi = (c - 0.5) * L, (c + 0.5) * L
        t = min(((hi[a] if v[a] > 0 else lo[a]) - r[a]) / v[a] for a in range(3) if v[a] != 0)
        assert abs(d - t) <= 1e-12
    for r in ([100.5, 0, 5], [0, -100.5, 5], [0, 0, 100.5], [0, 0, -1.5]):
        assert O.distance(O.GEOM_SLAB, L, r, [0, 0, 1])[0] == -1.0


def test_rotation_to_lab_frame():
    rng = np.random.default_rng(13)
    axes = _unit(rng, 400)
    for a in axes:
        cost = rng.uniform(-1, 1)
        phi = rng.uniform(0, 2 * np.pi)
        sint = np.sqrt(1 - cost * cost)
        loc = (sint * np.cos(phi), sint * np.sin(phi), cost)
        out = np.array(O.rotate(loc, a))
        assert abs(np.linalg.norm(out) - 1) < 1e-12
        assert abs(out @ a - cost) < 1e-12            # polar angle about the frame axis
        opp = np.array(O.rotate((-loc[0], -loc[1], cost), a))
        np.testing.assert_allclose(out + opp, 2 * cost * a, atol=1e-12)   # opposite azimuths mirror
    loc = (0.6, 0.0, 0.8)
    np.testing.assert_allclose(O.rotate(loc, (0, 0, 1)), loc, atol=0)     # identity frame
    np.testing.assert_allclose(O.rotate(loc, (0, 0, -1)), (-0.6, 0.0, -0.8), atol=0)


def test_table_spectrum_inverse_cdf():
    # histogram spectrum, one bin of weight 0: over the whole 2^-24 lattice of
    # uniforms each bin receives its weight's share of the lattice, energies
    # never fall in the empty bin, rise with u, and the boundaries of the
    # cumulative weights map to the edges of the bins they open
    edges = np.array([0.5, 1.0, 2.0, 4.0, 6.0, 15.0])
    w = np.array([1.0, 3.0, 0.0, 2.5, 0.5])
    u = np.arange(0, 1 << 24, 7) / float(1 << 24)
    e = O.table_energy(edges, w, u)
    assert np.all(np.diff(e) >= 0) and e[0] == 0.5 and e[-1] < 15.0
    counts = np.histogram(e, bins=edges)[0] / u.size
    np.testing.assert_allclose(counts, w / w.sum(), atol=2e-6)
    cdf = np.concatenate([[0], np.cumsum(w) / w.sum()])
    # u = cdf[2] = cdf[3] opens bin 3: the empty bin 2 is skipped
    np.testing.assert_allclose(O.table_energy(edges, w, cdf[[1, 2, 4]]), edges[[1, 3, 4]], rtol=1e-12)
    # uniform within a bin: the mean of the lattice energies is the histogram mean
    mean = np.sum(w / w.sum() * 0.5 * (edges[1:] + edges[:-1]))
    assert abs(e.mean() - mean) < 1e-5
