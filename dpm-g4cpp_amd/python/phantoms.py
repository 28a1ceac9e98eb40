"""Synthetic phantoms and beamlet sets of the benchmark configurations.

Material indices refer to the tables of generator configuration 3
(0 water, 1 Ti, 2 compact bone) or configuration 0 (0 water).

* water_cube(n): homogeneous water, density 1.0 (configuration 2: n = 150 at 2 mm).
* heterogeneous(nx, ny, nz, lbox, seed): water background, 40 compact-bone
  ellipsoids (1.85 g/cm3) and 4 Ti ellipsoids (4.5 g/cm3) with seeded centres
  inside the phantom and semi-axes U(5, 40) mm (scaled with the phantom size
  for larger phantoms), then every voxel density multiplied by
  lognormal(0, sigma=0.05) (configuration 3: 256x256x128 at 2 mm).
* beamlets(n, seed): configuration 4 -- pencil beamlets, type uniform
  {e-, gamma}; e- energies uniform in {6, 9, 12, 15, 18} MeV, photons U(1, 20)
  MeV; entry points drawn without replacement from the 2 mm grid over
  160x160 mm2; history counts lognormal(ln 1e5, sigma=1.0) clipped to
  [1e4, 1e6] (sigma is not given by the configuration; 1.0 is this choice).
* config5_fields(total, n_angles, seed): configuration 5's gantry plan --
  36 angles every 10 degrees around the phantom centre, history counts
  uneven in [1e7, 5e7] (uniform draws rescaled to the total, 1e9), each a
  10x10 cm2 photon field with a triangular (0.1, 2, 18) MeV spectrum
  (simulate_fields(..., emin=0.1, emode=2.0, emax=18.0, field=100.0)).
  The phantom is the heterogeneous recipe at 512x512x256, 1 mm, with the
  ellipsoid semi-axes scaled by 2 ("built like configuration 2 with scaled
  ellipsoids" read as the heterogeneous recipe with scaled ellipsoids).
"""
import numpy as np

WATER, TI, BONE = 0, 1, 2
NOMINAL_RHO = {WATER: 1.0, TI: 4.54, BONE: 1.85}


def water_cube(n):
    return np.zeros((n, n, n), dtype=np.uint8), np.ones((n, n, n), dtype=np.float32)


def heterogeneous(nx, ny, nz, lbox, seed=1234, n_bone=40, n_ti=4, axis_scale=1.0, jitter=0.05):
    rng = np.random.default_rng(seed)
    mat = np.zeros((nz, ny, nx), dtype=np.uint8)
    rho = np.ones((nz, ny, nx), dtype=np.float32)
    ext = np.array([nx, ny, nz]) * lbox
    zc = (np.arange(nz) + 0.5) * lbox - ext[2] / 2
    yc = (np.arange(ny) + 0.5) * lbox - ext[1] / 2
    xc = (np.arange(nx) + 0.5) * lbox - ext[0] / 2
    for k, (m, dens) in enumerate([(BONE, 1.85)] * n_bone + [(TI, 4.5)] * n_ti):
        c = rng.uniform(-ext / 2, ext / 2)
        ax = rng.uniform(5.0, 40.0, size=3) * axis_scale
        # bounding box in voxel indices
        lo = np.maximum(0, np.floor((c - ax + ext / 2) / lbox).astype(int))
        hi = np.minimum([nx, ny, nz], np.ceil((c + ax + ext / 2) / lbox).astype(int))
        if np.any(hi <= lo):
            continue
        z = zc[lo[2]:hi[2], None, None]
        y = yc[None, lo[1]:hi[1], None]
        x = xc[None, None, lo[0]:hi[0]]
        inside = ((x - c[0]) / ax[0]) ** 2 + ((y - c[1]) / ax[1]) ** 2 + ((z - c[2]) / ax[2]) ** 2 <= 1.0
        sub_m = mat[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        sub_r = rho[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        sub_m[inside] = m
        sub_r[inside] = dens
    if jitter > 0:
        rho *= rng.lognormal(0.0, jitter, size=rho.shape).astype(np.float32)
    return mat, rho


def beamlets(n=512, seed=4321, median=1e5, sigma=1.0, grid_mm=2.0, field_mm=160.0):
    rng = np.random.default_rng(seed)
    npts = int(round(field_mm / grid_mm))
    coords = (np.arange(npts) + 0.5) * grid_mm - field_mm / 2
    pick = rng.choice(npts * npts, size=n, replace=False)
    out = []
    for p in pick:
        x, y = coords[p % npts], coords[p // npts]
        if rng.uniform() < 0.5:
            particle, e = -1, float(rng.choice([6.0, 9.0, 12.0, 15.0, 18.0]))
        else:
            particle, e = 0, float(rng.uniform(1.0, 20.0))
        nh = int(np.clip(rng.lognormal(np.log(median), sigma), 1e4, 1e6))
        out.append((float(x), float(y), particle, e, nh))
    return out


def config5_fields(total=1e9, n_angles=36, seed=5678, lo=1e7, hi=5e7):
    """[(x0, y0, angle_deg, nhist)] of the gantry plan; counts in [lo, hi]
    summing to `total` (scaled down together when total < n_angles * lo)."""
    rng = np.random.default_rng(seed)
    w = rng.uniform(lo, hi, n_angles)
    if total >= n_angles * lo:
        for _ in range(100):   # rescale to the total, keeping every count in [lo, hi]
            w = np.clip(w * (total / w.sum()), lo, hi)
            if abs(w.sum() - total) < 1:
                break
    else:
        w = w * (total / w.sum())
    n = np.maximum(1, np.floor(w)).astype(np.int64)
    n[0] += int(total) - int(n.sum())
    return [(0.0, 0.0, 10.0 * i, int(n[i])) for i in range(n_angles)]

