"""Global gamma-index comparison of dose distributions (Low et al., Med. Phys.
25 (1998) 656): for each reference point above `cutoff` x max, gamma =
min over evaluated points r' of sqrt((D_e(r') - D_r(r))^2 / (dd * Dmax)^2 +
|r' - r|^2 / dta^2); a point passes when gamma <= 1. The evaluated
distribution is searched on a grid refined by linear interpolation so that a
distance-to-agreement below the voxel size is resolved: offsets of spacing /
sub within dta (1-D: spacing / 8; 3-D: at most dta / 2). The discrete search
can only over-estimate gamma (by at most half an offset step over dta), so a
reported pass rate is a lower bound of the exact one.
"""
import numpy as np
from scipy import ndimage


def _offsets(ndim, spacing_mm, dta_mm, sub):
    step = spacing_mm / sub
    r = int(np.ceil(dta_mm / step))
    rng = np.arange(-r, r + 1) * step
    grids = np.meshgrid(*([rng] * ndim), indexing="ij")
    off = np.stack([g.ravel() for g in grids], axis=1)
    d = np.sqrt((off ** 2).sum(axis=1))
    keep = d <= dta_mm + 1e-9
    return off[keep], d[keep]


def gamma_map(ref, ev, spacing_mm, dd=0.01, dta_mm=1.0, cutoff=0.1, sub=4):
    """Gamma at every reference point above the cutoff (NaN elsewhere) and the
    mask of those points. The evaluated distribution is interpolated
    (trilinear, edge values extended) at r + o for each search offset o, at the
    masked points only."""
    ref = np.asarray(ref, dtype=np.float64)
    ev = np.asarray(ev, dtype=np.float64)
    dmax = ref.max()
    if dmax <= 0:
        raise ValueError("reference distribution is empty")
    off, dist = _offsets(ref.ndim, spacing_mm, dta_mm, sub)
    mask = ref >= cutoff * dmax
    pts = np.array(np.nonzero(mask), dtype=np.float64)
    r = ref[mask]
    best = np.full(r.shape, np.inf)
    for o, d in zip(off, dist):
        if np.any(o):
            val = ndimage.map_coordinates(ev, pts + (o / spacing_mm)[:, None], order=1, mode="nearest")
        else:
            val = ev[mask]
        g2 = ((val - r) / (dd * dmax)) ** 2 + (d / dta_mm) ** 2
        np.minimum(best, g2, out=best)
    g = np.full(ref.shape, np.nan)
    g[mask] = np.sqrt(best)
    return g, mask


def gamma_pass_rate(ref, ev, spacing_mm, dd=0.01, dta_mm=1.0, cutoff=0.1, sub=4):
    g, mask = gamma_map(ref, ev, spacing_mm, dd, dta_mm, cutoff, sub)
    return float(np.mean(g[mask] <= 1.0))


def gamma_pass_rate_1d(ref, ev, spacing_mm, dd=0.01, dta_mm=1.0, cutoff=0.1):
    return gamma_pass_rate(ref, ev, spacing_mm, dd, dta_mm, cutoff, sub=8)


def gamma_pass_rate_3d(ref, ev, spacing_mm, dd=0.01, dta_mm=1.0, cutoff=0.1):
    # search offsets at most dta / 2 apart (2 mm voxels: 0.5 mm), so that a
    # sub-voxel distance to agreement is found; gamma is over-estimated by at
    # most a quarter of the spatial term's unit
    sub = max(2, int(np.ceil(2.0 * spacing_mm / dta_mm)))
    return gamma_pass_rate(ref, ev, spacing_mm, dd, dta_mm, cutoff, sub=sub)
