"""ctypes front end of the FP64 oracle (oracle/build/liboracle.so).

Test infrastructure only: used by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg as the checker, never as the thing measured.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "liboracle.so")
CLI_PATH = os.path.join(ROOT, "oracle", "build", "dpm_oracle_cli")

RNG_MT, RNG_PHILOX = 0, 1
GEOM_SLAB, GEOM_VOXEL = 0, 1
SRC_PENCIL, SRC_FIELD, SRC_GANTRY, SRC_PHSP = 0, 1, 2, 3
SPEC_FLAT, SPEC_TRIANGULAR, SPEC_TABLE = 0, 1, 2

NCNT = 9
COUNTER_NAMES = ("voxel_visits", "deposits", "sb", "moller", "kn", "gs", "secondaries",
                 "e_steps", "g_steps")


class Cfg(ctypes.Structure):
    _fields_ = [("rng_mode", ctypes.c_int), ("arg_rtl", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("geom_mode", ctypes.c_int), ("lbox", ctypes.c_double), ("geom_index", ctypes.c_int),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("vox_mat", ctypes.c_void_p), ("vox_rho", ctypes.c_void_p),
                ("majorant_scale", ctypes.c_double),
                ("src_kind", ctypes.c_int), ("ptype", ctypes.c_int),
                ("e0", ctypes.c_double), ("emin", ctypes.c_double), ("emax", ctypes.c_double),
                ("fx", ctypes.c_double), ("fy", ctypes.c_double), ("x0", ctypes.c_double), ("y0", ctypes.c_double),
                ("hist_begin", ctypes.c_uint64), ("nhist", ctypes.c_uint64), ("nbatch", ctypes.c_int),
                ("angle_deg", ctypes.c_double), ("spectrum", ctypes.c_int), ("emode", ctypes.c_double),
                ("spec_n", ctypes.c_int), ("spec_edges", ctypes.c_void_p), ("spec_weights", ctypes.c_void_p),
                ("phsp", ctypes.c_void_p),
                ("sad", ctypes.c_double), ("jaws", ctypes.c_int),
                ("jaw_x1", ctypes.c_double), ("jaw_x2", ctypes.c_double),
                ("jaw_y1", ctypes.c_double), ("jaw_y2", ctypes.c_double), ("jaw_tau", ctypes.c_double)]


class Out(ctypes.Structure):
    _fields_ = [("edep", ctypes.c_void_p), ("steps", ctypes.c_void_p),
                ("batch_sum", ctypes.c_void_p), ("batch_sum2", ctypes.c_void_p),
                ("hist_edep", ctypes.c_void_p),
                ("hist_cap", ctypes.c_int), ("hist_size", ctypes.c_int), ("max_iz", ctypes.c_int),
                ("e_escaped", ctypes.c_double), ("e_scored", ctypes.c_double),
                ("n_pair", ctypes.c_uint64), ("n_annih", ctypes.c_uint64),
                ("cnt", ctypes.c_uint64 * NCNT), ("e_primary", ctypes.c_double),
                ("prim", ctypes.c_void_p), ("n_jaw", ctypes.c_uint64)]


_lib = None
_loaded_dir = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_load.argtypes = [ctypes.c_char_p]
        L.oracle_error.restype = ctypes.c_char_p
        L.oracle_run.argtypes = [ctypes.POINTER(Cfg), ctypes.POINTER(Out)]
        for name, nr in (("oracle_sample_sb", 5), ("oracle_sample_moller", 4), ("oracle_sample_kn", 4)):
            f = getattr(L, name)
            f.restype = ctypes.c_double
            args = [ctypes.c_double] * nr
            if name == "oracle_sample_sb":
                args = [ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
            f.argtypes = args + [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.oracle_sample_gs.restype = ctypes.c_double
        L.oracle_sample_gs.argtypes = [ctypes.c_double] * 3 + [ctypes.POINTER(ctypes.c_int)] * 2
        L.oracle_sample_gs_omu.restype = ctypes.c_double
        L.oracle_sample_gs_omu.argtypes = [ctypes.c_double] * 3 + [ctypes.POINTER(ctypes.c_int)] * 2
        dptr = ctypes.POINTER(ctypes.c_double)
        L.oracle_sample_many.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, dptr, dptr, dptr,
                                         dptr]
        L.oracle_eval.restype = ctypes.c_double
        L.oracle_eval.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.oracle_philox.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_distance.restype = ctypes.c_double
        L.oracle_distance.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int)]
        L.oracle_table_energy.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_rotate.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def load(datadir):
    global _loaded_dir
    if _loaded_dir == os.path.abspath(datadir):
        return
    if lib().oracle_load(datadir.encode()) != 0:
        raise RuntimeError(lib().oracle_error().decode())
    _loaded_dir = os.path.abspath(datadir)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def run_slab(datadir, nhist, e0=15.0, ptype=-1, lbox=1.0, geom_index=3, rng=RNG_MT, seed=123,
             arg_rtl=1, hist_begin=0, per_history=False, hist_cap=4096, phsp=None):
    """Slab run; phsp: optional [nhist, 8] records (particle, E, x, y, z, u, v, w), one per history."""
    load(datadir)
    cfg = Cfg(rng_mode=rng, arg_rtl=arg_rtl, seed=seed, geom_mode=GEOM_SLAB, lbox=lbox, geom_index=geom_index,
              majorant_scale=1.0, src_kind=SRC_PENCIL, ptype=ptype, e0=e0, emin=0.0, emax=0.0,
              hist_begin=hist_begin, nhist=nhist)
    if phsp is not None:
        phsp = np.ascontiguousarray(phsp, dtype=np.float64)
        assert phsp.shape == (nhist, 8)
        cfg.src_kind, cfg.phsp = SRC_PHSP, _p(phsp)
    edep = np.zeros(hist_cap)
    steps = np.zeros(hist_cap)
    hed = np.zeros(nhist) if per_history else None
    out = Out(edep=_p(edep), steps=_p(steps), hist_edep=_p(hed), hist_cap=hist_cap)
    if lib().oracle_run(ctypes.byref(cfg), ctypes.byref(out)) != 0:
        raise RuntimeError(lib().oracle_error().decode())
    return dict(edep=edep, steps=steps, hist_edep=hed, hist_size=out.hist_size, max_iz=out.max_iz,
                e_escaped=out.e_escaped, e_scored=out.e_scored, n_pair=out.n_pair, n_annih=out.n_annih,
                counters=dict(zip(COUNTER_NAMES, list(out.cnt))))


def run_voxel(datadir, mat, rho, lbox, nhist, ptype=0, e0=10.0, emin=0.0, emax=0.0, field=0.0, x0=0.0, y0=0.0,
              seed=1, hist_begin=0, nbatch=0, per_history=False, majorant_scale=1.0, angle=None, emode=None,
              spectrum_table=None, field_y=None, primaries=False, phsp=None, sad=0.0, jaws=None, jaw_tau=0.0):
    """Voxel run in Philox mode. angle (degrees) selects the gantry source (a
    field, or a pencil when field == 0, whose ray enters the phantom from the
    side the beam direction (sin a, 0, cos a) points away from); emode selects
    the triangular spectrum (emin, emode, emax); field_y (default: field) is
    the field's y size; primaries=True returns each primary as generated
    (x y z u v w E per history) under "prim"; phsp: [nhist, 8] phase-space
    records (particle, E, x, y, z, u, v, w), one per history. sad, jaws
    (x1, x2, y1, y2) and jaw_tau collimate the gantry field."""
    load(datadir)
    mat = np.ascontiguousarray(mat, dtype=np.uint8)
    rho = np.ascontiguousarray(rho, dtype=np.float32)
    nz, ny, nx = mat.shape
    cfg = Cfg(rng_mode=RNG_PHILOX, arg_rtl=0, seed=seed, geom_mode=GEOM_VOXEL, lbox=lbox,
              nx=nx, ny=ny, nz=nz, vox_mat=_p(mat), vox_rho=_p(rho), majorant_scale=majorant_scale,
              src_kind=SRC_GANTRY if angle is not None else (SRC_FIELD if field > 0 else SRC_PENCIL),
              angle_deg=0.0 if angle is None else angle, spectrum=SPEC_TRIANGULAR if emode is not None else SPEC_FLAT,
              emode=0.0 if emode is None else emode, ptype=ptype, e0=e0, emin=emin, emax=emax,
              fx=field, fy=field if field_y is None else field_y, x0=x0, y0=y0, hist_begin=hist_begin,
              nhist=nhist, nbatch=nbatch, sad=sad, jaw_tau=jaw_tau)
    if jaws is not None:
        cfg.jaws = 1
        cfg.jaw_x1, cfg.jaw_x2, cfg.jaw_y1, cfg.jaw_y2 = (float(v) for v in jaws)
    if phsp is not None:
        phsp = np.ascontiguousarray(phsp, dtype=np.float64)
        assert phsp.shape == (nhist, 8)
        cfg.src_kind, cfg.phsp = SRC_PHSP, _p(phsp)
    if spectrum_table is not None:
        edges = np.ascontiguousarray(spectrum_table[0], dtype=np.float64)
        weights = np.ascontiguousarray(spectrum_table[1], dtype=np.float64)
        cfg.spectrum, cfg.spec_n, cfg.spec_edges, cfg.spec_weights = SPEC_TABLE, weights.size, _p(edges), _p(weights)
    edep = np.zeros(mat.size)
    bs = np.zeros(mat.size) if nbatch > 0 else None
    bs2 = np.zeros(mat.size) if nbatch > 0 else None
    hed = np.zeros(nhist) if per_history else None
    prim = np.zeros((nhist, 7)) if primaries else None
    out = Out(edep=_p(edep), batch_sum=_p(bs), batch_sum2=_p(bs2), hist_edep=_p(hed), prim=_p(prim))
    if lib().oracle_run(ctypes.byref(cfg), ctypes.byref(out)) != 0:
        raise RuntimeError(lib().oracle_error().decode())
    return dict(edep=edep.reshape(mat.shape), batch_sum=bs, batch_sum2=bs2, hist_edep=hed, prim=prim,
                e_escaped=out.e_escaped, e_scored=out.e_scored, n_pair=out.n_pair, n_annih=out.n_annih,
                counters=dict(zip(COUNTER_NAMES, list(out.cnt))), e_primary=out.e_primary, n_jaw=out.n_jaw)


def sample_many(kind, e, r, imat=0):
    """FP64 reference samples at energy e for uniforms r of shape (n, 3);
    kind in sb|mo|kn|gs (gs returns cos(theta))."""
    k = {"sb": 0, "mo": 1, "kn": 2, "gs": 3}[kind]
    r = np.ascontiguousarray(r, dtype=np.float64)
    cols = [np.ascontiguousarray(r[:, j]) for j in range(3)]
    out = np.zeros(r.shape[0])
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    if lib().oracle_sample_many(k, r.shape[0], e, imat, dp(cols[0]), dp(cols[1]), dp(cols[2]), dp(out)) != 0:
        raise RuntimeError(lib().oracle_error().decode())
    return out


def sample(kind, e, r, imat=0):
    """Reference-order sampler call with explicit uniforms; returns (value, ebin, abin).
    kind in sb|mo|kn|gs (cos theta)|gs_omu (1 - cos theta)."""
    eb, ab = ctypes.c_int(-1), ctypes.c_int(-1)
    L = lib()
    if kind == "sb":
        v = L.oracle_sample_sb(e, imat, r[0], r[1], r[2], ctypes.byref(eb), ctypes.byref(ab))
    elif kind == "mo":
        v = L.oracle_sample_moller(e, r[0], r[1], r[2], ctypes.byref(eb), ctypes.byref(ab))
    elif kind == "kn":
        v = L.oracle_sample_kn(e, r[0], r[1], r[2], ctypes.byref(eb), ctypes.byref(ab))
    elif kind == "gs":
        v = L.oracle_sample_gs(e, r[0], r[1], ctypes.byref(eb), ctypes.byref(ab))
    elif kind == "gs_omu":
        v = L.oracle_sample_gs_omu(e, r[0], r[1], ctypes.byref(eb), ctypes.byref(ab))
    else:
        raise ValueError(kind)
    return v, eb.value, ab.value


def evaluate(which, e, imat=0):
    return lib().oracle_eval(which, e, imat)


FAULT_NONE, FAULT_COMPTON_ELECTRON, FAULT_BATCH_ROW, FAULT_FIELD_WIDTH, FAULT_MAJORANT = 0, 1, 2, 3, 4
FAULT_JAW_OBLIQUITY = 5


def set_fault(f):
    """Plant a known mistake in the oracle: 1 drop the Compton electron, 2 batch
    rows shifted by one history, 3 field source half as wide, 4 Woodcock
    interaction probabilities from the unscaled majorant; 0 = none."""
    lib().oracle_set_fault(int(f))


class fault:
    """Context manager planting oracle mistake `f` for the duration of a block."""

    def __init__(self, f):
        self.f = f

    def __enter__(self):
        set_fault(self.f)

    def __exit__(self, *exc):
        set_fault(0)


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().oracle_philox(c, k, o)
    return list(o)


def distance(geom_mode, lbox, r, v, n=(0, 0, 0)):
    """The oracle's DistanceToBoundary: (distance or -1, box, position after the tolerance push)."""
    rr = (ctypes.c_double * 3)(*r)
    vv = (ctypes.c_double * 3)(*v)
    bb = (ctypes.c_int * 3)()
    d = lib().oracle_distance(int(geom_mode), float(lbox), int(n[0]), int(n[1]), int(n[2]), rr, vv, bb)
    return d, tuple(bb), tuple(rr)


def table_energy(edges, weights, u):
    """The oracle's table-spectrum energies for uniforms u."""
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    weights = np.ascontiguousarray(weights, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    e = np.zeros(u.size)
    lib().oracle_table_energy(weights.size, _p(edges), _p(weights), u.size, _p(u), _p(e))
    return e


def rotate(uvw, axis):
    """The oracle's RotateToLabFrame: the local vector uvw in the frame of `axis`, in the lab frame."""
    a = (ctypes.c_double * 3)(*uvw)
    lib().oracle_rotate(a, float(axis[0]), float(axis[1]), float(axis[2]))
    return tuple(a)


def slab_layer_material(iz, lbox, geom_index):
    """Material index of the reference slab layer iz (Geom.cc:133-164)."""
    inv = 1.0 / lbox
    if iz < 0:
        return -1
    if geom_index == 1:
        return 0 if iz < 10.0 * inv else (1 if iz < 30.0 * inv else 0)
    if geom_index == 2:
        return 0 if iz < 20.0 * inv else (1 if iz < 40.0 * inv else 0)
    if geom_index == 3:
        if iz < 10.0 * inv:
            return 0
        if iz < 15.0 * inv:
            return 1
        if iz < 25.0 * inv:
            return 2
        return 0
    return 0


def format_hist_sim(edep, steps, size, nprim, lbox, densities, geom_index):
    """hist.sim text exactly as Geom::Write (Geom.cc:167-192) prints it."""
    lines = []
    to_cm = 0.1
    norm = 1.0 / (nprim * lbox * to_cm)
    s = 0.0
    for i in range(size):
        s += edep[i]
    lines.append("# === Mean energy deposit in the target: %13.4e  [MeV/event]" % (s / nprim))
    lines.append("# === Energy deposit as a function of the depth: ")
    lines.append("# === Index       Depth [cm]    Edep [MeV cm2/g]  #Simulation-Steps ")
    for i in range(size):
        m = slab_layer_material(i, lbox, geom_index)
        rho = densities[m] if m > -1 else 1.0e-40
        lines.append(" %10d    %13.4e    %13.4e    %13.4e" % (i, (i + 0.5) * lbox * to_cm, edep[i] * norm / rho,
                                                         steps[i] / nprim))
    return "\n".join(lines) + "\n"
