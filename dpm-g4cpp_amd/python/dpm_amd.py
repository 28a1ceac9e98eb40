"""ctypes binding of the MI355X DPM engine C ABI (include/dpm_amd.h).

Loads ``dpm-g4cpp_amd/lib/libdpm_amd.so``, built in-tree by ``make -C
dpm-g4cpp_amd``; every call goes through the C ABI and raises
``DpmError`` carrying ``dpm_last_error()`` on failure. There is no CPU
fallback: without the library (or without a GPU) the calls fail.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libdpm_amd.so")
CLI_PATH = os.path.join(PKG, "bin", "dpm_Simulate")
VOXEL_CLI_PATH = os.path.join(PKG, "bin", "dpm_VoxelDose")

DPM_FP32, DPM_FP64 = 0, 1
ELECTRON, PHOTON, POSITRON = -1, 0, 1
SRC_PENCIL, SRC_FIELD, SRC_GANTRY = 0, 1, 2
SPEC_FLAT, SPEC_TRIANGULAR, SPEC_TABLE = 0, 1, 2
COUNTER_NAMES = ("voxel_visits", "deposits", "sb", "moller", "kn", "gs", "secondaries", "e_steps", "g_steps",
                 "tally_atomics")

EXPORTED = ("dpm_tables_create", "dpm_tables_load", "dpm_tables_destroy", "dpm_tables_num_materials",
            "dpm_phantom_create_slab", "dpm_phantom_create_voxel", "dpm_phantom_destroy", "dpm_phantom_num_bins",
            "dpm_simulate", "dpm_simulate_beamlets", "dpm_beamlets_fetch", "dpm_sample", "dpm_eval", "dpm_host_alloc", "dpm_host_free", "dpm_autotune",
            "dpm_autotune_beamlets", "dpm_tally_fetch", "dpm_perform", "dpm_simulate_fields", "dpm_last_error",
            "dpm_abi_version", "dpm_probe_atomics", "dpm_phantom_set_majorant", "dpm_phantom_get_majorant",
            "dpm_simulate_noalloc", "dpm_simulate_phsp")


class DpmError(RuntimeError):
    pass


class Source(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("particle", ctypes.c_int), ("energy", ctypes.c_double),
                ("emin", ctypes.c_double), ("emax", ctypes.c_double), ("x0", ctypes.c_double),
                ("y0", ctypes.c_double), ("field_x", ctypes.c_double), ("field_y", ctypes.c_double),
                ("angle_deg", ctypes.c_double), ("spectrum", ctypes.c_int), ("emode", ctypes.c_double),
                ("spec_n", ctypes.c_int), ("spec_edges", ctypes.c_void_p), ("spec_weights", ctypes.c_void_p),
                ("sad", ctypes.c_double), ("jaws", ctypes.c_int), ("jaw_x1", ctypes.c_double),
                ("jaw_x2", ctypes.c_double), ("jaw_y1", ctypes.c_double), ("jaw_y2", ctypes.c_double),
                ("jaw_tau", ctypes.c_double)]


class Field(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("angle_deg", ctypes.c_double),
                ("nhist", ctypes.c_uint64)]


class RunOptions(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("hist_begin", ctypes.c_uint64), ("precision", ctypes.c_int),
                ("nbatch", ctypes.c_int), ("chunk", ctypes.c_int64), ("waves_per_cu", ctypes.c_int),
                ("hist_edep", ctypes.c_void_p), ("refill_min", ctypes.c_int), ("tally_mode", ctypes.c_int),
                ("graph", ctypes.c_int), ("rounds_per_sync", ctypes.c_int), ("table_memory", ctypes.c_int),
                ("pipes", ctypes.c_int), ("tally_budget", ctypes.c_int64)]


class Stats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("e_scored", ctypes.c_double), ("e_escaped", ctypes.c_double),
                ("e_primary", ctypes.c_double), ("n_pair", ctypes.c_uint64), ("n_annih", ctypes.c_uint64), ("counters", ctypes.c_uint64 * 10),
                ("rounds", ctypes.c_int), ("hist_size", ctypes.c_int), ("kernel_ms", ctypes.c_double * 4),
                ("kernel_calls", ctypes.c_int * 4), ("counters_photon", ctypes.c_uint64 * 10),
                ("counters_electron", ctypes.c_uint64 * 10), ("graphs", ctypes.c_int), ("lds_tables", ctypes.c_int),
                ("lane_slots", ctypes.c_uint64 * 2), ("beamlet_groups", ctypes.c_int),
                ("guard_exits", ctypes.c_uint64)]

    def as_dict(self):
        return dict(seconds=self.seconds, e_scored=self.e_scored, e_escaped=self.e_escaped,
                    e_primary=self.e_primary, n_pair=self.n_pair,
                    n_annih=self.n_annih, counters=dict(zip(COUNTER_NAMES, list(self.counters))),
                    counters_photon=dict(zip(COUNTER_NAMES, list(self.counters_photon))),
                    counters_electron=dict(zip(COUNTER_NAMES, list(self.counters_electron))),
                    rounds=self.rounds, hist_size=self.hist_size, kernel_ms=list(self.kernel_ms),
                    kernel_calls=list(self.kernel_calls), graphs=self.graphs, lds_tables=self.lds_tables,
                    lane_slots=list(self.lane_slots), beamlet_groups=self.beamlet_groups,
                    guard_exits=self.guard_exits)


class Schedule(ctypes.Structure):
    _fields_ = [("chunk", ctypes.c_int64), ("waves_per_cu", ctypes.c_int), ("refill_min", ctypes.c_int),
                ("tally_mode", ctypes.c_int), ("candidates", ctypes.c_int), ("histories_per_s", ctypes.c_double),
                ("default_histories_per_s", ctypes.c_double), ("pipes", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PhspParticle(ctypes.Structure):
    _fields_ = [("particle", ctypes.c_int), ("energy", ctypes.c_double), ("x", ctypes.c_double),
                ("y", ctypes.c_double), ("z", ctypes.c_double), ("u", ctypes.c_double), ("v", ctypes.c_double),
                ("w", ctypes.c_double)]


class Beamlet(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("particle", ctypes.c_int),
                ("energy", ctypes.c_double), ("nhist", ctypes.c_uint64)]


_lib = None


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DpmError("libdpm_amd.so not built: run `make -C dpm-g4cpp_amd` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, i32, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_int32, ctypes.c_double
        L.dpm_last_error.restype = ctypes.c_char_p
        L.dpm_tables_load.argtypes = [ctypes.c_char_p, ip, ctypes.POINTER(vp)]
        L.dpm_tables_destroy.argtypes = [vp]
        L.dpm_tables_num_materials.argtypes = [vp]
        L.dpm_phantom_create_slab.argtypes = [vp, dp, ip, ctypes.POINTER(vp)]
        L.dpm_phantom_create_voxel.argtypes = [vp, ip, ip, ip, dp, vp, vp, ctypes.POINTER(vp)]
        L.dpm_phantom_destroy.argtypes = [vp]
        L.dpm_phantom_num_bins.argtypes = [vp]
        L.dpm_phantom_num_bins.restype = ctypes.c_int64
        L.dpm_simulate.argtypes = [vp, vp, ctypes.POINTER(Source), ctypes.c_uint64, ctypes.POINTER(RunOptions), vp,
                                   vp, ctypes.POINTER(Stats)]
        L.dpm_simulate_beamlets.argtypes = [vp, vp, ctypes.POINTER(Beamlet), ip, ctypes.POINTER(RunOptions), dp, vp,
                                            ctypes.POINTER(Stats)]
        L.dpm_beamlets_fetch.argtypes = [vp, vp, vp]
        L.dpm_sample.argtypes = [vp, ip, ip, ip, vp, vp, vp, vp, vp, vp, vp, vp]
        L.dpm_eval.argtypes = [vp, ip, ip, ip, vp, vp, vp]
        L.dpm_host_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(vp)]
        L.dpm_tally_fetch.argtypes = [vp, vp, vp]
        L.dpm_simulate_fields.argtypes = [vp, vp, ctypes.POINTER(Source), ctypes.POINTER(Field), ip,
                                          ctypes.POINTER(RunOptions), vp, vp, ctypes.POINTER(Stats)]
        L.dpm_perform.argtypes = [vp, ip, ip, ip, dp, ip, ctypes.c_uint64, vp, vp]
        L.dpm_autotune.argtypes = [vp, vp, ctypes.POINTER(Source), ctypes.c_uint64, ip, ip, ctypes.POINTER(Schedule)]
        L.dpm_autotune_beamlets.argtypes = [vp, vp, ctypes.POINTER(Beamlet), ip, ctypes.c_uint64, ip,
                                            ctypes.POINTER(Schedule)]
        L.dpm_host_free.argtypes = [vp]
        L.dpm_probe_atomics.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, vp, vp]
        L.dpm_phantom_set_majorant.argtypes = [vp, dp]
        L.dpm_simulate_phsp.argtypes = [vp, vp, vp, ctypes.c_uint64, ctypes.POINTER(RunOptions), vp, vp,
                                        ctypes.POINTER(Stats)]
        L.dpm_simulate_noalloc.argtypes = [vp, vp, ctypes.POINTER(Source), ctypes.c_uint64,
                                           ctypes.POINTER(RunOptions), vp, vp, ctypes.POINTER(Stats)]
        L.dpm_phantom_get_majorant.argtypes = [vp]
        L.dpm_phantom_get_majorant.restype = dp
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise DpmError("dpm error %d: %s" % (rc, lib().dpm_last_error().decode()))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Tables:
    """Device table pack loaded from a reference-format data directory."""

    def __init__(self, datadir, device=0):
        self.h = ctypes.c_void_p()
        _check(lib().dpm_tables_load(datadir.encode(), device, ctypes.byref(self.h)))
        self.nmat = lib().dpm_tables_num_materials(self.h)

    def close(self):
        if self.h:
            lib().dpm_tables_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sample(self, kind, energy, r1, r2, r3=None, imat=None, precision=DPM_FP32):
        """Sampler with explicit 24-bit integer uniforms; kind in sb|mo|kn|gs
        (cos theta)|gs_omu (1 - cos theta)."""
        k = {"sb": 0, "mo": 1, "kn": 2, "gs": 3, "gs_omu": 4}[kind]
        e = np.ascontiguousarray(energy, dtype=np.float64)
        n = e.size
        u1 = np.ascontiguousarray(r1, dtype=np.uint32)
        u2 = np.ascontiguousarray(r2, dtype=np.uint32)
        u3 = None if r3 is None else np.ascontiguousarray(r3, dtype=np.uint32)
        im = None if imat is None else np.ascontiguousarray(imat, dtype=np.int32)
        out = np.zeros(n)
        eb = np.zeros(n, dtype=np.int32)
        ab = np.zeros(n, dtype=np.int32)
        _check(lib().dpm_sample(self.h, k, precision, n, _ptr(e), _ptr(im), _ptr(u1), _ptr(u2), _ptr(u3), _ptr(out),
                                _ptr(eb), _ptr(ab)))
        return out, eb, ab

    def perform(self, kind, n, energy, imat=0, seed=1, precision=DPM_FP32):
        """n single interactions of a primary along +z (dpm_perform); kind in
        brem|moller|msc|compton. Returns (value, cos): photon/secondary energy,
        or 1-cos(theta) for msc, and cos(theta)."""
        k = {"brem": 0, "moller": 1, "msc": 2, "compton": 3}[kind]
        v = np.zeros(n)
        c = np.zeros(n)
        _check(lib().dpm_perform(self.h, k, precision, n, energy, imat, seed, _ptr(v), _ptr(c)))
        return v, c

    def eval(self, which, energy, imat=None, precision=DPM_FP32):
        e = np.ascontiguousarray(energy, dtype=np.float64)
        im = None if imat is None else np.ascontiguousarray(imat, dtype=np.int32)
        out = np.zeros(e.size)
        _check(lib().dpm_eval(self.h, which, precision, e.size, _ptr(e), _ptr(im), _ptr(out)))
        return out


class Phantom:
    def __init__(self, tables, h, shape):
        self.tables = tables
        self.h = h
        self.shape = shape
        self.nbins = lib().dpm_phantom_num_bins(h)

    @classmethod
    def slab(cls, tables, lbox, geom_index):
        h = ctypes.c_void_p()
        _check(lib().dpm_phantom_create_slab(tables.h, lbox, geom_index, ctypes.byref(h)))
        return cls(tables, h, None)

    @classmethod
    def voxel(cls, tables, mat, rho, lbox):
        mat = np.ascontiguousarray(mat, dtype=np.uint8)
        rho = np.ascontiguousarray(rho, dtype=np.float32)
        nz, ny, nx = mat.shape
        h = ctypes.c_void_p()
        _check(lib().dpm_phantom_create_voxel(tables.h, nx, ny, nz, lbox, _ptr(mat), _ptr(rho), ctypes.byref(h)))
        return cls(tables, h, mat.shape)

    @property
    def majorant(self):
        """Woodcock majorant scale of the photon steps (dpm_phantom_get_majorant)."""
        return lib().dpm_phantom_get_majorant(self.h)

    @majorant.setter
    def majorant(self, scale):
        _check(lib().dpm_phantom_set_majorant(self.h, float(scale)))

    def close(self):
        if self.h:
            lib().dpm_phantom_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Pinned:
    """Owner of one dpm_host_alloc block; frees it when the last array view goes."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _check(lib().dpm_host_alloc(nbytes, ctypes.byref(p)))
        self.ptr = p.value
        self.nbytes = nbytes

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.dpm_host_free(self.ptr)
            self.ptr = None


def probe_atomics(tables, ncells, nops):
    """(seconds, buffer sum) of nops FP64 atomic additions spread over ncells
    cells (dpm_probe_atomics); nops / seconds is the atomic throughput."""
    sec = ctypes.c_double(0.0)
    chk = ctypes.c_double(0.0)
    _check(lib().dpm_probe_atomics(tables.h, int(ncells), int(nops), ctypes.byref(sec), ctypes.byref(chk)))
    return sec.value, chk.value


def pinned_empty(n, dtype=np.float64):
    """A 1-D numpy array of n elements in page-locked host memory (fast tally copies)."""
    dtype = np.dtype(dtype)
    owner = _Pinned(max(1, n * dtype.itemsize))
    buf = (ctypes.c_char * owner.nbytes).from_address(owner.ptr)
    buf._dpm_owner = owner
    return np.frombuffer(buf, dtype=dtype, count=n)


def simulate(tables, phantom, nhist, particle=ELECTRON, energy=15.0, emin=0.0, emax=0.0, field=0.0, x0=0.0, y0=0.0,
             seed=1, hist_begin=0, precision=DPM_FP32, nbatch=0, chunk=0, waves_per_cu=0, per_history=False,
             out=None, refill_min=0, tally_mode=0, schedule=None, graph=False, rounds_per_sync=0, fetch=True,
             angle=None, emode=None, table_memory=0, pipes=0, spectrum_table=None, field_y=None, sad=0.0,
             jaws=None, jaw_tau=0.0):
    """One dpm_simulate call; returns (tally, steps_or_sum2, stats dict, per-history array or None).

    out: optional (tally, steps_or_sum2) pair of contiguous float64 arrays of
    phantom.nbins elements to write into (e.g. from pinned_empty), returned
    reshaped to the phantom's shape. fetch=False leaves the result on the
    device (tally and steps are returned as None; see tally_fetch). angle
    (degrees) selects the gantry source (voxel phantoms), emode the triangular
    spectrum (emin, emode, emax); field_y (default: field) the field's y size;
    sad, jaws (x1, x2, y1, y2) and jaw_tau collimate a gantry field."""
    src = _source(particle, energy, emin, emax, field, x0, y0, angle, emode, spectrum_table)
    if field_y is not None:
        src.field_y = field_y
    _collimate(src, sad, jaws, jaw_tau)
    he = np.zeros(nhist) if per_history else None
    if schedule is not None:
        chunk, waves_per_cu, refill_min, tally_mode = (schedule["chunk"], schedule["waves_per_cu"],
                                                      schedule["refill_min"], schedule["tally_mode"])
        pipes = schedule.get("pipes", 0)
    opt = RunOptions(seed=seed, hist_begin=hist_begin, precision=precision, nbatch=nbatch, chunk=chunk,
                     waves_per_cu=waves_per_cu, hist_edep=_ptr(he), refill_min=refill_min, tally_mode=tally_mode,
                     graph=int(bool(graph)), rounds_per_sync=rounds_per_sync, table_memory=table_memory,
                     pipes=pipes)
    if out is not None:
        tally, steps = out
        for a in (tally, steps):
            if a.dtype != np.float64 or a.size != phantom.nbins or not a.flags["C_CONTIGUOUS"]:
                raise ValueError("out arrays must be contiguous float64 with %d elements" % phantom.nbins)
    else:
        tally = np.zeros(phantom.nbins)
        steps = np.zeros(phantom.nbins)
    st = Stats()
    if not fetch:
        _check(lib().dpm_simulate(tables.h, phantom.h, ctypes.byref(src), nhist, ctypes.byref(opt), None, None,
                                  ctypes.byref(st)))
        return None, None, st.as_dict(), he
    _check(lib().dpm_simulate(tables.h, phantom.h, ctypes.byref(src), nhist, ctypes.byref(opt), _ptr(tally),
                              _ptr(steps), ctypes.byref(st)))
    if phantom.shape is not None:
        tally = tally.reshape(phantom.shape)
        steps = steps.reshape(phantom.shape)
    return tally, steps, st.as_dict(), he


def phsp_records(rows):
    """Phase-space records as the C array: rows of (particle, E, x, y, z, u, v, w)."""
    rows = np.asarray(rows, dtype=np.float64).reshape(-1, 8)
    rec = np.zeros(rows.shape[0], dtype=np.dtype([("particle", "<i4"), ("pad", "<i4"), ("energy", "<f8"),
                                                  ("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("u", "<f8"),
                                                  ("v", "<f8"), ("w", "<f8")]))
    assert rec.dtype.itemsize == ctypes.sizeof(PhspParticle)
    rec["particle"] = rows[:, 0].astype(np.int32)
    for j, f in enumerate(("energy", "x", "y", "z", "u", "v", "w")):
        rec[f] = rows[:, 1 + j]
    return rec


def simulate_phsp(tables, phantom, rows, seed=1, hist_begin=0, precision=DPM_FP32, nbatch=0, schedule=None,
                  per_history=False):
    """dpm_simulate_phsp: one history per phase-space record, rows of
    (particle, E, x, y, z, u, v, w). Returns (tally, steps_or_sum2, stats dict,
    per-history array or None); history ids put the photon records first."""
    rec = phsp_records(rows)
    n = rec.shape[0]
    he = np.zeros(n) if per_history else None
    opt = RunOptions(seed=seed, hist_begin=hist_begin, precision=precision, nbatch=nbatch, hist_edep=_ptr(he))
    if schedule is not None:
        opt.chunk, opt.waves_per_cu = schedule["chunk"], schedule["waves_per_cu"]
        opt.refill_min, opt.tally_mode = schedule["refill_min"], schedule["tally_mode"]
        opt.pipes = schedule.get("pipes", 0)
    tally, second = np.zeros(phantom.nbins), np.zeros(phantom.nbins)
    st = Stats()
    _check(lib().dpm_simulate_phsp(tables.h, phantom.h, rec.ctypes.data_as(ctypes.c_void_p), n, ctypes.byref(opt),
                                   _ptr(tally), _ptr(second), ctypes.byref(st)))
    if phantom.shape is not None:
        tally, second = tally.reshape(phantom.shape), second.reshape(phantom.shape)
    return tally, second, st.as_dict(), he


def simulate_noalloc(tables, phantom, nhist, d_tally, d_steps=None, particle=ELECTRON, energy=15.0, emin=0.0,
                     emax=0.0, field=0.0, x0=0.0, y0=0.0, seed=1, hist_begin=0, precision=DPM_FP32, chunk=0,
                     waves_per_cu=0, refill_min=0, tally_mode=0, angle=None, emode=None, spectrum_table=None):
    """dpm_simulate_noalloc: the run without any device allocation, into
    caller-owned device arrays of phantom.nbins doubles given as addresses
    (d_tally; d_steps for the slab). Returns the stats dict."""
    src = _source(particle, energy, emin, emax, field, x0, y0, angle, emode, spectrum_table)
    opt = RunOptions(seed=seed, hist_begin=hist_begin, precision=precision, chunk=chunk, waves_per_cu=waves_per_cu,
                     refill_min=refill_min, tally_mode=tally_mode)
    st = Stats()
    _check(lib().dpm_simulate_noalloc(tables.h, phantom.h, ctypes.byref(src), nhist, ctypes.byref(opt),
                                      ctypes.c_void_p(d_tally), None if d_steps is None else ctypes.c_void_p(d_steps),
                                      ctypes.byref(st)))
    return st.as_dict()


def _source(particle, energy, emin, emax, field, x0, y0, angle, emode, table=None):
    kind = SRC_GANTRY if angle is not None else (SRC_FIELD if field > 0 else SRC_PENCIL)
    src = Source(kind=kind, particle=particle, energy=energy, emin=emin, emax=emax, x0=x0, y0=y0, field_x=field,
                  field_y=field, angle_deg=0.0 if angle is None else angle,
                  spectrum=SPEC_TRIANGULAR if emode is not None else SPEC_FLAT, emode=0.0 if emode is None else emode)
    if table is not None:
        # a spectrum file: histogram bin edges [n + 1] (MeV) and weights [n];
        # the arrays stay referenced by the structure for the call
        edges = np.ascontiguousarray(table[0], dtype=np.float64)
        weights = np.ascontiguousarray(table[1], dtype=np.float64)
        if edges.size != weights.size + 1:
            raise ValueError("spectrum table: need len(edges) == len(weights) + 1")
        src.spectrum, src.spec_n = SPEC_TABLE, weights.size
        src.spec_edges, src.spec_weights = _ptr(edges), _ptr(weights)
        src._keep = (edges, weights)
    return src


def _collimate(src, sad, jaws, jaw_tau):
    src.sad = sad
    if jaws is not None:
        src.jaws = 1
        src.jaw_x1, src.jaw_x2, src.jaw_y1, src.jaw_y2 = (float(v) for v in jaws)
        src.jaw_tau = jaw_tau


def simulate_fields(tables, phantom, fields, particle=PHOTON, energy=10.0, emin=0.0, emax=0.0, emode=None,
                    field=100.0, gantry=True, seed=1, hist_begin=0, precision=DPM_FP32, nbatch=0, schedule=None,
                    out=None, fetch=True, table_memory=0, pipes=0, sad=0.0, jaws=None, jaw_tau=0.0):
    """fields: list of (x0, y0, angle_deg, nhist); one run into one tally
    (dpm_simulate_fields); sad, jaws (x1, x2, y1, y2) and jaw_tau collimate
    every field. Returns (tally, sum2_or_steps, stats dict)."""
    src = _source(particle, energy, emin, emax, field, 0.0, 0.0, 0.0 if gantry else None, emode)
    _collimate(src, sad, jaws, jaw_tau)
    arr = (Field * len(fields))(*[Field(x0=f[0], y0=f[1], angle_deg=f[2], nhist=int(f[3])) for f in fields])
    opt = RunOptions(seed=seed, hist_begin=hist_begin, precision=precision, nbatch=nbatch, table_memory=table_memory,
                     pipes=pipes)
    if schedule is not None:
        opt.chunk, opt.waves_per_cu = schedule["chunk"], schedule["waves_per_cu"]
        opt.refill_min, opt.tally_mode = schedule["refill_min"], schedule["tally_mode"]
        opt.pipes = schedule.get("pipes", 0)
    st = Stats()
    if not fetch:
        _check(lib().dpm_simulate_fields(tables.h, phantom.h, ctypes.byref(src), arr, len(fields), ctypes.byref(opt),
                                         None, None, ctypes.byref(st)))
        return None, None, st.as_dict()
    tally, second = out if out is not None else (np.zeros(phantom.nbins), np.zeros(phantom.nbins))
    _check(lib().dpm_simulate_fields(tables.h, phantom.h, ctypes.byref(src), arr, len(fields), ctypes.byref(opt),
                                     _ptr(tally), _ptr(second), ctypes.byref(st)))
    if phantom.shape is not None:
        tally, second = tally.reshape(phantom.shape), second.reshape(phantom.shape)
    return tally, second, st.as_dict()


def tally_fetch(tables, phantom, out=None):
    """The last run's (tally, steps_or_sum2) from the device (dpm_tally_fetch)."""
    tally, second = out if out is not None else (np.zeros(phantom.nbins), np.zeros(phantom.nbins))
    _check(lib().dpm_tally_fetch(tables.h, _ptr(tally), _ptr(second)))
    if phantom.shape is not None:
        return tally.reshape(phantom.shape), second.reshape(phantom.shape)
    return tally, second


def autotune(tables, phantom, nsample=2000000, particle=ELECTRON, energy=15.0, emin=0.0, emax=0.0, field=0.0,
             x0=0.0, y0=0.0, precision=DPM_FP32, angle=None, emode=None, nbatch=0):
    """Fastest launch schedule for this source on this phantom (dpm_autotune); a dict usable as schedule=.
    nbatch: time the candidates with the batch statistics of a run with that many batches."""
    src = _source(particle, energy, emin, emax, field, x0, y0, angle, emode)
    sch = Schedule()
    _check(lib().dpm_autotune(tables.h, phantom.h, ctypes.byref(src), int(nsample), precision, int(nbatch),
                              ctypes.byref(sch)))
    return sch.as_dict()


def autotune_beamlets(tables, phantom, beamlets, nsample=4000000, precision=DPM_FP32):
    """Fastest launch schedule for a beamlet set (history counts scaled to ~nsample in total)."""
    arr = (Beamlet * len(beamlets))(*[Beamlet(x0=b[0], y0=b[1], particle=b[2], energy=b[3], nhist=b[4])
                                       for b in beamlets])
    sch = Schedule()
    _check(lib().dpm_autotune_beamlets(tables.h, phantom.h, arr, len(beamlets), int(nsample), precision,
                                       ctypes.byref(sch)))
    return sch.as_dict()


def simulate_beamlets(tables, phantom, beamlets, rel_threshold=1e-3, seed=1, precision=DPM_FP32, chunk=0,
                      waves_per_cu=0, refill_min=0, tally_mode=0, schedule=None, hist_begin=0, graph=True,
                      table_memory=0, pipes=0, tally_budget=0):
    """beamlets: list of (x0, y0, particle, energy, nhist). Returns (row_ptr, col, val, stats).
    tally_budget: device bytes for the dense rows of the beamlets run together (0: 16 GiB)."""
    arr = (Beamlet * len(beamlets))(*[Beamlet(x0=b[0], y0=b[1], particle=b[2], energy=b[3], nhist=b[4])
                                       for b in beamlets])
    if schedule is not None:
        chunk, waves_per_cu, refill_min, tally_mode = (schedule["chunk"], schedule["waves_per_cu"],
                                                      schedule["refill_min"], schedule["tally_mode"])
        pipes = schedule.get("pipes", 0)
    opt = RunOptions(seed=seed, hist_begin=hist_begin, precision=precision, chunk=chunk, waves_per_cu=waves_per_cu,
                     refill_min=refill_min, tally_mode=tally_mode, graph=int(bool(graph)),
                     table_memory=table_memory, pipes=pipes, tally_budget=int(tally_budget))
    row_ptr = np.zeros(len(beamlets) + 1, dtype=np.int64)
    st = Stats()
    _check(lib().dpm_simulate_beamlets(tables.h, phantom.h, arr, len(beamlets), ctypes.byref(opt), rel_threshold,
                                       _ptr(row_ptr), ctypes.byref(st)))
    nnz = int(row_ptr[-1])
    col = np.zeros(nnz, dtype=np.int32)
    val = np.zeros(nnz)
    _check(lib().dpm_beamlets_fetch(tables.h, _ptr(col), _ptr(val)))
    return row_ptr, col, val, st.as_dict()


def read_voxel_dose(path):
    """A voxel-dose file of dpm_VoxelDose (host/DoseIO.hh): header, then dose
    [Gy per history] and relative standard deviation, each [nz][ny][nx] float64."""
    with open(path, "rb") as f:
        if not f.readline().startswith(b"# dpm-amd voxel dose v1"):
            raise ValueError("not a dpm-amd voxel dose file: %s" % path)
        nx, ny, nz, lbox, nhist, nbatch = f.readline().split()
        nx, ny, nz = int(nx), int(ny), int(nz)
        n = nx * ny * nz
        data = np.frombuffer(f.read(16 * n), dtype="<f8")
    if data.size != 2 * n:
        raise ValueError("truncated voxel dose file: %s" % path)
    shape = (nz, ny, nx)
    return dict(nx=nx, ny=ny, nz=nz, lbox=float(lbox), nhist=int(nhist), nbatch=int(nbatch),
                dose=data[:n].reshape(shape), rel_sigma=data[n:].reshape(shape))


def read_csr(path):
    """A dose-influence matrix of dpm_VoxelDose -B (host/DoseIO.hh): returns
    (row_ptr int64[nb+1], col int32[nnz], val float64[nnz] in MeV, nvox)."""
    with open(path, "rb") as f:
        if not f.readline().startswith(b"# dpm-amd csr v1"):
            raise ValueError("not a dpm-amd CSR file: %s" % path)
        nb, nvox, nnz = (int(x) for x in f.readline().split())
        row_ptr = np.frombuffer(f.read(8 * (nb + 1)), dtype="<i8")
        col = np.frombuffer(f.read(4 * nnz), dtype="<i4")
        val = np.frombuffer(f.read(8 * nnz), dtype="<f8")
    if row_ptr.size != nb + 1 or col.size != nnz or val.size != nnz:
        raise ValueError("truncated CSR file: %s" % path)
    return row_ptr, col, val, nvox

