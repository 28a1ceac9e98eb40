#!/usr/bin/env python3
"""Benchmark of the MI355X DPM engine on its headline workload.

Workload (one "step"): 1e7 photon histories in a homogeneous water cube of
150^3 voxels of 2 mm (density 1.0), parallel 10x10 cm^2 field on the -z face,
energies uniform in [0.5, 20] MeV, FP32 transport with an FP64 tally split in
100 history batches (per-voxel sum and sum of squares). Inputs (tables,
phantom) are resident on the GPU before the timed region; `value` times the
whole job with its result (dose, sum of squares) left in HBM, and
`pcie_inclusive` the same step returning both arrays to pinned host memory.
The launch schedule is the autotuner's (chosen before the timed region).

Printed: one JSON line with histories/s, the roofline of the dominant kernel
(algorithmic bytes from the engine's event counters over that kernel's device
time measured with HIP events on its stream; HBM traffic from the committed
rocprof summary), the 1 %/1 mm gamma pass rate of the FP32 dose against an
FP64-transport GPU run of the same histories, the 512-beamlet batch timed
plain vs batched, and a CPU baseline (the FP64 oracle on this host's cores, a
bounded sample).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dpm-g4cpp_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "dpm-g4cpp_amd", "tablegen"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per event (FP32 transport): voxel visit = u8 material +
# f32 density; tally update = FP64 read-modify-write (consecutive deposits of
# one particle into one voxel are combined in registers first, so this counts
# the global atomics, not the deposits); alias sample = 3 entries of 16 B; GS
# sample = 2 entries of 16 B + a(E) and E-grid pairs; secondary = one 44 B
# queue record written and read back
BYTES = dict(voxel_visits=5, tally_atomics=16, sb=48, moller=48, kn=48, gs=48, secondaries=88)


def algorithmic_bytes(c):
    return sum(BYTES[k] * c[k] for k in BYTES)


def atomic_roofline(dpm_amd, t, st, footprint):
    """FP64 tally atomics of the electron kernel per second of that kernel,
    against the atomic probe's throughput for the same number of additions
    spread over the tally's footprint (rows x voxels) and over an L2-resident
    buffer (2 MiB): the chip's ceilings for this access pattern."""
    n_at = st["counters_electron"]["tally_atomics"]
    ms_el = st["kernel_ms"][2]
    if n_at <= 0 or ms_el <= 0:
        return None
    achieved = n_at / (ms_el * 1e-3)
    sec_fp, _ = dpm_amd.probe_atomics(t, min(footprint, (1 << 32) - 1), n_at)
    sec_l2, _ = dpm_amd.probe_atomics(t, 1 << 18, n_at)
    return {"unit": "FP64 atomic adds/s", "achieved": achieved, "tally_atomics_per_step": n_at,
            "peak_same_footprint": n_at / sec_fp, "peak_l2_resident": n_at / sec_l2,
            "frac": achieved / (n_at / sec_fp), "footprint_cells": int(footprint)}


def stream_copy_gbs(gib=1.0, reps=3):
    """Device-to-device copy rate (read + write bytes / s) of a gib-GiB buffer:
    the measured HBM ceiling beside the 8 TB/s datasheet figure."""
    import torch
    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    best = 1e30
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    del a, b
    torch.cuda.empty_cache()
    return 2.0 * n * 4 / best / 1e9


# worker processes of the CPU baselines: every core this process may run on,
# at most the 16 host cores that a one-GPU share of the box provides
CPU_SHARE = 16


def host_cpus():
    """The host CPU the baselines run on: cores available to this process
    (affinity), the OS count, physical cores and model from lscpu, and the
    number of worker processes used."""
    info = {"affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict((a.strip(), b.strip()) for a, b in (l.split(":", 1) for l in out.splitlines() if ":" in l))
        info["model"] = kv.get("Model name")
        info["physical_cores"] = int(kv.get("Core(s) per socket", 0)) * int(kv.get("Socket(s)", 0)) or None
        info["threads_per_core"] = int(kv.get("Thread(s) per core", 0)) or None
    except Exception as e:   # lscpu missing: the affinity count still stands
        info["lscpu_error"] = str(e)
    info["workers"] = min(CPU_SHARE, info["affinity_cpus"])
    info["workers_rule"] = "min(%d, affinity CPUs): the host share of one GPU" % CPU_SHARE
    return info


def cpu_baseline(tables0, nx, lbox, field, emin, emax, nhist, seed, gpu_dose=None, gamma=None):
    """The FP64 oracle CLI on this host's cores (<= 16 processes) over exactly
    the histories of one timed step (seed, ids [0, nhist) split into disjoint
    ranges): the CPU rate, and the CPU dose of the same histories for an
    equal-history comparison with the GPU's FP32 dose."""
    import tempfile
    import numpy as np
    import oracle_lib  # the checker's CLI, timed as the CPU baseline
    cli = oracle_lib.CLI_PATH
    if not os.path.exists(cli):
        return None
    host = host_cpus()
    cores = host["workers"]
    args = [cli, "--data", tables0, "--mode", "philox", "--seed", str(seed), "--geom", "voxel", "--nx", str(nx),
            "--lbox", str(lbox), "--ptype", "0", "--emin", str(emin), "--emax", str(emax), "--field", str(field)]
    d = tempfile.mkdtemp(prefix="dpm_cpu_")
    files = [os.path.join(d, "dose%d.bin" % i) for i in range(cores)]
    bounds = [nhist * i // cores for i in range(cores + 1)]
    try:
        t0 = time.perf_counter()
        procs = [subprocess.Popen(args + ["--n", str(bounds[i + 1] - bounds[i]), "--hist-begin", str(bounds[i]),
                                          "--dose", files[i]], stdout=subprocess.PIPE, text=True)
                 for i in range(cores)]
        outs = [p.communicate()[0] for p in procs]
        wall = time.perf_counter() - t0
        if any(p.returncode != 0 for p in procs):
            return None
        per = [json.loads(o.strip().splitlines()[-1]) for o in outs]
        cpu_dose = None
        if gpu_dose is not None:
            cpu_dose = np.zeros(nx ** 3)
            for f in files:
                cpu_dose += np.fromfile(f, dtype=np.float64)
    finally:
        for f in files:
            if os.path.exists(f):
                os.remove(f)
        os.rmdir(d)
    # rate over the slowest process's transport time (each process reports
    # the time of its oracle_run; start-up and table loading are excluded, as
    # in the GPU's value, whose tables are resident before the timed region)
    slowest = max(p["seconds"] for p in per)
    r = {"value": nhist / slowest, "unit": "histories/s", "cores": cores, "kind": "port",
         "sample": "the timed step's %d histories (seed %d) in %d processes with disjoint history ranges (FP64 "
                   "oracle, Philox mode), slowest transport %.2f s; wall with process start-up and table loading "
                   "%.2f s; single-core %.0f hist/s"
                   % (nhist, seed, cores, slowest, wall, sum(p["hist_per_s"] for p in per) / cores),
         "value_wall_inclusive": nhist / wall, "host": host}
    if cpu_dose is not None:
        ref = cpu_dose.reshape(gpu_dose.shape)
        m = ref >= 0.1 * ref.max()
        r["equal_history_dose"] = {
            "gamma_pass_rate_gpu_fp32_vs_cpu": gamma.gamma_pass_rate_3d(ref, gpu_dose, spacing_mm=lbox, dd=0.01,
                                                                         dta_mm=1.0, cutoff=0.1),
            "max_abs_diff_above_cutoff_rel_to_max": float(np.max(np.abs(gpu_dose[m] - ref[m])) / ref.max()),
            "total_energy_rel_diff": float(abs(gpu_dose.sum() - ref.sum()) / ref.sum()),
            "criterion": "1%/1mm global, voxels > 10% of the CPU dose maximum; same histories and random streams"}
    return r


def _oracle_voxel_worker(spec):
    """One CPU-baseline process (bench.py --oracle-worker JSON): the FP64
    oracle on a phantom saved as .npy; prints its transport time and, when
    asked, saves its dose. Touches no GPU."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib  # the checker, timed as the CPU baseline
    a = json.loads(spec)
    mat = np.load(a["mat"], mmap_mode="r")
    rho = np.load(a["rho"], mmap_mode="r")
    oracle_lib.load(a["tables"])   # table loading stays outside the timed region
    t0 = time.perf_counter()
    r = oracle_lib.run_voxel(a["tables"], mat, rho, a["lbox"], **a["kw"])
    sec = time.perf_counter() - t0
    if a.get("dose"):
        np.save(a["dose"], r["edep"])
    print(json.dumps({"seconds": sec}), flush=True)


def cpu_baseline_voxel(tables, mat, rho, lbox, jobs, seconds_target=15.0, cal=4000, keep_dose=True):
    """The FP64 oracle on this host's cores over `jobs`, a list of run_voxel
    keyword sets (sources of the workload with their first history id);
    process i runs the (i // len(jobs))-th slice of n histories of job
    i % len(jobs), each process about seconds_target. The majorant scale is
    the one the GPU phantom uses, max(rho / rho_nominal). Returns the rate,
    the slices run [(job, hist_begin, n)] and, with keep_dose, the summed
    FP64 dose of all processes (for an equal-history comparison)."""
    import tempfile
    import numpy as np
    import phantoms
    host = host_cpus()
    cores = host["workers"]
    nominal = np.array([phantoms.NOMINAL_RHO[k] for k in sorted(phantoms.NOMINAL_RHO)])
    scale = float(np.max(rho / nominal[mat]))
    tmp = "/dev/shm" if os.path.isdir("/dev/shm") else None
    d = tempfile.mkdtemp(prefix="dpm_cpu_", dir=tmp)
    mp_, rp_ = os.path.join(d, "mat.npy"), os.path.join(d, "rho.npy")
    np.save(mp_, mat)
    np.save(rp_, rho)

    def spawn(kw, dose=None):
        spec = json.dumps({"tables": tables, "mat": mp_, "rho": rp_, "lbox": lbox,
                           "kw": dict(kw, majorant_scale=scale), "dose": dose})
        return subprocess.Popen([sys.executable, os.path.abspath(__file__), "--oracle-worker", spec],
                                stdout=subprocess.PIPE, text=True)

    def seconds(p):
        out = p.communicate()[0]
        if p.returncode != 0:
            raise RuntimeError("oracle worker failed")
        return json.loads(out.strip().splitlines()[-1])["seconds"]

    total = None
    try:
        sec = seconds(spawn(dict(jobs[0], nhist=cal)))
        n_each = max(cal, int(cal / sec * seconds_target))
        slices = [(i % len(jobs), jobs[i % len(jobs)].get("hist_begin", 0) + (i // len(jobs)) * n_each, n_each)
                  for i in range(cores)]
        files = [os.path.join(d, "dose%d.npy" % i) if keep_dose else None for i in range(cores)]
        t0 = time.perf_counter()
        procs = [spawn(dict(jobs[j], nhist=n, hist_begin=hb), files[i]) for i, (j, hb, n) in enumerate(slices)]
        secs = [seconds(p) for p in procs]
        wall = time.perf_counter() - t0
        if keep_dose:
            for f in files:
                part = np.load(f)
                total = part if total is None else total + part
                os.remove(f)
    finally:
        for f in os.listdir(d):
            os.remove(os.path.join(d, f))
        os.rmdir(d)
    # rate over the slowest process's transport time (Python start-up and
    # table loading excluded, as for the GPU, whose tables are resident)
    r = {"value": cores * n_each / max(secs), "unit": "histories/s", "cores": cores, "kind": "port",
         "sample": "%d concurrent processes x %d histories (FP64 oracle, Philox mode), slowest %.1f s "
                   "(wall with start-up %.1f s); single-core %.0f hist/s"
                   % (cores, n_each, max(secs), wall, n_each / (sum(secs) / cores)),
         "value_wall_inclusive": cores * n_each / wall, "host": host}
    return r, slices, total


def equal_history_report(gamma, cpu_dose, gpu_dose, rho, spacing):
    """GPU FP32 dose against the CPU FP64 dose of the same histories: 1 %/1 mm
    gamma over voxels above 10 % of the CPU maximum (dose = energy / density),
    largest difference, total energy."""
    import numpy as np
    ref, ev = cpu_dose / rho, gpu_dose / rho
    m = ref >= 0.1 * ref.max()
    return {"gamma_pass_rate_gpu_fp32_vs_cpu": gamma.gamma_pass_rate_3d(ref, ev, spacing_mm=spacing, dd=0.01,
                                                                         dta_mm=1.0, cutoff=0.1),
            "voxels_compared": int(m.sum()),
            "max_abs_diff_above_cutoff_rel_to_max": float(np.max(np.abs(ev[m] - ref[m])) / ref.max()),
            "total_energy_rel_diff": float(abs(gpu_dose.sum() - cpu_dose.sum()) / cpu_dose.sum()),
            "criterion": "1%/1mm global, voxels > 10% of the CPU dose maximum, full phantom; same histories and "
                         "random streams"}


def _read_hist_sim(path):
    """Mean deposit and the (depth, dose, steps) rows of a hist.sim file (Geom.cc:167-192 layout)."""
    import numpy as np
    mean, rows = None, []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                if "Mean energy deposit" in line:
                    mean = float(line.split(":")[1].split()[0])
                continue
            if line.strip():
                rows.append([float(v) for v in line.split()[1:4]])
    return mean, np.array(rows)


def config1_leg(dpm_amd, gamma, tables3, cpu=True, n=100_000):
    """Configuration 1, the reference's own case: slab of configuration index 3
    (Water/Ti/Bone/Water) at 1 mm, 15 MeV e- pencil beam, 1e5 histories. The
    GPU runs the plain drop-in path (FP32 transport, FP64 depth tally); the CPU
    baseline is the reference itself, compiled from its sources into
    oracle/_ref/ (single-threaded, mt19937 seed 123 fixed in its code), run on
    the same table files. Parity: 1 %/1 mm gamma of the two hist.sim depth-dose
    curves (independent random streams, equal history count)."""
    import tempfile
    import numpy as np
    t = dpm_amd.Tables(tables3)
    ph = dpm_amd.Phantom.slab(t, 1.0, 3)
    kw = dict(particle=dpm_amd.ELECTRON, energy=15.0, seed=1)
    dpm_amd.simulate(t, ph, 2000, **kw)   # warm-up
    walls = []
    for _ in range(5):
        t0 = time.perf_counter()
        dpm_amd.simulate(t, ph, n, fetch=False, **kw)
        walls.append(time.perf_counter() - t0)
    ph.close()
    t.close()
    wall = float(np.median(walls))
    r = {"workload": "config1: slab of configuration index 3 (Water/Ti/Bone/Water), 1 mm, 15 MeV e- pencil beam, "
                     "%d histories, FP32 transport, FP64 depth tally" % n,
         "histories_per_s": n / wall, "seconds": wall}
    d = tempfile.mkdtemp(prefix="dpm_c1_")
    gpu_out, ref_out = os.path.join(d, "gpu.sim"), os.path.join(d, "ref.sim")
    args = ["-p", "e-", "-e", "15", "-n", str(n), "-d", tables3, "-b", "1.0", "-c", "3"]
    try:
        # the drop-in command line (same flags as the reference's), end to end
        t0 = time.perf_counter()
        subprocess.check_call([os.path.join(ROOT, "dpm-g4cpp_amd", "bin", "dpm_Simulate")] + args +
                              ["-o", gpu_out, "-s", "1"], stdout=subprocess.DEVNULL)
        r["cli_seconds_end_to_end"] = time.perf_counter() - t0
        ref_bin = os.path.join(ROOT, "oracle", "_ref", "dpm_Simulate_ref")
        if not cpu:
            return r
        if not os.path.exists(ref_bin):
            r["cpu_baseline"] = {"error": "oracle/_ref/dpm_Simulate_ref not built (needs the reference sources)"}
            return r
        t0 = time.perf_counter()
        subprocess.check_call([ref_bin] + args + ["-o", ref_out], stdout=subprocess.DEVNULL)
        single = time.perf_counter() - t0
        # throughput of the host: one copy per core, side by side (the
        # reference has no seed option, so every copy runs the same histories)
        cores = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        procs = [subprocess.Popen([ref_bin] + args + ["-o", os.path.join(d, "ref%d.sim" % i)],
                                  stdout=subprocess.DEVNULL) for i in range(cores)]
        rcs = [p.wait() for p in procs]
        multi = time.perf_counter() - t0
        if any(rcs):
            raise RuntimeError("reference copies failed: %s" % rcs)
        # the port (the oracle used as the CPU baseline of configurations 2, 3
        # and 5, where the reference cannot run) on this same workload in its
        # mt19937 mode, which reproduces the reference's histories: both
        # single-threaded, wall time of the process incl. table loading
        port_s = None
        port_bin = os.path.join(ROOT, "oracle", "build", "dpm_oracle_cli")
        if os.path.exists(port_bin):
            t0 = time.perf_counter()
            subprocess.check_call([port_bin, "--data", tables3, "--mode", "mt", "--seed", "123", "--geom", "slab",
                                   "--lbox", "1.0", "--gidx", "3", "--ptype", "-1", "--e0", "15", "--n", str(n)],
                                  stdout=subprocess.DEVNULL)
            port_s = time.perf_counter() - t0
        mg, hg = _read_hist_sim(gpu_out)
        mr, hr = _read_hist_sim(ref_out)
        m = min(len(hg), len(hr))
        pr = gamma.gamma_pass_rate_1d(hr[:m, 1], hg[:m, 1], spacing_mm=1.0, dd=0.01, dta_mm=1.0, cutoff=0.1)
        r["cpu_baseline"] = {"value": cores * n / multi, "unit": "histories/s", "cores": cores, "kind": "reference",
                             "sample": "%d concurrent copies of the reference dpm_Simulate (compiled from its "
                                       "sources, -O3), each the full %d histories, wall %.1f s incl. table loading; "
                                       "single copy %.2f s = %.0f hist/s" % (cores, n, multi, single, n / single)}
        if port_s is not None:
            r["cpu_baseline"]["port_single_core_histories_per_s"] = n / port_s
            r["cpu_baseline"]["port_vs_reference_single_core"] = single / port_s
        r["vs_cpu"] = r["histories_per_s"] / r["cpu_baseline"]["value"]
        r["equal_history_depth_dose"] = {
            "gamma_pass_rate_gpu_fp32_vs_reference": pr,
            "criterion": "1%/1mm global, depth bins > 10% of the reference maximum, independent streams",
            "mean_edep_MeV_gpu": mg, "mean_edep_MeV_reference": mr,
            "max_abs_diff_rel_to_max": float(np.max(np.abs(hg[:m, 1] - hr[:m, 1])) / hr[:m, 1].max())}
    finally:
        for f in os.listdir(d):
            os.remove(os.path.join(d, f))
        os.rmdir(d)
    return r


def config3_leg(dpm_amd, gamma, phantoms, tables3, cpu=True, fixed=None):
    """Configuration 3: 256x256x128 heterogeneous phantom at 2 mm, 12 MeV e-
    10x10 cm2 field, 1e8 histories, FP32 transport, FP64 tally, 100 batches."""
    import numpy as np
    t = dpm_amd.Tables(tables3)
    mat, rho = phantoms.heterogeneous(256, 256, 128, 2.0, seed=1234)
    ph = dpm_amd.Phantom.voxel(t, mat, rho, 2.0)
    n, nb = 100_000_000, 100
    kw = dict(particle=dpm_amd.ELECTRON, energy=12.0, field=100.0, seed=3, nbatch=nb)
    if fixed is not None:
        sched = dict(fixed)
    else:
        # sample: 1e7 histories in batches of the real run's size (1e6)
        sched = dpm_amd.autotune(t, ph, nsample=10_000_000, particle=dpm_amd.ELECTRON, energy=12.0, field=100.0,
                                 nbatch=10)
    out = (dpm_amd.pinned_empty(ph.nbins), dpm_amd.pinned_empty(ph.nbins))
    dpm_amd.simulate(t, ph, 1_000_000, schedule=sched, out=out, fetch=False, **kw)   # warm-up
    t0 = time.perf_counter()
    _, _, st, _ = dpm_amd.simulate(t, ph, n, schedule=sched, out=out, fetch=False, **kw)
    wall = time.perf_counter() - t0
    d32, _ = dpm_amd.tally_fetch(t, ph, out=out)
    d32 = d32.copy()
    d64, _, st64, _ = dpm_amd.simulate(t, ph, n, precision=dpm_amd.DPM_FP64, **kw)
    pr = gamma.gamma_pass_rate_3d(d64 / rho, d32 / rho, spacing_mm=2.0, dd=0.01, dta_mm=1.0, cutoff=0.1)
    r = {"workload": "config3: 256x256x128 heterogeneous phantom (40 bone + 4 Ti ellipsoids, density x "
                     "lognormal(0.05)), 2 mm, 12 MeV e- 10x10 cm2 field, 1e8 histories, FP32 transport, FP64 "
                     "tally, 100 batches",
         "histories_per_s": n / wall, "seconds": wall, "schedule": sched,
         "gamma_fp32_vs_fp64": {"pass_rate": pr, "criterion": "1%/1mm global, voxels > 10% max, full phantom",
                                "fp64_histories_per_s": n / st64["seconds"]}}
    if cpu:
        base, slices, cpu_dose = cpu_baseline_voxel(tables3, mat, rho, 2.0, [dict(ptype=-1, e0=12.0, field=100.0,
                                                                                   seed=3, hist_begin=10_000_000)])
        r["cpu_baseline"] = base
        r["vs_cpu"] = r["histories_per_s"] / base["value"]
        # the GPU's FP32 dose of exactly the CPU's histories (contiguous ids)
        first, count = slices[0][1], sum(s[2] for s in slices)
        g32, _, _, _ = dpm_amd.simulate(t, ph, count, particle=dpm_amd.ELECTRON, energy=12.0, field=100.0, seed=3,
                                        hist_begin=first)
        base["equal_history_dose"] = equal_history_report(gamma, cpu_dose, g32, rho, 2.0)
        base["equal_history_dose"]["histories"] = count
    ph.close()
    t.close()
    return r


def config5_leg(dpm_amd, gamma, phantoms, tables3, cpu=True, fixed=None):
    """Configuration 5: 512x512x256 heterogeneous phantom at 1 mm, 36 gantry
    photon fields (10x10 cm2, triangular 0.1/2/18 MeV, uneven counts), 1e9
    histories in one run, FP32 transport, FP64 tally, 100 batches."""
    import numpy as np
    t = dpm_amd.Tables(tables3)
    mat, rho = phantoms.heterogeneous(512, 512, 256, 1.0, seed=2024, axis_scale=2.0)
    ph = dpm_amd.Phantom.voxel(t, mat, rho, 1.0)
    plan = phantoms.config5_fields(1e9)
    n = sum(f[3] for f in plan)
    kw = dict(particle=dpm_amd.PHOTON, emin=0.1, emode=2.0, emax=18.0, field=100.0, seed=5)
    # schedule tuned on one field of the plan (the 45-degree beam, whose rays
    # cross the most voxels): 1.6e7 histories in two batches of about the
    # real run's batch size (1e7), so that, as in the run, every chunk holds
    # one batch and one batch row
    if fixed is not None:
        sched = dict(fixed)
    else:
        sched = dpm_amd.autotune(t, ph, nsample=16_000_000, particle=dpm_amd.PHOTON, emin=0.1, emode=2.0, emax=18.0,
                                 field=100.0, angle=45.0, nbatch=2)
    dpm_amd.simulate_fields(t, ph, phantoms.config5_fields(1e7), fetch=False, schedule=sched, **kw)   # warm-up
    t0 = time.perf_counter()
    _, _, st = dpm_amd.simulate_fields(t, ph, plan, nbatch=100, fetch=False, schedule=sched, **kw)
    wall = time.perf_counter() - t0
    sub = phantoms.config5_fields(1e8)
    a32, _, _ = dpm_amd.simulate_fields(t, ph, sub, **kw)
    a64, _, st64 = dpm_amd.simulate_fields(t, ph, sub, precision=dpm_amd.DPM_FP64, **kw)
    pr = gamma.gamma_pass_rate_3d(a64 / rho, a32 / rho, spacing_mm=1.0, dd=0.01, dta_mm=1.0, cutoff=0.1)
    r = {"workload": "config5: 512x512x256 heterogeneous phantom at 1 mm, 36 gantry photon fields every 10 deg, "
                     "10x10 cm2, triangular (0.1, 2, 18) MeV, uneven counts in [1e7, 5e7], 1e9 histories in one "
                     "run, FP32 transport, FP64 tally, 100 batches",
         "histories_per_s": n / wall, "seconds": wall, "schedule": sched,
         "gamma_fp32_vs_fp64": {"pass_rate": pr, "criterion": "1%/1mm global, voxels > 10% max, 1e8 histories "
                                                              "of the plan, full phantom",
                                "fp64_histories_per_s": sum(f[3] for f in sub) / st64["seconds"]}}
    if cpu:
        offs = np.concatenate([[0], np.cumsum([f[3] for f in plan])])
        jobs = [dict(ptype=0, emin=0.1, emode=2.0, emax=18.0, field=100.0, angle=plan[i][2], seed=5,
                     hist_begin=int(offs[i])) for i in range(len(plan))]
        base, slices, cpu_dose = cpu_baseline_voxel(tables3, mat, rho, 1.0, jobs)
        r["cpu_baseline"] = base
        r["vs_cpu"] = r["histories_per_s"] / base["value"]
        # the GPU's FP32 dose of exactly the CPU's histories: each process ran
        # a slice of one gantry field (its own first history id)
        g32 = np.zeros(ph.nbins)
        for j, hb, cnt in slices:
            d, _, _, _ = dpm_amd.simulate(t, ph, cnt, particle=dpm_amd.PHOTON, emin=0.1, emode=2.0, emax=18.0,
                                          field=100.0, angle=plan[j][2], seed=5, hist_begin=hb)
            g32 += d.ravel()
        base["equal_history_dose"] = equal_history_report(gamma, cpu_dose, g32.reshape(mat.shape), rho, 1.0)
        base["equal_history_dose"]["histories"] = int(sum(s[2] for s in slices))
        base["equal_history_dose"]["fields"] = sorted(set(int(plan[j][2]) for j, _, _ in slices))
    ph.close()
    t.close()
    return r


def voxel_error_report(d32, sum2, d64, nbatch, cutoff=0.1):
    """Per-voxel error of the FP32-transport dose against the FP64-everywhere
    run of the same histories, next to the dose's own statistical uncertainty
    (standard error of the total from the batch sum and sum of squares), over
    voxels above `cutoff` of the FP64 maximum."""
    import numpy as np
    m = d64 >= cutoff * d64.max()
    diff = np.abs(d32[m] - d64[m])
    rel = diff / d64[m]
    var = np.maximum(sum2[m] - d32[m] * d32[m] / nbatch, 0.0) * nbatch / (nbatch - 1)
    sig = np.sqrt(var)
    z = diff / np.maximum(sig, 1e-300)
    return {"voxels": int(m.sum()), "rel_err_median": float(np.median(rel)), "rel_err_p99": float(np.quantile(rel, 0.99)),
            "rel_err_max": float(rel.max()), "stat_rel_sigma_median": float(np.median(sig / d32[m])),
            "err_over_sigma_median": float(np.median(z)), "err_over_sigma_p99": float(np.quantile(z, 0.99)),
            "frac_err_below_sigma": float(np.mean(z < 1.0)),
            "what": "|FP32 - FP64| per voxel (same histories and random streams) vs the FP32 dose's batch "
                    "standard error (%d batches)" % nbatch}


def beamlet_leg(dpm_amd, gamma, t, ph, nx, lbox, n_beamlets, seed=7, n_check=4, sched=None, plain=True):
    """Config-4 workload: beamlets on the water phantom, plain (one dpm_simulate
    per beamlet) vs batched (one dpm_simulate_beamlets call). Both FP32 with the
    same history ids. The plain form is timed three ways: with the default
    and with its own tuned schedule, each leaving its result on the device (no
    copy, no thresholding), and with the default schedule copying each dense
    result to pinned host memory. The headline ratio takes the fastest plain
    form."""
    import numpy as np
    import phantoms
    bl = phantoms.beamlets(n_beamlets)
    total = sum(b[4] for b in bl)
    if sched is None:
        sched = dpm_amd.autotune_beamlets(t, ph, bl, nsample=4000000)
        sched["source"] = "dpm_autotune_beamlets on ~4e6 histories of this beamlet set"
    # warm-up: same beamlets with few histories (allocates the per-beamlet rows)
    dpm_amd.simulate_beamlets(t, ph, [b[:4] + (100,) for b in bl], rel_threshold=1e-3, seed=seed)
    t0 = time.perf_counter()
    row_ptr, col, val, st = dpm_amd.simulate_beamlets(t, ph, bl, rel_threshold=1e-3, seed=seed, schedule=sched)
    t_batched = time.perf_counter() - t0
    out = (dpm_amd.pinned_empty(ph.nbins), dpm_amd.pinned_empty(ph.nbins))
    offs = np.concatenate([[0], np.cumsum([b[4] for b in bl])])
    t_plain = t_plain_fetch = plain_sched = None
    plain_times = {}
    if plain:
        # the plain form's schedule: tuned on the median beamlet (one pencil beam per run)
        mid = sorted(bl, key=lambda b: b[4])[len(bl) // 2]
        plain_sched = dpm_amd.autotune(t, ph, nsample=max(1000, int(mid[4])), particle=mid[2], energy=mid[3],
                                       x0=mid[0], y0=mid[1])
        plain_sched["source"] = "dpm_autotune on the median beamlet (%d histories)" % mid[4]
        for name, sch in (("default_schedule", None), ("tuned_schedule", plain_sched)):
            t0 = time.perf_counter()
            for i, (x, y, p, e, n) in enumerate(bl):
                dpm_amd.simulate(t, ph, n, particle=p, energy=e, x0=x, y0=y, seed=seed, hist_begin=int(offs[i]),
                                 schedule=sch, fetch=False)
            plain_times[name] = time.perf_counter() - t0
        t_plain = min(plain_times.values())
        t0 = time.perf_counter()
        for i, (x, y, p, e, n) in enumerate(bl):
            dpm_amd.simulate(t, ph, n, particle=p, energy=e, x0=x, y0=y, seed=seed, hist_begin=int(offs[i]),
                             out=out)
        t_plain_fetch = time.perf_counter() - t0
    # correctness: batched FP32 rows of the largest beamlets vs FP64 plain runs
    order = np.argsort([-b[4] for b in bl])[:n_check]
    rates = []
    for i in order:
        x, y, p, e, n = bl[i]
        d64, _, _, _ = dpm_amd.simulate(t, ph, n, particle=p, energy=e, x0=x, y0=y, seed=seed,
                                        hist_begin=int(offs[i]), precision=dpm_amd.DPM_FP64)
        row = np.zeros(ph.nbins)
        row[col[row_ptr[i]:row_ptr[i + 1]]] = val[row_ptr[i]:row_ptr[i + 1]]
        rates.append(gamma.gamma_pass_rate_3d(d64, row.reshape(d64.shape), spacing_mm=lbox, dd=0.01, dta_mm=1.0,
                                              cutoff=0.1))
    return {"workload": "config4: %d beamlets on the 150^3 water phantom (e- 6-18 MeV / gamma U(1,20) MeV pencils "
                        "on a 2 mm grid over 160x160 mm2, lognormal history counts), FP32 transport, CSR of voxels "
                        "> 0.1%% of each beamlet's max" % n_beamlets,
            "histories": int(total), "batched_s": t_batched, "plain_s": t_plain,
            "speedup": t_plain / t_batched if t_plain else None,
            "batched_histories_per_s": total / t_batched,
            "plain_histories_per_s": total / t_plain if t_plain else None,
            "plain_what": "one dpm_simulate per beamlet, result left on the device: the faster of the default and "
                          "the tuned schedule",
            "plain_s_by_schedule": plain_times,
            "plain_fetch_s": t_plain_fetch,
            "speedup_vs_plain_fetch": t_plain_fetch / t_batched if t_plain_fetch else None,
            "plain_fetch_what": "one dpm_simulate per beamlet, default schedule, dense result copied to pinned host "
                                "memory",
            "batched_what": "one dpm_simulate_beamlets call and dpm_beamlets_fetch: transport, CSR build, CSR "
                            "copied to host memory",
            "nnz": int(row_ptr[-1]), "batched_schedule": sched, "plain_schedule": plain_sched,
            "gamma_batched_fp32_vs_plain_fp64": {"beamlets": [int(i) for i in order], "pass_rates": rates,
                                                 "criterion": "1%/1mm global, voxels > 10% of the beamlet max"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--histories", type=float, default=1e7)
    ap.add_argument("--nbatch", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-gamma", action="store_true")
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--beamlets", type=int, default=512, help="beamlets of the batched leg (0: skip it)")
    ap.add_argument("--no-tune", action="store_true", help="default launch schedule instead of the autotuner's")
    ap.add_argument("--refill-min", type=int, default=0, help="fix the refill threshold (skips the autotuner)")
    ap.add_argument("--tally-mode", type=int, default=-1, help="fix the voxel-tally mode (skips the autotuner)")
    ap.add_argument("--pipes", type=int, default=0, help="fix the number of pipes, 1 or 2 (skips the autotuner)")
    ap.add_argument("--no-plain", action="store_true", help="beamlet leg: skip the plain per-beamlet runs")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the configuration 3 and 5 legs (full-size heterogeneous runs)")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="first run a plain device-to-device copy of this many GiB (counter calibration)")
    ap.add_argument("--oracle-worker", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.oracle_worker is not None:
        _oracle_voxel_worker(args.oracle_worker)
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; the multi-GPU split (history ranges per GPU, one FP64 "
              "all-reduce of the dose) is not built, so --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import __graft_entry__
    import dpm_amd
    import gamma

    if args.copy_probe > 0:
        import torch
        n = int(args.copy_probe * (1 << 30)) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        # an elementwise kernel that reads n floats and writes n floats
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4}), flush=True)
        del src, dst
    # torch measures the copy rate before the engine opens the device
    copy_gbs = stream_copy_gbs()
    tables0 = __graft_entry__.tables_dir(0)
    nx, lbox, field, emin, emax = 150, 2.0, 100.0, 0.5, 20.0
    nhist = int(args.histories)
    t = dpm_amd.Tables(tables0)
    mat = np.zeros((nx, nx, nx), dtype=np.uint8)
    rho = np.ones((nx, nx, nx), dtype=np.float32)
    ph = dpm_amd.Phantom.voxel(t, mat, rho, lbox)

    # result arrays in page-locked host memory, allocated once like any
    # caller that repeats runs would
    outs = {prec: (dpm_amd.pinned_empty(ph.nbins), dpm_amd.pinned_empty(ph.nbins))
            for prec in (dpm_amd.DPM_FP32, dpm_amd.DPM_FP64)}

    # launch schedule: the autotuner's choice for this source (outside the
    # timed region), unless fixed on the command line
    fixed = (args.no_tune or args.waves_per_cu or args.chunk or args.refill_min or args.tally_mode >= 0 or
             args.pipes)
    if fixed:
        sched = {"chunk": args.chunk, "waves_per_cu": args.waves_per_cu, "refill_min": args.refill_min,
                 "tally_mode": max(0, args.tally_mode), "pipes": args.pipes, "source": "command line / defaults"}
    else:
        ns = min(nhist, 1 << 24)
        sched = dpm_amd.autotune(t, ph, nsample=ns, particle=dpm_amd.PHOTON, emin=emin, emax=emax, field=field,
                                 nbatch=args.nbatch)
        sched["source"] = "dpm_autotune on %d histories of this workload (%d batches)" % (ns, args.nbatch)

    def step(precision=dpm_amd.DPM_FP32, seed=1, fetch=True):
        return dpm_amd.simulate(t, ph, nhist, particle=dpm_amd.PHOTON, emin=emin, emax=emax, field=field, seed=seed,
                                precision=precision, nbatch=args.nbatch, out=outs[precision], schedule=sched,
                                fetch=fetch)

    for _ in range(args.warmup):
        step(fetch=False)
    # value: the whole job with the dose and its batch statistics left in HBM
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, _, st, _ = step(fetch=False)
    wall = time.perf_counter() - t0
    value = args.steps * nhist / wall
    dose, sum2 = dpm_amd.tally_fetch(t, ph, out=outs[dpm_amd.DPM_FP32])
    # the same job including the copy of dose and sum of squares to (pinned) host memory
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    wall_pcie = time.perf_counter() - t0

    # roofline of the dominant kernel (device time from HIP events on the run stream)
    ms_ph, ms_el = st["kernel_ms"][1], st["kernel_ms"][2]
    dom = "electrons" if ms_el >= ms_ph else "photons"
    cnt = st["counters_electron"] if dom == "electrons" else st["counters_photon"]
    launches = st["kernel_calls"][2 if dom == "electrons" else 1]
    ms_dom = ms_el if dom == "electrons" else ms_ph
    bytes_per_launch = algorithmic_bytes(cnt) / max(1, launches)
    avg_ms = ms_dom / max(1, launches)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    # HBM traffic per launch is not measured in this run: it comes from the
    # committed rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE) of this
    # workload, named in roofline.traffic_source
    traffic = traffic_source = None
    prof = os.path.join(ROOT, "profiles", "traffic_k_electrons_water150_photons.json")
    if os.path.exists(prof):
        with open(prof) as f:
            tr = json.load(f)
        traffic = tr.get("bytes_per_launch")
        traffic_source = tr.get("source", "profiles/traffic_k_electrons_water150_photons.json")
    atomics = atomic_roofline(dpm_amd, t, st, ph.nbins * max(1, args.nbatch))
    # the dominant kernel on one pipe (no launch shares the GPU with another
    # stream's): a few more steps of the same workload after the timed region
    one_pipe = None
    if int(sched.get("pipes") or 1) > 1:
        s1 = dict(sched, pipes=1, waves_per_cu=12, chunk=0)
        kw1 = dict(particle=dpm_amd.PHOTON, emin=emin, emax=emax, field=field, seed=1, nbatch=args.nbatch,
                   out=outs[dpm_amd.DPM_FP32], schedule=s1, fetch=False)
        dpm_amd.simulate(t, ph, nhist, **kw1)
        ms1 = calls1 = b1 = 0.0
        for _ in range(3):
            _, _, st1, _ = dpm_amd.simulate(t, ph, nhist, **kw1)
            k = 2 if dom == "electrons" else 1
            ms1 += st1["kernel_ms"][k]
            calls1 += st1["kernel_calls"][k]
            b1 += algorithmic_bytes(st1["counters_electron"] if dom == "electrons" else st1["counters_photon"])
        avg1 = ms1 / max(1, calls1)
        ach1 = b1 / max(1, calls1) / (avg1 * 1e-3) / 1e9
        one_pipe = {"avg_launch_ms": avg1, "launches": int(calls1), "bytes_per_launch": b1 / max(1, calls1),
                    "achieved": ach1, "frac": ach1 / HBM_PEAK_GBS,
                    "what": "3 more steps of this workload on one pipe at 12 waves/CU, after the timed region"}

    result = {
        "metric": "histories/s", "value": value, "unit": "histories/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "config2: 150^3 water cube, 2 mm voxels, 10x10 cm2 parallel photon field, "
                               "E ~ U(0.5, 20) MeV, %d histories per step, FP32 transport, FP64 tally, %d batches"
                               % (nhist, args.nbatch),
                   "histories": nhist, "voxels": nx ** 3, "nbatch": args.nbatch},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_source,
                     "kernel": "k_" + dom,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms, "launches": launches,
                     # with two pipes, launches of the two streams share the GPU: the
                     # per-launch rate above counts each launch over its whole duration;
                     # this one divides the kernel's bytes of a step by the step's device time
                     "pipes": int(sched.get("pipes") or 1),
                     "achieved_per_device_second": algorithmic_bytes(cnt) / st["seconds"] / 1e9,
                     "one_pipe": one_pipe,
                     "measured_copy_GBs": copy_gbs, "atomics": atomics},
        "kernel_ms": {"source": st["kernel_ms"][0], "photons": ms_ph, "electrons": ms_el, "fold": st["kernel_ms"][3]},
        "device_seconds_per_step": st["seconds"],
        "pcie_inclusive": {"value": args.steps * nhist / wall_pcie, "ms_per_step": wall_pcie / args.steps * 1e3,
                           "what": "the same step returning dose and sum of squares (2 x %d FP64) to pinned "
                                   "host memory" % ph.nbins},
        "schedule": sched,
        "counters": st["counters"],
    }
    if not args.no_gamma:
        d64, _, st64, _ = step(precision=dpm_amd.DPM_FP64)
        pr = gamma.gamma_pass_rate_3d(d64, dose, spacing_mm=lbox, dd=0.01, dta_mm=1.0, cutoff=0.1)
        m = d64 >= 0.1 * d64.max()
        result["gamma_fp32_vs_fp64"] = {"pass_rate": pr, "criterion": "1%/1mm global, voxels > 10% max",
                                        "max_rel_dose_diff_above_cutoff":
                                            float(np.max(np.abs(dose[m] - d64[m]) / d64.max())),
                                        "fp64_histories_per_s": nhist / st64["seconds"],
                                        "voxel_error": (voxel_error_report(dose, sum2, d64, args.nbatch)
                                                        if args.nbatch > 1 else None)}
    if args.beamlets > 0:
        result["beamlets"] = beamlet_leg(dpm_amd, gamma, t, ph, nx, lbox, args.beamlets,
                                         sched=None if not fixed else sched, plain=not args.no_plain)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(tables0, nx, lbox, field, emin, emax, nhist, 1, gpu_dose=dose,
                                              gamma=gamma)
    if not args.no_extra:
        import phantoms
        ph.close()
        t.close()
        tables3 = __graft_entry__.tables_dir(3)
        try:
            result["config1"] = config1_leg(dpm_amd, gamma, tables3, cpu=not args.no_cpu)
        except Exception as e:
            result["config1"] = {"error": "%s: %s" % (type(e).__name__, e)}
        for key, leg in (("config3", config3_leg), ("config5", config5_leg)):
            try:
                result[key] = leg(dpm_amd, gamma, phantoms, tables3, cpu=not args.no_cpu,
                                  fixed=sched if fixed else None)
            except Exception as e:   # a failing extra leg is reported, the headline line still prints
                result[key] = {"error": "%s: %s" % (type(e).__name__, e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
