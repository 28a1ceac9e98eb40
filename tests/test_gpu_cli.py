"""The command-line programs end to end on the GPU.

* dpm_Simulate (the reference's command line): hist.sim in the layout of
  Geom::Write (Geom.cc:167-192), FP64 transport matching the FP64 oracle on
  the same Philox histories;
* dpm_VoxelDose: voxel dose [Gy/history] and batch uncertainty equal to the
  library run converted by hand; beamlet mode writes the library's CSR.
"""
import os
import subprocess

import numpy as np
import pytest

import dpm_amd
import oracle_lib

MEV_J = 1.602176634e-13


def read_hist_sim(path):
    rows, header = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                header.append(line.split(":")[0])
            else:
                rows.append([float(x) for x in line.split()])
    return header, np.array(rows)


@pytest.mark.gpu
def test_dpm_simulate_cli_writes_reference_hist_sim(tables3, tmp_path):
    out = tmp_path / "hist.sim"
    n = 20000
    subprocess.run([dpm_amd.CLI_PATH, "-p", "e-", "-e", "15", "-n", str(n), "-d", tables3, "-b", "1.0", "-c", "3",
                    "-f", "-s", "5", "-o", str(out)], check=True, capture_output=True, timeout=120)
    header, rows = read_hist_sim(out)
    ref_header, ref_rows = read_hist_sim(os.path.join(os.path.dirname(__file__), "golden",
                                                      "ref_hist_cfg3_e15_n10000.sim"))
    assert header == ref_header and rows.shape[1] == ref_rows.shape[1] == 4
    assert np.array_equal(rows[:, 0], np.arange(len(rows)))
    np.testing.assert_allclose(rows[:, 1], (rows[:, 0] + 0.5) * 0.1, rtol=1e-4)
    ref = oracle_lib.run_slab(tables3, n, e0=15.0, ptype=-1, rng=oracle_lib.RNG_PHILOX, seed=5)
    ref_txt = oracle_lib.format_hist_sim(ref["edep"], ref["steps"], len(rows), n, 1.0, [1.0, 4.54, 1.85], 3)
    lines = ref_txt.splitlines() if isinstance(ref_txt, str) else ref_txt
    ref_vals = np.array([[float(x) for x in l.split()] for l in lines if l.strip() and not l.startswith("#")])
    m = ref_vals[:, 2] > 0.1 * ref_vals[:, 2].max()
    np.testing.assert_allclose(rows[m, 2], ref_vals[m, 2], rtol=2e-3)


REFMAIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                       "dpm_Simulate_refmain")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(REFMAIN), reason="the reference's main is built only where its tree exists")
def test_reference_main_unchanged_runs_on_the_engine(tables3, tmp_path):
    # the reference's own dpm_Simulate.cc, compiled unchanged against the host
    # data layer (oracle/Makefile target refmain): its Simulate() call runs the
    # GPU engine with the defaults (FP32, Philox seed 1, device 0), so its
    # hist.sim equals the engine's own command line with those settings
    args = ["-p", "e-", "-e", "15", "-n", "20000", "-d", tables3, "-b", "1.0", "-c", "3"]
    a, b = tmp_path / "ref_main.sim", tmp_path / "engine.sim"
    subprocess.run([REFMAIN] + args + ["-o", str(a)], check=True, capture_output=True, timeout=120)
    subprocess.run([dpm_amd.CLI_PATH] + args + ["-s", "1", "-o", str(b)], check=True, capture_output=True,
                   timeout=120)
    ha, ra = read_hist_sim(a)
    hb, rb = read_hist_sim(b)
    assert ha == hb and ra.shape == rb.shape
    np.testing.assert_allclose(ra, rb, rtol=1e-9, atol=1e-12)
    assert 13.0 < float(open(a).readline().split(":")[1].split()[0]) < 14.5


@pytest.mark.gpu
def test_dpm_voxeldose_cli_dose_and_uncertainty(gpu_tables0, tables0, tmp_path):
    nx, ny, nz, lbox, n, nb = 16, 14, 12, 2.0, 20000, 8
    mat = np.zeros((nz, ny, nx), np.uint8)
    rho = np.random.default_rng(3).uniform(0.8, 1.2, (nz, ny, nx)).astype(np.float32)
    (tmp_path / "mat.u8").write_bytes(mat.tobytes())
    (tmp_path / "rho.f32").write_bytes(rho.tobytes())
    out = tmp_path / "dose.bin"
    subprocess.run([dpm_amd.VOXEL_CLI_PATH, "-d", tables0, "-x", str(nx), "-y", str(ny), "-z", str(nz), "-l",
                    str(lbox), "-m", str(tmp_path / "mat.u8"), "-r", str(tmp_path / "rho.f32"), "-p", "gamma",
                    "-E", "0.5,20", "-F", "20", "-n", str(n), "-k", str(nb), "-s", "9", "-f", "-o", str(out)],
                   check=True, capture_output=True, timeout=120)
    d = dpm_amd.read_voxel_dose(str(out))
    assert (d["nx"], d["ny"], d["nz"], d["nhist"], d["nbatch"]) == (nx, ny, nz, n, nb)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, lbox)
    edep, sum2, _, _ = dpm_amd.simulate(gpu_tables0, ph, n, particle=0, emin=0.5, emax=20.0, field=20.0, seed=9,
                                        nbatch=nb, precision=dpm_amd.DPM_FP64)
    mass_kg = rho.astype(np.float64) * (lbox * 0.1) ** 3 * 1e-3
    np.testing.assert_allclose(d["dose"], edep * MEV_J / mass_kg / n, rtol=1e-9, atol=1e-30)
    mean = edep / nb
    with np.errstate(invalid="ignore", divide="ignore"):
        var = np.maximum(sum2 / nb - mean * mean, 0.0) / (nb - 1)
        rel = np.where(edep > 0, np.sqrt(var) / mean, 0.0)
    np.testing.assert_allclose(d["rel_sigma"], rel, rtol=1e-6, atol=1e-12)
    assert 0 < np.median(d["rel_sigma"][edep > 0.1 * edep.max()]) < 0.5


@pytest.mark.gpu
def test_dpm_voxeldose_cli_beamlet_csr(gpu_tables0, tables0, tmp_path):
    nx = ny = nz = 16
    bl = [(-6.0, 4.0, -1, 9.0, 3000), (10.0, -2.0, 0, 12.0, 5000)]
    (tmp_path / "bl.txt").write_text("# x y particle energy histories\n" + "".join(
        "%g %g %s %g %d\n" % (x, y, {-1: "e-", 0: "gamma", 1: "e+"}[p], e, h) for x, y, p, e, h in bl))
    out = tmp_path / "bl.csr"
    subprocess.run([dpm_amd.VOXEL_CLI_PATH, "-d", tables0, "-x", str(nx), "-y", str(ny), "-z", str(nz), "-l", "2",
                    "-B", str(tmp_path / "bl.txt"), "-s", "4", "-f", "-o", str(out)],
                   check=True, capture_output=True, timeout=120)
    row_ptr, col, val, nvox = dpm_amd.read_csr(str(out))
    assert nvox == nx * ny * nz
    ph = dpm_amd.Phantom.voxel(gpu_tables0, np.zeros((nz, ny, nx), np.uint8), np.ones((nz, ny, nx), np.float32), 2.0)
    r2, c2, v2, _ = dpm_amd.simulate_beamlets(gpu_tables0, ph, bl, seed=4, precision=dpm_amd.DPM_FP64)
    assert np.array_equal(row_ptr, r2) and np.array_equal(col, c2)
    np.testing.assert_allclose(val, v2, rtol=1e-9)


@pytest.mark.gpu
def test_dpm_voxeldose_cli_spectrum_file(gpu_tables0, tables0, tmp_path):
    # -S spectrum file ("E_low E_high weight" per bin): the same dose as the
    # library call with the same table spectrum
    nx = ny = nz = 12
    lbox, n = 2.0, 5000
    spec = tmp_path / "spectrum.txt"
    spec.write_text("# E_low E_high weight (MeV)\n0.5 1.0 1.0\n1.0 2.0 3.0\n2.0 4.0 0.0\n4.0 6.0 2.5\n6.0 15.0 0.5\n")
    out = tmp_path / "dose.bin"
    subprocess.run([dpm_amd.VOXEL_CLI_PATH, "-d", tables0, "-x", str(nx), "-y", str(ny), "-z", str(nz), "-l",
                    str(lbox), "-p", "gamma", "-S", str(spec), "-F", "16", "-n", str(n), "-k", "4", "-s", "5", "-f",
                    "-o", str(out)], check=True, capture_output=True, timeout=120)
    d = dpm_amd.read_voxel_dose(str(out))
    mat = np.zeros((nz, ny, nx), np.uint8)
    rho = np.ones((nz, ny, nx), np.float32)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, lbox)
    table = (np.array([0.5, 1.0, 2.0, 4.0, 6.0, 15.0]), np.array([1.0, 3.0, 0.0, 2.5, 0.5]))
    edep, _, _, _ = dpm_amd.simulate(gpu_tables0, ph, n, particle=0, field=16.0, seed=5, nbatch=4,
                                     precision=dpm_amd.DPM_FP64, spectrum_table=table)
    mass_kg = rho.astype(np.float64) * (lbox * 0.1) ** 3 * 1e-3
    np.testing.assert_allclose(d["dose"], edep * MEV_J / mass_kg / n, rtol=1e-9, atol=1e-30)
    bad = tmp_path / "gap.txt"
    bad.write_text("0.5 1.0 1.0\n1.5 2.0 1.0\n")
    r = subprocess.run([dpm_amd.VOXEL_CLI_PATH, "-d", tables0, "-x", "4", "-y", "4", "-z", "4", "-S", str(bad),
                        "-n", "10", "-o", str(tmp_path / "x.bin")], capture_output=True, timeout=60)
    assert r.returncode != 0


@pytest.mark.gpu
def test_dpm_voxeldose_cli_phase_space_file(gpu_tables0, tables0, tmp_path):
    # -P phase-space file: one history per record; the dose equals the
    # library's dpm_simulate_phsp on the same records
    nx = ny = nz = 12
    lbox = 2.0
    rng = np.random.default_rng(4)
    n = 3000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    kinds = rng.choice(["gamma", "e-", "e+"], n)
    rows = np.column_stack([rng.uniform(1, 10, n), rng.uniform(-10, 10, (n, 3)), d])
    lines = ["# particle E x y z u v w"] + ["%s %.17g %.17g %.17g %.17g %.17g %.17g %.17g" % ((k,) + tuple(r))
                                            for k, r in zip(kinds, rows)]
    ps = tmp_path / "ps.txt"
    ps.write_text("\n".join(lines) + "\n")
    out = tmp_path / "dose.bin"
    subprocess.run([dpm_amd.VOXEL_CLI_PATH, "-d", tables0, "-x", str(nx), "-y", str(ny), "-z", str(nz), "-l",
                    str(lbox), "-P", str(ps), "-k", "5", "-s", "7", "-f", "-o", str(out)], check=True,
                   capture_output=True, timeout=120)
    v = dpm_amd.read_voxel_dose(str(out))
    assert v["nhist"] == n
    mat = np.zeros((nz, ny, nx), np.uint8)
    rho = np.ones((nz, ny, nx), np.float32)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, lbox)
    code = {"gamma": 0, "e-": -1, "e+": 1}
    recs = np.column_stack([[code[k] for k in kinds], rows])
    edep, _, _, _ = dpm_amd.simulate_phsp(gpu_tables0, ph, recs, seed=7, nbatch=5, precision=dpm_amd.DPM_FP64)
    mass_kg = rho.astype(np.float64) * (lbox * 0.1) ** 3 * 1e-3
    np.testing.assert_allclose(v["dose"], edep * MEV_J / mass_kg / n, rtol=1e-9, atol=1e-30)
    ph.close()


@pytest.mark.gpu
def test_dpm_voxeldose_cli_collimated_field(gpu_tables0, tables0, tmp_path):
    # -A gantry field collimated by -D (point source) and -J/-K (jaws): the
    # same dose as the library call with the same collimation
    nx = ny = nz = 14
    lbox, n = 2.0, 6000
    out = tmp_path / "dose.bin"
    subprocess.run([dpm_amd.VOXEL_CLI_PATH, "-d", tables0, "-x", str(nx), "-y", str(ny), "-z", str(nz), "-l",
                    str(lbox), "-p", "gamma", "-e", "6", "-F", "24", "-A", "30", "-D", "120", "-J", "-8,5,-4,9",
                    "-K", "1.2", "-n", str(n), "-k", "4", "-s", "13", "-f", "-o", str(out)],
                   check=True, capture_output=True, timeout=120)
    d = dpm_amd.read_voxel_dose(str(out))
    mat = np.zeros((nz, ny, nx), np.uint8)
    rho = np.ones((nz, ny, nx), np.float32)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, lbox)
    edep, _, _, _ = dpm_amd.simulate(gpu_tables0, ph, n, particle=0, energy=6.0, field=24.0, angle=30.0, seed=13,
                                     nbatch=4, precision=dpm_amd.DPM_FP64, sad=120.0, jaws=(-8.0, 5.0, -4.0, 9.0),
                                     jaw_tau=1.2)
    mass_kg = rho.astype(np.float64) * (lbox * 0.1) ** 3 * 1e-3
    np.testing.assert_allclose(d["dose"], edep * MEV_J / mass_kg / n, rtol=1e-9, atol=1e-30)
    r = subprocess.run([dpm_amd.VOXEL_CLI_PATH, "-d", tables0, "-x", "4", "-y", "4", "-z", "4", "-D", "100",
                        "-n", "10", "-o", str(tmp_path / "x.bin")], capture_output=True, timeout=60)
    assert r.returncode != 0          # collimation needs a gantry field
