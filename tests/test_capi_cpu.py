"""The C-ABI library loads and exports every entry point include/dpm_amd.h
declares; error paths that do not need a GPU return the documented codes.
No compute call is made here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import dpm_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dpm_amd.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dpm_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = dpm_amd.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(dpm_amd.EXPORTED)


def test_abi_version():
    assert dpm_amd.lib().dpm_abi_version() == 4


def test_struct_layouts_match_the_header(tmp_path):
    # sizes and member offsets of the ctypes mirrors against the C compiler's
    # view of include/dpm_amd.h
    structs = {"dpm_stats": dpm_amd.Stats, "dpm_run_options": dpm_amd.RunOptions, "dpm_source": dpm_amd.Source,
               "dpm_schedule": dpm_amd.Schedule, "dpm_beamlet": dpm_amd.Beamlet, "dpm_field": dpm_amd.Field,
               "dpm_phsp_particle": dpm_amd.PhspParticle}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "dpm_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append('  printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    c = dict(l.split() for l in subprocess.check_output([str(exe)], text=True).splitlines())
    for cname, py in structs.items():
        assert int(c[cname]) == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert int(c["%s.%s" % (cname, f[0])]) == getattr(py, f[0]).offset, (cname, f[0])


def test_missing_table_directory_is_an_io_error(tmp_path):
    h = ctypes.c_void_p()
    rc = dpm_amd.lib().dpm_tables_load(str(tmp_path / "nope").encode(), 0, ctypes.byref(h))
    assert rc == -2
    assert b"not found" in dpm_amd.lib().dpm_last_error()


def test_truncated_table_file_is_an_io_error(tmp_path, tables3):
    import shutil
    d = tmp_path / "tables"
    shutil.copytree(tables3, d)
    with open(d / "el_GSDtrData.dat", "r+") as f:
        f.truncate(5000)
    h = ctypes.c_void_p()
    rc = dpm_amd.lib().dpm_tables_load(str(d).encode(), 0, ctypes.byref(h))
    assert rc == -2
    assert b"malformed" in dpm_amd.lib().dpm_last_error()


def test_cli_help_lists_reference_flags():
    out = subprocess.run([dpm_amd.CLI_PATH, "-h"], capture_output=True, text=True)
    for flag in ("-p", "-e", "-n", "-d", "-b", "-c", "-o"):
        assert "\t%s  --" % flag in out.stdout


def test_null_arguments_are_rejected():
    L = dpm_amd.lib()
    assert L.dpm_simulate(None, None, None, 0, None, None, None, None) == -1
    assert L.dpm_sample(None, 0, 0, 0, None, None, None, None, None, None, None, None) == -1


def test_voxeldose_cli_help():
    out = subprocess.run([dpm_amd.VOXEL_CLI_PATH, "-h"], capture_output=True, text=True)
    assert out.returncode == 0
    for flag in ("-d", "-x", "-m", "-r", "-E", "-F", "-k", "-B", "-o"):
        assert "  %s " % flag in out.stdout
    bad = subprocess.run([dpm_amd.VOXEL_CLI_PATH], capture_output=True, text=True)
    assert bad.returncode != 0


def test_dose_file_readers_round_trip(tmp_path):
    nx, ny, nz = 3, 4, 5
    rng = np.random.default_rng(0)
    dose, rel = rng.random((nz, ny, nx)), rng.random((nz, ny, nx))
    p = tmp_path / "d.bin"
    p.write_bytes(b"# dpm-amd voxel dose v1\n3 4 5 2 1000 10\n" + dose.tobytes() + rel.tobytes())
    d = dpm_amd.read_voxel_dose(str(p))
    assert (d["nx"], d["ny"], d["nz"], d["lbox"], d["nhist"], d["nbatch"]) == (3, 4, 5, 2.0, 1000, 10)
    assert np.array_equal(d["dose"], dose) and np.array_equal(d["rel_sigma"], rel)
    p.write_bytes(b"# dpm-amd voxel dose v1\n3 4 5 2 1000 10\n" + dose.tobytes())
    with pytest.raises(ValueError):
        dpm_amd.read_voxel_dose(str(p))
    rp = np.array([0, 2, 3], np.int64)
    col = np.array([1, 7, 4], np.int32)
    val = np.array([0.5, 1.5, 2.5])
    q = tmp_path / "m.csr"
    q.write_bytes(b"# dpm-amd csr v1\n2 60 3\n" + rp.tobytes() + col.tobytes() + val.tobytes())
    r2, c2, v2, nvox = dpm_amd.read_csr(str(q))
    assert nvox == 60 and np.array_equal(r2, rp) and np.array_equal(c2, col) and np.array_equal(v2, val)


@pytest.mark.parametrize("text", ["0.5 1.0 1.0\n1.0 2.0 0,5\n", "0.5 1.0 1.0\n1.0 2.0\n", "0.5 1.0 1.0 7\n",
                                  "0.5 1.0 1.0\n1.0 2.0 x\n"])
def test_voxeldose_rejects_a_malformed_spectrum_line(tmp_path, tables0, text):
    # every non-blank, non-comment line must be exactly three numbers; the
    # file is checked before the device is touched, so this runs without a GPU
    spec = tmp_path / "s.txt"
    spec.write_text("# E_low E_high weight\n\n" + text)
    r = subprocess.run([dpm_amd.VOXEL_CLI_PATH, "-d", tables0, "-x", "4", "-y", "4", "-z", "4", "-S", str(spec),
                        "-n", "10", "-o", str(tmp_path / "x.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "expected 'E_low E_high weight'" in r.stderr


@pytest.mark.parametrize("text", ["gamma 5 0 0 -10 0 0 1\ne- 5 0 0 -10 0 0\n", "muon 5 0 0 -10 0 0 1\n",
                                  "gamma 5 0 0 -10 0 0 1 9\n"])
def test_voxeldose_rejects_a_malformed_phase_space_line(tmp_path, tables0, text):
    f = tmp_path / "ps.txt"
    f.write_text("# particle E x y z u v w\n" + text)
    r = subprocess.run([dpm_amd.VOXEL_CLI_PATH, "-d", tables0, "-x", "4", "-y", "4", "-z", "4", "-P", str(f),
                        "-o", str(tmp_path / "x.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "expected 'particle E x y z u v w'" in r.stderr
