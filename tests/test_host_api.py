"""The reference-compatible host API (dpm-g4cpp_amd/host): the reference's
data-layer classes with its Load(dataDir, verbose) and getters, and its
Simulate() signature.

* The reference's own command-line main (dpm_Simulate.cc), unchanged,
  compiles against these headers and runs (where the reference tree is
  present; its transport is then the GPU engine).
* The getters and samplers of SimElectronData / SimPhotonData give the
  reference's values bit for bit on the fixture's queries
  (tests/golden/samplers_cfg3.npz, written by the reference's own classes).
* Load(dataDir) of a missing directory ends the process with EXIT_FAILURE,
  as the reference's loaders do.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "dpm-g4cpp_amd", "host")
REF_MAIN = "/root/reference/dpm_Simulate.cc"
GOLD = os.path.join(ROOT, "tests", "golden", "samplers_cfg3.npz")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = tmp_path_factory.mktemp("host") / "host_api_driver"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", HOST, "-I",
                           os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "host_api_driver.cc"),
                           os.path.join(HOST, "SimData.cc"), os.path.join(HOST, "SimTables.cc"), "-o", str(exe)])
    return str(exe)


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="the reference tree is not present")
def test_reference_main_compiles_unchanged_and_runs(tmp_path):
    lib = os.path.join(ROOT, "dpm-g4cpp_amd", "lib")
    exe = tmp_path / "dpm_Simulate"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", HOST, "-I", os.path.join(ROOT, "include"), REF_MAIN,
                           os.path.join(HOST, "SimDPMLike.cc"), os.path.join(HOST, "SimData.cc"),
                           os.path.join(HOST, "SimTables.cc"), "-L", lib, "-ldpm_amd", "-Wl,-rpath," + lib,
                           "-o", str(exe)])
    out = subprocess.run([str(exe), "-h"], capture_output=True, text=True)
    # the reference's -h prints the usage and ends with exit(-1)
    assert out.returncode == 255 and "Usage: dpm_Simulate" in out.stdout
    for flag in ("-p", "-e", "-n", "-d", "-b", "-c", "-o"):
        assert "\t%s  --" % flag in out.stdout


def _ask(driver, datadir, lines):
    out = subprocess.run([driver, datadir], input="\n".join(lines) + "\n", capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    return np.array([float(v) for v in out.stdout.split()])


def test_getters_and_samplers_match_the_reference_bit_for_bit(driver, tables3):
    from dpm_tablegen import tables_digest
    with open(os.path.join(ROOT, "tests", "golden", "tables_cfg3.sha256")) as f:
        assert tables_digest(tables3) == f.read().strip()
    g = np.load(GOLD)
    u = 1.0 / (1 << 24)
    for kind in ("sb", "mo", "kn", "gs"):
        e, k, im = g[kind + "_e"], g[kind + "_k"], g[kind + "_imat"]
        lines = []
        for i in range(e.size):
            r = ["%.17g" % (x * u) for x in k[i]]
            if kind == "sb":
                lines.append("sb %.17g %d %s" % (e[i], im[i], " ".join(r)))
            else:
                lines.append("%s %.17g %s" % (kind, e[i], " ".join(r)))
        got = _ask(driver, tables3, lines)
        np.testing.assert_array_equal(got, g[kind + "_ref"], err_msg=kind)
    rows = g["ev_rows"]
    got = _ask(driver, tables3, ["ev %d %.17g %d" % (int(w), e, int(m)) for w, e, m in rows])
    np.testing.assert_array_equal(got, g["ev_ref"])


def test_load_of_a_missing_directory_exits_like_the_reference(driver, tmp_path):
    out = subprocess.run([driver, str(tmp_path / "nope")], input="", capture_output=True, text=True)
    assert out.returncode == 1
    assert "not found" in out.stderr
