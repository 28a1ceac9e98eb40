"""The C-ABI library loads and exports every entry point include/dpm_amd.h
declares; error paths that do not need a GPU return the documented codes.
No compute call is made here."""
import ctypes
import os
import re
import subprocess

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
    assert dpm_amd.lib().dpm_abi_version() == 1


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
