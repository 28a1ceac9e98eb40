#!/usr/bin/env python3
"""Writes sb_nist_brem.npz: the Seltzer-Berger scaled bremsstrahlung DCS of
the elements dpm_tablegen knows, taken from the reference's data files.

The reference ships the NIST tables (Seltzer & Berger, At. Data Nucl. Data
Tables 35 (1986) 345) as text under SBTables/data/NIST_BREM: a grid file
(``nist_brems_grid``: the counts 99 113 32, then 113 electron kinetic
energies [MeV] and 32 reduced photon energies kappa = k/T) and one file per
Z (``nist_brems_Z``: 113 x 32 values of chi = (beta^2/Z^2) k dsigma/dk in
mb, energy-major), read by SBTableBuilder::LoadDCSData
(DataInit/src/SBTableBuilder.cc:244-293). This script keeps the values of
the elements in dpm_tablegen.ELEMENTS unchanged (float64 parses of the text),
so the generator has them where the reference tree is absent.

    python make_sb_data.py [NIST_BREM directory]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import dpm_tablegen  # noqa: E402

DEFAULT_SRC = "/root/reference/SBTables/data/NIST_BREM"


def read_nist_brem(src, zs):
    """(energy[113], kappa[32], chi[len(zs), 113, 32]) from the text files."""
    with open(os.path.join(src, "nist_brems_grid")) as f:
        tok = f.read().split()
    nz, ne, nk = (int(t) for t in tok[:3])
    vals = np.array([float(t) for t in tok[3:3 + ne + nk]])
    if len(vals) != ne + nk:
        raise ValueError("nist_brems_grid: expected %d values" % (ne + nk))
    energy, kappa = vals[:ne], vals[ne:]
    chi = np.empty((len(zs), ne, nk))
    for i, z in enumerate(zs):
        if not 1 <= z <= nz:
            raise ValueError("Z=%d outside the tables" % z)
        with open(os.path.join(src, "nist_brems_%d" % z)) as f:
            v = np.array([float(t) for t in f.read().split()])
        if len(v) != ne * nk:
            raise ValueError("nist_brems_%d: %d values, expected %d" % (z, len(v), ne * nk))
        chi[i] = v.reshape(ne, nk)
    return energy, kappa, chi


def main(argv):
    src = argv[1] if len(argv) > 1 else DEFAULT_SRC
    zs = sorted({z for z, _, _ in dpm_tablegen.ELEMENTS.values()})
    energy, kappa, chi = read_nist_brem(src, zs)
    out = os.path.join(HERE, "sb_nist_brem.npz")
    np.savez_compressed(out, Z=np.array(zs, dtype=np.int32), energy=energy, kappa=kappa, chi=chi)
    print("wrote %s: Z=%s, %d energies x %d kappas" % (out, zs, len(energy), len(kappa)))


if __name__ == "__main__":
    main(sys.argv)
