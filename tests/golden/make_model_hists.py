#!/usr/bin/env python3
"""Writes tests/golden/model_hists_cfg3.npz: expected histograms of single
interactions, the quantities the reference's model tests histogram
(testAPI.cc:92-199: BremTest, MollerTest, MSCAngularDeflectionTest, plus the
Compton final state), computed with the FP64 oracle samplers -- which are
bit-exact to the reference's SimSBTables / SimMollerTables / SimKNTables /
SimGSTables (tests/test_oracle_reference.py) -- from independent numpy
uniforms, on the tables of dpm_tablegen configuration 3.

Per case: the histogram variable, its bin edges, the counts of N samples.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "dpm-g4cpp_amd", "tablegen"))
sys.path.insert(0, ROOT)

import oracle_lib  # noqa: E402
from dpm_tablegen import tables_digest  # noqa: E402

EMC2 = 0.510991
ECUT, GCUT = 0.2, 0.05          # configuration 3 tracking cuts (MeV)
N = 4_000_000
NBINS = 60

# name: (dpm_perform kind, oracle sampler, energy MeV, material, variable)
CASES = {
    "brem_e10_water": ("brem", "sb", 10.0, 0, "log10_k_over_e"),
    "brem_e5_ti": ("brem", "sb", 5.0, 1, "log10_k_over_e"),
    "moller_e10": ("moller", "mo", 10.0, 0, "log10_t_over_e"),
    "msc_e5": ("msc", "gs", 5.0, 0, "log10_one_minus_cos"),
    "msc_e0.5": ("msc", "gs", 0.5, 0, "log10_one_minus_cos"),
    "compton_e2": ("compton", "kn", 2.0, 0, "eps"),
}


def edges(name):
    kind, _, e, _, _ = CASES[name]
    if kind == "brem":
        return np.linspace(np.log10(GCUT / e) - 0.02, 0.0, NBINS + 1)
    if kind == "moller":
        return np.linspace(np.log10(ECUT / e) - 0.02, np.log10(0.5) + 0.02, NBINS + 1)
    if kind == "msc":
        return np.linspace(-9.0, np.log10(2.0) + 0.01, NBINS + 1)
    kappa = e / EMC2
    return np.linspace(1.0 / (1.0 + 2.0 * kappa) - 0.01, 1.0 + 1e-9, NBINS + 1)


def variable(name, v):
    """The histogram variable from the sampler output (photon or secondary
    energy; cos(theta) for GS; eps for KN) or from dpm_perform's value."""
    kind, _, e, _, _ = CASES[name]
    if kind in ("brem", "moller"):
        return np.log10(v / e)
    if kind == "msc":
        return np.log10(np.maximum(v, 1e-300))
    return v


def main():
    import __graft_entry__
    datadir = __graft_entry__.tables_dir(3)
    with open(os.path.join(HERE, "tables_cfg3.sha256")) as f:
        assert tables_digest(datadir) == f.read().strip(), "generated tables differ from the fixtures' tables"
    oracle_lib.load(datadir)
    out = {}
    for i, (name, (kind, smp, e, imat, var)) in enumerate(CASES.items()):
        r = np.random.default_rng(1000 + i).random((N, 3))
        v = oracle_lib.sample_many(smp, e, r, imat)
        if smp == "gs":
            v = 1.0 - v                     # 1 - cos(theta), as dpm_perform returns it
        elif smp == "kn":
            pass                            # eps
        x = variable(name, v)
        ed = edges(name)
        out[name + "_edges"] = ed
        out[name + "_counts"] = np.histogram(x, bins=ed)[0].astype(np.int64)
        out[name + "_under_over"] = np.array([np.sum(x < ed[0]), np.sum(x >= ed[-1])])
        print(name, "in range %.6f" % (out[name + "_counts"].sum() / N), flush=True)
    out["n"] = np.array(N)
    np.savez_compressed(os.path.join(HERE, "model_hists_cfg3.npz"), **out)


if __name__ == "__main__":
    main()
