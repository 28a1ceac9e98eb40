"""The oracle is pinned against the reference itself.

Fixtures in tests/golden/ were written by the reference compiled from its own
sources (tests/golden/make_golden.py): its dpm_Simulate hist.sim and the
outputs of its Sim*Tables samplers and interpolation getters, on the tables
of dpm_tablegen configuration 3, and its hist.sim of the other pre-defined
slab geometries (0: water; 1: water/air/water; 2: water/bone/water) on the
tables of their configurations. The oracle must reproduce them bit for bit.
"""
import os

import numpy as np
import pytest

import oracle_lib

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
DENS_CFG3 = [1.0, 4.54, 1.85]


@pytest.fixture(scope="module")
def checked_tables(tables3):
    from dpm_tablegen import tables_digest
    with open(os.path.join(GOLD, "tables_cfg3.sha256")) as f:
        assert tables_digest(tables3) == f.read().strip(), "generated tables differ from the fixtures' tables"
    return tables3


@pytest.mark.parametrize("n", [10000, 100000])
def test_hist_sim_identical_to_reference(checked_tables, n):
    r = oracle_lib.run_slab(checked_tables, n, e0=15.0, ptype=-1, lbox=1.0, geom_index=3, rng=oracle_lib.RNG_MT,
                            seed=123, arg_rtl=1)
    txt = oracle_lib.format_hist_sim(r["edep"], r["steps"], r["hist_size"], n, 1.0, DENS_CFG3, 3)
    with open(os.path.join(GOLD, "ref_hist_cfg3_e15_n%d.sim" % n)) as f:
        ref = f.read()
    assert txt == ref


# the reference's other pre-defined slab geometries (Geom.cc:133-164): the
# case of each fixture (tests/golden/make_golden.py) and its materials' densities
SLAB_CASES = {0: (-1, 10.0, 1.0, "ref_hist_geom0_e10_b1_n10000.sim", [1.0]),
              1: (0, 12.0, 2.0, "ref_hist_geom1_g12_b2_n10000.sim", [1.0, 0.00120479]),
              2: (-1, 6.0, 0.5, "ref_hist_geom2_e6_b0.5_n10000.sim", [1.0, 1.92])}


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_other_slab_geometries_identical_to_reference(request, idx):
    from dpm_tablegen import tables_digest
    d = request.getfixturevalue("tables%d" % idx)
    with open(os.path.join(GOLD, "tables_cfg%d.sha256" % idx)) as f:
        assert tables_digest(d) == f.read().strip(), "generated tables differ from the fixtures' tables"
    ptype, e0, lbox, name, dens = SLAB_CASES[idx]
    n = 10000
    r = oracle_lib.run_slab(d, n, e0=e0, ptype=ptype, lbox=lbox, geom_index=idx, rng=oracle_lib.RNG_MT, seed=123,
                            arg_rtl=1)
    txt = oracle_lib.format_hist_sim(r["edep"], r["steps"], r["hist_size"], n, lbox, dens, idx)
    with open(os.path.join(GOLD, name)) as f:
        assert txt == f.read()


def test_argument_order_matters(checked_tables):
    # the other uniform-consumption order must NOT reproduce the reference:
    # the bit-exact match above is a real constraint
    r = oracle_lib.run_slab(checked_tables, 10000, rng=oracle_lib.RNG_MT, seed=123, arg_rtl=0)
    txt = oracle_lib.format_hist_sim(r["edep"], r["steps"], r["hist_size"], 10000, 1.0, DENS_CFG3, 3)
    with open(os.path.join(GOLD, "ref_hist_cfg3_e15_n10000.sim")) as f:
        assert txt != f.read()


@pytest.mark.parametrize("kind", ["sb", "mo", "kn", "gs"])
def test_samplers_bit_exact(checked_tables, kind):
    g = np.load(os.path.join(GOLD, "samplers_cfg3.npz"))
    oracle_lib.load(checked_tables)
    e, k, im, ref = g[kind + "_e"], g[kind + "_k"], g[kind + "_imat"], g[kind + "_ref"]
    out = np.array([oracle_lib.sample(kind, float(e[i]), k[i] / float(1 << 24), int(im[i]))[0] for i in range(e.size)])
    assert np.array_equal(out, ref)


def test_interpolations_bit_exact(checked_tables):
    g = np.load(os.path.join(GOLD, "samplers_cfg3.npz"))
    oracle_lib.load(checked_tables)
    out = np.array([oracle_lib.evaluate(int(w), x, int(m)) for (w, x, m) in g["ev_rows"]])
    assert np.array_equal(out, g["ev_ref"])


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10
    assert oracle_lib.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle_lib.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle_lib.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_depth_dose_sanity_vs_reference_fixture(checked_tables):
    """Physics sanity of the Geant4-free tables against the reference's own
    Geant4-table result (tests/WaterTiBoneWater_15MeV/hist.sim, copied values
    of its first rows): not a parity check, the tables differ."""
    ref_fixture_mean_edep = 13.661
    r = oracle_lib.run_slab(checked_tables, 20000, rng=oracle_lib.RNG_PHILOX, seed=2)
    mean = r["edep"].sum() / 20000
    assert abs(mean - ref_fixture_mean_edep) / ref_fixture_mean_edep < 0.03
