"""GPU sampler and interpolation parity (through the C ABI).

Reference values: tests/golden/samplers_cfg3.npz, written by the reference's
own Sim*Tables classes (compiled from its sources) on the configuration-3
tables. Bar (north star): FP32 samplers within 1e-6 relative of the FP64
reference with bit-exact energy-bin and alias-bin selection; the FP64 path
within 1e-12.
"""
import hashlib
import os

import numpy as np
import pytest

import dpm_amd
import oracle_lib

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "samplers_cfg3.npz")
KINDS = ("sb", "mo", "kn", "gs")


@pytest.fixture(scope="module")
def gold(tables3):
    from dpm_tablegen import tables_digest
    with open(os.path.join(HERE, "golden", "tables_cfg3.sha256")) as f:
        assert tables_digest(tables3) == f.read().strip(), "generated tables differ from the fixtures' tables"
    return np.load(GOLD)


def _gpu_sample(t, g, kind, precision):
    e = g[kind + "_e"]
    k = g[kind + "_k"].astype(np.uint32)
    r3 = k[:, 2] if k.shape[1] > 2 else None
    imat = g[kind + "_imat"] if kind == "sb" else None
    return t.sample(kind, e, k[:, 0], k[:, 1], r3, imat=imat, precision=precision)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", KINDS)
def test_fp64_samplers_match_reference(gpu_tables3, gold, kind):
    out, eb, ab = _gpu_sample(gpu_tables3, gold, kind, dpm_amd.DPM_FP64)
    ref = gold[kind + "_ref"]
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-15)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", KINDS)
def test_fp32_samplers_match_reference(gpu_tables3, tables3, gold, kind):
    out, eb, ab = _gpu_sample(gpu_tables3, gold, kind, dpm_amd.DPM_FP32)
    ref = gold[kind + "_ref"]
    e = gold[kind + "_e"]
    k = gold[kind + "_k"]
    im = gold[kind + "_imat"]
    oracle_lib.load(tables3)
    # bin selection identical to the FP64 oracle (which equals the reference bit for bit)
    for i in range(e.size):
        r = k[i] / float(1 << 24)
        v, reb, rab = oracle_lib.sample(kind, float(np.float32(e[i])), r, int(im[i]))
        assert eb[i] == reb and ab[i] == rab, (kind, i, e[i], eb[i], reb, ab[i], rab)
    # FP32 evaluation of E: compare against the FP64 sampler at the same (float) energy
    ref32 = np.array([oracle_lib.sample(kind, float(np.float32(e[i])), k[i] / float(1 << 24), int(im[i]))[0]
                      for i in range(e.size)])
    if kind == "gs":
        # cos(theta) is bounded in [-1, 1]: absolute bound
        assert np.max(np.abs(out - ref32)) <= 1e-6
    else:
        assert np.max(np.abs(out - ref32) / np.abs(ref32)) <= 1e-6
    # and the float energy itself only moves the result within float resolution
    assert np.max(np.abs(out - ref) / np.maximum(np.abs(ref), 1.0)) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("kind", KINDS)
def test_fp32_samplers_random_uniforms(gpu_tables3, tables3, gold, kind):
    """20000 random (energy, uniforms) per sampler over the fixture's energy
    range against the FP64 oracle at the same float energy: bins identical,
    values within 1e-6 (relative; absolute for cos(theta) of GS)."""
    rng = np.random.default_rng(2024 + KINDS.index(kind))
    e_fix = gold[kind + "_e"]
    n = 20000
    e = np.exp(rng.uniform(np.log(e_fix.min()), np.log(e_fix.max()), n)).astype(np.float32).astype(np.float64)
    k = rng.integers(0, 1 << 24, size=(n, 3), dtype=np.uint32)
    imat = rng.integers(0, 3, n).astype(np.int32) if kind == "sb" else None
    out, eb, ab = gpu_tables3.sample(kind, e, k[:, 0], k[:, 1], None if kind == "gs" else k[:, 2], imat=imat,
                                     precision=dpm_amd.DPM_FP32)
    oracle_lib.load(tables3)
    ref = np.empty(n)
    for i in range(n):
        v, reb, rab = oracle_lib.sample(kind, float(e[i]), k[i] / float(1 << 24), int(imat[i]) if kind == "sb" else 0)
        assert eb[i] == reb and ab[i] == rab, (kind, i, e[i], eb[i], reb, ab[i], rab)
        ref[i] = v
    err = np.abs(out - ref) if kind == "gs" else np.abs(out - ref) / np.abs(ref)
    assert err.max() <= 1e-6, (kind, err.max(), int(np.argmax(err)))


@pytest.mark.gpu
def test_eval_matches_reference(gpu_tables3, gold):
    rows = gold["ev_rows"]
    ref = gold["ev_ref"]
    for which in range(9):
        sel = rows[:, 0] == which
        e = rows[sel, 1]
        im = rows[sel, 2].astype(np.int32)
        out64 = gpu_tables3.eval(which, e, imat=im, precision=dpm_amd.DPM_FP64)
        np.testing.assert_allclose(out64, ref[sel], rtol=2e-13, atol=0)
        out32 = gpu_tables3.eval(which, e, imat=im, precision=dpm_amd.DPM_FP32)
        # FP32 tables and arithmetic: the value at the rounded energy within float accuracy
        assert np.max(np.abs(out32 - ref[sel]) / np.abs(ref[sel])) < 2e-5, which


@pytest.mark.gpu
def test_sampler_rejects_bad_uniforms(gpu_tables3):
    with pytest.raises(dpm_amd.DpmError):
        gpu_tables3.sample("kn", np.array([1.0]), np.array([1 << 24]), np.array([0]), np.array([0]))


@pytest.mark.gpu
@pytest.mark.parametrize("precision,bound", [(dpm_amd.DPM_FP32, 1e-6), (dpm_amd.DPM_FP64, 1e-12)])
def test_gs_small_deflections_relative(gpu_tables3, tables3, gold, precision, bound):
    """1 - cos(theta) of the GS sampler, the quantity the MSC hinge applies,
    within a relative bound of the FP64 oracle's 2 a U / (1 - U + a) at the
    same (float) energy -- including the smallest deflections (uniforms in the
    first cumulative bins, high energies), where an absolute bound on cos(theta)
    says nothing."""
    rng = np.random.default_rng(77)
    e_fix = gold["gs_e"]
    n = 20000
    e = np.exp(rng.uniform(np.log(max(e_fix.min(), 1.0)), np.log(e_fix.max()), n)).astype(np.float32).astype(np.float64)
    k1 = rng.integers(0, 1 << 24, n, dtype=np.uint32)
    # half of the draws in the first 1/64 of the cumulative range (smallest angles)
    k2 = np.where(rng.uniform(size=n) < 0.5, rng.integers(0, 1 << 18, n), rng.integers(0, 1 << 24, n)).astype(np.uint32)
    out, eb, ab = gpu_tables3.sample("gs_omu", e, k1, k2, None, precision=precision)
    oracle_lib.load(tables3)
    ref = np.empty(n)
    for i in range(n):
        v, reb, rab = oracle_lib.sample("gs_omu", float(e[i]), (k1[i] / float(1 << 24), k2[i] / float(1 << 24)))
        assert eb[i] == reb and ab[i] == rab, (i, e[i], eb[i], reb, ab[i], rab)
        ref[i] = v
    assert np.all(ref > 0.0)
    rel = np.abs(out - ref) / ref
    print("GS 1-cos(theta): smallest %.3e, largest relative error %.3e" % (ref.min(), rel.max()))
    assert ref.min() < 1e-5          # the small-angle region is reached
    assert rel.max() <= bound, (rel.max(), int(np.argmax(rel)))
