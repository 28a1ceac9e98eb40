"""GPU transport parity (through the C ABI) against the FP64 oracle.

1. FP64 GPU transport vs the oracle in Philox mode (same per-particle streams):
   per-history scored energy agrees to 1e-9 relative for (almost) every
   history -- only transcendental-function last-bit differences between the
   GPU and glibc can make a history take another branch.
2. FP32 GPU transport vs the oracle at equal history count: 1 %/1 mm gamma
   pass rate >= 99 % over bins above 10 % of the maximum (the acceptance bar).
3. Energy conservation per run: scored + escaped = primaries + 2 mc^2
   (annihilations - pairs), to 1e-9 (FP64) / 1e-5 (FP32).
"""
import numpy as np
import pytest

import dpm_amd
import gamma
import oracle_lib

EMC2 = 0.510991


def _energy_balance(nhist, e0, scored, escaped, n_pair, n_annih):
    return (scored + escaped) - (nhist * e0 + 2 * EMC2 * (n_annih - n_pair))


@pytest.mark.gpu
@pytest.mark.parametrize("ptype,e0", [(-1, 15.0), (0, 10.0)])
def test_slab_fp64_history_by_history(gpu_tables3, tables3, ptype, e0):
    n = 2000
    ph = dpm_amd.Phantom.slab(gpu_tables3, 1.0, 3)
    tally, steps, st, he = dpm_amd.simulate(gpu_tables3, ph, n, particle=ptype, energy=e0, seed=11,
                                            precision=dpm_amd.DPM_FP64, per_history=True)
    ref = oracle_lib.run_slab(tables3, n, e0=e0, ptype=ptype, rng=oracle_lib.RNG_PHILOX, seed=11, per_history=True)
    rel = np.abs(he - ref["hist_edep"]) / np.maximum(np.abs(ref["hist_edep"]), 1e-300)
    assert np.mean(rel < 1e-9) >= 0.99, np.mean(rel < 1e-9)
    # depth tallies: sums over the agreeing histories dominate
    nb = min(ph.nbins, 4096)
    np.testing.assert_allclose(tally[:nb].sum(), ref["edep"].sum(), rtol=1e-3)
    bal = _energy_balance(n, e0, st["e_scored"], st["e_escaped"], st["n_pair"], st["n_annih"])
    assert abs(bal) < 1e-9 * n * e0, bal


@pytest.mark.gpu
def test_slab_fp32_gamma_vs_oracle(gpu_tables3, tables3):
    n = 100000
    ph = dpm_amd.Phantom.slab(gpu_tables3, 1.0, 3)
    tally, steps, st, _ = dpm_amd.simulate(gpu_tables3, ph, n, particle=-1, energy=15.0, seed=5)
    ref = oracle_lib.run_slab(tables3, n, e0=15.0, ptype=-1, rng=oracle_lib.RNG_MT, seed=123)
    m = min(len(tally), 105)
    dens = np.array([[1.0, 4.54, 1.85][max(oracle_lib.slab_layer_material(i, 1.0, 3), 0)] for i in range(m)])
    d_gpu = tally[:m] / dens
    d_ref = ref["edep"][:m] / dens
    pr = gamma.gamma_pass_rate_1d(d_ref, d_gpu, spacing_mm=1.0, dd=0.01, dta_mm=1.0, cutoff=0.1)
    assert pr >= 0.99, pr
    bal = _energy_balance(n, 15.0, st["e_scored"], st["e_escaped"], st["n_pair"], st["n_annih"])
    assert abs(bal) < 1e-5 * n * 15.0, bal


@pytest.mark.gpu
@pytest.mark.parametrize("idx,ptype,e0,lbox", [(0, -1, 10.0, 1.0), (1, 0, 12.0, 2.0), (1, -1, 15.0, 1.0),
                                                (2, -1, 6.0, 0.5), (2, 0, 8.0, 2.0)])
def test_other_slab_geometries_fp64_history_by_history(request, idx, ptype, e0, lbox):
    # the reference's pre-defined slabs 0 (water), 1 (water/air/water) and 2
    # (water/bone/water) on the tables of their configurations, at several voxel sizes
    d = request.getfixturevalue("tables%d" % idx)
    t = dpm_amd.Tables(d)
    try:
        n = 2000
        ph = dpm_amd.Phantom.slab(t, lbox, idx)
        tally, _, st, he = dpm_amd.simulate(t, ph, n, particle=ptype, energy=e0, seed=17 + idx,
                                            precision=dpm_amd.DPM_FP64, per_history=True)
        ref = oracle_lib.run_slab(d, n, e0=e0, ptype=ptype, lbox=lbox, geom_index=idx, rng=oracle_lib.RNG_PHILOX,
                                  seed=17 + idx, per_history=True)
        rel = np.abs(he - ref["hist_edep"]) / np.maximum(np.abs(ref["hist_edep"]), 1e-300)
        ok = (rel < 1e-9) | ((np.abs(he) < 1e-300) & (np.abs(ref["hist_edep"]) < 1e-300))
        assert np.mean(ok) >= 0.99, np.mean(ok)
        np.testing.assert_allclose(tally.sum(), ref["edep"].sum(), rtol=1e-3)
        bal = _energy_balance(n, e0, st["e_scored"], st["e_escaped"], st["n_pair"], st["n_annih"])
        assert abs(bal) < 1e-9 * n * e0, bal
        # FP32 depth dose against the oracle's, same histories and streams
        n32 = 50000
        t32, _, _, _ = dpm_amd.simulate(t, ph, n32, particle=ptype, energy=e0, seed=3)
        r32 = oracle_lib.run_slab(d, n32, e0=e0, ptype=ptype, lbox=lbox, geom_index=idx, rng=oracle_lib.RNG_PHILOX,
                                  seed=3)
        m = min(len(t32), len(r32["edep"]))
        pr = gamma.gamma_pass_rate_1d(r32["edep"][:m], t32[:m], spacing_mm=lbox, dd=0.01, dta_mm=1.0, cutoff=0.1)
        assert pr >= 0.99, pr
    finally:
        t.close()


def water_cube(n):
    mat = np.zeros((n, n, n), dtype=np.uint8)
    rho = np.ones((n, n, n), dtype=np.float32)
    return mat, rho


@pytest.mark.gpu
def test_voxel_fp64_history_by_history(gpu_tables0, tables0):
    mat, rho = water_cube(40)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    n = 2000
    tally, _, st, he = dpm_amd.simulate(gpu_tables0, ph, n, particle=0, emin=0.5, emax=20.0, field=60.0, seed=3,
                                        precision=dpm_amd.DPM_FP64, per_history=True)
    ref = oracle_lib.run_voxel(tables0, mat, rho, 2.0, n, ptype=0, emin=0.5, emax=20.0, field=60.0, seed=3,
                               per_history=True)
    rel = np.abs(he - ref["hist_edep"]) / np.maximum(np.abs(ref["hist_edep"]), 1e-300)
    ok = (rel < 1e-9) | ((np.abs(he) < 1e-300) & (np.abs(ref["hist_edep"]) < 1e-300))
    assert np.mean(ok) >= 0.99, np.mean(ok)
    diff = np.abs(tally - ref["edep"]).sum() / ref["edep"].sum()
    assert diff < 0.02, diff
    # the event counters behind the algorithmic-byte model: interactions,
    # secondaries, deposits and steps agree with the oracle's count of the
    # same histories (only the rare histories that took another branch
    # differ). Voxel visits are counted where each side reads a voxel: the
    # GPU's track start in a voxel phantom is a bounds test, not a read.
    g, o = st["counters"], ref["counters"]
    print("GPU counters", g, "\noracle counters", o)
    for k in ("sb", "moller", "kn", "gs", "secondaries", "deposits", "e_steps", "g_steps"):
        assert abs(g[k] - o[k]) <= 0.02 * o[k] + 2, (k, g[k], o[k])


@pytest.mark.gpu
def test_voxel_batches_and_energy_balance(gpu_tables0):
    mat, rho = water_cube(30)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    n = 20000
    tot, sum2, st, _ = dpm_amd.simulate(gpu_tables0, ph, n, particle=0, emin=0.5, emax=20.0, field=40.0, seed=9,
                                        nbatch=10, chunk=4000)
    tot1, _, st1, _ = dpm_amd.simulate(gpu_tables0, ph, n, particle=0, emin=0.5, emax=20.0, field=40.0, seed=9)
    # same histories, same streams: batching only changes the tally layout
    np.testing.assert_allclose(tot.sum(), tot1.sum(), rtol=1e-4)
    assert np.all(sum2 >= 0)
    # variance estimator: sum2 >= sum^2 / nbatch (Cauchy-Schwarz) voxel by voxel
    assert np.all(sum2 + 1e-12 >= tot * tot / 10.0 * (1 - 1e-9))
    # primary energies are drawn per history: the engine reports their sum
    assert 0.5 * n < st1["e_primary"] / 10.25 < 1.5 * n
    bal = _energy_balance(1, st1["e_primary"], st1["e_scored"], st1["e_escaped"], st1["n_pair"], st1["n_annih"])
    assert abs(bal) < 1e-5 * st1["e_primary"], bal
    np.testing.assert_allclose(st["e_primary"], st1["e_primary"], rtol=1e-9)


@pytest.mark.gpu
def test_voxel_fp32_vs_fp64_gamma(gpu_tables0):
    mat, rho = water_cube(50)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    n = 200000
    d32, _, _, _ = dpm_amd.simulate(gpu_tables0, ph, n, particle=0, emin=0.5, emax=20.0, field=50.0, seed=21)
    d64, _, _, _ = dpm_amd.simulate(gpu_tables0, ph, n, particle=0, emin=0.5, emax=20.0, field=50.0, seed=21,
                                    precision=dpm_amd.DPM_FP64)
    pr = gamma.gamma_pass_rate_3d(d64, d32, spacing_mm=2.0, dd=0.01, dta_mm=1.0, cutoff=0.1)
    assert pr >= 0.99, pr


@pytest.mark.gpu
def test_runs_are_reproducible_and_chunk_independent(gpu_tables3):
    ph = dpm_amd.Phantom.slab(gpu_tables3, 1.0, 3)
    a, _, _, ha = dpm_amd.simulate(gpu_tables3, ph, 5000, seed=4, per_history=True)
    b, _, _, hb = dpm_amd.simulate(gpu_tables3, ph, 5000, seed=4, per_history=True, chunk=777)
    # per-history results do not depend on scheduling (per-particle streams);
    # only the order of FP64 tally additions can differ
    np.testing.assert_allclose(ha, hb, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9)


@pytest.mark.gpu
def test_invalid_inputs_raise(gpu_tables3):
    ph = dpm_amd.Phantom.slab(gpu_tables3, 1.0, 3)
    with pytest.raises(dpm_amd.DpmError):
        dpm_amd.simulate(gpu_tables3, ph, 10, energy=50.0)
    with pytest.raises(dpm_amd.DpmError):
        dpm_amd.simulate(gpu_tables3, ph, 10, energy=0.1)
    with pytest.raises(dpm_amd.DpmError):
        dpm_amd.Phantom.voxel(gpu_tables3, np.full((2, 2, 2), 7, np.uint8), np.ones((2, 2, 2), np.float32), 1.0)
    # the limits of a call: history count and ids in 32 bits, batches, pipes, precision
    for bad in (dict(nhist=0), dict(nhist=1 << 32), dict(nhist=10, hist_begin=(1 << 32) - 5),
                dict(nhist=10, nbatch=11), dict(nhist=10, pipes=3), dict(nhist=10, precision=7)):
        n = bad.pop("nhist")
        with pytest.raises(dpm_amd.DpmError):
            dpm_amd.simulate(gpu_tables3, ph, n, energy=15.0, **bad)
    # the largest history id is accepted
    _, _, st, _ = dpm_amd.simulate(gpu_tables3, ph, 10, energy=15.0, hist_begin=(1 << 32) - 10)
    assert st["e_primary"] == pytest.approx(150.0)


@pytest.mark.gpu
def test_field_source_reproducible_and_chunk_independent(gpu_tables0):
    # field sources consume uniforms before transport: the resumed stream
    # must continue where the source left it
    mat, rho = water_cube(20)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    kw = dict(particle=0, emin=0.5, emax=20.0, field=30.0, seed=13, per_history=True)
    _, _, _, h1 = dpm_amd.simulate(gpu_tables0, ph, 3000, **kw)
    _, _, _, h2 = dpm_amd.simulate(gpu_tables0, ph, 3000, chunk=500, **kw)
    np.testing.assert_allclose(h1, h2, rtol=1e-9, atol=1e-9)


@pytest.mark.gpu
def test_beamlets_equal_plain_runs(gpu_tables0):
    mat, rho = water_cube(24)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    bl = [(-6.0, 4.0, -1, 9.0, 3000), (10.0, -2.0, 0, 12.0, 5000), (0.0, 0.0, -1, 15.0, 2000)]
    row_ptr, col, val, st = dpm_amd.simulate_beamlets(gpu_tables0, ph, bl, rel_threshold=1e-3, seed=8,
                                                      precision=dpm_amd.DPM_FP64)
    off = 0
    for b, (x, y, p, e, n) in enumerate(bl):
        d, _, _, _ = dpm_amd.simulate(gpu_tables0, ph, n, particle=p, energy=e, x0=x, y0=y, seed=8, hist_begin=off,
                                      precision=dpm_amd.DPM_FP64)
        off += n
        d = d.ravel()
        keep = np.nonzero((d > 1e-3 * d.max()) & (d > 0))[0]
        cols = col[row_ptr[b]:row_ptr[b + 1]]
        vals = val[row_ptr[b]:row_ptr[b + 1]]
        assert np.array_equal(np.sort(cols), keep)
        np.testing.assert_allclose(vals, d[cols], rtol=1e-9)


@pytest.mark.gpu
def test_beamlet_groups_give_the_single_group_csr(gpu_tables0):
    # a tally budget of 3 dense rows runs 7 beamlets in groups of 3, 3 and 1:
    # the same history ids, so the same CSR as one group (to the order of FP64
    # atomic additions) and the same event counts
    mat, rho = water_cube(24)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    bl = [(-6.0, 4.0, -1, 9.0, 3000), (10.0, -2.0, 0, 12.0, 5000), (0.0, 0.0, -1, 15.0, 2000),
          (4.0, 4.0, 0, 2.0, 4000), (-8.0, -8.0, -1, 6.0, 2500), (2.0, -9.0, 0, 18.0, 1500), (7.0, 7.0, -1, 3.0, 3500)]
    one = dpm_amd.simulate_beamlets(gpu_tables0, ph, bl, rel_threshold=1e-3, seed=12)
    grp = dpm_amd.simulate_beamlets(gpu_tables0, ph, bl, rel_threshold=1e-3, seed=12, tally_budget=3 * 8 * ph.nbins)
    assert one[3]["beamlet_groups"] == 1 and grp[3]["beamlet_groups"] == 3
    np.testing.assert_array_equal(grp[0], one[0])
    np.testing.assert_array_equal(grp[1], one[1])
    np.testing.assert_allclose(grp[2], one[2], rtol=1e-12)
    assert grp[3]["counters"] == one[3]["counters"]
    np.testing.assert_allclose(grp[3]["e_scored"], one[3]["e_scored"], rtol=1e-12)
    np.testing.assert_allclose(grp[3]["e_primary"], one[3]["e_primary"], rtol=1e-12)
    with pytest.raises(dpm_amd.DpmError):
        dpm_amd.simulate_beamlets(gpu_tables0, ph, bl, tally_budget=8 * ph.nbins - 8)
    ph.close()


@pytest.mark.gpu
def test_schedules_do_not_change_results(gpu_tables0):
    # per-particle streams: chunk, waves per CU, refill threshold and tally
    # mode change only the order of FP64 tally additions
    mat, rho = water_cube(30)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    kw = dict(particle=0, emin=0.5, emax=20.0, field=40.0, seed=17, per_history=True)
    d0, _, _, h0 = dpm_amd.simulate(gpu_tables0, ph, 4000, **kw)
    for sched in (dict(chunk=1000, waves_per_cu=8, refill_min=16, tally_mode=1),
                  dict(chunk=0, waves_per_cu=12, refill_min=4, tally_mode=0),
                  dict(chunk=0, waves_per_cu=12, refill_min=4, tally_mode=2),
                  dict(chunk=0, waves_per_cu=16, refill_min=64, tally_mode=1),
                  dict(chunk=0, waves_per_cu=8, refill_min=4, tally_mode=0, pipes=2),
                  dict(chunk=700, waves_per_cu=12, refill_min=1, tally_mode=2, pipes=2)):
        d, _, _, h = dpm_amd.simulate(gpu_tables0, ph, 4000, schedule=sched, **kw)
        np.testing.assert_allclose(h, h0, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(d, d0, rtol=1e-9, atol=1e-12)


@pytest.mark.gpu
def test_autotune_returns_a_timed_candidate(gpu_tables0):
    mat, rho = water_cube(30)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    s = dpm_amd.autotune(gpu_tables0, ph, nsample=20000, particle=0, emin=0.5, emax=20.0, field=40.0)
    assert s["waves_per_cu"] in (8, 12, 16) and s["refill_min"] in (1, 4, 16) and s["tally_mode"] in (0, 1, 2)
    # 54 schedules with the automatic chunk plan on one and on two pipes; a
    # sample below 2^21 histories leaves no chunk size to try (chunk 0)
    assert s["candidates"] == 54 and s["chunk"] == 0 and s["pipes"] in (1, 2)
    # a sample of 4.2e6 histories: chunks of 2^21 and 2^22 histories are
    # timed at the best schedule, and the one returned runs
    n4 = 4_200_000
    s4 = dpm_amd.autotune(gpu_tables0, ph, nsample=n4, particle=0, emin=0.5, emax=20.0, field=40.0, nbatch=4)
    assert s4["candidates"] == 56 and s4["chunk"] in (0, 1 << 21, 1 << 22)
    d4, _, st4, _ = dpm_amd.simulate(gpu_tables0, ph, n4, particle=0, emin=0.5, emax=20.0, field=40.0, nbatch=4,
                                     schedule=s4)
    assert d4.sum() > 0 and st4["e_primary"] > 0
    assert s["histories_per_s"] >= s["default_histories_per_s"] > 0
    bl = [(-6.0, 4.0, -1, 9.0, 3000), (10.0, -2.0, 0, 12.0, 5000)]
    sb = dpm_amd.autotune_beamlets(gpu_tables0, ph, bl, nsample=5000)
    assert sb["candidates"] == 54 and sb["histories_per_s"] > 0
    with pytest.raises(dpm_amd.DpmError):
        dpm_amd.autotune(gpu_tables0, ph, nsample=10, particle=0, emin=0.5, emax=20.0)
    # with the batch rows and folds of a 10-batch run
    s10 = dpm_amd.autotune(gpu_tables0, ph, nsample=20000, particle=0, emin=0.5, emax=20.0, field=40.0, nbatch=10)
    assert s10["candidates"] == 54 and s10["histories_per_s"] >= s10["default_histories_per_s"] > 0


@pytest.mark.gpu
def test_two_pipes_equal_one_pipe(gpu_tables0):
    # two pipes: each keeps its own batch rows and folds them one at a time
    # into the shared sum and sum of squares; chunks alternate between them
    mat, rho = water_cube(30)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    kw = dict(particle=0, emin=0.5, emax=20.0, field=40.0, seed=41, per_history=True, nbatch=10)
    d1, s1, st1, h1 = dpm_amd.simulate(gpu_tables0, ph, 20000, **kw)
    for extra in (dict(pipes=2), dict(pipes=2, chunk=3000), dict(pipes=2, chunk=3000, graph=True, rounds_per_sync=2)):
        d2, s2, st2, h2 = dpm_amd.simulate(gpu_tables0, ph, 20000, **kw, **extra)
        np.testing.assert_allclose(h2, h1, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(d2, d1, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(s2, s1, rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(st2["e_primary"], st1["e_primary"], rtol=1e-12)
        assert st2["counters"]["e_steps"] == st1["counters"]["e_steps"]
    # FP64 transport, and a plain (unbatched) tally both pipes add into
    kw64 = dict(particle=-1, energy=12.0, field=40.0, seed=5, per_history=True, precision=dpm_amd.DPM_FP64)
    a1, _, _, g1 = dpm_amd.simulate(gpu_tables0, ph, 4000, **kw64)
    a2, _, _, g2 = dpm_amd.simulate(gpu_tables0, ph, 4000, pipes=2, chunk=900, **kw64)
    np.testing.assert_allclose(g2, g1, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(a2, a1, rtol=1e-12, atol=1e-12)
    # beamlets: rows chosen per history, shared by both pipes
    bl = [(-6.0, 4.0, -1, 9.0, 3000), (10.0, -2.0, 0, 12.0, 5000), (0.0, 0.0, 1, 15.0, 2000)]
    ra, ca, va, _ = dpm_amd.simulate_beamlets(gpu_tables0, ph, bl, seed=8)
    rb, cb, vb, _ = dpm_amd.simulate_beamlets(gpu_tables0, ph, bl, seed=8, pipes=2)
    assert np.array_equal(ra, rb) and np.array_equal(ca, cb)
    np.testing.assert_allclose(va, vb, rtol=1e-9)
    with pytest.raises(dpm_amd.DpmError):
        dpm_amd.simulate(gpu_tables0, ph, 100, particle=0, energy=5.0, pipes=3)


@pytest.mark.gpu
def test_graph_replays_match_eager(gpu_tables0):
    # 2 rounds per captured block and 3 chunks: the drain graph is replayed
    # several times per call; eager runs in between, every replay checked
    mat, rho = water_cube(30)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    kw = dict(particle=0, emin=0.5, emax=20.0, field=40.0, seed=23, per_history=True, chunk=2000)
    d0, _, st0, h0 = dpm_amd.simulate(gpu_tables0, ph, 6000, **kw)
    assert st0["graphs"] == 0
    for rep in range(3):
        d, _, st, h = dpm_amd.simulate(gpu_tables0, ph, 6000, graph=True, rounds_per_sync=2, **kw)
        assert st["graphs"] >= 4 and st["rounds"] >= 6, st
        np.testing.assert_allclose(h, h0, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(d, d0, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(st["e_primary"], st0["e_primary"], rtol=1e-12)
        dpm_amd.simulate(gpu_tables0, ph, 500, particle=-1, energy=12.0, seed=rep)   # eager work in between
    with pytest.raises(dpm_amd.DpmError):
        dpm_amd.simulate(gpu_tables0, ph, 100, particle=0, energy=5.0, rounds_per_sync=3)
    bl = [(-6.0, 4.0, -1, 9.0, 3000), (10.0, -2.0, 0, 12.0, 5000), (0.0, 0.0, 1, 15.0, 2000)]
    ra, ca, va, _ = dpm_amd.simulate_beamlets(gpu_tables0, ph, bl, seed=8, graph=False)
    for rep in range(2):
        rb, cb, vb, sb = dpm_amd.simulate_beamlets(gpu_tables0, ph, bl, seed=8, graph=True)
        assert sb["graphs"] >= 1
        assert np.array_equal(ra, rb) and np.array_equal(ca, cb)
        np.testing.assert_allclose(va, vb, rtol=1e-9)


@pytest.mark.gpu
def test_result_left_on_device_then_fetched(gpu_tables0):
    mat, rho = water_cube(20)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    kw = dict(particle=0, emin=0.5, emax=20.0, field=30.0, seed=31, nbatch=10)
    d0, s0, st0, _ = dpm_amd.simulate(gpu_tables0, ph, 3000, **kw)
    none, none2, st1, _ = dpm_amd.simulate(gpu_tables0, ph, 3000, fetch=False, **kw)
    assert none is None and none2 is None
    d1, s1 = dpm_amd.tally_fetch(gpu_tables0, ph)
    np.testing.assert_allclose(d1, d0, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(s1, s0, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(st1["e_scored"], d1.sum(), rtol=1e-12)
    # a beamlet run reuses the tally buffers: nothing left to fetch
    dpm_amd.simulate_beamlets(gpu_tables0, ph, [(0.0, 0.0, -1, 9.0, 500)], seed=2)
    with pytest.raises(dpm_amd.DpmError):
        dpm_amd.tally_fetch(gpu_tables0, ph)


def small_heterogeneous(shape=(24, 20, 16), seed=0):
    rng = np.random.default_rng(seed)
    mat = rng.integers(0, 3, size=shape).astype(np.uint8)
    rho = (np.array([1.0, 4.54, 1.85], dtype=np.float32)[mat] *
           rng.uniform(0.9, 1.1, mat.shape).astype(np.float32))
    return mat, rho


@pytest.mark.gpu
@pytest.mark.parametrize("ptype,e0", [(-1, 12.0), (0, 8.0)])
def test_heterogeneous_fp64_history_by_history(gpu_tables3, tables3, ptype, e0):
    # water / Ti / bone voxels with density jitter (majorant scaled above 1)
    mat, rho = small_heterogeneous()
    ph = dpm_amd.Phantom.voxel(gpu_tables3, mat, rho, 2.0)
    scale = float(np.max(rho / np.array([1.0, 4.54, 1.85])[mat]))
    n = 1500
    tally, _, st, he = dpm_amd.simulate(gpu_tables3, ph, n, particle=ptype, energy=e0, field=20.0, seed=41,
                                        precision=dpm_amd.DPM_FP64, per_history=True)
    ref = oracle_lib.run_voxel(tables3, mat, rho, 2.0, n, ptype=ptype, e0=e0, field=20.0, seed=41,
                               per_history=True, majorant_scale=scale)
    rel = np.abs(he - ref["hist_edep"]) / np.maximum(np.abs(ref["hist_edep"]), 1e-300)
    ok = (rel < 1e-9) | ((np.abs(he) < 1e-300) & (np.abs(ref["hist_edep"]) < 1e-300))
    assert np.mean(ok) >= 0.99, np.mean(ok)
    bal = _energy_balance(n, e0, st["e_scored"], st["e_escaped"], st["n_pair"], st["n_annih"])
    assert abs(bal) < 1e-9 * n * e0, bal


@pytest.mark.gpu
def test_config3_recipe_fp32_vs_fp64_gamma(gpu_tables3):
    # the heterogeneous recipe (bone and Ti ellipsoids, lognormal density
    # jitter) at reduced size, 12 MeV e- 10x10 cm2 field
    import phantoms
    mat, rho = phantoms.heterogeneous(64, 64, 48, 2.0, seed=1234, axis_scale=0.5)
    ph = dpm_amd.Phantom.voxel(gpu_tables3, mat, rho, 2.0)
    n = 1000000
    kw = dict(particle=-1, energy=12.0, field=100.0, seed=43, nbatch=10)
    d32, s32, st32, _ = dpm_amd.simulate(gpu_tables3, ph, n, **kw)
    d64, _, st64, _ = dpm_amd.simulate(gpu_tables3, ph, n, precision=dpm_amd.DPM_FP64, **kw)
    dose32, dose64 = d32 / rho, d64 / rho
    pr = gamma.gamma_pass_rate_3d(dose64, dose32, spacing_mm=2.0, dd=0.01, dta_mm=1.0, cutoff=0.1)
    assert pr >= 0.99, pr
    bal = _energy_balance(n, 12.0, st32["e_scored"], st32["e_escaped"], st32["n_pair"], st32["n_annih"])
    assert abs(bal) < 1e-5 * n * 12.0, bal


@pytest.mark.gpu
def test_gantry_field_fp64_history_by_history(gpu_tables3, tables3):
    # 30-degree photon field, triangular (0.1, 2, 18) MeV spectrum, heterogeneous voxels
    mat, rho = small_heterogeneous((20, 22, 24), seed=2)
    ph = dpm_amd.Phantom.voxel(gpu_tables3, mat, rho, 2.0)
    scale = float(np.max(rho / np.array([1.0, 4.54, 1.85])[mat]))
    n = 2000
    kw = dict(emin=0.1, emax=18.0, field=20.0, seed=47)
    tally, _, st, he = dpm_amd.simulate(gpu_tables3, ph, n, particle=0, angle=30.0, emode=2.0,
                                        precision=dpm_amd.DPM_FP64, per_history=True, **kw)
    ref = oracle_lib.run_voxel(tables3, mat, rho, 2.0, n, ptype=0, angle=30.0, emode=2.0, per_history=True,
                               majorant_scale=scale, **kw)
    rel = np.abs(he - ref["hist_edep"]) / np.maximum(np.abs(ref["hist_edep"]), 1e-300)
    ok = (rel < 1e-9) | ((np.abs(he) < 1e-300) & (np.abs(ref["hist_edep"]) < 1e-300))
    assert np.mean(ok) >= 0.99, np.mean(ok)
    np.testing.assert_allclose(st["e_primary"], ref["e_primary"], rtol=1e-12)
    bal = _energy_balance(1, st["e_primary"], st["e_scored"], st["e_escaped"], st["n_pair"], st["n_annih"])
    assert abs(bal) < 1e-9 * st["e_primary"], bal


@pytest.mark.gpu
def test_fields_run_equals_the_sum_of_single_fields(gpu_tables0):
    mat, rho = water_cube(24)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    fields = [(0.0, 0.0, 0.0, 1500), (5.0, -3.0, 100.0, 2500), (-4.0, 2.0, 250.0, 1000)]
    kw = dict(particle=0, emin=0.1, emax=18.0, emode=2.0, field=30.0, seed=51, precision=dpm_amd.DPM_FP64)
    d, _, st = dpm_amd.simulate_fields(gpu_tables0, ph, fields, **kw)
    total, off = np.zeros_like(d), 0
    for x0, y0, a, n in fields:
        kw1 = dict(kw)
        kw1.pop("emode")
        di, _, _, _ = dpm_amd.simulate(gpu_tables0, ph, n, x0=x0, y0=y0, angle=a, emode=2.0, hist_begin=off, **kw1)
        total += di
        off += n
    np.testing.assert_allclose(d, total, rtol=1e-9, atol=1e-12)
    # batch statistics over a multi-field run
    d8, s8, _ = dpm_amd.simulate_fields(gpu_tables0, ph, fields, nbatch=8, **kw)
    np.testing.assert_allclose(d8, d, rtol=1e-9, atol=1e-12)
    assert np.all(s8 + 1e-12 >= d8 * d8 / 8 * (1 - 1e-9))


@pytest.mark.gpu
def test_gantry_rays_missing_the_phantom(gpu_tables0):
    mat, rho = water_cube(10)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    d, _, st, _ = dpm_amd.simulate(gpu_tables0, ph, 500, particle=0, energy=5.0, x0=50.0, angle=45.0, seed=1)
    assert d.sum() == 0.0 and abs(st["e_escaped"] - 2500.0) < 1e-6
    with pytest.raises(dpm_amd.DpmError):
        dpm_amd.simulate(gpu_tables0, dpm_amd.Phantom.slab(gpu_tables0, 1.0, 0), 10, particle=0, energy=5.0,
                         angle=10.0)


@pytest.mark.gpu
def test_config5_recipe_fp32_vs_fp64_gamma(gpu_tables3):
    # configuration 5 at reduced size: heterogeneous phantom, 36 gantry angles
    # with uneven history counts, triangular (0.1, 2, 18) MeV photon fields
    import phantoms
    mat, rho = phantoms.heterogeneous(80, 80, 64, 2.0, seed=99, axis_scale=0.6)
    ph = dpm_amd.Phantom.voxel(gpu_tables3, mat, rho, 2.0)
    fields = phantoms.config5_fields(total=1.8e6, lo=3e4, hi=9e4)
    kw = dict(particle=0, emin=0.1, emax=18.0, emode=2.0, field=100.0, seed=53, nbatch=10)
    d32, _, st32 = dpm_amd.simulate_fields(gpu_tables3, ph, fields, **kw)
    d64, _, _ = dpm_amd.simulate_fields(gpu_tables3, ph, fields, precision=dpm_amd.DPM_FP64, **kw)
    pr = gamma.gamma_pass_rate_3d(d64 / rho, d32 / rho, spacing_mm=2.0, dd=0.01, dta_mm=1.0, cutoff=0.1)
    assert pr >= 0.99, pr
    bal = _energy_balance(1, st32["e_primary"], st32["e_scored"], st32["e_escaped"], st32["n_pair"],
                          st32["n_annih"])
    assert abs(bal) < 1e-5 * st32["e_primary"], bal


@pytest.mark.gpu
@pytest.mark.parametrize("hetero,particle", [(False, -1), (False, 0), (True, -1), (True, 0)])
def test_lds_tables_match_device_tables(gpu_tables0, gpu_tables3, hetero, particle):
    # FP32 electron splines and photon IMFP tables read from the block's LDS
    # copy or from the device pack: the same arithmetic on the same values, so
    # every history scores the same energy (up to the order of its FP64 tally
    # additions). Both kernels use LDS tables for water alone (the photon
    # kernel's small copy) and for three materials (its large copy).
    if hetero:
        t = gpu_tables3
        mat, rho = small_heterogeneous()
    else:
        t = gpu_tables0
        mat, rho = water_cube(30)
    ph = dpm_amd.Phantom.voxel(t, mat, rho, 2.0)
    kw = dict(particle=particle, energy=12.0, field=20.0, seed=23, per_history=True, nbatch=4)
    d_lds, s_lds, st_lds, h_lds = dpm_amd.simulate(t, ph, 3000, table_memory=0, **kw)
    d_dev, s_dev, st_dev, h_dev = dpm_amd.simulate(t, ph, 3000, table_memory=1, **kw)
    assert st_lds["lds_tables"] == 3 and st_dev["lds_tables"] == 0
    np.testing.assert_allclose(h_lds, h_dev, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(d_lds, d_dev, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(s_lds, s_dev, rtol=1e-12, atol=1e-12)
    assert st_lds["counters"] == st_dev["counters"]
    # the FP64 path never uses the LDS copy
    _, _, st64, _ = dpm_amd.simulate(t, ph, 200, precision=dpm_amd.DPM_FP64, **kw)
    assert st64["lds_tables"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("pencil", [True, False])
def test_hot_voxel_lds_tally_equals_global_atomics(gpu_tables0, pencil):
    # tally mode 2 (per-wave LDS table of hot cells) only regroups the FP64
    # additions of mode 0; a pencil beam keeps the table hot (hits, evictions,
    # lanes of one wave racing for one slot), a broad field mostly misses it
    mat, rho = water_cube(40)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    kw = dict(particle=-1, energy=15.0, field=0.0 if pencil else 60.0, seed=29, nbatch=5, per_history=True)
    n = 20000
    d0, q0, st0, h0 = dpm_amd.simulate(gpu_tables0, ph, n, tally_mode=0, **kw)
    d2, q2, st2, h2 = dpm_amd.simulate(gpu_tables0, ph, n, tally_mode=2, **kw)
    np.testing.assert_allclose(h2, h0, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(d2, d0, rtol=1e-11, atol=1e-12 * d0.max())
    np.testing.assert_allclose(q2, q0, rtol=1e-11, atol=1e-12 * q0.max())
    assert abs(st2["e_scored"] - st0["e_scored"]) < 1e-10 * st0["e_scored"]
    c0, c2 = st0["counters_electron"], st2["counters_electron"]
    assert c2["deposits"] == c0["deposits"]
    # the table absorbs part of a pencil beam's write-combined sums
    print("tally atomics, mode 2 / mode 0: %.3f" % (c2["tally_atomics"] / c0["tally_atomics"]), flush=True)
    if pencil:
        assert c2["tally_atomics"] < c0["tally_atomics"], (c2["tally_atomics"], c0["tally_atomics"])


@pytest.mark.gpu
def test_atomic_probe_counts_every_addition(gpu_tables0):
    # the measurement helper behind the bench's atomic-throughput figure: every
    # one of its additions lands (timed pass + warm-up pass of <= 4096 per CU)
    ncells, nops = 1 << 20, 1 << 24
    sec, total = dpm_amd.probe_atomics(gpu_tables0, ncells, nops)
    assert sec > 0
    assert nops < total <= nops + 4096 * 1024 and total == int(total)
    with pytest.raises(dpm_amd.DpmError):
        dpm_amd.probe_atomics(gpu_tables0, 1 << 32, 10)


@pytest.mark.gpu
def test_batch_rows_reused_clean_across_runs(gpu_tables0):
    # a batch run leaves its rows cleared by the folds and the next batch run
    # skips clearing them; runs of other kinds in between invalidate that
    mat, rho = water_cube(30)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    y = dict(particle=0, emin=0.5, emax=20.0, field=40.0, seed=31, nbatch=8)
    d1, q1, _, _ = dpm_amd.simulate(gpu_tables0, ph, 20000, **y)
    for other in (dict(particle=-1, energy=9.0, seed=5, nbatch=5),      # batch run of another source
                  dict(particle=-1, energy=9.0, seed=6),                # single-row run: tally left in row 0
                  None):                                                # beamlets
        if other is None:
            dpm_amd.simulate_beamlets(gpu_tables0, ph, [(0.0, 0.0, -1, 9.0, 3000), (8.0, 2.0, 0, 5.0, 3000)])
        else:
            dpm_amd.simulate(gpu_tables0, ph, 7000, **other)
        d2, q2, _, _ = dpm_amd.simulate(gpu_tables0, ph, 20000, **y)
        np.testing.assert_allclose(d2, d1, rtol=1e-11, atol=1e-12 * d1.max())
        np.testing.assert_allclose(q2, q1, rtol=1e-11, atol=1e-12 * q1.max())
        d3, q3, _, _ = dpm_amd.simulate(gpu_tables0, ph, 20000, **y)   # straight after a batch run
        np.testing.assert_allclose(d3, d1, rtol=1e-11, atol=1e-12 * d1.max())
        np.testing.assert_allclose(q3, q1, rtol=1e-11, atol=1e-12 * q1.max())


@pytest.mark.gpu
def test_table_spectrum_source(gpu_tables0, tables0):
    # a spectrum file (histogram bins of relative weight, one with weight 0):
    # FP64 GPU transport follows the oracle history by history, and the
    # primaries' mean energy is the histogram's mean
    mat, rho = water_cube(30)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    edges = np.array([0.5, 1.0, 2.0, 4.0, 6.0, 15.0])
    weights = np.array([1.0, 3.0, 0.0, 2.5, 0.5])
    n = 2000
    kw = dict(particle=0, field=40.0, seed=17, spectrum_table=(edges, weights))
    _, _, st, he = dpm_amd.simulate(gpu_tables0, ph, n, precision=dpm_amd.DPM_FP64, per_history=True, **kw)
    ref = oracle_lib.run_voxel(tables0, mat, rho, 2.0, n, ptype=0, field=40.0, seed=17, per_history=True,
                               spectrum_table=(edges, weights))
    rel = np.abs(he - ref["hist_edep"]) / np.maximum(np.abs(ref["hist_edep"]), 1e-300)
    ok = (rel < 1e-9) | ((np.abs(he) < 1e-300) & (np.abs(ref["hist_edep"]) < 1e-300))
    assert np.mean(ok) >= 0.99, np.mean(ok)
    np.testing.assert_allclose(st["e_primary"], ref["e_primary"], rtol=1e-12)
    # mean primary energy over many histories (FP32) against the histogram's mean
    m = 200000
    _, _, st32, _ = dpm_amd.simulate(gpu_tables0, ph, m, **kw)
    p = weights / weights.sum()
    mean = np.sum(p * 0.5 * (edges[1:] + edges[:-1]))
    sd = np.sqrt(np.sum(p * (edges[1:] ** 2 + edges[1:] * edges[:-1] + edges[:-1] ** 2) / 3) - mean ** 2)
    assert abs(st32["e_primary"] / m - mean) < 5 * sd / np.sqrt(m), (st32["e_primary"] / m, mean)
    for bad in ((edges, -weights), (edges[::-1].copy(), weights), (edges, 0 * weights),
                (np.array([0.01, 1.0]), np.array([1.0]))):     # below the photon cut
        with pytest.raises(dpm_amd.DpmError):
            dpm_amd.simulate(gpu_tables0, ph, 10, particle=0, field=40.0, spectrum_table=bad)
