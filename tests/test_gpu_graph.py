"""HIP-graph capture of a batched synthesis + analysis: every replay (with
eager work and synchronisations in between) equals the eager result
bitwise."""
import numpy as np
import pytest

from helpers import rand_cilm

pytestmark = pytest.mark.gpu


def test_batched_step_graph_replays(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from shtools_amd.batch import DHBatch
    rng = np.random.default_rng(21)
    Ls = [17, 64, 5, 100, 64, 33]
    batch = DHBatch(Ls, sampling=2, norm=1, csphase=1)
    cilms = [torch.from_numpy(rand_cilm(rng, L)).cuda() for L in Ls]
    grids = [torch.empty(batch.grid_shape(i), dtype=torch.float64, device="cuda") for i in range(len(Ls))]
    outs = [torch.empty(batch.cilm_shape(i), dtype=torch.float64, device="cuda") for i in range(len(Ls))]
    ref_g = [g.clone() for g in batch.makegrid(cilms)]
    ref_c = [c.clone() for c in batch.expand(ref_g)]
    graph = batch.capture(cilms, grids, outs)
    for rep in range(4):
        for g, c in zip(grids, outs):
            g.fill_(np.nan)
            c.fill_(np.nan)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for i in range(len(Ls)):
            assert torch.equal(grids[i], ref_g[i]), (rep, i)
            assert torch.equal(outs[i], ref_c[i]), (rep, i)
        # eager work between replays
        _ = [x * 2.0 for x in cilms]
        torch.cuda.synchronize()


def test_batch_tune_keeps_results():
    """Autotuning changes only workgroup sizes: results stay bitwise equal."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from shtools_amd.batch import DHBatch
    rng = np.random.default_rng(22)
    Ls = [300, 31, 700, 129, 64]
    batch = DHBatch(Ls, sampling=2, norm=2, csphase=-1)
    cilms = [torch.from_numpy(rand_cilm(rng, L)).cuda() for L in Ls]
    ref_g = [g.clone() for g in batch.makegrid(cilms)]
    ref_c = [c.clone() for c in batch.expand(ref_g)]
    grids = [torch.empty_like(g) for g in ref_g]
    outs = [torch.empty_like(c) for c in ref_c]
    t = batch.tune(cilms, grids, outs, reps=2)
    shapes = t["fft_threads"]
    assert shapes and all(x in (256, 384, 512, 768, 1024) for pair in shapes for x in pair)
    assert 1 <= t["synth_kmax"] <= 4 and 1 <= t["anal_kmax"] <= 12 and t["mfma_min"] in (0, 2, 4, 8, 16)
    for i in range(len(Ls)):
        assert torch.equal(grids[i], ref_g[i])
        assert torch.equal(outs[i], ref_c[i])
    g2 = batch.makegrid(cilms)
    for i in range(len(Ls)):
        assert torch.equal(g2[i], ref_g[i])


def test_graph_keeps_its_arrays_after_rebinding():
    """A captured step keeps reading and writing the arrays it was captured
    with after the plan has been called eagerly on other arrays."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from shtools_amd.batch import DHBatch
    rng = np.random.default_rng(23)
    Ls = [40, 7, 200, 129, 64, 300]
    batch = DHBatch(Ls, sampling=2, norm=1, csphase=1)
    cilms = [torch.from_numpy(rand_cilm(rng, L)).cuda() for L in Ls]
    other = [torch.from_numpy(rand_cilm(rng, L)).cuda() for L in Ls]
    ref_g = [g.clone() for g in batch.makegrid(cilms)]
    ref_c = [c.clone() for c in batch.expand(ref_g)]
    grids = [torch.empty_like(g) for g in ref_g]
    outs = [torch.empty_like(c) for c in ref_c]
    graph = batch.capture(cilms, grids, outs)
    for rep in range(3):
        # eager calls on other arrays rebind every slot of the plan
        og = batch.makegrid(other)
        oc = batch.expand(og)
        torch.cuda.synchronize()
        for g, c in zip(grids, outs):
            g.fill_(np.nan)
            c.fill_(np.nan)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for i in range(len(Ls)):
            assert torch.equal(grids[i], ref_g[i]), (rep, i)
            assert torch.equal(outs[i], ref_c[i]), (rep, i)
            assert not torch.equal(og[i], ref_g[i])
        del og, oc


def test_tune_legendre_shapes_mfma_and_cache(tmp_path):
    """The tuner's Legendre choices (synthesis ring pairs per lane, MFMA
    buckets) on a plan with same-degree buckets keep every result bitwise;
    with analysis=True the analysis regrouping changes coefficients by
    rounding only.  A tuning stored in the cache file is applied to a new
    plan of the same degrees without timing it again."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from shtools_amd.batch import DHBatch
    print("case tune legendre", flush=True)
    rng = np.random.default_rng(24)
    Ls = [200] * 12 + [64] * 5 + [500, 31]
    cilms = [torch.from_numpy(rand_cilm(rng, L)).cuda() for L in Ls]
    cache = str(tmp_path / "tune.json")
    batch = DHBatch(Ls, sampling=2)
    ref_g = [g.clone() for g in batch.makegrid(cilms)]
    ref_c = [c.clone() for c in batch.expand(ref_g)]
    grids = [torch.empty_like(g) for g in ref_g]
    outs = [torch.empty_like(c) for c in ref_c]
    t = batch.tune(cilms, grids, outs, reps=1, cache=cache)
    print(f"  tuned {t}", flush=True)
    for i in range(len(Ls)):
        assert torch.equal(grids[i], ref_g[i]), i
        assert torch.equal(outs[i], ref_c[i]), i
    # forcing each choice keeps the results bitwise too
    for sk, mm in ((2, 0), (3, 2), (4, 8), (1, 16)):
        batch.set_tuning(dict(t, synth_kmax=sk, mfma_min=mm))
        g = batch.makegrid(cilms)
        torch.cuda.synchronize()
        for i in range(len(Ls)):
            assert torch.equal(g[i], ref_g[i]), (sk, mm, i)
    # the cache: a new plan of the same degrees takes the stored choice
    other = DHBatch(Ls, sampling=2)
    t2 = other.tune(cilms, grids, outs, reps=1, cache=cache)
    assert t2 == t
    for i in range(len(Ls)):
        assert torch.equal(grids[i], ref_g[i]), i
    # the analysis waves too: coefficients equal to rounding
    ta = other.tune(cilms, grids, outs, reps=1, analysis=True)
    assert ta["anal_kmax"] in (4, 6, 8, 12)
    for i in range(len(Ls)):
        assert torch.equal(grids[i], ref_g[i]), i
        assert float(torch.linalg.norm(outs[i] - ref_c[i]) / torch.linalg.norm(ref_c[i])) <= 1e-14, i
