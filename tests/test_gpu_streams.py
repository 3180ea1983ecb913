"""Plain transforms on several streams and under graph capture: the
library's per-device intermediates are reused across calls, so calls on
another stream wait for the previous user on the device, a larger call grows
the buffer in stream order, and captured calls take their own block."""
import numpy as np
import pytest

from helpers import rand_cilm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def test_alternating_streams_and_growth(torch):
    from shtools_amd import backend as B
    print("case streams", flush=True)
    rng = np.random.default_rng(31)
    Ls = [40, 300, 17, 511, 129, 700, 64, 255]
    cs = [torch.from_numpy(rand_cilm(rng, L)).cuda() for L in Ls]
    ref_g = [B.MakeGridDH(c, sampling=2) for c in cs]
    ref_c = [B.SHExpandDH(g, sampling=2) for g in ref_g]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = []
    for rep in range(2):
        for i, c in enumerate(cs):
            s = streams[(i + rep) % 2]
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                g = B.MakeGridDH(c, sampling=2)
                outs.append((i, g, B.SHExpandDH(g, sampling=2)))
            torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    for i, g, c2 in outs:
        assert torch.equal(g, ref_g[i]), i
        assert torch.equal(c2, ref_c[i]), i


def test_captured_plain_calls_replay(torch):
    from shtools_amd import backend as B
    print("case plain capture", flush=True)
    rng = np.random.default_rng(32)
    L = 200
    c = torch.from_numpy(rand_cilm(rng, L)).cuda()
    ref_g = B.MakeGridDH(c, sampling=2)
    ref_c = B.SHExpandDH(ref_g, sampling=2)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        g = B.MakeGridDH(c, sampling=2)
        c2 = B.SHExpandDH(g, sampling=2)
    big = torch.from_numpy(rand_cilm(rng, 900)).cuda()
    for rep in range(3):
        g.fill_(np.nan)
        c2.fill_(np.nan)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(g, ref_g), rep
        assert torch.equal(c2, ref_c), rep
        # eager calls in between, the larger ones growing the shared buffer
        B.SHExpandDH(B.MakeGridDH(big, sampling=2), sampling=2)
        torch.cuda.synchronize()


def test_host_threads_on_their_own_streams(torch):
    """Four host threads, each on its own stream, call the plain transforms
    at interleaved degrees (the library locks only its table lookups, and a
    call finding the shared intermediates held by another thread's call takes
    its own): every result equals the single-threaded one bitwise."""
    import threading
    from shtools_amd import backend as B
    print("case host threads", flush=True)
    rng = np.random.default_rng(33)
    Ls = [31, 255, 64, 600, 5, 129, 400, 90]
    cs = [torch.from_numpy(rand_cilm(rng, L)).cuda() for L in Ls]
    ref_g = [B.MakeGridDH(c, sampling=2) for c in cs]
    ref_c = [B.SHExpandDH(g, sampling=2) for g in ref_g]
    torch.cuda.synchronize()
    outs, errors = [[] for _ in range(4)], []

    def work(k):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for rep in range(3):
                    for j in range(len(Ls)):
                        i = (j + 3 * k + rep) % len(Ls)
                        g = B.MakeGridDH(cs[i], sampling=2)
                        outs[k].append((i, g, B.SHExpandDH(g, sampling=2)))
            s.synchronize()
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    torch.cuda.synchronize()
    for k in range(4):
        assert len(outs[k]) == 3 * len(Ls)
        for i, g, c2 in outs[k]:
            assert torch.equal(g, ref_g[i]), (k, i)
            assert torch.equal(c2, ref_c[i]), (k, i)


def test_host_threads_cold_tables_and_own_plans(torch):
    """Host threads that start on tables nobody has built (a normalisation
    not used before, degrees that make the recursion tables grow while other
    threads' launches are in flight) and threads that each run their own
    batched plan (DH and GLQ): every result equals the one computed
    afterwards on one thread, bitwise."""
    import threading
    from shtools_amd import backend as B
    from shtools_amd.batch import DHBatch, GLQBatch
    print("case host threads cold tables + plans", flush=True)
    rng = np.random.default_rng(34)
    Ls = [5, 700, 64, 1100, 300, 17, 900, 129]
    cs = [torch.from_numpy(rand_cilm(rng, L)).cuda() for L in Ls]
    plan_Ls = [[40, 200, 7], [300, 31, 31, 90], [128, 12, 500]]
    plan_cs = [[torch.from_numpy(rand_cilm(rng, L)).cuda() for L in pl] for pl in plan_Ls]
    plans = [DHBatch(plan_Ls[0], sampling=2, norm=2, csphase=-1),
             GLQBatch(plan_Ls[1], norm=2, csphase=-1), DHBatch(plan_Ls[2], sampling=1, norm=2, csphase=-1)]
    outs, pouts, errors = [[] for _ in range(4)], [None] * 3, []

    def plain(k):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for rep in range(2):
                    for j in range(len(Ls)):
                        i = (j + 2 * k + rep) % len(Ls)
                        g = B.MakeGridDH(cs[i], sampling=2, norm=2, csphase=-1)
                        outs[k].append((i, g, B.SHExpandDH(g, sampling=2, norm=2, csphase=-1)))
            s.synchronize()
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    def batched(k):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                res = []
                for rep in range(3):
                    g = plans[k].makegrid(plan_cs[k])
                    res.append((g, plans[k].expand(g)))
                pouts[k] = res
            s.synchronize()
        except Exception as e:
            errors.append(repr(e))

    threads = [threading.Thread(target=plain, args=(k,)) for k in range(4)]
    threads += [threading.Thread(target=batched, args=(k,)) for k in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    torch.cuda.synchronize()
    ref_g = [B.MakeGridDH(c, sampling=2, norm=2, csphase=-1) for c in cs]
    ref_c = [B.SHExpandDH(g, sampling=2, norm=2, csphase=-1) for g in ref_g]
    torch.cuda.synchronize()
    for k in range(4):
        assert len(outs[k]) == 2 * len(Ls)
        for i, g, c2 in outs[k]:
            assert torch.equal(g, ref_g[i]), (k, i)
            assert torch.equal(c2, ref_c[i]), (k, i)
    for k in range(3):
        g1 = plans[k].makegrid(plan_cs[k])
        c1 = plans[k].expand(g1)
        torch.cuda.synchronize()
        for g, c2 in pouts[k]:
            assert all(torch.equal(a, b) for a, b in zip(g, g1)), k
            assert all(torch.equal(a, b) for a, b in zip(c2, c1)), k


def test_batched_calls_on_a_side_stream_keep_their_arrays(torch):
    """Back-to-back batched calls on one plan with different arrays on a
    non-blocking stream, without any synchronisation in between: each call
    reads and writes the arrays it was given (its pointer table is written in
    stream order, never under work still queued), more calls than the ring of
    tables holds included."""
    from shtools_amd import backend as B
    from shtools_amd.batch import DHBatch
    print("case batched side stream rebinding", flush=True)
    rng = np.random.default_rng(34)
    Ls = [700, 31, 255, 5, 500, 64]
    batch = DHBatch(Ls, sampling=2)
    sets = [[torch.from_numpy(rand_cilm(rng, L)).cuda() for L in Ls] for _ in range(7)]
    refs = [[B.MakeGridDH(c, sampling=2) for c in cs] for cs in sets]
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    grids, backs = [], []
    with torch.cuda.stream(s):
        for cs in sets:
            g = batch.makegrid(cs)          # new output arrays each call
            grids.append(g)
            backs.append(batch.expand(g))
    s.synchronize()
    for k, cs in enumerate(sets):
        for i in range(len(Ls)):
            assert torch.equal(grids[k][i], refs[k][i]), (k, i)
            assert float(torch.linalg.norm(backs[k][i] - cs[i]) / torch.linalg.norm(cs[i])) <= 1e-11, (k, i)


def test_cold_call_under_capture_raises_cleanly(torch):
    """A plain call whose grid tables do not exist yet (degree 337, sampling
    1: no other test uses it) cannot build them under graph capture: it
    raises (status 6) without enqueuing anything; after one eager call at
    that degree the same call captures and replays."""
    from shtools_amd import backend as B
    print("case cold capture", flush=True)
    rng = np.random.default_rng(35)
    B.MakeGridDH(torch.from_numpy(rand_cilm(rng, 20)).cuda(), sampling=2)  # library initialised
    torch.cuda.synchronize()
    L = 337
    c = torch.from_numpy(rand_cilm(rng, L)).cuda()
    graph = torch.cuda.CUDAGraph()
    with pytest.raises(B.SHToolsError, match="graph capture"):
        with torch.cuda.graph(graph):
            B.MakeGridDH(c, sampling=1, norm=2)
    torch.cuda.synchronize()
    ref = B.MakeGridDH(c, sampling=1, norm=2)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        g = B.MakeGridDH(c, sampling=1, norm=2)
    for _ in range(2):
        g.fill_(np.nan)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(g, ref)
