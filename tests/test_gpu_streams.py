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
    at interleaved degrees (the library serialises the calls on the host and
    orders its shared intermediates on the device): every result equals the
    single-threaded one bitwise."""
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
