"""GPU parity at the benchmark workloads' own shapes, bit-exact against the
oracle (the CPU restatement of src/decoder/vp9_reconstruct.c:194-370 with the
reference pool rule of src/decoder/vp9_decoder.c:224-231):

  * the 64-stream ragged batch (16 each of 640x360, 854x480, 1280x720 and
    1920x1080, generator preset 3, one seed per stream), a KEY frame and two
    inter frames per stream, as one batch per step and as two interleaved
    batches per step on the engine's two queues;
  * the serving loop over a mix of the same four resolutions (preset 4),
    streams arriving over time under an HBM budget that makes some wait,
    every frame's NV12 output checked;
  * the 3840x2160 sequence (preset 2) in the fused persistent kernel, run
    directly and replayed from a HIP graph, and the 1080p sequence in the
    fused kernel.
"""
import concurrent.futures as cf

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIX = [(640, 360), (854, 480), (1280, 720), (1920, 1080)]


def oracle_all(orc, seqs, nv12=False):
    """Oracle decode of every sequence (each alone, fresh pool), in threads."""
    def one(q):
        out = orc.run_oracle(q)
        if nv12:
            w, h = q[0].pf.hdr.width, q[0].pf.hdr.height
            out = [orc.nv12_pack(img, w, h) for img in out]
        return out
    with cf.ThreadPoolExecutor(max_workers=16) as ex:
        return list(ex.map(one, seqs))


def gen_all(pkg, specs):
    """specs: (preset, w, h, seed, frames) per stream."""
    def one(s):
        preset, w, h, seed, n = s
        p = pkg.synth_params(preset, w, h, seed)
        return [pkg.SynthFrame(p, i) for i in range(n)]
    with cf.ThreadPoolExecutor(max_workers=16) as ex:
        return list(ex.map(one, specs))


def first_diff(got, want, w, h):
    bad = np.nonzero(got != want)[0]
    return f"{bad.size} bytes differ, first at {bad[0]} (luma {w * h})" if bad.size else ""


@pytest.fixture(scope="module")
def mix64(pkg, orc):
    layout = [wh for wh in MIX for _ in range(16)]
    seqs = gen_all(pkg, [(3, w, h, 7000 + sid, 3) for sid, (w, h) in enumerate(layout)])
    assert seqs[0][0].pf.hdr.frame_type == pkg.VP9_FRAME_KEY
    assert all(q[1].pf.hdr.frame_type != pkg.VP9_FRAME_KEY for q in seqs)
    return layout, seqs, oracle_all(orc, seqs)


@pytest.mark.parametrize("groups", [1, 2])
def test_config3_64_stream_mix(pkg, gpu_ready, mix64, groups):
    """The 64-stream workload: each step one frame of every stream, as one
    batch (one launch per stage over all 64 frames) or as two batches of
    interleaved streams, each mixing every resolution, on the two queues."""
    layout, seqs, want = mix64
    print(f"[case] config 3: 64 streams, {groups} batch(es) per step", flush=True)
    with pkg.Engine(0) as eng:
        sids = [eng.open_stream() for _ in layout]
        preps = [[eng.prepare(f) for f in q] for q in seqs]
        for i in range(3):
            for g in range(groups):
                mine = [s for s in range(len(sids)) if s % groups == g]
                eng.run([sids[s] for s in mine], [preps[s][i] for s in mine])
            eng.sync()
            for s, (w, h) in enumerate(layout):
                got = eng.read(sids[s], w, h)
                assert np.array_equal(got, want[s][i]), \
                    f"stream {s} ({w}x{h}) frame {i}: {first_diff(got, want[s][i], w, h)}"


def test_config3_back_to_back_steps(pkg, gpu_ready, mix64):
    """The same steps enqueued with no synchronisation between them (frame
    pipelining overlaps a step's first phase with the previous step's
    second); the last frame of every stream depends on every frame before."""
    layout, seqs, want = mix64
    print("[case] config 3: 64 streams, 3 steps back to back, 2 groups", flush=True)
    with pkg.Engine(0) as eng:
        sids = [eng.open_stream() for _ in layout]
        preps = [[eng.prepare(f) for f in q] for q in seqs]
        for i in range(3):
            for g in range(2):
                mine = [s for s in range(len(sids)) if s % 2 == g]
                eng.run([sids[s] for s in mine], [preps[s][i] for s in mine])
        eng.sync()
        for s, (w, h) in enumerate(layout):
            got = eng.read(sids[s], w, h)
            assert np.array_equal(got, want[s][2]), f"stream {s} ({w}x{h})"


def test_config4_serving_mix(pkg, orc, gpu_ready):
    """A shortened serving run: 16 streams, 4 each of 360p/480p/720p/1080p
    (preset 4), 3-8 frames each, arriving one every other loop turn, frames
    submitted as they "arrive", an HBM budget for about three streams of the
    mix so that later streams wait for earlier ones to close. Each frame's
    NV12 output equals the oracle's decode of its stream alone."""
    rng = np.random.default_rng(404)
    sizes = [MIX[s % 4] for s in range(16)]
    lengths = [int(x) for x in rng.integers(3, 9, 16)]
    seqs = gen_all(pkg, [(4, w, h, 9100 + s, n) for s, ((w, h), n) in enumerate(zip(sizes, lengths))])
    want = oracle_all(orc, seqs, nv12=True)
    cost = [pkg.stream_cost(w, h) for w, h in sizes]
    budget = 3 * max(cost)
    print(f"[case] config 4: 16 streams, budget {budget >> 20} MiB", flush=True)
    with pkg.Engine(0) as eng:
        preps = [[eng.prepare(f) for f in q] for q in seqs]
        with pkg.Server(eng, budget, max_batch=4096, max_inflight=2) as srv:
            handle, waiting = {}, []
            nxt, done = [0] * 16, [0] * 16
            checked, waited, turn, arrived = 0, False, 0, 0
            while checked < sum(lengths):
                turn += 1
                assert turn < 20000, "serving loop made no progress"
                for _ in range(2):  # two arrivals per loop turn, in stream order
                    if arrived < 16:
                        waiting.append(arrived)
                        arrived += 1
                while waiting:
                    h = srv.open(*sizes[waiting[0]])
                    if h is None:
                        waited = True
                        break
                    handle[waiting.pop(0)] = h
                assert sum(cost[s] for s in handle) <= budget
                for s, h in handle.items():
                    if nxt[s] < lengths[s] and rng.random() < 0.7:
                        srv.submit(h, preps[s][nxt[s]], s * 1000 + nxt[s])
                        nxt[s] += 1
                srv.pump()
                for tag, _ in srv.poll():
                    s, f = divmod(int(tag), 1000)
                    w, hh = sizes[s]
                    got = srv.read_nv12(handle[s], f, w, hh)
                    assert np.array_equal(got, want[s][f]), \
                        f"stream {s} ({w}x{hh}) frame {f}: {first_diff(got, want[s][f], w, hh)}"
                    assert f == done[s], "frames of a stream complete in order"
                    done[s] += 1
                    checked += 1
                    if done[s] == lengths[s]:
                        srv.close(handle.pop(s))
            srv.drain()
    assert waited, "the budget held streams back"


def test_config2_fused_and_graph_2160p(pkg, orc, gpu_ready):
    """3840x2160 preset 2 (32x32 transforms at p=0.5, MVs +-256 px): a KEY
    frame and an inter frame in the fused persistent kernel, then the same
    two frames captured as one HIP graph and replayed twice, with eager work
    and synchronisation between the replays; every result bit-exact."""
    W, H = 3840, 2160
    seq = gen_all(pkg, [(2, W, H, 2026, 2)])[0]
    want = orc.run_oracle(seq)
    print("[case] config 2: 2160p fused, then graph replays", flush=True)
    with pkg.Engine(0) as eng:
        eng.set_fused(True)
        sid = eng.open_stream()
        preps = [eng.prepare(f) for f in seq]
        for i, p in enumerate(preps):
            eng.run([sid], [p])
            eng.sync()
            got = eng.read(sid, W, H)
            assert np.array_equal(got, want[i]), f"fused frame {i}: {first_diff(got, want[i], W, H)}"
        eng.capture_begin()
        for p in preps:
            eng.run([sid], [p])
        g = eng.capture_end()
        other = eng.open_stream()
        for rep in range(2):
            eng.graph_launch(g)
            eng.sync()
            got = eng.read(sid, W, H)
            assert np.array_equal(got, want[1]), f"replay {rep}: {first_diff(got, want[1], W, H)}"
            eng.run([other], [preps[0]])  # eager work between the replays
            eng.sync()
        eng.graph_free(g)


def test_config1_fused_1080p(pkg, orc, gpu_ready):
    """1920x1080 preset 1 in the fused persistent kernel: a KEY frame and
    two inter frames, each bit-exact."""
    W, H = 1920, 1080
    seq = gen_all(pkg, [(1, W, H, 2024, 3)])[0]
    want = orc.run_oracle(seq)
    print("[case] config 1: 1080p fused", flush=True)
    with pkg.Engine(0) as eng:
        eng.set_fused(True)
        sid = eng.open_stream()
        for i, f in enumerate(seq):
            eng.run([sid], [eng.prepare(f)])
            eng.sync()
            got = eng.read(sid, W, H)
            assert np.array_equal(got, want[i]), f"frame {i}: {first_diff(got, want[i], W, H)}"


@pytest.mark.parametrize("lanes,workers", [(64, 1024), (64, 300)])
def test_config1_chain_workers_1080p(pkg, orc, gpu_ready, lanes, workers):
    """1920x1080 preset 1 with chain workers (the 256 most critical chains
    of intra blocks each on a wave of its own, the rest dealt by stride),
    frames pipelined back to back: a KEY frame and four inter frames, each
    bit-exact; with 300 workers only 44 deal the rest."""
    W, H = 1920, 1080
    seq = gen_all(pkg, [(1, W, H, 2031, 5)])[0]
    want = orc.run_oracle(seq)
    print(f"[case] config 1: 1080p chain workers, {workers} workers", flush=True)
    with pkg.Engine(0) as eng:
        eng.set_tune([1, 1, 4, 4], lanes, workers, 0, 0, 1)
        sid = eng.open_stream()
        preps = [eng.prepare(f) for f in seq]
        for i in range(len(seq)):
            eng.run([sid], [preps[i]])
        eng.sync()
        assert eng.stream_error(sid) == 0
        got = eng.read(sid, W, H)
        assert np.array_equal(got, want[-1]), first_diff(got, want[-1], W, H)
    with pkg.Engine(0) as eng:
        eng.set_tune([1, 1, 4, 4], lanes, workers, 0, 0, 1)
        sid = eng.open_stream()
        for i, f in enumerate(seq):
            eng.run([sid], [eng.prepare(f)])
            eng.sync()
            got = eng.read(sid, W, H)
            assert np.array_equal(got, want[i]), f"frame {i}: {first_diff(got, want[i], W, H)}"
