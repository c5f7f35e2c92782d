"""GPU: the fused persistent kernel, the level-synchronous intra stage,
graph replay, and block lists in arbitrary order with holes (intra
neighbours owned by later blocks or by no block read the initial 128), all
bit-exact against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def shuffled_with_holes(pkg, frame, rng, drop=0.1):
    """Same blocks and coefficients, random list order, some blocks removed."""
    b = frame.blocks()
    keep = rng.random(b.size) >= drop
    b = b[keep][rng.permutation(int(keep.sum()))]
    h = frame.pf.hdr
    return pkg.OwnedFrame(h.width, h.height, b, frame.coeffs(), frame_type=h.frame_type,
                          filter_level=h.filter_level, sharpness=h.sharpness_level)


def set_mode(eng, mode):
    """plain: per-stage kernels with persistent intra workers; fused: one
    persistent kernel; levels: per-stage kernels, one intra launch per
    dependency level (the first 64; persistent workers after); *_defer: the
    same with the blocks' samples written by the parallel pass after the
    ordered stage."""
    eng.set_fused(mode == "fused")
    if mode == "levels":
        eng.set_tune([1, 1, 2, 4], 64, 512, 1)
    elif mode == "levels_defer":
        eng.set_tune([1, 1, 2, 4], 64, 512, 1, 1)
    elif mode == "plain_defer":
        eng.set_tune([1, 1, 2, 4], 64, 512, 0, 1)
    elif mode == "chains":
        eng.set_tune([1, 1, 2, 4], 64, 512, 0, 0, 1)


def run_engine(pkg, frames, mode, w, h):
    with pkg.Engine(0) as eng:
        set_mode(eng, mode)
        sid = eng.open_stream()
        out = []
        for f in frames:
            eng.run([sid], [eng.prepare(f)])
            eng.sync()
            out.append(eng.read(sid, w, h))
        return out


@pytest.mark.parametrize("preset,w,h", [(1, 320, 184), (0, 264, 136), (2, 200, 120)])
def test_fused_bit_exact(pkg, orc, gpu_ready, preset, w, h):
    frames = [pkg.SynthFrame(pkg.synth_params(preset, w, h, 60 + preset), i) for i in range(4)]
    want = orc.run_oracle(frames)
    print(f"[case] fused preset{preset} {w}x{h}", flush=True)
    got = run_engine(pkg, frames, "fused", w, h)
    for i, (g, x) in enumerate(zip(got, want)):
        assert np.array_equal(g, x), f"frame {i}"


@pytest.mark.parametrize("mode", ["plain", "fused", "levels", "plain_defer", "levels_defer",
                                  "chains"])
def test_shuffled_blocks_with_holes(pkg, orc, gpu_ready, mode):
    rng = np.random.default_rng(11)
    w, h = 200, 136
    base = [pkg.SynthFrame(pkg.synth_params(1, w, h, 71), i) for i in range(4)]
    frames = [shuffled_with_holes(pkg, f, rng) for f in base]
    want = orc.run_oracle(frames)
    print(f"[case] shuffled holes mode={mode}", flush=True)
    got = run_engine(pkg, frames, mode, w, h)
    for i, (g, x) in enumerate(zip(got, want)):
        assert np.array_equal(g, x), f"frame {i}"


@pytest.mark.parametrize("mode", ["fused", "levels"])
def test_multi_stream_batch(pkg, orc, gpu_ready, mode):
    sizes = [(136, 72), (320, 200), (64, 64)]
    seqs = [[pkg.SynthFrame(pkg.synth_params(1, w, h, 80 + s), i) for i in range(3)]
            for s, (w, h) in enumerate(sizes)]
    want = [orc.run_oracle(q) for q in seqs]
    print(f"[case] multi-stream mode={mode}", flush=True)
    with pkg.Engine(0) as eng:
        set_mode(eng, mode)
        sids = [eng.open_stream() for _ in sizes]
        preps = [[eng.prepare(f) for f in q] for q in seqs]
        for i in range(3):
            eng.run(sids, [preps[s][i] for s in range(len(sizes))])
            eng.sync()
            for s, (w, h) in enumerate(sizes):
                assert np.array_equal(eng.read(sids[s], w, h), want[s][i]), f"stream {s} frame {i}"


@pytest.mark.parametrize("mode", ["plain", "fused", "levels"])
def test_graph_replay_bit_exact(pkg, orc, gpu_ready, mode):
    sizes = [(264, 136), (136, 72)]
    n = 4
    seqs = [[pkg.SynthFrame(pkg.synth_params(2, w, h, 90 + s), i) for i in range(n)]
            for s, (w, h) in enumerate(sizes)]
    want = [orc.run_oracle(q)[-1] for q in seqs]
    print(f"[case] graph replay mode={mode}", flush=True)
    with pkg.Engine(0) as eng:
        set_mode(eng, mode)
        sids = [eng.open_stream() for _ in sizes]
        preps = [[eng.prepare(f) for f in q] for q in seqs]
        for i in range(n):  # sizes the scratch and the frame pool
            eng.run(sids, [preps[s][i] for s in range(len(sizes))])
        eng.sync()
        eng.capture_begin()
        for i in range(n):
            eng.run(sids, [preps[s][i] for s in range(len(sizes))])
        g = eng.capture_end()
        for rep in range(2):
            eng.graph_launch(g)
            eng.sync()
            for s, (w, h) in enumerate(sizes):
                assert np.array_equal(eng.read(sids[s], w, h), want[s]), f"replay {rep} stream {s}"
        eng.graph_free(g)
