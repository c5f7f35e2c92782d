"""GPU: several decoder contexts in one process decoding at once (each
drop-in context has an engine of its own, with two queues that may each run
a persistent intra launch): the launches share the device's resident
capacity, every frame of every context stays bit-exact to the oracle, and
no dependency wait gives up."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_three_contexts_interleaved(pkg, orc, gpu_ready):
    seqs = [[pkg.SynthFrame(pkg.synth_params(1, 1920, 1080, 4300 + k), i) for i in range(3)]
            for k in range(3)]
    want = [orc.run_oracle(s) for s in seqs]
    print("[case] 3 contexts x 3 frames 1080p, interleaved", flush=True)
    bes = [pkg.HipBackend() for _ in seqs]
    try:
        got = [[] for _ in seqs]
        for i in range(3):
            for k, be in enumerate(bes):
                rc = be.decode(seqs[k][i], pts=i)
                assert rc == pkg.CVP9_OK, pkg.ERR_NAMES.get(rc, rc)
        for k, be in enumerate(bes):
            for _ in range(3):
                code, img, _ = be.get_frame(wait=True)
                assert code == pkg.CVP9_OK, pkg.ERR_NAMES.get(code, code)
                got[k].append(img)
        for k in range(3):
            for i in range(3):
                assert np.array_equal(got[k][i], want[k][i]), f"context {k} frame {i}"
    finally:
        for be in bes:
            be.close()


def test_engines_side_by_side(pkg, orc, gpu_ready):
    """Four engines, each with a batch of two streams per step on its two
    queues, stepped in turn."""
    specs = [(1, 854, 480), (3, 1280, 720), (1, 640, 360), (2, 642, 362)]
    seqs = [[[pkg.SynthFrame(pkg.synth_params(p, w, h, 4400 + 10 * e + s), i) for i in range(3)]
             for s in range(2)] for e, (p, w, h) in enumerate(specs)]
    print("[case] 4 engines x 2 streams x 3 frames", flush=True)
    engs = [pkg.Engine(0) for _ in specs]
    try:
        sids = [[eng.open_stream() for _ in range(2)] for eng in engs]
        # prepared up front: a prep freed while in use would synchronise its engine
        preps = [[[eng.prepare(f) for f in seqs[e][s]] for s in range(2)]
                 for e, eng in enumerate(engs)]
        for i in range(3):
            for e, eng in enumerate(engs):
                for s in range(2):
                    eng.run([sids[e][s]], [preps[e][s][i]])
        for e, (eng, (p, w, h)) in enumerate(zip(engs, specs)):
            eng.sync()
            for s in range(2):
                assert eng.stream_error(sids[e][s]) == 0
                want = orc.run_oracle(seqs[e][s])[-1]
                assert np.array_equal(eng.read(sids[e][s], w, h), want), f"engine {e} stream {s}"
    finally:
        for eng in engs:
            eng.close()
