"""GPU: the continuous-batching server. Streams of different sizes arrive
over time under an HBM budget small enough that some must wait for others to
close; frames are submitted as they "arrive", batched across streams while
earlier batches are still on the GPU, and every frame's NV12 output equals
the oracle's decode of that stream alone, packed as NV12."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pipeline", [True, False])
def test_server_mix_equals_alone(pkg, orc, gpu_ready, pipeline):
    rng = np.random.default_rng(7)
    sizes = [(64, 64), (136, 72), (200, 120), (96, 64), (136, 72), (64, 64), (200, 120), (96, 64)]
    lengths = [5, 7, 4, 6, 3, 8, 5, 6]
    seqs, want = [], []
    for s, ((w, h), n) in enumerate(zip(sizes, lengths)):
        p = pkg.synth_params(1 + s % 2, w, h, 300 + s)
        fr = [pkg.SynthFrame(p, i) for i in range(n)]
        seqs.append(fr)
        want.append([orc.nv12_pack(img, w, h) for img in orc.run_oracle(fr)])
    budget = 3 * max(pkg.stream_cost(w, h) for w, h in sizes)
    print(f"[case] server mix, budget for 3 streams, pipeline {pipeline}", flush=True)
    with pkg.Engine(0) as eng:
        eng.set_pipeline(pipeline)
        preps = [[eng.prepare(f) for f in fr] for fr in seqs]
        with pkg.Server(eng, budget, max_batch=64, max_inflight=2) as srv:
            handle = {}
            next_frame = [0] * len(sizes)
            waiting = list(range(len(sizes)))
            done = [0] * len(sizes)
            checked = 0
            waited = False
            cost = [pkg.stream_cost(w, h) for w, h in sizes]
            for _ in range(10000):
                # admissions (in arrival order) while the budget allows
                while waiting:
                    h = srv.open(*sizes[waiting[0]])
                    if h is None:
                        waited = True
                        break
                    handle[waiting.pop(0)] = h
                active = [s for s in handle if done[s] < lengths[s]]
                assert sum(cost[s] for s in active) <= budget
                # frames arrive a few at a time, irregularly
                for s, h in handle.items():
                    if next_frame[s] < lengths[s] and rng.random() < 0.6:
                        srv.submit(h, preps[s][next_frame[s]], s * 1000 + next_frame[s])
                        next_frame[s] += 1
                srv.pump()
                for tag, _ in srv.poll():
                    s, f = divmod(int(tag), 1000)
                    w, hh = sizes[s]
                    got = srv.read_nv12(handle[s], f, w, hh)
                    assert np.array_equal(got, want[s][f]), f"stream {s} frame {f}"
                    assert f == done[s], "frames of a stream complete in order"
                    done[s] += 1
                    checked += 1
                    if done[s] == lengths[s]:
                        srv.close(handle[s])
                if checked == sum(lengths):
                    break
            srv.drain()
    assert checked == sum(lengths)
    assert waited   # the budget held streams back until others closed
