"""The engine's multi-queue schedule, checked on the CPU.

tests/schedule_model builds the engine's own host code (engine.cpp and
engine_api.cpp, unchanged) against a trace-recording stand-in for the HIP
runtime and the kernel launches (fake_hip.cpp), drives it through
multi-stream schedules (driver.cpp) and checks the happens-before order of
everything it enqueued (check.py): no two unordered nodes touch one
allocation with a write, the host never rewrites a pinned buffer or frees an
allocation that may still be in use, and the persistent intra launches that
can run at once never need more than the device holds resident.

Schedules: two to four stream groups alternating over the two queues (the
four-group form is the one that once stalled on the GPU), groups that change
every step, one batch and several batches alternating, level launches, frame
pipelining on and off, one queue, one-job batches (whose second phase the host
queues without a wait once it has seen the first finish), and a switch to the
fused kernel; under hipEventQuery answers that are always busy, always done, or
pseudo-random.
The engine as built serialises a regrouped batch behind the other queue's
work; the "overlap" build lets it overlap, which is checked as well.

The checker itself is tested by planting mistakes in a recorded trace: waits
taken out, host synchronisations taken out, persistent launches made larger.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL = os.path.join(HERE, "schedule_model")


def plain_env(**extra):
    """The environment for the model's own programs (make, the driver): a
    sanitizer runtime preloaded into the test process is not theirs."""
    env = {k: v for k, v in os.environ.items() if k not in ("LD_PRELOAD", "ASAN_OPTIONS", "UBSAN_OPTIONS")}
    env.update(extra)
    return env


@pytest.fixture(scope="module")
def model():
    subprocess.run(["make", "-s", "-C", MODEL, "all", "overlap"], check=True, env=plain_env())
    import importlib.util
    spec = importlib.util.spec_from_file_location("schedule_check", os.path.join(MODEL, "check.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def trace(tmp_path, binary, query, args):
    out = tmp_path / f"{binary}_{query}_{'_'.join(map(str, args))}.txt"
    env = plain_env(CVP9_MODEL_TRACE=str(out), CVP9_MODEL_QUERY=query)
    r = subprocess.run([os.path.join(MODEL, "_build", binary), *map(str, args)], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return out


# (streams, groups, steps, pattern, pipeline, queues, level_min, fused_step)
SCHEDULES = [
    (8, 2, 6, "alt", 1, 2, 0, -1),
    (8, 3, 7, "alt", 1, 2, 0, -1),
    (8, 4, 7, "alt", 1, 2, 0, -1),
    (16, 4, 9, "alt", 1, 2, 64, -1),
    (8, 3, 7, "rotate", 1, 2, 0, -1),
    (8, 4, 7, "switch", 1, 2, 0, -1),
    (8, 4, 6, "alt", 0, 2, 0, -1),
    (8, 2, 6, "alt", 1, 1, 0, -1),
    (8, 2, 6, "alt", 1, 2, 0, 3),
    (1, 1, 8, "alt", 1, 2, 0, -1),   # one stream: one-job batches, host-side wait for phase A
    (2, 2, 8, "alt", 1, 2, 0, -1),   # one stream per queue
]


def check(model, path):
    rep = model.analyse(str(path))
    detail = "\n".join(rep.races[:5] + rep.host_reuse[:5] + rep.frees[:5])
    assert not rep.races, detail
    assert not rep.host_reuse, detail
    assert not rep.frees, detail
    assert rep.dangling_waits == 0
    assert rep.max_concurrent_share <= 1.0 + 1e-9, rep.max_concurrent
    return rep


@pytest.mark.parametrize("query", ["busy", "done", "random7"])
@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "-".join(map(str, s)))
def test_engine_schedule(model, tmp_path, sched, query):
    rep = check(model, trace(tmp_path, "model_driver", query, sched))
    assert rep.persistent > 0
    if sched[5] == 1:  # one queue: its launches run one after another
        assert rep.max_concurrent_share == pytest.approx(0.5)


@pytest.mark.parametrize("sched", [s for s in SCHEDULES if s[1] >= 3 and s[5] == 2],
                         ids=lambda s: "-".join(map(str, s)))
def test_regrouped_batches_overlapping(model, tmp_path, sched):
    """The form that stalled once: a batch of a new stream set overlaps the
    other queue. Its graph is as clean as the serialised one's."""
    rep = check(model, trace(tmp_path, "model_driver_overlap", "busy", sched))
    if sched[3] == "alt":
        assert rep.max_concurrent_share == pytest.approx(1.0)


def test_four_groups_serialised_when_busy(model, tmp_path):
    """As built, with the other queue busy, four alternating groups never
    run two persistent intra launches at once."""
    rep = check(model, trace(tmp_path, "model_driver", "busy", (8, 4, 7, "alt", 1, 2, 0, -1)))
    assert rep.max_concurrent_share == pytest.approx(0.5)


@pytest.mark.parametrize("sched", [(8, 4, 7, "alt", 1, 2, 0, -1), (1, 1, 8, "alt", 1, 2, 0, -1),
                                   (8, 2, 6, "alt", 1, 2, 0, 3)], ids=lambda s: "-".join(map(str, s)))
def test_engine_host_code_under_asan(model, tmp_path, sched):
    """The same host code under AddressSanitizer and UndefinedBehaviorSanitizer
    (leaks included), through multi-group, one-stream and fused schedules."""
    subprocess.run(["make", "-s", "-C", MODEL, "asan"], check=True, env=plain_env())
    env = plain_env(CVP9_MODEL_TRACE=str(tmp_path / "t.txt"), CVP9_MODEL_QUERY="random3",
                    ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([os.path.join(MODEL, "_build", "model_driver_asan"), *map(str, sched)], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]


def _mutate(src, dst, keep):
    with open(src) as f, open(dst, "w") as g:
        for line in f:
            out = keep(line)
            if out is not None:
                g.write(out)


def test_checker_finds_planted_mistakes(model, tmp_path):
    base = trace(tmp_path, "model_driver", "busy", (8, 2, 6, "alt", 1, 2, 0, -1))
    assert check(model, base).max_concurrent_share == pytest.approx(1.0)

    # every queue wait taken out: phases and queues race
    nowait = tmp_path / "nowait.txt"
    _mutate(base, nowait, lambda l: None if l.startswith("W ") else l)
    assert model.analyse(str(nowait)).races

    # only the first phases' waits for the last reader of their buffer set
    # taken out (one batch per step, where no descriptor-slot wait covers it)
    one = trace(tmp_path, "model_driver", "busy", (8, 1, 8, "alt", 1, 2, 0, -1))
    with open(one) as f:
        first = {l.split()[2] for l in f if l.startswith("K ") and l.split()[3] == "owner"}
    noreader = tmp_path / "noreader.txt"
    _mutate(one, noreader, lambda l: None if l.startswith("W ") and l.split()[2] in first else l)
    assert check(model, one) and model.analyse(str(noreader)).races

    # host synchronisations taken out: descriptor slots and frees unprotected
    nohost = tmp_path / "nohost.txt"
    _mutate(base, nohost, lambda l: None if l.startswith("H ") else l)
    rep = model.analyse(str(nohost))
    assert rep.host_reuse or rep.frees

    # persistent launches a little over half the device: two at once no longer fit
    big = tmp_path / "big.txt"
    _mutate(base, big, lambda l: l.replace(" intra 0.500000 ", " intra 0.600000 "))
    assert model.analyse(str(big)).max_concurrent_share > 1.0
