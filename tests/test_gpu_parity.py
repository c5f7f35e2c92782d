"""GPU parity: the HIP backend against the oracle, bit-exact per frame and plane.

Every case decodes a frame sequence through the C-ABI (hip_decode_frame /
hip_get_frame) and compares each packed I420 frame with the CPU restatement
of the reference path (oracle/cvp9_oracle.c), which keeps the reference pool
rule of vp9_decoder.c:224-231.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def decode_all(P, frames):
    out = []
    with P.HipBackend() as be:
        for i, f in enumerate(frames):
            rc = be.decode(f, pts=i)
            if rc == P.CVP9_ERR_AGAIN:
                code, img, _ = be.get_frame(wait=True)
                assert code == P.CVP9_OK
                out.append(img)
                rc = be.decode(f, pts=i)
            assert rc == P.CVP9_OK, P.ERR_NAMES.get(rc, rc)
        while True:
            code, img, _ = be.get_frame(wait=True)
            if code == P.CVP9_ERR_UNSUPPORTED:
                break
            assert code == P.CVP9_OK, P.ERR_NAMES.get(code, code)
            out.append(img)
    return out


def compare(P, orc, frames, tag):
    print(f"[case] {tag}", flush=True)
    want = orc.run_oracle(frames)
    got = decode_all(P, frames)
    assert len(got) == len(want)
    for i, (g, w) in enumerate(zip(got, want)):
        if not np.array_equal(g, w):
            W, H = frames[i].pf.hdr.width, frames[i].pf.hdr.height
            ny = W * H
            bad = np.nonzero(g != w)[0]
            plane = ["Y" if b < ny else "UV" for b in bad[:5]]
            first = bad[0]
            where = (first // W, first % W) if first < ny else ("uv", first - ny)
            pytest.fail(f"{tag}: frame {i} differs at {len(bad)} bytes, first {where} "
                        f"planes {plane} got {g[first]} want {w[first]}")


SMALL = [
    (0, 136, 72, 5), (1, 136, 72, 5), (2, 136, 72, 4),
    (1, 64, 64, 3), (1, 320, 200, 4), (1, 854, 480, 3), (2, 642, 362, 3),
]


@pytest.mark.parametrize("preset,w,h,n", SMALL)
def test_small_sequences(pkg, orc, gpu_ready, preset, w, h, n):
    p = pkg.synth_params(preset, w, h, 1000 + preset * 7 + w)
    frames = [pkg.SynthFrame(p, i) for i in range(n)]
    compare(pkg, orc, frames, f"preset{preset} {w}x{h} n={n}")


# Frames smaller than a block: chroma planes narrower than the 5 samples one
# motion-compensation quad reads (its source window is clamped into the row),
# planes one sample high, widths that are not multiples of 4, and motion
# vectors that point far outside such planes (preset 2: +-256 px).
TINY = [(2, 8, 8, 4), (0, 6, 10, 4), (1, 18, 2, 3), (2, 2, 2, 3), (1, 10, 6, 4), (2, 26, 14, 4)]


@pytest.mark.parametrize("preset,w,h,n", TINY)
def test_tiny_frames(pkg, orc, gpu_ready, preset, w, h, n):
    p = pkg.synth_params(preset, w, h, 3000 + preset * 7 + w)
    frames = [pkg.SynthFrame(p, i) for i in range(n)]
    compare(pkg, orc, frames, f"tiny preset{preset} {w}x{h} n={n}")


def _random_shapes(count, seed=20261203):
    """Frame sizes drawn from a fixed generator: 2..300 x 2..200, none 1 mod 8
    (the host validator sends those to the CPU path), with a preset, a seed
    and a KEY-frame interval each."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        w, h = int(rng.integers(2, 301)), int(rng.integers(2, 201))
        if w % 8 == 1 or h % 8 == 1:
            continue
        out.append((int(rng.integers(0, 3)), w, h, int(rng.integers(1, 1 << 20)),
                    int(rng.choice([2, 3, 1000]))))
    return out


@pytest.mark.parametrize("preset,w,h,seed,key", _random_shapes(12))
def test_random_sizes(pkg, orc, gpu_ready, preset, w, h, seed, key):
    p = pkg.synth_params(preset, w, h, seed, key_interval=key)
    frames = [pkg.SynthFrame(p, i) for i in range(4)]
    compare(pkg, orc, frames, f"random preset{preset} {w}x{h} seed{seed} key/{key}")


def test_skip_blocks_and_sharpness(pkg, orc, gpu_ready):
    for sharp in (0, 3, 6):
        p = pkg.synth_params(1, 200, 136, 55 + sharp, skip_permille=300, sharpness=sharp)
        frames = [pkg.SynthFrame(p, i) for i in range(3)]
        compare(pkg, orc, frames, f"skip30% sharp{sharp}")


def test_level_zero_and_key_interval(pkg, orc, gpu_ready):
    p = pkg.synth_params(1, 256, 128, 99, filter_level=0, key_interval=2)
    frames = [pkg.SynthFrame(p, i) for i in range(4)]
    compare(pkg, orc, frames, "level0 key/2")


def test_inter_before_first_key(pkg, orc, gpu_ready):
    # no KEY frame yet: every reference slot is empty -> flat 127 prediction
    p = pkg.synth_params(1, 128, 96, 5, key_interval=1000)
    frames = [pkg.SynthFrame(p, i) for i in range(1, 4)]
    compare(pkg, orc, frames, "inter before key")


def test_1080p_config1(pkg, orc, gpu_ready):
    p = pkg.synth_params(1, 1920, 1080, 2024)
    frames = [pkg.SynthFrame(p, i) for i in range(3)]
    compare(pkg, orc, frames, "1080p preset1")


def test_720p_config0(pkg, orc, gpu_ready):
    p = pkg.synth_params(0, 1280, 720, 2025)
    frames = [pkg.SynthFrame(p, i) for i in range(4)]
    compare(pkg, orc, frames, "720p preset0")


def test_2160p_config2(pkg, orc, gpu_ready):
    p = pkg.synth_params(2, 3840, 2160, 2026)
    frames = [pkg.SynthFrame(p, i) for i in range(2)]
    compare(pkg, orc, frames, "2160p preset2")
