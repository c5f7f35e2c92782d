"""GPU parity with compact coefficient staging (CVP9_HIP_STAGING=compact):
the drop-in boundary sends each frame's coefficients as nonzero masks and
values and the device expands them to the dense pool before the
transforms; every frame stays bit-exact to the oracle, with dense and sparse
pools, pools of all zeros, frames of every preset and odd sizes."""
import numpy as np
import pytest

from test_gpu_parity import compare

pytestmark = pytest.mark.gpu


@pytest.fixture
def compact(monkeypatch):
    monkeypatch.setenv("CVP9_HIP_STAGING", "compact")


CASES = [(0, 1280, 720, 4), (1, 1920, 1080, 3), (2, 642, 362, 3), (3, 854, 480, 3),
         (4, 136, 72, 4), (1, 2, 2, 3), (2, 26, 14, 4), (0, 264, 200, 5)]


@pytest.mark.parametrize("preset,w,h,n", CASES)
def test_compact_staging_bit_exact(pkg, orc, gpu_ready, compact, preset, w, h, n):
    p = pkg.synth_params(preset, w, h, 4100 + preset * 13 + w)
    frames = [pkg.SynthFrame(p, i) for i in range(n)]
    compare(pkg, orc, frames, f"compact preset{preset} {w}x{h} n={n}")


def test_compact_staging_zero_and_dense_pools(pkg, orc, gpu_ready, compact):
    """A pool of zeros and a pool with every entry nonzero through the same
    context, frame after frame."""
    base = [pkg.SynthFrame(pkg.synth_params(1, 200, 136, 4200), i) for i in range(3)]
    rng = np.random.default_rng(17)
    frames = []
    for i, f in enumerate(base):
        h = f.pf.hdr
        c = f.coeffs()
        if i == 1:
            c = np.zeros_like(c)
        elif i == 2:
            c = rng.integers(-300, 300, c.size).astype(np.int16)
            c[c == 0] = 7
        frames.append(pkg.OwnedFrame(h.width, h.height, f.blocks(), c, frame_type=h.frame_type,
                                     filter_level=h.filter_level, sharpness=h.sharpness_level))
    compare(pkg, orc, frames, "compact zero/dense pools")
