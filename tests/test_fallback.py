"""CPU: the backend's routing of frames between itself and the caller's CPU
path (src/decoder/vp9_decoder.c:209-247). A frame the validator declines is
reconstructed by the CPU path, so the device reference pool lacks it: the
backend then declines every frame until the next KEY frame it can take,
which writes all 8 slots. Checked through cvp9_hip_route_check, which runs
hip_decode_frame's decision without a device."""
import numpy as np
import pytest


def overlapping(pkg, f):
    """The same frame with its first block listed twice (overlapping blocks:
    the CPU path reconstructs it, the HIP validator declines it)."""
    b = f.blocks()
    h = f.pf.hdr
    return pkg.OwnedFrame(h.width, h.height, np.concatenate([b, b[:1]]), f.coeffs(),
                          frame_type=h.frame_type, filter_level=h.filter_level,
                          sharpness=h.sharpness_level)


@pytest.fixture(scope="module")
def seq(pkg):
    p = pkg.synth_params(1, 136, 72, 7, key_interval=4)
    return [pkg.SynthFrame(p, i) for i in range(9)]


def test_declined_inter_frame_until_key(pkg, seq):
    OK, NO = pkg.CVP9_OK, pkg.CVP9_ERR_UNSUPPORTED
    assert [f.pf.hdr.frame_type for f in seq[:5]] == [0, 1, 1, 1, 0]
    bad = overlapping(pkg, seq[2])
    assert pkg.lib().cvp9_hip_plan_check(bad.ptr) == NO
    frames = seq[:2] + [bad] + seq[3:]
    assert pkg.route_check(frames) == [OK, OK, NO, NO, OK, OK, OK, OK, OK]


def test_declined_key_frame(pkg, seq):
    OK, NO = pkg.CVP9_OK, pkg.CVP9_ERR_UNSUPPORTED
    # a KEY frame that is declined keeps the stream on the CPU path until
    # the next KEY frame the validator takes
    frames = [overlapping(pkg, seq[0])] + seq[1:4] + [overlapping(pkg, seq[4])] + seq[5:]
    assert pkg.route_check(frames) == [NO] * 8 + [OK]  # frame 8 is the next KEY frame
    frames = seq[:4] + [overlapping(pkg, seq[4])] + seq[5:8] + [seq[0]]
    assert pkg.route_check(frames) == [OK] * 4 + [NO] * 4 + [OK]


def test_declined_frame_size(pkg):
    OK, NO = pkg.CVP9_OK, pkg.CVP9_ERR_UNSUPPORTED
    good = [pkg.SynthFrame(pkg.synth_params(1, 136, 72, 8, key_interval=3), i) for i in range(4)]
    # W = 1 mod 8: the CPU deblock reads past the plane; declined
    odd = [pkg.SynthFrame(pkg.synth_params(1, 137, 72, 8, key_interval=3), i) for i in range(4)]
    assert pkg.route_check(good[:2] + odd[2:3] + good[3:] + good[:1]) == [OK, OK, NO, OK, OK]
    assert pkg.route_check(good[:2] + odd[2:]) == [OK, OK, NO, NO]


def test_invalid_data_is_not_a_decline(pkg, seq):
    OK, BAD = pkg.CVP9_OK, pkg.CVP9_ERR_INVALID_DATA
    f = seq[1]
    b = f.blocks().copy()
    b["coeff_offset"][0] = f.pf.num_coeffs + 1000
    h = f.pf.hdr
    broken = pkg.OwnedFrame(h.width, h.height, b, f.coeffs(), frame_type=h.frame_type,
                            filter_level=h.filter_level)
    assert pkg.route_check([seq[0], broken, seq[2]]) == [OK, BAD, OK]
