"""CPU: the host planner on several threads lays out exactly the device
image it lays out on one (block records, motion-compensation units, intra
order, transform tables, coefficients), and reports the same code for a
frame it declines (the first offending block in list order)."""
import numpy as np
import pytest


CASES = [(1, 1920, 1080, 31, 2), (2, 3840, 2160, 32, 0), (2, 3840, 2160, 33, 1),
         (3, 1280, 720, 34, 3), (0, 1280, 720, 35, 4), (1, 2000, 1130, 36, 1)]


@pytest.mark.parametrize("preset,w,h,seed,idx", CASES)
def test_parallel_image_equals_serial(pkg, preset, w, h, seed, idx):
    f = pkg.SynthFrame(pkg.synth_params(preset, w, h, seed), idx)
    assert f.pf.num_blocks >= 2048, "the frame takes the parallel path"
    one = pkg.plan_image(f, threads=1)
    for t in (2, 3, 8, 0):
        assert np.array_equal(pkg.plan_image(f, threads=t), one), f"threads={t}"


def test_parallel_validation_reports_the_first_error(pkg):
    f = pkg.SynthFrame(pkg.synth_params(1, 1920, 1080, 40), 1)
    b = f.blocks().copy()
    h = f.pf.hdr
    n = b.size

    def frame(blocks):
        return pkg.OwnedFrame(h.width, h.height, blocks, f.coeffs(), frame_type=h.frame_type,
                              filter_level=h.filter_level)
    # an invalid coefficient range late in the list, an overlap earlier:
    # the overlap (the earlier offending block) decides
    def corrupt(blocks, i):
        """block i gets 4x4 transforms starting past the coefficient pool"""
        blocks["skip"][i] = 0
        blocks["tx_size"][i] = 0
        blocks["coeff_offset"][i] = f.pf.num_coeffs
    bad = b.copy()
    corrupt(bad, n - 10)
    over = np.concatenate([bad[:n // 3], bad[:1], bad[n // 3:]])
    for t in (1, 8):
        assert pkg.lib().cvp9_hip_plan_image(frame(over).ptr, t, None, 0) == pkg.CVP9_ERR_UNSUPPORTED
    # the invalid range alone
    for t in (1, 8):
        assert pkg.lib().cvp9_hip_plan_image(frame(bad).ptr, t, None, 0) == pkg.CVP9_ERR_INVALID_DATA
    # an overlap of a late block with an early one
    late = np.concatenate([b, b[5:6]])
    for t in (1, 8):
        assert pkg.lib().cvp9_hip_plan_image(frame(late).ptr, t, None, 0) == pkg.CVP9_ERR_UNSUPPORTED
    # an invalid range before an overlap: the invalid range decides
    bad2 = b.copy()
    corrupt(bad2, 3)
    both = np.concatenate([bad2, bad2[n - 2:n - 1]])
    for t in (1, 8):
        assert pkg.lib().cvp9_hip_plan_image(frame(both).ptr, t, None, 0) == pkg.CVP9_ERR_INVALID_DATA
