"""Pins the oracle (CPU restatement of the reference path) on the CPU.

Checks, in order of strength:
  - golden vectors written from the reference's own compiled C
    (tests/golden/, made by tests/golden/make_golden.py);
  - the reference's C compiled in place (oracle/_ref), where present;
  - closed forms the mathematics fixes for each stage, chosen so a dropped
    term, a wrong sign or index, or a transposed operand fails one of them.
"""
import hashlib
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_hashes():
    with open(os.path.join(GOLD, "golden_hashes.json")) as f:
        return json.load(f)


def gen(pkg, case):
    p = pkg.synth_params(case["preset"], case["width"], case["height"], case["seed"],
                         **case["overrides"])
    return [pkg.SynthFrame(p, i) for i in range(case["frames"])]


# ---- golden vectors ---------------------------------------------------------

@pytest.mark.parametrize("name", list(load_hashes()["hash_cases"]))
def test_oracle_matches_golden_hashes(pkg, orc, name):
    case = load_hashes()["hash_cases"][name]
    frames = gen(pkg, case)
    for f, h in zip(frames, case["input"]):
        got = sha(np.concatenate([f.blocks().view(np.uint8), f.coeffs().view(np.uint8)]))
        assert got == h, f"{name}: generator drifted"
    out = orc.run_oracle(frames)
    w, hh = case["width"], case["height"]
    cw, ch = (w + 1) // 2, (hh + 1) // 2
    for i, (img, want) in enumerate(zip(out, case["planes"])):
        got = [sha(img[:w * hh]), sha(img[w * hh:w * hh + cw * ch]), sha(img[w * hh + cw * ch:])]
        assert got == want, f"{name}: frame {i} plane hashes differ"


def test_oracle_matches_golden_frames(pkg, orc):
    meta = load_hashes()["frame_cases"]
    data = np.load(os.path.join(GOLD, "golden_frames.npz"))
    for name, case in meta.items():
        out = np.stack(orc.run_oracle(gen(pkg, case)))
        assert np.array_equal(out, data[name]), name


def test_idct_tables_fixture(orc, pkg):
    import ctypes as C
    with open(os.path.join(GOLD, "idct_tables.json")) as f:
        tables = json.load(f)
    for n in (4, 8, 16, 32):
        want = np.array([float.fromhex(x) for x in tables[str(n)]]).reshape(n, n)
        got = orc.idct_matrix(n)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), n
        lib_tab = np.zeros(n * n)
        assert pkg.lib().cvp9_hip_idct_table(n, lib_tab.ctypes.data_as(C.POINTER(C.c_double))) == 0
        assert np.array_equal(lib_tab.view(np.uint64), want.ravel().view(np.uint64)), n
        # the matrix is orthonormal: M M^T = I to rounding
        assert np.allclose(want @ want.T, np.eye(n), atol=1e-14)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(GOLD), "..", "oracle", "_ref",
                                                    "libcvp9_ref.so")),
                    reason="reference C not built here (oracle/_ref)")
@pytest.mark.parametrize("preset,w,h,n,over", [
    (0, 320, 184, 4, {}), (1, 208, 136, 4, {"skip_permille": 300}),
    (2, 264, 138, 3, {"sharpness": 7}), (1, 138, 66, 5, {"key_interval": 3}),
    (1, 96, 64, 3, {"filter_level": 63, "sharpness": 2}),
])
def test_oracle_matches_compiled_reference(pkg, orc, preset, w, h, n, over):
    frames = [pkg.SynthFrame(pkg.synth_params(preset, w, h, 77 + n, **over), i) for i in range(n)]
    a = orc.run_oracle(frames)
    b = orc.run_reference(frames)
    for i, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), f"frame {i}"


# ---- IDCT closed forms --------------------------------------------------------

@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_idct_dc_only_is_rounded_mean(orc, n):
    # DC-only input: out[i][j] = (s*c)*s with s = sqrt(1/N), i.e. c/N rounded
    # half away from zero, for every |c| <= 4000
    for c in list(range(-4000, 4001, 37)) + [1, -1, n // 2, -n // 2, 3 * n // 2, -(3 * n // 2)]:
        blk = np.zeros((n, n), np.int16)
        blk[0, 0] = c
        out = orc.idct_2d(blk)
        q = abs(c) / n
        want = int(np.floor(q + 0.5)) * (1 if c >= 0 else -1)
        assert (out == want).all(), (n, c, out[0, 0], want)


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_idct_zero_and_roundtrip(orc, n):
    assert (orc.idct_2d(np.zeros((n, n), np.int16)) == 0).all()
    rng = np.random.default_rng(n)
    x = rng.integers(-60, 61, (n, n))
    m = orc.idct_matrix(n)
    fwd = np.rint(m @ x @ m.T).astype(np.int16)   # forward DCT of x
    back = orc.idct_2d(fwd).astype(np.int64)
    err = np.abs(back - x)
    assert err.mean() < 0.6 and err.max() <= 3


def test_idct_is_not_transposed(orc):
    # a single coefficient in row 1 (vertical frequency) must vary along rows only
    blk = np.zeros((8, 8), np.int16)
    blk[1, 0] = 400
    out = orc.idct_2d(blk).astype(int)
    assert (out == out[:, :1]).all() and out[0, 0] > 0 > out[7, 0]


# ---- intra closed forms -----------------------------------------------------

def test_intra_constant_neighbours(orc):
    for mode in (0, 1, 2, 3, 9):
        out = orc.intra_predict(8, 8, np.full(8, 90), np.full(8, 90), 90, mode)
        assert (out == 90).all(), mode


def test_intra_dc_divisor_rectangle(orc):
    # DC over w above + h left samples: (8*10 + 4*20 + 6) // 12 = 13
    out = orc.intra_predict(8, 4, np.full(8, 10), np.full(4, 20), 0, 0)
    assert (out == 13).all()


def test_intra_tm_clamps_and_modes_above_two_are_tm(orc):
    above = np.array([250, 10, 128, 0])
    left = np.array([200, 5, 128, 255])
    for mode in (3, 4, 8, 9):
        out = orc.intra_predict(4, 4, above, left, 100, mode).astype(int)
        want = np.clip(above[None, :] + left[:, None] - 100, 0, 255)
        assert (out == want).all(), mode
    v = orc.intra_predict(4, 4, above, left, 100, 1)
    h = orc.intra_predict(4, 4, above, left, 100, 2)
    assert (v == above[None, :]).all() and (h == left[:, None]).all()


# ---- motion compensation closed forms -----------------------------------------

def test_mc_zero_and_integer_mv(orc):
    rng = np.random.default_rng(3)
    ref = rng.integers(0, 256, (40, 48)).astype(np.uint8)
    assert (orc.motion_compensate(ref, 16, 8, 8, 16, 0, 0) == ref[16:24, 8:24]).all()
    # +3 px right, -2 px up: shifted copy
    assert (orc.motion_compensate(ref, 16, 8, 8, 16, 24, -16) == ref[14:22, 11:27]).all()
    # far outside: pure edge replicate of the corner sample
    far = orc.motion_compensate(ref, 8, 8, 0, 0, -2048, -2048)
    assert (far == ref[0, 0]).all()
    far = orc.motion_compensate(ref, 8, 8, 40, 32, 2048, 2048)
    assert (far == ref[-1, -1]).all()


def test_mc_half_pel_on_ramp(orc):
    ref = np.tile(np.arange(0, 200, 10, dtype=np.uint8), (8, 1))  # x ramp step 10
    out = orc.motion_compensate(ref, 4, 2, 4, 2, 4, 0)  # +1/2 px
    # (a*4 + b*4 + 4) >> 3 between a and a+10 -> a + 5, then vertical no-op
    assert (out[0] == ref[0, 4:8] + 5).all()
    out = orc.motion_compensate(ref, 4, 2, 4, 2, 1, 0)  # +1/8 px: (7a + a+10 + 4)>>3 = a+1
    assert (out[0] == ref[0, 4:8] + 1).all()
    out = orc.motion_compensate(ref, 4, 2, 4, 2, -1, 0)  # -1/8: pixel left, phase 7
    assert (out[0] == ref[0, 3:7] + 9).all()


# ---- loop filter closed forms -------------------------------------------------

def test_lf_threshold_table(orc):
    assert orc.lf_threshold(32, 0) == 32
    assert orc.lf_threshold(32, 1) == 8      # 16, capped at 9-1
    assert orc.lf_threshold(32, 5) == 4      # 32>>2 = 8, capped at 4
    assert orc.lf_threshold(1, 7) == 1       # 0 -> 1
    assert orc.lf_threshold(6, 3) == 3       # 6>>1
    assert orc.lf_threshold(63, 0) == 63
    for lvl in range(1, 64):
        for sh in range(8):
            t = orc.lf_threshold(lvl, sh)
            assert 1 <= t <= (lvl if sh == 0 else 9 - sh)


def test_lf_flat_level0_and_step_edge(orc):
    flat = np.full((32, 32), 77, np.uint8)
    assert (orc.loop_filter(flat, 40, 0) == flat).all()
    step = np.zeros((32, 32), np.uint8)
    step[:8] = 100
    step[8:] = 104
    assert (orc.loop_filter(step, 0, 0) == step).all()
    out = orc.loop_filter(step, 32, 0)
    # edge y=8: p1=p0=100, q0=q1=104 -> delta=(3*4+4)>>3=2
    assert (out[6] == 101).all() and (out[7] == 102).all()
    assert (out[8] == 102).all() and (out[9] == 103).all()
    # a step above threshold is left alone
    big = np.zeros((16, 16), np.uint8)
    big[8:] = 200
    assert (orc.loop_filter(big, 32, 0) == big).all()


def test_lf_vertical_pass_reads_horizontal_output(orc):
    img = np.zeros((16, 16), np.uint8)
    img[:8, :8] = 100
    img[:8, 8:] = 104
    img[8:, :8] = 104
    img[8:, 8:] = 108
    out = orc.loop_filter(img, 32, 0).astype(int)
    # Horizontal pass (edge y=8): columns < 8 step 100 -> 104, columns >= 8
    # step 104 -> 108; both give delta (3*4+4)>>3 = 2, so rows 6..9 become
    # 101,102,102,103 (left) and 105,106,106,107 (right). The vertical pass
    # (edge x=8) then reads those rows: row 7 steps 102 -> 106 (delta 2), so
    # (7,6..9) = 103,104,104,105; row 6 steps 101 -> 105, so (6,7) = 103.
    # A vertical pass that read the unfiltered image would give (7,7) = 102.
    assert out[7, 6] == 103 and out[7, 7] == 104 and out[7, 8] == 104 and out[7, 9] == 105
    assert out[6, 7] == 103 and out[8, 7] == 104
    assert out[0, 7] == 102 and out[7, 0] == 102


# ---- NV12 ------------------------------------------------------------------------

def test_nv12_interleave(orc):
    w, h = 6, 4
    cw, ch = 3, 2
    i420 = np.arange(w * h + 2 * cw * ch, dtype=np.uint8)
    out = orc.nv12_pack(i420, w, h)
    assert (out[:w * h] == i420[:w * h]).all()
    u = i420[w * h:w * h + cw * ch]
    v = i420[w * h + cw * ch:]
    assert (out[w * h::2] == u).all() and (out[w * h + 1::2] == v).all()


# ---- reference pool rule -----------------------------------------------------------

def _blocks(pkg, spec):
    b = np.zeros(len(spec), pkg.BLOCK_DTYPE)
    for i, s in enumerate(spec):
        for k, v in s.items():
            b[i][k] = v
    return b


def test_ref_pool_key_refreshes_all_slots(pkg, orc):
    W = H = 16
    # KEY: one intra 16x16 DC block over 127 neighbours, residual +40 via DC coeff
    key = pkg.OwnedFrame(W, H, _blocks(pkg, [dict(x=0, y=0, width=16, height=16, tx_size=2,
                                                  is_intra=1)]),
                         np.r_[[40 * 16], np.zeros(255)], frame_type=pkg.VP9_FRAME_KEY)
    # inter from LAST with zero MV, residual +10
    inter = pkg.OwnedFrame(W, H, _blocks(pkg, [dict(x=0, y=0, width=16, height=16, tx_size=2,
                                                    ref_frame=[1, 0])]),
                           np.r_[[10 * 16], np.zeros(255)])
    # inter from GOLDEN (slot 1) with zero MV, no residual: copies the KEY frame
    gold = pkg.OwnedFrame(W, H, _blocks(pkg, [dict(x=0, y=0, width=16, height=16,
                                                   ref_frame=[2, 0], skip=1)]), [])
    out = orc.run_oracle([key, inter, inter, gold])
    y = [o[:W * H] for o in out]
    assert (y[0] == 127 + 40).all()
    assert (y[1] == 177).all() and (y[2] == 187).all()
    assert (y[3] == 167).all()       # slot 1 still holds the KEY frame
    assert (out[3][W * H:] == 127).all()  # chroma: DC of 127 neighbours, copied
