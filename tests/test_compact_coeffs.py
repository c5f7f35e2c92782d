"""CPU: compact coefficient staging of the drop-in boundary. The frame's
device image carries the parser's coefficient pool as a nonzero mask per
chunk of 64 entries, each chunk's position among the packed values and the
nonzero values; expanded the way the device stage does it, they give the
pool back entry for entry, the tables before them are the dense image's,
and nothing depends on the number of planner threads."""
import numpy as np
import pytest

CASES = [(1, 1920, 1080, 51, 2), (0, 1280, 720, 52, 3), (2, 3840, 2160, 53, 1),
         (3, 854, 480, 54, 0), (1, 264, 200, 55, 1), (4, 136, 72, 56, 2)]


def expand(img, info):
    """Dense pool from the compact image, entry by entry as the device
    expansion computes it: chunk start + nonzeros below the entry."""
    nc = info["n_chunks"]
    mask = img[info["off_cmask"]:info["off_cmask"] + 8 * nc].view(np.uint64)
    start = img[info["off_cstart"]:info["off_cstart"] + 4 * nc].view(np.uint32)
    vals = img[info["off_cvals"]:info["off_cvals"] + 2 * max(info["nnz"], 1)].view(np.int16)
    out = np.zeros(nc * 64, np.int16)
    bits = ((mask[:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool)
    below = np.cumsum(bits, axis=1) - bits  # nonzeros before each entry of its chunk
    pos = start[:, None].astype(np.int64) + below
    out[bits.reshape(-1)] = vals[pos[bits]]
    return out, bits


@pytest.mark.parametrize("preset,w,h,seed,idx", CASES)
def test_compact_pool_expands_to_the_parsers(pkg, preset, w, h, seed, idx):
    f = pkg.SynthFrame(pkg.synth_params(preset, w, h, seed), idx)
    img, info = pkg.plan_image_compact(f)
    pool = f.coeffs()
    assert info["n_chunks"] == (pool.size + 63) // 64
    dense, bits = expand(img, info)
    assert np.array_equal(dense[:pool.size], pool)
    assert not dense[pool.size:].any()
    assert info["nnz"] == int(np.count_nonzero(pool)) == int(bits.sum())
    # packed values are exactly the nonzero entries, in pool order
    vals = img[info["off_cvals"]:info["off_cvals"] + 2 * info["nnz"]].view(np.int16)
    assert np.array_equal(vals, pool[pool != 0])
    # what crosses to the device ends with the values
    assert info["upload"] == info["off_cvals"] + 2 * max(info["nnz"], 1)
    # the tables in front are the dense image's
    full = pkg.plan_image(f)
    assert np.array_equal(img[:info["off_cmask"]], full[:info["off_cmask"]])


def test_compact_image_does_not_depend_on_threads(pkg):
    f = pkg.SynthFrame(pkg.synth_params(1, 1920, 1080, 57), 1)
    one, info1 = pkg.plan_image_compact(f, threads=1)
    for t in (2, 3, 8, 0):
        img, info = pkg.plan_image_compact(f, threads=t)
        assert info == info1, f"threads={t}"
        assert np.array_equal(img[:info["upload"]], one[:info1["upload"]]), f"threads={t}"


def test_compact_edge_pools(pkg):
    """No coefficients, a pool of all zeros, a pool not a multiple of 64 and
    a dense pool (every entry nonzero, negatives included)."""
    base = pkg.SynthFrame(pkg.synth_params(1, 264, 200, 58), 1)
    h = base.pf.hdr
    b = base.blocks()
    pool = base.coeffs()
    rng = np.random.default_rng(5)
    for name, p in [("zeros", np.zeros_like(pool)),
                    ("dense", rng.integers(-32768, 32767, pool.size, endpoint=True).astype(np.int16)),
                    ("ragged", np.concatenate([pool, rng.integers(-3, 3, 37).astype(np.int16)]))]:
        p[p == 0] = 1 if name == "dense" else 0
        fr = pkg.OwnedFrame(h.width, h.height, b, p, frame_type=h.frame_type,
                            filter_level=h.filter_level)
        img, info = pkg.plan_image_compact(fr)
        dense, _ = expand(img, info)
        assert np.array_equal(dense[:p.size], p), name
        assert info["nnz"] == int(np.count_nonzero(p)), name
    # frame without a coefficient pool: every block skipped
    sk = b.copy()
    sk["skip"] = 1
    fr = pkg.OwnedFrame(h.width, h.height, sk, np.zeros(0, np.int16), frame_type=h.frame_type)
    img, info = pkg.plan_image_compact(fr)
    assert info["n_chunks"] == 0 and info["nnz"] == 0
