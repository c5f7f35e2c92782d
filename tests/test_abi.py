"""CPU checks of the C ABI: the library loads and exports every declared
entry point, struct layouts match the reference's, host validation sends
unsupported layouts to the CPU path, and the generator is deterministic and
well-formed. No GPU compute is called here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include", "cvp9_hip")
REFERENCE = "/root/reference"


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(pkg):
    names = declared_functions()
    assert {"hip_backend_init", "hip_decode_frame", "hip_get_frame", "hip_get_frame_dmabuf",
            "hip_export_buffer_alloc", "hip_export_buffer_free", "hip_set_render_target",
            "hip_backend_destroy"} <= names
    L = pkg.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # and the Python mirror binds all of them
    assert names <= set(pkg.SIGNATURES), sorted(names - set(pkg.SIGNATURES))


def test_struct_sizes(pkg):
    assert C.sizeof(pkg.FrameHeader) == 76
    assert C.sizeof(pkg.BlockInfo) == 32
    assert C.sizeof(pkg.ParsedFrame) == 160
    assert C.sizeof(pkg.FrameInfo) == 48
    assert C.sizeof(pkg.DmabufFrame) == 56
    assert C.sizeof(pkg.ExportBuffer) == 40
    assert pkg.FrameHeader.filter_level.offset == 63
    assert pkg.ParsedFrame.coeffs.offset == 96


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include HDR1
#include HDR2
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
  printf("sizes %zu %zu %zu %zu %zu %zu\n", sizeof(vp9_frame_header_t),
         sizeof(vp9_macroblock_info_t), sizeof(vp9_parsed_frame_t),
         sizeof(cvp9_frame_info_t), sizeof(cvp9_dmabuf_frame_t), sizeof(cvp9_export_buffer_t));
  P(vp9_frame_header_t, frame_type); P(vp9_frame_header_t, width); P(vp9_frame_header_t, height);
  P(vp9_frame_header_t, base_qindex); P(vp9_frame_header_t, filter_level);
  P(vp9_frame_header_t, sharpness_level); P(vp9_frame_header_t, refresh_frame_flags);
  P(vp9_macroblock_info_t, x); P(vp9_macroblock_info_t, tx_size); P(vp9_macroblock_info_t, mv);
  P(vp9_macroblock_info_t, skip); P(vp9_macroblock_info_t, coeff_offset);
  P(vp9_parsed_frame_t, blocks); P(vp9_parsed_frame_t, num_blocks);
  P(vp9_parsed_frame_t, coeffs); P(vp9_parsed_frame_t, num_coeffs);
  P(cvp9_dmabuf_frame_t, offsets); P(cvp9_dmabuf_frame_t, pts);
  P(cvp9_export_buffer_t, priv_mem); P(cvp9_frame_info_t, pts);
  return 0; }
"""


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "src", "decoder")),
                    reason="reference headers not mounted")
def test_layouts_match_reference_headers(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    outs = []
    for h1, h2, inc in [
        ('"compute_vp9/decoder.h"', '"decoder/vp9_parsed_frame.h"',
         [f"-I{REFERENCE}/include", f"-I{REFERENCE}/src"]),
        ('"cvp9_hip/abi_types.h"', '"cvp9_hip/hip_backend.h"', [f"-I{ROOT}/include"]),
    ]:
        exe = tmp_path / f"probe{len(outs)}"
        subprocess.run(["gcc", "-std=c11", f"-DHDR1={h1}", f"-DHDR2={h2}", *inc, str(src),
                        "-o", str(exe)], check=True)
        outs.append(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    assert outs[0] == outs[1]


# ---- host validation (what goes to the CPU path) ------------------------------

def blocks(pkg, specs):
    b = np.zeros(len(specs), pkg.BLOCK_DTYPE)
    for i, s in enumerate(specs):
        for k, v in s.items():
            b[i][k] = v
    return b


def plan(pkg, frame):
    return pkg.lib().cvp9_hip_plan_check(frame.ptr)


def test_plan_accepts_generated_frames(pkg):
    for preset, w, h in [(0, 1280, 720), (1, 854, 480), (2, 642, 362), (1, 64, 64)]:
        p = pkg.synth_params(preset, w, h, 3)
        for i in range(2):
            assert plan(pkg, pkg.SynthFrame(p, i)) == pkg.CVP9_OK


def test_plan_rejections(pkg):
    ok = dict(x=0, y=0, width=16, height=16, tx_size=1, is_intra=1)
    coef = np.zeros(256)
    f = pkg.OwnedFrame(16, 16, blocks(pkg, [ok]), coef)
    assert plan(pkg, f) == pkg.CVP9_OK
    # off the 4x4 grid -> CPU path
    f = pkg.OwnedFrame(16, 16, blocks(pkg, [dict(ok, x=2, width=8, height=8)]), coef)
    assert plan(pkg, f) == pkg.CVP9_ERR_UNSUPPORTED
    # overlapping blocks -> CPU path
    f = pkg.OwnedFrame(16, 16, blocks(pkg, [dict(ok, width=8, height=8, skip=1),
                                            dict(ok, x=4, y=4, width=8, height=8, skip=1)]), coef)
    assert plan(pkg, f) == pkg.CVP9_ERR_UNSUPPORTED
    # transform size class above 32x32 -> CPU path
    f = pkg.OwnedFrame(16, 16, blocks(pkg, [dict(ok, tx_size=4)]), coef)
    assert plan(pkg, f) == pkg.CVP9_ERR_UNSUPPORTED
    # ... unless the block has no residual
    f = pkg.OwnedFrame(16, 16, blocks(pkg, [dict(ok, tx_size=4, skip=1)]), coef)
    assert plan(pkg, f) == pkg.CVP9_OK
    # coefficients referenced past the pool -> invalid data
    f = pkg.OwnedFrame(16, 16, blocks(pkg, [dict(ok, coeff_offset=200)]), coef)
    assert plan(pkg, f) == pkg.CVP9_ERR_INVALID_DATA
    # W or H = 1 mod 8: the CPU deblock reads past the plane -> CPU path
    f = pkg.OwnedFrame(17, 16, blocks(pkg, [ok]), coef)
    assert plan(pkg, f) == pkg.CVP9_ERR_UNSUPPORTED
    f = pkg.OwnedFrame(16, 9, blocks(pkg, [ok]), coef)
    assert plan(pkg, f) == pkg.CVP9_ERR_UNSUPPORTED
    # blocks wholly outside the frame are ignored, like the reference
    f = pkg.OwnedFrame(16, 16, blocks(pkg, [ok, dict(ok, x=3, y=40)]), coef)
    assert plan(pkg, f) == pkg.CVP9_OK
    assert pkg.lib().cvp9_hip_plan_check(None) == pkg.CVP9_ERR_INVALID_DATA


def test_backend_without_gpu_reports_no_backend(pkg):
    if pkg.lib().cvp9_hip_device_count() > 0:
        pytest.skip("a GPU is present")
    ctx = C.c_void_p()
    assert pkg.lib().hip_backend_init(C.byref(ctx)) == pkg.CVP9_ERR_NO_BACKEND
    assert not ctx.value


# ---- generator -------------------------------------------------------------------

def test_generator_deterministic(pkg):
    p = pkg.synth_params(1, 640, 360, 42)
    a, b = pkg.SynthFrame(p, 3), pkg.SynthFrame(p, 3)
    assert a.blocks().tobytes() == b.blocks().tobytes()
    assert np.array_equal(a.coeffs(), b.coeffs())
    c = pkg.SynthFrame(pkg.synth_params(1, 640, 360, 43), 3)
    assert a.blocks().tobytes() != c.blocks().tobytes()


@pytest.mark.parametrize("preset,w,h", [(0, 1280, 720), (1, 854, 480), (2, 1000, 562)])
def test_generator_layout(pkg, preset, w, h):
    p = pkg.synth_params(preset, w, h, 9)
    for idx in range(2):
        f = pkg.SynthFrame(p, idx)
        b = f.blocks()
        owner = -np.ones(((h + 3) // 4, (w + 3) // 4), np.int64)
        ncoef = 0
        for i, k in enumerate(b):
            x, y, s = int(k["x"]), int(k["y"]), int(k["width"])
            assert x % 8 == 0 and y % 8 == 0 and s in (8, 16, 32, 64)
            assert x < w and y < h
            bw, bh = min(s, w - x), min(s, h - y)
            cells = owner[y // 4:(y + bh + 3) // 4, x // 4:(x + bw + 3) // 4]
            assert (cells == -1).all(), "overlap"
            cells[:] = i
            # Z order: the block above-left, above and left were emitted earlier
            if y > 0:
                assert owner[(y - 1) // 4, x // 4] < i
            if x > 0:
                assert owner[y // 4, (x - 1) // 4] < i
            T = 4 << int(k["tx_size"])
            assert T <= s
            if idx == 0:
                assert k["is_intra"] == 1
            if not k["skip"]:
                assert k["coeff_offset"] == ncoef
                ncoef += (bw // T) * (bh // T) * T * T
        assert (owner >= 0).all(), "frame not covered"
        assert f.pf.num_coeffs == ncoef
        assert f.pf.hdr.frame_type == (0 if idx == 0 else 1)


def test_generator_coefficient_density(pkg):
    # preset 1: 25% of each transform nonzero at the lowest zigzag positions
    f = pkg.SynthFrame(pkg.synth_params(1, 256, 256, 5), 1)
    c = f.coeffs()
    frac = np.count_nonzero(c) / c.size
    assert 0.22 < frac <= 0.25
    # preset 0: about 6 nonzero per transform, at zigzag 0..5
    f = pkg.SynthFrame(pkg.synth_params(0, 256, 256, 5), 1)
    for k in f.blocks():
        T = 4 << int(k["tx_size"])
        x, y, s = int(k["x"]), int(k["y"]), int(k["width"])
        blk = f.coeffs()[k["coeff_offset"]:k["coeff_offset"] + T * T].reshape(T, T)
        nz = np.argwhere(blk != 0)
        assert len(nz) <= 6 and all(r + c <= 2 for r, c in nz)
        break


# ---- intra work order (what keeps the persistent intra workers deadlock-free) ----

def mode_reads(y_mode, uv_mode):
    """(above, left, above-left) read by the predictions of a block's modes
    (vp9_reconstruct.c:63-91: DC above and left, V above, H left, any other
    value TM: all three), the union over luma and chroma."""
    def one(m):
        m = m if m <= 2 else 3
        return (m in (0, 1, 3), m in (0, 2, 3), m == 3)
    a, b = one(int(y_mode)), one(int(uv_mode))
    return tuple(x or y for x, y in zip(a, b))


def intra_dependencies(b, w, h):
    """For each intra block i (index into b): the earlier intra blocks owning
    a neighbour sample its prediction reads, luma and chroma, sample by
    sample (vp9_reconstruct.c:224-247, :272-311), for the samples its modes
    read (above row, left column, above-left sample)."""
    owner = -np.ones(((h + 3) // 4, (w + 3) // 4), np.int64)
    for i, k in enumerate(b):
        x, y = int(k["x"]), int(k["y"])
        if x >= w or y >= h:
            continue
        bw, bh = min(int(k["width"]), w - x), min(int(k["height"]), h - y)
        owner[y // 4:(y + bh + 3) // 4, x // 4:(x + bw + 3) // 4] = i
    deps = {}
    for i, k in enumerate(b):
        x, y = int(k["x"]), int(k["y"])
        if not k["is_intra"] or x >= w or y >= h:
            continue
        bw, bh = min(int(k["width"]), w - x), min(int(k["height"]), h - y)
        above, left, corner = mode_reads(k["y_mode"], k["uv_mode"])
        samples = []
        for sh, (px, py, pw, ph) in ((2, (x, y, bw, bh)),
                                     (1, (x >> 1, y >> 1, (bw + 1) >> 1, (bh + 1) >> 1))):
            if py > 0 and above:
                samples += [(px + e, py - 1, sh) for e in range(pw)]
            if px > 0 and left:
                samples += [(px - 1, py + r, sh) for r in range(ph)]
            if px > 0 and py > 0 and corner:
                samples.append((px - 1, py - 1, sh))
        d = set()
        for sx, sy, sh in samples:
            o = int(owner[sy >> sh, sx >> sh])
            if 0 <= o < i and b[o]["is_intra"]:
                d.add(o)
        deps[i] = d
    return deps


def check_order(pkg, frame, b):
    w, h = frame.pf.hdr.width, frame.pf.hdr.height
    deps = intra_dependencies(b, w, h)
    buf = np.zeros(max(len(deps), 1), np.uint32)
    n = pkg.lib().cvp9_hip_intra_order(frame.ptr, buf.ctypes.data_as(pkg.C.POINTER(pkg.C.c_uint32)),
                                       buf.size)
    assert n == len(deps)
    order = buf[:n].tolist()
    assert sorted(order) == sorted(deps), "not a permutation of the intra blocks"
    pos = {blk: p for p, blk in enumerate(order)}
    for blk, d in deps.items():
        assert all(pos[o] < pos[blk] for o in d), f"block {blk} handed out before a dependency"
    # level 0 (no intra dependency) comes first: small blocks, then large, each by
    # the height of the chains they head, then in list order
    free = [blk for blk, d in deps.items() if not d]
    assert order[:len(free)] == by_class(b, w, h, free, brute_heights(deps))
    return n


def small(k, w, h):
    """At most 16x16 once clamped to the frame (one pass of a 64-thread worker)."""
    return min(int(k["width"]), w - int(k["x"])) <= 16 and min(int(k["height"]), h - int(k["y"])) <= 16


def by_class(b, w, h, blks, hgt):
    """The planner's order inside one level: small blocks, then large, each
    with the blocks heading the longest dependency chains first, then in list
    order."""
    key = lambda x: (-hgt[x], x)
    return (sorted((x for x in blks if small(b[x], w, h)), key=key) +
            sorted((x for x in blks if not small(b[x], w, h)), key=key))


def brute_heights(deps):
    """Height of each intra block: 1 without a later intra block depending on
    it, else one more than the highest height among its dependents."""
    users = {blk: [] for blk in deps}
    for blk, d in deps.items():
        for o in d:
            users[o].append(blk)
    hgt = {}
    for blk in sorted(deps, reverse=True):
        hgt[blk] = 1 + max((hgt[u] for u in users[blk]), default=0)
    return hgt


@pytest.mark.parametrize("preset,w,h", [(1, 320, 184), (2, 643, 362), (0, 136, 72)])
def test_intra_order_respects_dependencies(pkg, preset, w, h):
    p = pkg.synth_params(preset, w, h, 21)
    for idx in range(3):
        f = pkg.SynthFrame(p, idx)
        check_order(pkg, f, f.blocks())


def test_intra_order_shuffled_lists(pkg):
    rng = np.random.default_rng(8)
    p = pkg.synth_params(1, 264, 136, 4)
    for idx in range(3):
        f = pkg.SynthFrame(p, idx)
        b = f.blocks()
        keep = rng.random(b.size) >= 0.1
        b = b[keep][rng.permutation(int(keep.sum()))]
        h = f.pf.hdr
        g = pkg.OwnedFrame(h.width, h.height, b, f.coeffs(), frame_type=h.frame_type)
        assert check_order(pkg, g, b) > 0


def brute_levels(deps):
    """Level of each intra block from its dependency sets: 0 without an
    intra dependency, else one more than the highest level among them
    (dependencies are earlier blocks, so list order resolves them)."""
    lev = {}
    for blk in sorted(deps):
        d = deps[blk]
        lev[blk] = 0 if not d else 1 + max(lev[o] for o in d)
    return lev


@pytest.mark.parametrize("preset,w,h,shuffle", [(1, 320, 184, False), (2, 643, 362, False),
                                                (1, 264, 136, True)])
def test_intra_levels_match_dependencies(pkg, preset, w, h, shuffle):
    rng = np.random.default_rng(5)
    p = pkg.synth_params(preset, w, h, 33)
    for idx in range(3):
        f = pkg.SynthFrame(p, idx)
        b = f.blocks()
        if shuffle:
            keep = rng.random(b.size) >= 0.1
            b = b[keep][rng.permutation(int(keep.sum()))]
            f = pkg.OwnedFrame(w, h, b, f.coeffs(), frame_type=f.pf.hdr.frame_type)
        deps = intra_dependencies(b, w, h)
        lev = brute_levels(deps)
        hgt = brute_heights(deps)
        starts = pkg.intra_levels(f)
        order = np.zeros(max(len(lev), 1), np.uint32)
        n = pkg.lib().cvp9_hip_intra_order(f.ptr, order.ctypes.data_as(pkg.C.POINTER(pkg.C.c_uint32)),
                                           order.size)
        assert n == len(lev)
        depth = (max(lev.values()) + 1) if lev else 0
        assert len(starts) == (depth + 1 if depth else 0)
        for L in range(depth):
            got = order[starts[L]:starts[L + 1]].tolist()
            want = by_class(b, w, h, [blk for blk, v in lev.items() if v == L], hgt)
            assert got == want, f"frame {idx} level {L}"


# ---- chains of intra blocks (chain workers) ----

@pytest.mark.parametrize("preset,w,h,shuffle", [(1, 320, 184, False), (2, 643, 362, False),
                                                (1, 1920, 1080, False), (1, 264, 136, True),
                                                (3, 854, 480, False)])
def test_intra_chains(pkg, preset, w, h, shuffle):
    """Each listed chain follows dependencies (every block depends on the one
    before it, so levels rise along it), the chains are the most critical
    ones (level + height of the first block, then length), at most 256, and
    every other entry follows in list order: a chain worker and the workers
    dealing the rest each meet levels in non-decreasing order."""
    rng = np.random.default_rng(9)
    p = pkg.synth_params(preset, w, h, 44)
    for idx in range(1, 3):
        f = pkg.SynthFrame(p, idx)
        b = f.blocks()
        if shuffle:
            keep = rng.random(b.size) >= 0.1
            b = b[keep][rng.permutation(int(keep.sum()))]
            f = pkg.OwnedFrame(w, h, b, f.coeffs(), frame_type=f.pf.hdr.frame_type)
        deps = intra_dependencies(b, w, h)
        lev, hgt = brute_levels(deps), brute_heights(deps)
        n = len(deps)
        order = np.zeros(max(n, 1), np.uint32)
        assert pkg.lib().cvp9_hip_intra_order(f.ptr, order.ctypes.data_as(pkg.C.POINTER(pkg.C.c_uint32)),
                                              order.size) == n
        buf = np.zeros(256 + 1 + max(n, 1), np.uint32)
        K = pkg.lib().cvp9_hip_intra_chains(f.ptr, buf.ctypes.data_as(pkg.C.POINTER(pkg.C.c_uint32)),
                                            buf.size)
        assert 0 <= K <= 256
        off = buf[:K + 1].tolist()
        ent = buf[K + 1:K + 1 + n].tolist()
        assert sorted(ent) == list(range(n)), "not a permutation of the entries"
        assert off[0] == 0 and all(off[c] < off[c + 1] - 1 for c in range(K)), "chains of 2+ blocks"
        rest = ent[off[K]:]
        assert rest == sorted(rest)
        crit = []
        for c in range(K):
            blks = [int(order[e]) for e in ent[off[c]:off[c + 1]]]
            for a, z in zip(blks, blks[1:]):
                assert a in deps[z], f"chain {c}: {z} does not depend on {a}"
                assert lev[a] < lev[z]
            crit.append((lev[blks[0]] + hgt[blks[0]], len(blks)))
        assert crit == sorted(crit, reverse=True)
