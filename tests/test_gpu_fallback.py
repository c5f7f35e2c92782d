"""GPU: boundary semantics of the drop-in backend that the reference's
dispatch and its VA-API driver rely on (src/decoder/vp9_decoder.c:200-259;
src/backend/vulkan/vulkan_backend.c:1037-1062, :1391-1455;
src/vaapi/vaapi_entrypoint.c:459-484):

  * a frame declined mid-stream (CVP9_ERR_UNSUPPORTED) and every frame after
    it until the next KEY frame go to the caller's CPU path; every frame the
    backend does deliver is bit-exact, and the device reference pool can be
    read out for the CPU path to take over;
  * hip_get_frame(wait=0): CVP9_ERR_AGAIN while the oldest frame is on the
    GPU, blocking when the ring is full;
  * plane pointers stay valid until their slot is reused;
  * the AGAIN / drain / retry cycle with an armed NV12 render target;
  * a device error is reported for its frame and every later one until a
    KEY frame, which decodes bit-exact again.
"""
import ctypes as C
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def overlapping(pkg, f):
    b = f.blocks()
    h = f.pf.hdr
    return pkg.OwnedFrame(h.width, h.height, np.concatenate([b, b[:1]]), f.coeffs(),
                          frame_type=h.frame_type, filter_level=h.filter_level,
                          sharpness=h.sharpness_level)


def test_declined_frame_mid_stream(pkg, orc, gpu_ready):
    W, H = 200, 136
    p = pkg.synth_params(1, W, H, 41, key_interval=4)
    base = [pkg.SynthFrame(p, i) for i in range(7)]
    seq = base[:2] + [overlapping(pkg, base[2])] + base[3:]
    want = orc.run_oracle(seq)  # the reference's CPU path decodes every frame
    print("[case] declined frame mid-stream", flush=True)
    with pkg.HipBackend() as be:
        for i in range(2):
            assert be.decode(seq[i], i) == pkg.CVP9_OK
        for i in range(2):
            rc, img, info = be.get_frame(wait=True)
            assert rc == pkg.CVP9_OK and info.pts == i
            assert np.array_equal(img, want[i]), f"frame {i}"
        # the pool the CPU path takes over: KEY frame 0 in slots 1-7, frame 1 in slot 0
        assert np.array_equal(be.read_ref(0), want[1])
        for k in range(1, 8):
            assert np.array_equal(be.read_ref(k), want[0]), f"slot {k}"
        codes = [be.decode(seq[i], i) for i in range(2, 4)]
        assert codes == [pkg.CVP9_ERR_UNSUPPORTED] * 2
        assert be.get_frame(wait=False)[0] == pkg.CVP9_ERR_UNSUPPORTED  # nothing queued
        for i in range(4, 7):  # KEY frame 4 and what follows it
            assert be.decode(seq[i], i) == pkg.CVP9_OK, f"frame {i}"
        for i in range(4, 7):
            rc, img, info = be.get_frame(wait=True)
            assert rc == pkg.CVP9_OK and info.pts == i
            assert np.array_equal(img, want[i]), f"frame {i}"


def test_read_ref_empty_pool(pkg, gpu_ready):
    with pkg.HipBackend() as be:
        assert all(be.read_ref(k) is None for k in range(8))


def test_get_frame_nonblocking(pkg, orc, gpu_ready):
    """wait=0 returns AGAIN while the oldest frame is still on the GPU, and
    blocks instead when the ring is full."""
    W, H = 3840, 2160
    p = pkg.synth_params(2, W, H, 77)
    seq = [pkg.SynthFrame(p, i) for i in range(3)]
    want = orc.run_oracle(seq)
    print("[case] non-blocking get", flush=True)
    with pkg.HipBackend() as be:
        assert be.decode(seq[0], 0) == pkg.CVP9_OK
        saw_again, t0 = 0, time.monotonic()
        while True:
            rc, img, _ = be.get_frame(wait=False)
            if rc != pkg.CVP9_ERR_AGAIN:
                break
            saw_again += 1
            assert time.monotonic() - t0 < 30, "frame never completed"
        assert rc == pkg.CVP9_OK and np.array_equal(img, want[0])
        # a 2160p frame (16 MB of coefficients over PCIe, then the kernels)
        # cannot be ready within the few microseconds of the first poll
        assert saw_again > 0
        # ring full: wait=0 blocks and delivers (frames 10-12 decode the
        # sequence again from its KEY frame)
        for i in range(3):
            assert be.decode(seq[i], 10 + i) == pkg.CVP9_OK
        assert be.decode(seq[0], 99) == pkg.CVP9_ERR_AGAIN
        rc, img, info = be.get_frame(wait=False)
        assert rc == pkg.CVP9_OK and info.pts == 10
        assert np.array_equal(img, want[0])
        for i in (1, 2):
            rc, img, info = be.get_frame(wait=True)
            assert rc == pkg.CVP9_OK and info.pts == 10 + i and np.array_equal(img, want[i])


def test_plane_pointers_live_until_slot_reuse(pkg, orc, gpu_ready):
    """A frame's plane pointers stay valid for the next CVP9_HIP_INFLIGHT - 1
    decodes (decoder.h:102-106)."""
    W, H = 320, 184
    seq = [pkg.SynthFrame(pkg.synth_params(1, W, H, 88), i) for i in range(3)]
    want = orc.run_oracle(seq)
    n = pkg.i420_size(W, H)
    with pkg.HipBackend() as be:
        assert be.decode(seq[0], 0) == pkg.CVP9_OK
        rc, _, info0 = be.get_frame(wait=True)
        assert rc == pkg.CVP9_OK
        for i in (1, 2):  # two more decodes, each retrieved
            assert be.decode(seq[i], i) == pkg.CVP9_OK
            rc, img, _ = be.get_frame(wait=True)
            assert rc == pkg.CVP9_OK and np.array_equal(img, want[i])
        again = pkg.frame_info_to_i420(info0)
        assert again.size == n and np.array_equal(again, want[0])


def test_again_drain_retry_with_render_target(pkg, orc, gpu_ready):
    """VA-API EndPicture: decode into an armed render target; on AGAIN drain
    one frame and retry (vaapi_entrypoint.c:459-471). The AGAIN leaves the
    target armed, and the retried decode packs the frame into it."""
    W, H = 320, 184
    seq = [pkg.SynthFrame(pkg.synth_params(1, W, H, 93), i) for i in range(5)]
    want = orc.run_oracle(seq)
    n = pkg.i420_size(W, H)
    print("[case] AGAIN / drain / retry with a render target", flush=True)
    with pkg.HipBackend() as be:
        rc, buf = be.export_buffer_alloc(n)
        if rc == pkg.CVP9_ERR_UNSUPPORTED:
            pytest.skip("no DMA-BUF export on this device")
        pkg.check(rc, "export_buffer_alloc")
        try:
            got = []
            for i in range(3):
                assert be.decode(seq[i], i) == pkg.CVP9_OK
            pkg.check(be.set_render_target(buf), "set_render_target")
            for i in (3, 4):
                rc = be.decode(seq[i], i)
                if rc == pkg.CVP9_ERR_AGAIN:  # drain one, retry
                    r2, img, _ = be.get_frame(wait=True)
                    assert r2 == pkg.CVP9_OK
                    got.append(img)
                    rc = be.decode(seq[i], i)
                assert rc == pkg.CVP9_OK, f"frame {i}"
            while True:
                r2, img, _ = be.get_frame(wait=True)
                if r2 == pkg.CVP9_ERR_UNSUPPORTED:
                    break
                assert r2 == pkg.CVP9_OK
                got.append(img)
            assert len(got) == 5
            for i in range(5):
                assert np.array_equal(got[i], want[i]), f"frame {i}"
            # the target (armed once) received frame 3, the retried decode
            assert np.array_equal(be.export_read(buf, n), orc.nv12_pack(want[3], W, H))
        finally:
            be.export_buffer_free(buf)


def test_device_error_until_key_frame(pkg, orc, gpu_ready):
    """A dependency wait forced to give up (one poll per wait) makes the KEY
    frame report CVP9_ERR_GPU; the next inter frame is refused with
    CVP9_ERR_GPU (its pool holds a broken frame); the next KEY frame starts
    afresh and it and its successor are bit-exact."""
    W, H = 320, 184
    p = pkg.synth_params(1, W, H, 61, key_interval=2)
    seq = [pkg.SynthFrame(p, i) for i in range(4)]
    want = orc.run_oracle(seq[2:])
    print("[case] device error until KEY", flush=True)
    with pkg.HipBackend() as be:
        be.set_debug(pkg.CVP9_HIP_DEBUG_SPIN_ONE)
        assert be.decode(seq[0], 0) == pkg.CVP9_OK
        rc, _, _ = be.get_frame(wait=True)
        assert rc == pkg.CVP9_ERR_GPU
        be.set_debug(0)
        assert be.decode(seq[1], 1) == pkg.CVP9_ERR_GPU
        for i in (2, 3):
            assert be.decode(seq[i], i) == pkg.CVP9_OK
        for i in (2, 3):
            rc, img, info = be.get_frame(wait=True)
            assert rc == pkg.CVP9_OK and info.pts == i
            assert np.array_equal(img, want[i - 2]), f"frame {i}"
