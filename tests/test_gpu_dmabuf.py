"""GPU: the DMA-BUF entry points of the drop-in backend (hip_export_buffer_*,
hip_set_render_target, hip_get_frame_dmabuf; vulkan_backend.c:213-331,
:1457-1544 semantics). A render target receives the frame as NV12 (layout of
vaapi_entrypoint.c:131-153) once, and the fd of a decoded frame maps its
packed I420 planes; both bit-exact against the oracle."""
import ctypes as C
import mmap
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 320, 184
CW, CH = (W + 1) // 2, (H + 1) // 2
I420 = W * H + 2 * CW * CH


def frames(pkg, n, seed):
    p = pkg.synth_params(1, W, H, seed)
    return [pkg.SynthFrame(p, i) for i in range(n)]


def test_render_target_receives_nv12_once(pkg, orc, gpu_ready):
    seq = frames(pkg, 3, 501)
    want = orc.run_oracle(seq)
    nv12 = [orc.nv12_pack(x, W, H) for x in want]
    print("[case] render target NV12 320x184", flush=True)
    with pkg.HipBackend() as be:
        rc, buf = be.export_buffer_alloc(I420)
        if rc == pkg.CVP9_ERR_UNSUPPORTED:
            pytest.skip("no DMA-BUF export on this device (hip_export_buffer_alloc: UNSUPPORTED)")
        pkg.check(rc, "export_buffer_alloc")
        try:
            assert buf.fd >= 0 and buf.size == I420
            # frame 0 into the target
            pkg.check(be.set_render_target(buf), "set_render_target")
            pkg.check(be.decode(seq[0], 0), "decode")
            rc, i420, _ = be.get_frame(wait=True)
            pkg.check(rc, "get_frame")
            assert np.array_equal(i420, want[0])
            assert np.array_equal(be.export_read(buf, I420), nv12[0]), "NV12 via HIP"
            if buf.mapped:
                view = np.frombuffer(C.string_at(buf.mapped, I420), np.uint8)
                assert np.array_equal(view, nv12[0]), "NV12 via the CPU mapping"
            # the target is one-shot: frame 1 leaves it alone
            pkg.check(be.decode(seq[1], 1), "decode")
            rc, i420, _ = be.get_frame(wait=True)
            pkg.check(rc, "get_frame")
            assert np.array_equal(i420, want[1])
            assert np.array_equal(be.export_read(buf, I420), nv12[0]), "target rewritten"
            # set again: frame 2 lands in it
            pkg.check(be.set_render_target(buf), "set_render_target")
            pkg.check(be.decode(seq[2], 2), "decode")
            rc, i420, _ = be.get_frame(wait=True)
            pkg.check(rc, "get_frame")
            assert np.array_equal(i420, want[2])
            assert np.array_equal(be.export_read(buf, I420), nv12[2])
        finally:
            be.export_buffer_free(buf)
        assert buf.fd == -1 and not buf.priv_buf


def test_get_frame_dmabuf_maps_i420(pkg, orc, gpu_ready):
    seq = frames(pkg, 3, 502)
    want = orc.run_oracle(seq)
    print("[case] get_frame_dmabuf 320x184", flush=True)
    with pkg.HipBackend() as be:
        for i, f in enumerate(seq):
            pkg.check(be.decode(f, 100 + i), "decode")
            rc, d = be.get_frame_dmabuf()
            if rc == pkg.CVP9_ERR_UNSUPPORTED:
                pytest.skip("no DMA-BUF export on this device (hip_get_frame_dmabuf: UNSUPPORTED)")
            pkg.check(rc, "get_frame_dmabuf")
            try:
                assert (d.width, d.height, d.pts) == (W, H, 100 + i)
                o = d.offsets[0]  # the buffer's place inside the DMA-BUF
                assert list(d.offsets) == [o, o + W * H, o + W * H + CW * CH]
                assert list(d.pitches) == [W, CW, CW]
                assert d.size >= d.offsets[0] + I420
                got = None
                try:  # a CPU mapping of the buffer, where the driver allows it
                    m = mmap.mmap(d.fd, d.size, mmap.MAP_SHARED, mmap.PROT_READ)
                    try:
                        got = np.frombuffer(m[d.offsets[0]:d.offsets[0] + I420], np.uint8).copy()
                    finally:
                        m.close()
                except OSError:
                    pass
                if got is None:  # else imported by HIP, as another GPU API would
                    got = np.zeros(I420, np.uint8)
                    rc = pkg.lib().cvp9_hip_dmabuf_read(d.fd, d.size, d.offsets[0],
                                                        got.ctypes.data_as(C.POINTER(C.c_uint8)), I420)
                    if rc == pkg.CVP9_ERR_UNSUPPORTED:
                        pytest.skip("the DMA-BUF can be neither mapped nor imported on this box")
                    pkg.check(rc, "dmabuf_read")
                assert np.array_equal(got, want[i]), f"frame {i}"
            finally:
                os.close(d.fd)
