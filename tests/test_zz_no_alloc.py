"""GPU: the allocation-free entry point cvp9_engine_run_static against the
existing entry point (cvp9_engine_run + cvp9_engine_read) and the oracle,
bit-exact: two calls in a row, the second with a smaller input, with
reference frames and outputs held by the caller and every per-frame scratch
buffer taken from the library's fixed device array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def engine_frames(pkg, seq, w, h):
    with pkg.Engine(0) as eng:
        sid = eng.open_stream()
        out = []
        for f in seq:
            eng.run([sid], [eng.prepare(f)])
            eng.sync()
            out.append(eng.read(sid, w, h))
        return out


def test_static_entry_matches_engine_run(pkg, orc, gpu_ready):
    big = [pkg.SynthFrame(pkg.synth_params(1, 320, 184, 501), i) for i in range(2)]
    small = [pkg.SynthFrame(pkg.synth_params(2, 136, 72, 502), i) for i in range(2)]
    want_big = engine_frames(pkg, big, 320, 184)
    want_small = engine_frames(pkg, small, 136, 72)
    for got, want in zip(want_big + want_small, orc.run_oracle(big) + orc.run_oracle(small)):
        assert np.array_equal(got, want)
    print("[case] static entry: 320x184 KEY, inter, then 136x72 KEY, inter", flush=True)
    lib = pkg.lib()
    frames = []
    status = lib.cvp9_hip_dev_alloc(256)
    nv12 = lib.cvp9_hip_dev_alloc(pkg.i420_size(136, 72))
    try:
        with pkg.Engine(0) as eng:  # prepares the inputs (the static entry takes preps)
            pb = [eng.prepare(f) for f in big]
            ps = [eng.prepare(f) for f in small]
            A, B = pkg.DeviceFrame(320, 184), pkg.DeviceFrame(320, 184)
            Cf, D = pkg.DeviceFrame(136, 72), pkg.DeviceFrame(136, 72)
            frames = [A, B, Cf, D]
            # KEY frame, empty pool; its inter frame reads it from all 8
            # slots (the pool rule of a KEY frame); then a smaller stream
            assert pkg.run_static(pb[0], None, A, status=status) == pkg.CVP9_OK
            assert pkg.run_static(pb[1], [A] * 8, B, status=status) == pkg.CVP9_OK
            assert pkg.run_static(ps[0], None, Cf, status=status) == pkg.CVP9_OK
            assert pkg.run_static(ps[1], [Cf] * 8, D, nv12=nv12, status=status) == pkg.CVP9_OK
            assert np.array_equal(A.read_i420(), want_big[0]), "call 1"
            assert np.array_equal(B.read_i420(), want_big[1]), "call 2"
            assert np.array_equal(Cf.read_i420(), want_small[0]), "call 3 (smaller)"
            assert np.array_equal(D.read_i420(), want_small[1]), "call 4"
            got = np.zeros(pkg.i420_size(136, 72), np.uint8)
            pkg.check(lib.cvp9_hip_dev_read(nv12, got.ctypes.data_as(C.POINTER(C.c_uint8)), got.size),
                      "dev_read")
            assert np.array_equal(got, orc.nv12_pack(want_small[1], 136, 72)), "NV12"
            st = np.zeros(4, np.uint8)
            pkg.check(lib.cvp9_hip_dev_read(status, st.ctypes.data_as(C.POINTER(C.c_uint8)), 4),
                      "dev_read")
            assert not st.any(), "device error flags"
    finally:
        for f in frames:
            f.free()
        lib.cvp9_hip_dev_free(status)
        lib.cvp9_hip_dev_free(nv12)


def test_static_entry_rejects_frames_beyond_its_scratch(pkg, gpu_ready):
    """A 3840x2160 frame of 8x8 blocks only (129600 blocks, above the 49152
    the fixed edge region holds) is refused before anything is enqueued."""
    f = pkg.SynthFrame(pkg.synth_params(2, 3840, 2160, 7, split_permille=1000), 1)
    assert f.pf.num_blocks > 49152
    with pkg.Engine(0) as eng:
        p = eng.prepare(f)
        out = pkg.DeviceFrame(3840, 2160)
        try:
            assert pkg.run_static(p, None, out) == pkg.CVP9_ERR_UNSUPPORTED
        finally:
            out.free()
