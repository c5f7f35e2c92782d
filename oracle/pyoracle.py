"""ctypes access to the parity oracle. TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this
module, and only as the checker. It loads
  oracle/_build/libcvp9_oracle.so  C restatement of the reference CPU path
  oracle/_ref/libcvp9_ref.so       the reference's own C, compiled from its
                                    sources (present only where it was built)
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CVP9_ORACLE_LIB names another build of the oracle (the sanitizer build)
ORACLE_LIB = os.environ.get("CVP9_ORACLE_LIB") or os.path.join(HERE, "_build", "libcvp9_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libcvp9_ref.so")

_orc = None
_ref = None


class FrameInfo(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("stride_y", C.c_uint32),
        ("stride_uv", C.c_uint32), ("plane_y", C.POINTER(C.c_uint8)),
        ("plane_u", C.POINTER(C.c_uint8)), ("plane_v", C.POINTER(C.c_uint8)),
        ("pts", C.c_int64),
    ]


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_LIB)
        L.orc_decoder_new.restype = C.c_void_p
        L.orc_decoder_free.argtypes = [C.c_void_p]
        L.orc_decoder_run.restype = C.c_int
        L.orc_decoder_run.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_int,
                                      C.POINTER(C.c_uint8)]
        L.orc_idct_matrix.argtypes = [C.c_int, C.POINTER(C.c_double)]
        L.orc_idct_2d.argtypes = [C.POINTER(C.c_int16), C.POINTER(C.c_int16), C.c_int]
        L.orc_intra_predict.argtypes = [C.POINTER(C.c_uint8), C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_uint8), C.POINTER(C.c_uint8),
                                        C.c_uint8, C.c_int]
        L.orc_motion_compensate.argtypes = [C.POINTER(C.c_uint8), C.c_int,
                                            C.POINTER(C.c_uint8), C.c_int,
                                            C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_lf_threshold.restype = C.c_int
        L.orc_lf_threshold.argtypes = [C.c_uint8, C.c_uint8]
        L.orc_loop_filter.argtypes = [C.POINTER(FrameInfo), C.c_uint8, C.c_uint8]
        L.orc_nv12_pack.argtypes = [C.POINTER(FrameInfo), C.POINTER(C.c_uint8)]
        _orc = L
    return _orc


def have_ref():
    return os.path.exists(REF_LIB)


def reference():
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError(f"{REF_LIB} missing: run `make -C oracle ref`")
        L = C.CDLL(REF_LIB)
        L.ref_decoder_new.restype = C.c_void_p
        L.ref_decoder_free.argtypes = [C.c_void_p]
        L.ref_decoder_run.restype = C.c_int
        L.ref_decoder_run.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_int,
                                      C.POINTER(C.c_uint8)]
        _ref = L
    return _ref


def i420_size(w, h):
    return w * h + 2 * ((w + 1) // 2) * ((h + 1) // 2)


def _run(L, prefix, frames):
    """frames: objects with .ptr (POINTER(ParsedFrame)) and .pf.hdr."""
    n = len(frames)
    sizes = [i420_size(f.pf.hdr.width, f.pf.hdr.height) for f in frames]
    out = np.zeros(sum(sizes), np.uint8)
    arr = (C.c_void_p * n)(*[C.cast(f.ptr, C.c_void_p) for f in frames])
    d = getattr(L, prefix + "_decoder_new")()
    try:
        done = getattr(L, prefix + "_decoder_run")(
            d, arr, n, out.ctypes.data_as(C.POINTER(C.c_uint8)))
    finally:
        getattr(L, prefix + "_decoder_free")(d)
    if done != n:
        raise RuntimeError(f"{prefix} decoded {done}/{n} frames")
    res, o = [], 0
    for s in sizes:
        res.append(out[o:o + s])
        o += s
    return res


def time_per_frame(frames, use_ref=True):
    """Seconds per frame of a sequence decoded one call per frame on one
    decoder (the reference's own C, or the restatement), the reference pool
    carried from frame to frame."""
    import time
    L, prefix = (reference(), "ref") if use_ref else (oracle(), "orc")
    d = getattr(L, prefix + "_decoder_new")()
    run = getattr(L, prefix + "_decoder_run")
    lat = []
    try:
        for f in frames:
            out = np.zeros(i420_size(f.pf.hdr.width, f.pf.hdr.height), np.uint8)
            arr = (C.c_void_p * 1)(C.cast(f.ptr, C.c_void_p))
            t0 = time.perf_counter()
            done = run(d, arr, 1, out.ctypes.data_as(C.POINTER(C.c_uint8)))
            lat.append(time.perf_counter() - t0)
            if done != 1:
                raise RuntimeError(f"{prefix} decode failed")
    finally:
        getattr(L, prefix + "_decoder_free")(d)
    return lat


def run_oracle(frames):
    """Packed I420 output of every frame, CPU restatement, fresh pool."""
    return _run(oracle(), "orc", frames)


def run_reference(frames):
    """Same through the reference's own compiled reconstruction."""
    return _run(reference(), "ref", frames)


class Decoder:
    """Incremental oracle decoder (keeps the reference pool across calls)."""

    def __init__(self):
        self.L = oracle()
        self.d = self.L.orc_decoder_new()

    def decode(self, frame):
        return self.decode_many([frame])[0]

    def decode_many(self, frames):
        n = len(frames)
        sizes = [i420_size(f.pf.hdr.width, f.pf.hdr.height) for f in frames]
        out = np.zeros(sum(sizes), np.uint8)
        arr = (C.c_void_p * n)(*[C.cast(f.ptr, C.c_void_p) for f in frames])
        done = self.L.orc_decoder_run(self.d, arr, n, out.ctypes.data_as(C.POINTER(C.c_uint8)))
        if done != n:
            raise RuntimeError("oracle decode failed")
        res, o = [], 0
        for s in sizes:
            res.append(out[o:o + s])
            o += s
        return res

    def close(self):
        if self.d:
            self.L.orc_decoder_free(self.d)
            self.d = None

    def __del__(self):
        self.close()


def idct_matrix(n):
    m = np.zeros(n * n, np.float64)
    oracle().orc_idct_matrix(n, m.ctypes.data_as(C.POINTER(C.c_double)))
    return m.reshape(n, n)


def idct_2d(block):
    n = block.shape[0]
    src = np.ascontiguousarray(block, np.int16).ravel()
    dst = np.zeros(n * n, np.int16)
    oracle().orc_idct_2d(src.ctypes.data_as(C.POINTER(C.c_int16)),
                         dst.ctypes.data_as(C.POINTER(C.c_int16)), n)
    return dst.reshape(n, n)


def lf_threshold(level, sharp):
    return oracle().orc_lf_threshold(level, sharp)


def loop_filter(y, level, sharp):
    """Deblock a luma plane (2-D uint8 array) in place semantics; returns copy."""
    y = np.ascontiguousarray(y, np.uint8).copy()
    h, w = y.shape
    cw, ch = (w + 1) // 2, (h + 1) // 2
    u = np.full(cw * ch, 128, np.uint8)
    v = np.full(cw * ch, 128, np.uint8)
    f = FrameInfo(w, h, w, cw, y.ctypes.data_as(C.POINTER(C.c_uint8)),
                  u.ctypes.data_as(C.POINTER(C.c_uint8)), v.ctypes.data_as(C.POINTER(C.c_uint8)), 0)
    oracle().orc_loop_filter(C.byref(f), level, sharp)
    return y


def intra_predict(w, h, above, left, top_left, mode):
    dst = np.zeros((h, w), np.uint8)
    a = np.ascontiguousarray(above, np.uint8)
    l = np.ascontiguousarray(left, np.uint8)
    oracle().orc_intra_predict(dst.ctypes.data_as(C.POINTER(C.c_uint8)), w, w, h,
                               a.ctypes.data_as(C.POINTER(C.c_uint8)),
                               l.ctypes.data_as(C.POINTER(C.c_uint8)), top_left, mode)
    return dst


def motion_compensate(ref, w, h, bx, by, mvx, mvy, ref_w=None, ref_h=None):
    ref = np.ascontiguousarray(ref, np.uint8)
    rh, rw = ref.shape
    dst = np.zeros((h, w), np.uint8)
    oracle().orc_motion_compensate(dst.ctypes.data_as(C.POINTER(C.c_uint8)), w,
                                   ref.ctypes.data_as(C.POINTER(C.c_uint8)), rw,
                                   w, h, bx, by, mvx, mvy,
                                   rw if ref_w is None else ref_w,
                                   rh if ref_h is None else ref_h)
    return dst


def nv12_pack(i420, w, h):
    cw, ch = (w + 1) // 2, (h + 1) // 2
    buf = np.ascontiguousarray(i420, np.uint8)
    base = buf.ctypes.data
    f = FrameInfo(w, h, w, cw, C.cast(base, C.POINTER(C.c_uint8)),
                  C.cast(base + w * h, C.POINTER(C.c_uint8)),
                  C.cast(base + w * h + cw * ch, C.POINTER(C.c_uint8)), 0)
    out = np.zeros(w * h + 2 * cw * ch, np.uint8)
    oracle().orc_nv12_pack(C.byref(f), out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out
