"""compute_vp9_amd — Python bindings of the MI355X VP9 reconstruction backend.

The product is the native library ``libcvp9_hip.so`` (HIP kernels for gfx950
plus the C-ABI host side, declared in ``include/cvp9_hip/*.h``). This module
only loads it with ctypes and mirrors its structs, so tests and the benchmark
drive exactly the entry points a C caller (the reference decoder's
``src/decoder/vp9_decoder.c`` dispatch) would use.

The directory name is not a Python identifier; load it with
``load_package()`` from ``_pkgload.py`` at the repository root, or import it
after registering it under the name ``compute_vp9_amd``.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# CVP9_HIP_LIB names another build of the library (the sanitizer build of
# the CPU test run); by default the in-tree one
LIB_PATH = os.environ.get("CVP9_HIP_LIB") or os.path.join(PKG_DIR, "libcvp9_hip.so")

# ---- error codes and enums (include/compute_vp9/decoder.h:21-37) ---------
CVP9_OK = 0
CVP9_ERR_NOMEM = -1
CVP9_ERR_NO_BACKEND = -2
CVP9_ERR_INVALID_DATA = -3
CVP9_ERR_UNSUPPORTED = -4
CVP9_ERR_GPU = -5
CVP9_ERR_AGAIN = -6
ERR_NAMES = {0: "OK", -1: "NOMEM", -2: "NO_BACKEND", -3: "INVALID_DATA",
             -4: "UNSUPPORTED", -5: "GPU", -6: "AGAIN"}

VP9_FRAME_KEY = 0
VP9_FRAME_NON_KEY = 1
CVP9_HIP_INFLIGHT = 3
CVP9_HIP_DEBUG_NO_DEBLOCK = 1
CVP9_HIP_DEBUG_SPIN_ONE = 2


class Cvp9Error(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"{what}: {ERR_NAMES.get(code, code)}")
        self.code = code


# ---- structs (layouts pinned against the reference headers in tests) ------
class FrameHeader(C.Structure):
    """vp9_frame_header_t (src/decoder/vp9_bitstream.h:39-87), 76 bytes."""
    _fields_ = [
        ("frame_type", C.c_int), ("width", C.c_uint32), ("height", C.c_uint32),
        ("render_width", C.c_uint32), ("render_height", C.c_uint32),
        ("profile", C.c_uint8), ("bit_depth", C.c_uint8), ("color_space", C.c_int),
        ("color_range", C.c_bool), ("subsampling_x", C.c_uint8),
        ("subsampling_y", C.c_uint8), ("show_frame", C.c_bool),
        ("show_existing_frame", C.c_bool), ("frame_to_show_map_idx", C.c_uint8),
        ("intra_only", C.c_bool), ("reset_frame_context", C.c_uint8),
        ("error_resilient", C.c_bool), ("allow_high_precision_mv", C.c_bool),
        ("interp_filter", C.c_int), ("refresh_frame_context", C.c_bool),
        ("frame_parallel", C.c_bool), ("frame_context_idx", C.c_uint8),
        ("log2_tile_cols", C.c_uint8), ("log2_tile_rows", C.c_uint8),
        ("uncompressed_header_bytes", C.c_uint32), ("first_partition_size", C.c_uint16),
        ("base_qindex", C.c_int16), ("y_dc_delta_q", C.c_int8),
        ("uv_dc_delta_q", C.c_int8), ("uv_ac_delta_q", C.c_int8),
        ("filter_level", C.c_uint8), ("sharpness_level", C.c_uint8),
        ("refresh_frame_flags", C.c_uint8), ("ref_frame_idx", C.c_uint8 * 3),
        ("ref_frame_sign_bias", C.c_uint8 * 4), ("segmentation_enabled", C.c_bool),
    ]


class BlockInfo(C.Structure):
    """vp9_macroblock_info_t (src/decoder/vp9_parsed_frame.h:20-33), 32 bytes."""
    _fields_ = [
        ("x", C.c_uint16), ("y", C.c_uint16), ("width", C.c_uint8),
        ("height", C.c_uint8), ("tx_size", C.c_uint8), ("is_intra", C.c_uint8),
        ("y_mode", C.c_uint8), ("uv_mode", C.c_uint8), ("ref_frame", C.c_uint8 * 2),
        ("mv", (C.c_int16 * 2) * 2), ("skip", C.c_uint8),
        ("coeff_offset", C.c_uint32), ("eob", C.c_uint16),
    ]


BLOCK_DTYPE = np.dtype({
    "names": ["x", "y", "width", "height", "tx_size", "is_intra", "y_mode", "uv_mode",
              "ref_frame", "mv", "skip", "coeff_offset", "eob"],
    "formats": ["<u2", "<u2", "u1", "u1", "u1", "u1", "u1", "u1", ("u1", 2),
                ("<i2", (2, 2)), "u1", "<u4", "<u2"],
    "offsets": [0, 2, 4, 5, 6, 7, 8, 9, 10, 12, 20, 24, 28],
    "itemsize": 32,
})


class ParsedFrame(C.Structure):
    """vp9_parsed_frame_t (src/decoder/vp9_parsed_frame.h:36-63), 160 bytes."""
    _fields_ = [
        ("hdr", FrameHeader), ("blocks", C.POINTER(BlockInfo)),
        ("num_blocks", C.c_uint32), ("blocks_capacity", C.c_uint32),
        ("coeffs", C.POINTER(C.c_int16)), ("num_coeffs", C.c_uint32),
        ("coeffs_capacity", C.c_uint32), ("mv_grid", C.c_void_p),
        ("mv_grid_width", C.c_uint32), ("mv_grid_height", C.c_uint32),
        ("mi_width_grid", C.c_void_p), ("mi_height_grid", C.c_void_p),
        ("mi_grid_width", C.c_uint32), ("mi_grid_height", C.c_uint32),
        ("mi_block_grid", C.c_void_p),
    ]


class FrameInfo(C.Structure):
    """cvp9_frame_info_t (include/compute_vp9/decoder.h:47-56)."""
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("stride_y", C.c_uint32),
        ("stride_uv", C.c_uint32), ("plane_y", C.POINTER(C.c_uint8)),
        ("plane_u", C.POINTER(C.c_uint8)), ("plane_v", C.POINTER(C.c_uint8)),
        ("pts", C.c_int64),
    ]


class DmabufFrame(C.Structure):
    """cvp9_dmabuf_frame_t (include/compute_vp9/decoder.h:126-134)."""
    _fields_ = [
        ("fd", C.c_int), ("size", C.c_uint64), ("width", C.c_uint32),
        ("height", C.c_uint32), ("offsets", C.c_uint32 * 3),
        ("pitches", C.c_uint32 * 3), ("pts", C.c_int64),
    ]


class ExportBuffer(C.Structure):
    """cvp9_export_buffer_t (include/compute_vp9/decoder.h:152-158)."""
    _fields_ = [
        ("fd", C.c_int), ("mapped", C.c_void_p), ("size", C.c_uint64),
        ("priv_buf", C.c_void_p), ("priv_mem", C.c_void_p),
    ]


class FrameBuf(C.Structure):
    """cvp9_frame_buf_t (include/cvp9_hip/engine.h): a device frame buffer."""
    _fields_ = [("base", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32)]


class Tune(C.Structure):
    """cvp9_tune_t (include/cvp9_hip/engine.h)."""
    _fields_ = [("res_split", C.c_uint8 * 4), ("intra_lanes", C.c_uint16),
                ("intra_workers", C.c_uint32), ("intra_level_min", C.c_uint32),
                ("intra_defer", C.c_uint32), ("intra_chains", C.c_uint32)]

    def as_dict(self):
        return {"res_split": list(self.res_split), "intra_lanes": self.intra_lanes,
                "intra_workers": self.intra_workers, "intra_level_min": self.intra_level_min,
                "intra_defer": self.intra_defer, "intra_chains": self.intra_chains}


class SynthParams(C.Structure):
    """cvp9_synth_params_t (include/cvp9_hip/synth.h)."""
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("seed", C.c_uint64),
        ("split_permille", C.c_uint32), ("intra_permille", C.c_uint32),
        ("skip_permille", C.c_uint32), ("tx32_permille", C.c_uint32),
        ("coef_fixed", C.c_uint32), ("coef_permille", C.c_uint32),
        ("lap_scale", C.c_uint32), ("mv_range", C.c_int32),
        ("filter_level", C.c_int32), ("sharpness", C.c_uint32),
        ("key_interval", C.c_uint32), ("qindex_lo", C.c_int32), ("qindex_hi", C.c_int32),
    ]


PF_P = C.POINTER(ParsedFrame)

# ---- exported symbols, as declared in include/cvp9_hip/*.h ----------------
SIGNATURES = {
    # hip_backend.h
    "hip_backend_init": (C.c_int, [C.POINTER(C.c_void_p)]),
    "hip_backend_destroy": (None, [C.c_void_p]),
    "hip_decode_frame": (C.c_int, [C.c_void_p, PF_P, C.c_int64]),
    "hip_get_frame": (C.c_int, [C.c_void_p, C.POINTER(FrameInfo), C.c_int]),
    "hip_get_frame_dmabuf": (C.c_int, [C.c_void_p, C.POINTER(DmabufFrame)]),
    "hip_export_buffer_alloc": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(ExportBuffer)]),
    "hip_export_buffer_free": (None, [C.c_void_p, C.POINTER(ExportBuffer)]),
    "hip_set_render_target": (C.c_int, [C.c_void_p, C.POINTER(ExportBuffer)]),
    "hip_backend_read_ref": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(FrameInfo)]),
    # synth.h
    "cvp9_synth_preset": (C.c_int, [C.POINTER(SynthParams), C.c_int, C.c_uint32,
                                    C.c_uint32, C.c_uint64]),
    "cvp9_synth_frame": (PF_P, [C.POINTER(SynthParams), C.c_uint32]),
    "cvp9_synth_free": (None, [PF_P]),
    "cvp9_synth_algo_bytes": (C.c_uint64, [PF_P, C.c_int]),
    # engine.h
    "cvp9_engine_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "cvp9_engine_destroy": (None, [C.c_void_p]),
    "cvp9_engine_stream_open": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "cvp9_engine_stream_close": (None, [C.c_void_p, C.c_int]),
    "cvp9_engine_stream_bytes": (C.c_uint64, [C.c_void_p, C.c_int]),
    "cvp9_engine_stream_error": (C.c_uint32, [C.c_void_p, C.c_int]),
    "cvp9_engine_prepare": (C.c_int, [C.c_void_p, PF_P, C.POINTER(C.c_void_p)]),
    "cvp9_engine_prep_free": (None, [C.c_void_p, C.c_void_p]),
    "cvp9_engine_prep_bytes": (C.c_uint64, [C.c_void_p]),
    "cvp9_engine_run": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                  C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "cvp9_engine_sync": (C.c_int, [C.c_void_p]),
    "cvp9_engine_read": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint8), C.c_uint64,
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "cvp9_engine_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "cvp9_engine_get_tune": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cvp9_engine_set_tune": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cvp9_engine_compute_units": (C.c_int, [C.c_void_p]),
    "cvp9_engine_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                       C.POINTER(C.c_void_p), C.c_int, C.c_void_p]),
    "cvp9_engine_stage_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double),
                                          C.POINTER(C.c_uint32), C.c_int]),
    "cvp9_engine_stream_reset": (C.c_int, [C.c_void_p, C.c_int]),
    "cvp9_engine_set_fused": (C.c_int, [C.c_void_p, C.c_int]),
    "cvp9_engine_set_queues": (C.c_int, [C.c_void_p, C.c_int]),
    "cvp9_engine_set_pipeline": (C.c_int, [C.c_void_p, C.c_int]),
    "cvp9_engine_set_spin_limit": (C.c_int, [C.c_void_p, C.c_uint32]),
    "cvp9_engine_capture_begin": (C.c_int, [C.c_void_p]),
    "cvp9_engine_capture_end": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "cvp9_engine_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cvp9_engine_graph_free": (None, [C.c_void_p, C.c_void_p]),
    "cvp9_engine_hip_stream": (C.c_void_p, [C.c_void_p]),
    # server.h
    "cvp9_server_create": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "cvp9_server_destroy": (None, [C.c_void_p]),
    "cvp9_server_stream_cost": (C.c_uint64, [C.c_uint32, C.c_uint32]),
    "cvp9_server_open": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_int)]),
    "cvp9_server_close": (None, [C.c_void_p, C.c_int]),
    "cvp9_server_submit": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64]),
    "cvp9_server_submit_n": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                       C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]),
    "cvp9_server_pump": (C.c_int, [C.c_void_p]),
    "cvp9_server_poll": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                   C.c_int]),
    "cvp9_server_pending": (C.c_uint32, [C.c_void_p]),
    "cvp9_server_read_nv12": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64, C.POINTER(C.c_uint8),
                                        C.c_uint64]),
    "cvp9_server_drain": (C.c_int, [C.c_void_p]),
    # testing.h
    "cvp9_hip_version": (C.c_char_p, []),
    "cvp9_hip_device_count": (C.c_int, []),
    "cvp9_hip_idct_table": (C.c_int, [C.c_int, C.POINTER(C.c_double)]),
    "cvp9_hip_plan_check": (C.c_int, [PF_P]),
    "cvp9_hip_plan_image": (C.c_int64, [PF_P, C.c_int, C.POINTER(C.c_uint8), C.c_uint64]),
    "cvp9_hip_plan_image_compact": (C.c_int64, [PF_P, C.c_int, C.POINTER(C.c_uint8), C.c_uint64,
                                                C.POINTER(C.c_uint64)]),
    "cvp9_hip_route_check": (C.c_int, [C.POINTER(PF_P), C.c_int, C.POINTER(C.c_int)]),
    "cvp9_hip_intra_order": (C.c_int, [PF_P, C.POINTER(C.c_uint32), C.c_uint32]),
    "cvp9_hip_intra_levels": (C.c_int, [PF_P, C.POINTER(C.c_uint32), C.c_uint32]),
    "cvp9_hip_intra_chains": (C.c_int, [PF_P, C.POINTER(C.c_uint32), C.c_uint32]),
    "cvp9_hip_export_read": (C.c_int, [C.c_void_p, C.POINTER(ExportBuffer), C.POINTER(C.c_uint8),
                                       C.c_uint64]),
    "cvp9_hip_dmabuf_read": (C.c_int, [C.c_int, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint8),
                                       C.c_uint64]),
    "hip_backend_set_debug": (C.c_int, [C.c_void_p, C.c_uint32]),
    "cvp9_hip_copy_probe": (C.c_double, [C.c_uint64, C.c_int]),
    "cvp9_hip_dev_alloc": (C.c_void_p, [C.c_uint64]),
    "cvp9_hip_dev_free": (None, [C.c_void_p]),
    "cvp9_hip_dev_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint8), C.c_uint64]),
    "cvp9_frame_bytes": (C.c_uint64, [C.c_uint32, C.c_uint32]),
    "cvp9_engine_run_static": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]),
}

_lib = None


def lib():
    """The native library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code, what):
    if code != CVP9_OK:
        raise Cvp9Error(code, what)


# ---- generator -------------------------------------------------------------
def synth_params(preset, width, height, seed, **over):
    p = SynthParams()
    if lib().cvp9_synth_preset(C.byref(p), preset, width, height, seed) != 0:
        raise ValueError(f"unknown preset {preset}")
    for k, v in over.items():
        setattr(p, k, v)
    return p


class SynthFrame:
    """Owned vp9_parsed_frame_t from the generator."""

    def __init__(self, params, index):
        self.ptr = lib().cvp9_synth_frame(C.byref(params), index)
        if not self.ptr:
            raise MemoryError("cvp9_synth_frame failed")

    @property
    def pf(self):
        return self.ptr.contents

    def blocks(self):
        n = self.pf.num_blocks
        if n == 0:
            return np.zeros(0, BLOCK_DTYPE)
        buf = (C.c_uint8 * (32 * n)).from_address(C.addressof(self.pf.blocks.contents))
        # copy the raw bytes (a structured-array copy drops padding bytes)
        return np.frombuffer(bytearray(buf), BLOCK_DTYPE)

    def coeffs(self):
        n = self.pf.num_coeffs
        if n == 0:
            return np.zeros(0, np.int16)
        return np.ctypeslib.as_array(self.pf.coeffs, shape=(n,)).copy()

    def algo_bytes(self, nv12=False):
        return int(lib().cvp9_synth_algo_bytes(self.ptr, 1 if nv12 else 0))

    def close(self):
        if self.ptr:
            lib().cvp9_synth_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


class OwnedFrame:
    """A parsed frame built from numpy arrays (hand-made test cases)."""

    def __init__(self, width, height, blocks, coeffs, frame_type=VP9_FRAME_NON_KEY,
                 filter_level=0, sharpness=0):
        self._blocks = np.ascontiguousarray(blocks, dtype=BLOCK_DTYPE)
        self._coeffs = np.ascontiguousarray(coeffs, dtype=np.int16)
        if self._coeffs.size == 0:
            self._coeffs = np.zeros(1, np.int16)
            ncoef = 0
        else:
            ncoef = self._coeffs.size
        f = ParsedFrame()
        f.hdr.frame_type = frame_type
        f.hdr.width = width
        f.hdr.height = height
        f.hdr.filter_level = filter_level
        f.hdr.sharpness_level = sharpness
        f.blocks = self._blocks.ctypes.data_as(C.POINTER(BlockInfo))
        f.num_blocks = f.blocks_capacity = self._blocks.size
        f.coeffs = self._coeffs.ctypes.data_as(C.POINTER(C.c_int16))
        f.num_coeffs = ncoef
        f.coeffs_capacity = self._coeffs.size
        self._f = f
        self.ptr = C.pointer(f)

    @property
    def pf(self):
        return self._f


def intra_levels(frame):
    """Order positions where each intra dependency level of a parsed frame
    starts, plus the intra block count (numpy uint32, depth + 1 entries)."""
    n = lib().cvp9_hip_intra_levels(frame.ptr, None, 0)
    check(min(n, 0), "intra_levels")
    buf = np.zeros(n + 1, np.uint32)
    if n:
        check(min(lib().cvp9_hip_intra_levels(frame.ptr, buf.ctypes.data_as(C.POINTER(C.c_uint32)),
                                               buf.size), 0), "intra_levels")
    return buf if n else np.zeros(0, np.uint32)


def frame_info_to_i420(info):
    """Packed I420 bytes of a cvp9_frame_info_t (honours the strides)."""
    w, h = info.width, info.height
    cw, ch = (w + 1) // 2, (h + 1) // 2

    def plane(p, stride, pw, ph):
        a = np.ctypeslib.as_array(p, shape=(stride * (ph - 1) + pw,))
        out = np.empty((ph, pw), np.uint8)
        for r in range(ph):
            out[r] = a[r * stride:r * stride + pw]
        return out

    y = plane(info.plane_y, info.stride_y, w, h)
    u = plane(info.plane_u, info.stride_uv, cw, ch)
    v = plane(info.plane_v, info.stride_uv, cw, ch)
    return np.concatenate([y.ravel(), u.ravel(), v.ravel()])


def i420_size(w, h):
    return w * h + 2 * ((w + 1) // 2) * ((h + 1) // 2)


class DeviceFrame:
    """A zero-filled device frame buffer in the engine's layout (test use)."""

    def __init__(self, w, h):
        self.w, self.h = w, h
        self.nbytes = int(lib().cvp9_frame_bytes(w, h))
        self.ptr = lib().cvp9_hip_dev_alloc(self.nbytes)
        if not self.ptr:
            raise MemoryError("cvp9_hip_dev_alloc failed")
        self.buf = FrameBuf(self.ptr, w, h)

    def read_i420(self):
        """Packed I420 (Y W*H, U, V ceil(W/2)*ceil(H/2)) of the buffer."""
        raw = np.zeros(self.nbytes, np.uint8)
        check(lib().cvp9_hip_dev_read(self.ptr, raw.ctypes.data_as(C.POINTER(C.c_uint8)),
                                      self.nbytes), "cvp9_hip_dev_read")
        w, h = self.w, self.h
        cw, ch = (w + 1) // 2, (h + 1) // 2
        sy, suv = (w + 63) // 64 * 64, (cw + 63) // 64 * 64
        y = raw[:sy * h].reshape(h, sy)[:, :w]
        u = raw[sy * h:sy * h + suv * ch].reshape(ch, suv)[:, :cw]
        v = raw[sy * h + suv * ch:sy * h + 2 * suv * ch].reshape(ch, suv)[:, :cw]
        return np.concatenate([y.ravel(), u.ravel(), v.ravel()])

    def free(self):
        if self.ptr:
            lib().cvp9_hip_dev_free(self.ptr)
            self.ptr = None


def run_static(prep, refs, out, nv12=None, status=None, stream=None):
    """cvp9_engine_run_static: refs is a list of 8 DeviceFrame or None."""
    arr = (FrameBuf * 8)()
    for i, r in enumerate(refs or []):
        if r is not None:
            arr[i] = r.buf
    return lib().cvp9_engine_run_static(prep, C.cast(arr, C.c_void_p), C.cast(C.pointer(out.buf), C.c_void_p),
                                        nv12, status, stream)


def plan_image(frame, threads=0):
    """The planner's device image of a frame (cvp9_hip_plan_image)."""
    n = lib().cvp9_hip_plan_image(frame.ptr, threads, None, 0)
    if n < 0:
        raise Cvp9Error(int(n), "cvp9_hip_plan_image")
    out = np.zeros(n, np.uint8)
    check(int(lib().cvp9_hip_plan_image(frame.ptr, threads, out.ctypes.data_as(C.POINTER(C.c_uint8)), n)
              >= 0) - 1, "cvp9_hip_plan_image")
    return out


def plan_image_compact(frame, threads=0):
    """The host part of the frame's device image with compact coefficients
    (cvp9_hip_plan_image_compact): (image bytes, info dict)."""
    info = (C.c_uint64 * 8)()
    n = lib().cvp9_hip_plan_image_compact(frame.ptr, threads, None, 0, info)
    if n < 0:
        raise Cvp9Error(int(n), "cvp9_hip_plan_image_compact")
    out = np.zeros(n, np.uint8)
    r = lib().cvp9_hip_plan_image_compact(frame.ptr, threads, out.ctypes.data_as(C.POINTER(C.c_uint8)),
                                          n, info)
    if r < 0:
        raise Cvp9Error(int(r), "cvp9_hip_plan_image_compact")
    keys = ("off_cmask", "off_cstart", "off_cvals", "nnz", "n_chunks", "upload", "off_coeffs",
            "bytes")
    return out, dict(zip(keys, (int(x) for x in info)))


def route_check(frames):
    """Codes hip_decode_frame would return for a sequence before any device
    work (validator + the decline-until-KEY rule); needs no device."""
    n = len(frames)
    arr = (PF_P * n)(*[f.ptr for f in frames])
    codes = (C.c_int * n)()
    r = lib().cvp9_hip_route_check(arr, n, codes)
    if r != n:
        raise Cvp9Error(r, "cvp9_hip_route_check")
    return list(codes)


# ---- HIP backend (include/cvp9_hip/hip_backend.h) ----------------------------
class HipBackend:
    """Thin mirror of the eight hip_* entry points."""

    def __init__(self):
        ctx = C.c_void_p()
        check(lib().hip_backend_init(C.byref(ctx)), "hip_backend_init")
        self.ctx = ctx

    def decode(self, frame, pts=0):
        """Returns the cvp9_err_t code (AGAIN/UNSUPPORTED are not raised)."""
        return lib().hip_decode_frame(self.ctx, frame.ptr, pts)

    def get_frame(self, wait=True):
        """(code, packed I420 numpy array or None, FrameInfo)."""
        info = FrameInfo()
        rc = lib().hip_get_frame(self.ctx, C.byref(info), 1 if wait else 0)
        if rc != CVP9_OK:
            return rc, None, info
        return rc, frame_info_to_i420(info), info

    def set_debug(self, flags):
        check(lib().hip_backend_set_debug(self.ctx, flags), "hip_backend_set_debug")

    def read_ref(self, slot):
        """Reference slot `slot` of the device pool as packed I420, or None
        when the slot is empty (hip_backend_read_ref)."""
        info = FrameInfo()
        check(lib().hip_backend_read_ref(self.ctx, slot, C.byref(info)), "hip_backend_read_ref")
        if info.width == 0:
            return None
        w, h = info.width, info.height
        cw, ch = (w + 1) // 2, (h + 1) // 2
        out = np.zeros(w * h + 2 * cw * ch, np.uint8)
        base = out.ctypes.data
        info.stride_y, info.stride_uv = w, cw
        info.plane_y = C.cast(base, C.POINTER(C.c_uint8))
        info.plane_u = C.cast(base + w * h, C.POINTER(C.c_uint8))
        info.plane_v = C.cast(base + w * h + cw * ch, C.POINTER(C.c_uint8))
        check(lib().hip_backend_read_ref(self.ctx, slot, C.byref(info)), "hip_backend_read_ref")
        return out

    def export_buffer_alloc(self, size):
        """(code, ExportBuffer): hip_export_buffer_alloc."""
        buf = ExportBuffer()
        return lib().hip_export_buffer_alloc(self.ctx, size, C.byref(buf)), buf

    def export_buffer_free(self, buf):
        lib().hip_export_buffer_free(self.ctx, C.byref(buf))

    def set_render_target(self, buf):
        """hip_set_render_target (None clears a pending target)."""
        return lib().hip_set_render_target(self.ctx, C.byref(buf) if buf is not None else None)

    def get_frame_dmabuf(self):
        """(code, DmabufFrame): the caller closes the returned fd."""
        f = DmabufFrame()
        return lib().hip_get_frame_dmabuf(self.ctx, C.byref(f)), f

    def export_read(self, buf, n):
        """First n bytes of an export buffer, read back through HIP."""
        out = np.zeros(n, np.uint8)
        check(lib().cvp9_hip_export_read(self.ctx, C.byref(buf),
                                         out.ctypes.data_as(C.POINTER(C.c_uint8)), n), "export_read")
        return out

    def close(self):
        if self.ctx:
            lib().hip_backend_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- multi-stream engine (include/cvp9_hip/engine.h) ------------------------------
STAGE_NAMES = ["fill", "owner", "mc", "residual4", "residual8", "residual16", "residual32",
               "neighbours", "intra", "deblock", "nv12", "fused", "expand"]


class Engine:
    """Mirror of the cvp9_engine_* entry points."""

    def __init__(self, device=0):
        e = C.c_void_p()
        check(lib().cvp9_engine_create(device, C.byref(e)), "cvp9_engine_create")
        self.e = e
        self._preps = []

    def open_stream(self):
        sid = C.c_int()
        check(lib().cvp9_engine_stream_open(self.e, C.byref(sid)), "stream_open")
        return sid.value

    def close_stream(self, sid):
        lib().cvp9_engine_stream_close(self.e, sid)

    def reset_stream(self, sid):
        check(lib().cvp9_engine_stream_reset(self.e, sid), "stream_reset")

    def stream_bytes(self, sid):
        return int(lib().cvp9_engine_stream_bytes(self.e, sid))

    def stream_error(self, sid):
        return int(lib().cvp9_engine_stream_error(self.e, sid))

    def prepare(self, frame):
        p = C.c_void_p()
        check(lib().cvp9_engine_prepare(self.e, frame.ptr, C.byref(p)), "cvp9_engine_prepare")
        return p

    def free_prep(self, p):
        lib().cvp9_engine_prep_free(self.e, p)

    @staticmethod
    def prep_bytes(p):
        return int(lib().cvp9_engine_prep_bytes(p))

    def run(self, streams, preps, nv12=None):
        n = len(streams)
        s = (C.c_int * n)(*streams)
        p = (C.c_void_p * n)(*[x.value if isinstance(x, C.c_void_p) else x for x in preps])
        nv = None if nv12 is None else (C.c_void_p * n)(*nv12)
        check(lib().cvp9_engine_run(self.e, n, s, p, nv), "cvp9_engine_run")

    def sync(self):
        check(lib().cvp9_engine_sync(self.e), "cvp9_engine_sync")

    def read(self, sid, w, h):
        out = np.zeros(i420_size(w, h), np.uint8)
        ww, hh = C.c_uint32(), C.c_uint32()
        check(lib().cvp9_engine_read(self.e, sid, out.ctypes.data_as(C.POINTER(C.c_uint8)),
                                     out.size, C.byref(ww), C.byref(hh)), "cvp9_engine_read")
        assert (ww.value, hh.value) == (w, h)
        return out

    def get_tune(self):
        t = Tune()
        check(lib().cvp9_engine_get_tune(self.e, C.byref(t)), "get_tune")
        return t

    def set_tune(self, res_split, intra_lanes, intra_workers, intra_level_min=0, intra_defer=0,
                 intra_chains=0):
        t = Tune((C.c_uint8 * 4)(*res_split), intra_lanes, intra_workers, intra_level_min,
                 intra_defer, intra_chains)
        check(lib().cvp9_engine_set_tune(self.e, C.byref(t)), "set_tune")

    def autotune(self, streams, preps, iters=3):
        n = len(streams)
        s = (C.c_int * n)(*streams)
        p = (C.c_void_p * n)(*[x.value if isinstance(x, C.c_void_p) else x for x in preps])
        t = Tune()
        check(lib().cvp9_engine_autotune(self.e, n, s, p, iters, C.byref(t)), "autotune")
        return t

    def set_timing(self, on):
        check(lib().cvp9_engine_set_timing(self.e, 1 if on else 0), "set_timing")

    def stage_stats(self, reset=True):
        k = len(STAGE_NAMES)
        ms = (C.c_double * k)()
        n = (C.c_uint32 * k)()
        check(lib().cvp9_engine_stage_stats(self.e, ms, n, 1 if reset else 0), "stage_stats")
        return {STAGE_NAMES[i]: (ms[i], n[i]) for i in range(k)}

    def set_fused(self, on):
        check(lib().cvp9_engine_set_fused(self.e, 1 if on else 0), "set_fused")

    def set_queues(self, n):
        check(lib().cvp9_engine_set_queues(self.e, n), "set_queues")

    def set_pipeline(self, on):
        check(lib().cvp9_engine_set_pipeline(self.e, int(on)), "set_pipeline")

    def set_spin_limit(self, polls):
        check(lib().cvp9_engine_set_spin_limit(self.e, polls), "set_spin_limit")

    def capture_begin(self):
        check(lib().cvp9_engine_capture_begin(self.e), "capture_begin")

    def capture_end(self):
        g = C.c_void_p()
        check(lib().cvp9_engine_capture_end(self.e, C.byref(g)), "capture_end")
        return g

    def graph_launch(self, g):
        check(lib().cvp9_engine_graph_launch(self.e, g), "graph_launch")

    def graph_free(self, g):
        lib().cvp9_engine_graph_free(self.e, g)

    def close(self):
        if self.e:
            lib().cvp9_engine_destroy(self.e)
            self.e = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ---- continuous-batching server (include/cvp9_hip/server.h) ----------------------
class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_batch", C.c_uint32),
                ("max_inflight", C.c_uint32)]


def stream_cost(w, h):
    return int(lib().cvp9_server_stream_cost(w, h))


class Server:
    """Mirror of the cvp9_server_* entry points."""

    def __init__(self, engine, hbm_budget, max_batch=4096, max_inflight=2):
        s = C.c_void_p()
        cfg = ServerConfig(hbm_budget, max_batch, max_inflight)
        check(lib().cvp9_server_create(engine.e, C.byref(cfg), C.byref(s)), "server_create")
        self.s = s
        self._tags = (C.c_uint64 * 4096)()
        self._ns = (C.c_uint64 * 4096)()

    def open(self, w, h):
        """Stream handle, or None when the HBM budget is full."""
        sid = C.c_int()
        rc = lib().cvp9_server_open(self.s, w, h, C.byref(sid))
        if rc == CVP9_ERR_AGAIN:
            return None
        check(rc, "server_open")
        return sid.value

    def close(self, sid):
        lib().cvp9_server_close(self.s, sid)

    def submit(self, sid, prep, tag):
        check(lib().cvp9_server_submit(self.s, sid, prep, tag), "server_submit")

    def submit_many(self, sids, preps, tags):
        """cvp9_server_submit_n: frames of several streams in one call."""
        n = len(sids)
        if n == 0:
            return
        check(lib().cvp9_server_submit_n(self.s, n, (C.c_int * n)(*sids),
                                         (C.c_void_p * n)(*[p.value if hasattr(p, "value") else p
                                                             for p in preps]),
                                         (C.c_uint64 * n)(*tags)), "server_submit_n")

    def pump(self):
        r = lib().cvp9_server_pump(self.s)
        if r < 0:
            raise Cvp9Error(r, "server_pump")
        return r

    def poll(self):
        n = lib().cvp9_server_poll(self.s, self._tags, self._ns, 4096)
        if n < 0:
            raise Cvp9Error(n, "server_poll")
        return [(self._tags[i], self._ns[i]) for i in range(n)]

    def pending(self):
        return int(lib().cvp9_server_pending(self.s))

    def read_nv12(self, sid, frame, w, h):
        out = np.zeros(i420_size(w, h), np.uint8)
        check(lib().cvp9_server_read_nv12(self.s, sid, frame,
                                          out.ctypes.data_as(C.POINTER(C.c_uint8)), out.size),
              "server_read_nv12")
        return out

    def drain(self):
        check(lib().cvp9_server_drain(self.s), "server_drain")

    def close_server(self):
        if self.s:
            lib().cvp9_server_destroy(self.s)
            self.s = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close_server()
