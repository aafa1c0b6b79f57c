"""Python face of the transcoder, mirroring crnd's API (inc/crn_defs.h:139-221).

    ctx = unpack_begin(data)                 # crnd_unpack_begin
    faces = unpack_level(ctx, level)         # crnd_unpack_level (host buffers)
    unpack_end(ctx)                          # crnd_unpack_end

Functions return False / None where crnd returns false / NULL; the device
entry points raise on failure.  All transcoding runs in libcrn2's HIP kernels.
"""
import ctypes as C

import numpy as np

from ._lib import lib, LevelInfo, TextureInfo, CRN2_OK


def _buf(data):
    if isinstance(data, np.ndarray):
        return data.ctypes.data_as(C.c_void_p), data.nbytes
    b = bytes(data)
    return C.c_char_p(b), len(b)


def validate_file(data):
    p, n = _buf(data)
    return bool(lib().crn2_validate_file(p, n))


def get_texture_info(data):
    p, n = _buf(data)
    t = TextureInfo()
    if lib().crn2_get_texture_info(p, n, C.byref(t)) != CRN2_OK:
        return None
    return {k: getattr(t, k) for k, _ in TextureInfo._fields_}


def get_level_info(data, level):
    p, n = _buf(data)
    t = LevelInfo()
    if lib().crn2_get_level_info(p, n, level, C.byref(t)) != CRN2_OK:
        return None
    return {k: getattr(t, k) for k, _ in LevelInfo._fields_}


class UnpackContext:
    """Holds the crn2 context and keeps the file bytes alive (crnd requires the
    caller's data to stay valid until unpack_end)."""

    def __init__(self, handle, keep):
        self.handle = handle
        self._keep = keep


def unpack_begin(data):
    b = np.frombuffer(bytes(data), np.uint8).copy()
    h = lib().crn2_unpack_begin(b.ctypes.data_as(C.c_void_p), b.nbytes)
    if not h:
        return None
    return UnpackContext(h, b)


def unpack_level(ctx, level, row_pitch=0, dst=None, dst_size=None):
    """Transcodes `level` into host buffers; returns the list of per-face
    numpy uint8 arrays, or False when crnd_unpack_level would return false."""
    if ctx is None:
        return False
    info = get_level_info(ctx._keep, level)
    if info is None:
        return False
    pitch = row_pitch or info["blocks_x"] * info["bytes_per_block"]
    size = pitch * info["blocks_y"]
    if dst is None:
        dst = [np.zeros(size, np.uint8) for _ in range(info["faces"])]
    if dst_size is None:
        dst_size = min(d.nbytes for d in dst)
    ptrs = (C.c_void_p * 6)(*[d.ctypes.data for d in dst])
    rc = lib().crn2_unpack_level(C.c_void_p(ctx.handle), ptrs, dst_size, row_pitch, level)
    return dst if rc == CRN2_OK else False


def unpack_level_segmented(ctx, src, level, row_pitch=0):
    if ctx is None:
        return False
    info = get_level_info(ctx._keep, level)
    if info is None:
        return False
    pitch = row_pitch or info["blocks_x"] * info["bytes_per_block"]
    size = pitch * info["blocks_y"]
    dst = [np.zeros(size, np.uint8) for _ in range(info["faces"])]
    ptrs = (C.c_void_p * 6)(*[d.ctypes.data for d in dst])
    s = np.frombuffer(bytes(src), np.uint8).copy()
    rc = lib().crn2_unpack_level_segmented(C.c_void_p(ctx.handle), s.ctypes.data_as(C.c_void_p), s.nbytes, ptrs,
                                           size, row_pitch, level)
    return dst if rc == CRN2_OK else False


def unpack_end(ctx):
    if ctx is None or not ctx.handle:
        return False
    rc = lib().crn2_unpack_end(C.c_void_p(ctx.handle))
    ctx.handle = None
    return rc == CRN2_OK


def levels_size(ctx, first, count):
    return int(lib().crn2_levels_size(C.c_void_p(ctx.handle), first, count))


def unpack_levels_device(ctx, dst_ptr, dst_size, first, count, stream=0):
    """Queues the transcode of levels [first, first+count) into device memory
    at dst_ptr (levels back to back, faces back to back, minimal pitch)."""
    rc = lib().crn2_unpack_levels_device(C.c_void_p(ctx.handle), C.c_void_p(dst_ptr), dst_size, first, count,
                                         C.c_void_p(stream))
    if rc != CRN2_OK:
        raise RuntimeError("crn2_unpack_levels_device failed: %d" % rc)


def crn_decompress(data, first_level, level_count):
    """emscripten/crunch_lib.cpp's crn_decompress: levels packed back to back."""
    b = np.frombuffer(bytes(data), np.uint8).copy()
    L = lib()
    size = sum(L.crn_get_uncompressed_size(b.ctypes.data_as(C.c_void_p), b.nbytes, l)
               for l in range(first_level, first_level + level_count))
    out = np.zeros(max(size, 1), np.uint8)
    L.crn_decompress(b.ctypes.data_as(C.c_void_p), b.nbytes, out.ctypes.data_as(C.c_void_p), size, first_level,
                     level_count)
    return out[:size]
