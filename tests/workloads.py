"""Test-side description of the benchmark workloads (axis requests per workload) and
a restatement of the reference's integer conversion constants.

Request derivation follows the reference plugin: luma requests use src_width =
width, src_left = 0 (vs_plugin.cpp:368-372); 4:2:0 chroma requests use
shift = 0.25 - 0.25 * dw / sw + src_left * cw / sw and active = src_width * cw / sw
horizontally, shift = src_top * ch / sh vertically (vs_plugin.cpp:656-705).
Integer conversion: vs_plugin.cpp:761-781.
"""
import math

from support import (BICUBIC, BILINEAR, F64_AUTO, LANCZOS, SPLINE36, SPLINE64, SYMMETRIC,
                     Conversion, make_request)


def round_half_up(x):
    return int(math.floor(x + 0.5))


def width_for_height(h, even=False):
    w = round_half_up(h * 16.0 / 9.0)
    if even and w % 2:
        w += 1
    return w


def luma_requests(sw, sh, dw, dh, kernel, taps=3, b=0.0, c=0.5, blur=1.0, border=SYMMETRIC,
                  f64_mode=F64_AUTO):
    kw = dict(kernel=kernel, taps=taps, b=b, c=c, blur=blur, border=border, f64_mode=f64_mode)
    return [make_request(sw, dw, float(dw), 0.0, **kw), make_request(sh, dh, float(dh), 0.0, **kw)]


def chroma420_requests(sw, sh, dw, dh, kernel, taps=3, b=0.0, c=0.5, blur=1.0,
                       border=SYMMETRIC, f64_mode=F64_AUTO):
    kw = dict(kernel=kernel, taps=taps, b=b, c=c, blur=blur, border=border, f64_mode=f64_mode)
    cw, ch = sw >> 1, sh >> 1
    shift_h = 0.25 - 0.25 * float(dw) / float(sw) + 0.0 * float(cw) / float(sw)
    active_w = float(dw) * float(cw) / float(sw)
    shift_v = 0.0 * float(ch) / float(sh)
    active_h = float(dh) * float(ch) / float(sh)
    return [make_request(cw, dw >> 1, active_w, shift_h, **kw),
            make_request(ch, dh >> 1, active_h, shift_v, **kw)]


def workload_plan_requests(workload):
    if workload == "debilinear_1080p_720p":
        return luma_requests(1920, 1080, 1280, 720, BILINEAR)
    if workload == "yuv420p16_debicubic_1080p_720p":
        return (luma_requests(1920, 1080, 1280, 720, BICUBIC, b=0.0, c=0.5)
                + chroma420_requests(1920, 1080, 1280, 720, BICUBIC, b=0.0, c=0.5))
    if workload == "near_identity_blur09":
        out = []
        for kernel, taps in ((LANCZOS, 4), (SPLINE64, 3)):
            for dw, dh in ((1904, 1071), (960, 540)):
                out += luma_requests(1920, 1080, dw, dh, kernel, taps=taps, blur=0.9)
        return out
    if workload in ("mitchell_501_heights", "mitchell_21_heights"):
        heights = range(500, 1001) if workload == "mitchell_501_heights" else range(500, 1001, 25)
        out = []
        for h in heights:
            out += luma_requests(1920, 1080, width_for_height(h), h, BICUBIC, b=1 / 3, c=1 / 3)
        return out
    if workload in ("served_mix_sample",):
        out = []
        kernels = [(BILINEAR, 3, 0.0, 0.5), (BICUBIC, 3, 0.0, 0.5), (LANCZOS, 3, 0.0, 0.5),
                   (SPLINE36, 3, 0.0, 0.5)]
        for (sw, sh), frac, (kernel, taps, b, c), mode in zip(
                [(3840, 2160), (1920, 1080), (1280, 720), (1920, 1080)],
                [0.5, 0.71, 0.95, 0.83], kernels, [0, 1, 2, 0]):
            dh = int(sh * frac) & ~1
            dw = width_for_height(dh, even=True)
            out += luma_requests(sw, sh, dw, dh, kernel, taps=taps, b=b, c=c, f64_mode=mode)
            out += chroma420_requests(sw, sh, dw, dh, kernel, taps=taps, b=b, c=c,
                                      f64_mode=mode)
        return out
    raise KeyError(workload)


def integer_conversion(bits, chroma, full_range):
    """vs_plugin.cpp:761-781 (YUV clip)."""
    maximum = (1 << bits) - 1
    offset, scale = 0, maximum
    if not full_range:
        depth = 1 << (bits - 8)
        offset = (128 if chroma else 16) * depth
        scale = (224 if chroma else 219) * depth
    elif chroma:
        offset = 1 << (bits - 1)
    import numpy as np
    return Conversion(float(offset), float(np.float32(1.0) / np.float32(scale)), float(scale),
                      float(offset), maximum)
