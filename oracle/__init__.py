"""CPU oracle for the spimagine ray caster -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import this package, and only as the checker or the timed CPU baseline.  The
product path (``spimagine_amd``) never imports it.

``spimref.c`` restates the reference kernels (``spimagine/volumerender/kernels/*.cl``)
with a fixed float32 operation sequence; this module binds it with ctypes and adds
``render_max_project`` / ``render_iso`` helpers that mirror the launch sequence of
``spimagine/volumerender/volumerender.py:327-506``.

Parity status of each function (how it is pinned, see tests/test_oracle_*.py):
  sample (LINEAR/NEAREST)     scipy.ndimage.map_coordinates, closed forms
  max_project (f32/u16/u8)    constant-volume closed form, linear-ramp fixture KAT,
                              nearest+ortho == numpy max reduction, float64 brute force
  alpha (opacity) mode        step-volume geometric closed form incl. the inner-loop
                              break quirk
  multipass                   numParts=1 identity, fmax merge closed form
  iso_surface                 radial-field analytic hit depth (+1.1..1.2 dt bias) and
                              anisotropic (8,8,6) gradient normal
  conv_x/y, conv_vec_x/y      scipy.ndimage.correlate1d normalized convolution
  occlusion                   LCG restated with Python integers; constant/step depth maps
  shading                     closed-form Phong special cases
u16/u8 with a LINEAR sampler is spec-undefined in OpenCL 1.2; this oracle interpolates
integer volumes in float32 -- that choice is the build's own (parity unpinned against the
reference's OpenCL runtime, which cannot run here).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libspimref.so")

DTYPE_CODES = {np.dtype(np.float32): 0, np.dtype(np.uint16): 1, np.dtype(np.uint8): 2}
FILTERS = {"linear": 0, "nearest": 1}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        fp = ctypes.POINTER(ctypes.c_float)
        vp = ctypes.c_void_p
        i = ctypes.c_int
        f = ctypes.c_float
        u = ctypes.c_uint32
        L.spimref_max_project.argtypes = [vp, i, i, i, i, i, fp, fp, fp, f, f, f, f, i, i, i,
                                          i, i, i, i, fp, fp, i]
        L.spimref_max_project.restype = i
        L.spimref_iso_surface.argtypes = [vp, i, i, i, i, i, fp, fp, fp, f, f, i, i, i, i, i,
                                          fp, fp, fp, fp, i]
        L.spimref_iso_surface.restype = i
        L.spimref_render_iso.argtypes = [vp, i, i, i, i, i, fp, fp, fp, f, f, i, f, i, i, i, i,
                                         fp, fp, fp, fp, fp, fp, fp, i]
        L.spimref_render_iso.restype = i
        L.spimref_render_iso_rows.argtypes = [vp, i, i, i, i, i, fp, fp, fp, f, f, i, f, i, i, i,
                                              i, i, i, fp, fp, fp, fp, fp, fp, fp, i]
        L.spimref_render_iso_rows.restype = i
        for n in ("spimref_conv_x", "spimref_conv_y", "spimref_conv_vec_x", "spimref_conv_vec_y"):
            getattr(L, n).argtypes = [fp, fp, i, i, i]
            getattr(L, n).restype = None
        L.spimref_conv_weights.argtypes = [i, f, fp]
        L.spimref_occlusion.argtypes = [fp, i, i, i, i, fp]
        L.spimref_shading.argtypes = [fp, i, i, fp, fp, f, fp, fp, fp]
        L.spimref_random.argtypes = [u, u]
        L.spimref_random.restype = f
        L.spimref_rand_int.argtypes = [u, u, i, i]
        L.spimref_rand_int.restype = f
        L.spimref_sincos.argtypes = [f, fp, fp]
        L.spimref_sample.argtypes = [vp, i, i, i, i, i, f, f, f]
        L.spimref_sample.restype = f
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())
    if n is not None:
        assert a.size == n
    return a


def _vol(data):
    data = np.ascontiguousarray(data)
    if data.dtype not in DTYPE_CODES:
        raise TypeError("oracle supports float32/uint16/uint8, got %s" % data.dtype)
    nz, ny, nx = data.shape
    return data, DTYPE_CODES[data.dtype], nx, ny, nz


def max_project(data, invP, invM, box=(-1, 1, -1, 1, -1, 1), min_val=0.0, max_val=0.0,
                gamma=1.0, alpha_pow=0.0, max_steps=200, num_parts=1, current_part=0,
                width=64, height=64, interpolation="linear", rows=None, out=None, alpha=None,
                nthreads=0):
    """volume_kernel.cl max_project_float / max_project_short on the CPU.

    ``out``/``alpha`` (H, W) float32 are updated in place when given (multipass merges
    into their previous contents, volume_kernel.cl:172-183)."""
    data, dt, nx, ny, nz = _vol(data)
    invP = _f32(invP, 16)
    invM = _f32(invM, 16)
    box = _f32(box, 6)
    if out is None:
        out = np.zeros((height, width), np.float32)
    if alpha is None:
        alpha = np.zeros((height, width), np.float32)
    y0, y1 = (0, height) if rows is None else rows
    rc = lib().spimref_max_project(data.ctypes.data, dt, nx, ny, nz, FILTERS[interpolation],
                                   _fp(invP), _fp(invM), _fp(box), min_val, max_val, gamma,
                                   alpha_pow, max_steps, num_parts, current_part, width, height,
                                   y0, y1, _fp(out), _fp(alpha), nthreads)
    if rc:
        raise ValueError("invalid max_project arguments (max_steps=%d, num_parts=%d)"
                         % (max_steps, num_parts))
    return out, alpha


def iso_surface(data, invP, invM, box=(-1, 1, -1, 1, -1, 1), iso_val=0.0, gamma=1.0,
                max_steps=200, width=64, height=64, interpolation="linear", rows=None,
                nthreads=0):
    """iso_kernel.cl:17-225 alone (no post passes); rows=(y0, y1) marches only those rows
    (the others stay 0)."""
    data, dt, nx, ny, nz = _vol(data)
    invP, invM, box = _f32(invP, 16), _f32(invM, 16), _f32(box, 6)
    out = np.zeros((height, width), np.float32)
    alpha = np.zeros_like(out)
    depth = np.zeros_like(out)
    normals = np.zeros((height, width, 3), np.float32)
    y0, y1 = (0, height) if rows is None else rows
    rc = lib().spimref_iso_surface(data.ctypes.data, dt, nx, ny, nz, FILTERS[interpolation],
                                   _fp(invP), _fp(invM), _fp(box), iso_val, gamma, max_steps,
                                   width, height, y0, y1, _fp(out), _fp(alpha), _fp(depth),
                                   _fp(normals), nthreads)
    if rc:
        raise ValueError("invalid iso_surface arguments")
    return out, alpha, depth, normals


def render_iso(data, invP, invM, box=(-1, 1, -1, 1, -1, 1), iso_val=0.0, gamma=1.0,
               max_steps=200, occ_strength=0.1, occ_radius=21, occ_n_points=30, width=64,
               height=64, interpolation="linear", nthreads=0):
    """volumerender.py:446-506: iso_surface -> conv_vec(7) -> occlusion -> conv(5) -> shading.
    Returns (output, alpha, depth, normals, occlusion)."""
    data, dt, nx, ny, nz = _vol(data)
    invP, invM, box = _f32(invP, 16), _f32(invM, 16), _f32(box, 6)
    out = np.zeros((height, width), np.float32)
    alpha = np.zeros_like(out)
    depth = np.zeros_like(out)
    occ = np.zeros_like(out)
    tmp = np.zeros_like(out)
    normals = np.zeros((height, width, 3), np.float32)
    tmpv = np.zeros_like(normals)
    rc = lib().spimref_render_iso(data.ctypes.data, dt, nx, ny, nz, FILTERS[interpolation],
                                  _fp(invP), _fp(invM), _fp(box), iso_val, gamma, max_steps,
                                  occ_strength, occ_radius, occ_n_points, width, height, _fp(out),
                                  _fp(alpha), _fp(depth), _fp(normals), _fp(occ), _fp(tmpv),
                                  _fp(tmp), nthreads)
    if rc:
        raise ValueError("invalid iso arguments")
    return out, alpha, depth, normals, occ


def render_iso_rows(data, invP, invM, rows, box=(-1, 1, -1, 1, -1, 1), iso_val=0.0,
                    gamma=1.0, max_steps=200, occ_strength=0.1, occ_radius=21,
                    occ_n_points=30, width=64, height=64, interpolation="linear", nthreads=0):
    """render_iso's five outputs on rows=(y0, y1) only: the iso march runs on the rows
    those rows depend on (within max(3, occ_radius + 3)); the rest of each frame holds
    misses.  Returns full-size (output, alpha, depth, normals, occlusion), valid on the
    rows asked for."""
    data, dt, nx, ny, nz = _vol(data)
    invP, invM, box = _f32(invP, 16), _f32(invM, 16), _f32(box, 6)
    out = np.zeros((height, width), np.float32)
    alpha = np.zeros_like(out)
    depth = np.full_like(out, np.inf)
    occ = np.zeros_like(out)
    tmp = np.zeros_like(out)
    normals = np.zeros((height, width, 3), np.float32)
    tmpv = np.zeros_like(normals)
    rc = lib().spimref_render_iso_rows(data.ctypes.data, dt, nx, ny, nz, FILTERS[interpolation],
                                       _fp(invP), _fp(invM), _fp(box), iso_val, gamma, max_steps,
                                       occ_strength, occ_radius, occ_n_points, width, height,
                                       int(rows[0]), int(rows[1]), _fp(out), _fp(alpha),
                                       _fp(depth), _fp(normals), _fp(occ), _fp(tmpv), _fp(tmp),
                                       nthreads)
    if rc:
        raise ValueError("invalid iso arguments")
    return out, alpha, depth, normals, occ


def conv(img, Nh, axis, vec=False):
    """convolve_2d.cl conv_x/conv_y (scalar) and conv_vec_x/conv_vec_y (3-vector)."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    H, W = img.shape[:2]
    out = np.zeros_like(img)
    name = "spimref_conv_%s%s" % ("vec_" if vec else "", "x" if axis == "x" else "y")
    getattr(lib(), name)(_fp(img), _fp(out), W, H, Nh)
    return out


def conv_weights(Nh, k):
    w = np.zeros(Nh, np.float32)
    lib().spimref_conv_weights(Nh, k, _fp(w))
    return w


def occlusion(depth, radius=21, n_points=30):
    depth = np.ascontiguousarray(depth, dtype=np.float32)
    H, W = depth.shape
    out = np.zeros_like(depth)
    lib().spimref_occlusion(_fp(out), W, H, radius, n_points, _fp(depth))
    return out


def shading(normals, depth, occ, invP, invM, occ_strength=0.1):
    normals = np.ascontiguousarray(normals, dtype=np.float32)
    depth = np.ascontiguousarray(depth, dtype=np.float32)
    occ = np.ascontiguousarray(occ, dtype=np.float32)
    H, W = depth.shape
    out = np.zeros_like(depth)
    invP, invM = _f32(invP, 16), _f32(invM, 16)
    lib().spimref_shading(_fp(out), W, H, _fp(invP), _fp(invM), occ_strength, _fp(normals),
                          _fp(depth), _fp(occ))
    return out


def random(x, y):
    return lib().spimref_random(x, y)


def rand_int(x, y, start, end):
    return lib().spimref_rand_int(x, y, start, end)


def sincos(x):
    s = ctypes.c_float()
    c = ctypes.c_float()
    lib().spimref_sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def sample(data, sx, sy, sz, interpolation="linear"):
    data, dt, nx, ny, nz = _vol(data)
    return lib().spimref_sample(data.ctypes.data, dt, nx, ny, nz, FILTERS[interpolation],
                                sx, sy, sz)
