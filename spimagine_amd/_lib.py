"""ctypes binding of libspimrender.so (C ABI declared in include/spimrender.h).

The library is the only compute path: when it is missing or no HIP device is visible,
the calls raise -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspimrender.so")

SPR_OK = 0
SPR_ERR_INVALID_ARG = -1
SPR_ERR_NO_VOLUME = -2
SPR_ERR_UNSUPPORTED_DTYPE = -3
SPR_ERR_OUT_OF_MEMORY = -4
SPR_ERR_HIP = -5
SPR_ERR_NO_DEVICE = -6

DTYPE_F32, DTYPE_U16, DTYPE_U8, DTYPE_U16X2 = 0, 1, 2, 3
DTYPE_CODES = {np.dtype(np.float32): DTYPE_F32, np.dtype(np.uint16): DTYPE_U16,
               np.dtype(np.uint8): DTYPE_U8}
FILTER_LINEAR, FILTER_NEAREST = 0, 1
METHOD_MAX_PROJECT, METHOD_ISO_SURFACE = 0, 1
LAYOUT_AUTO, LAYOUT_LINEAR, LAYOUT_PADDED = 0, 1, 2
NOALLOC_MAX_JOBS = 4096
TUNE_TILE_MAP, TUNE_SLAB_KB = 0, 1
TUNE_STAGING, TUNE_CHUNK, TUNE_ROW_PAD, TUNE_WG_WAVES, TUNE_SKIP = 2, 3, 4, 5, 6
TUNING_KEYS = {"tile_map": TUNE_TILE_MAP, "slab_kb": TUNE_SLAB_KB, "staging": TUNE_STAGING,
               "chunk": TUNE_CHUNK, "row_pad": TUNE_ROW_PAD, "wg_waves": TUNE_WG_WAVES,
               "skip": TUNE_SKIP}

# every symbol include/spimrender.h declares
EXPORTS = (
    "spr_abi_version", "spr_device_count", "spr_device_memory", "spr_create", "spr_destroy",
    "spr_last_error", "spr_set_layout", "spr_get_layout", "spr_set_tuning", "spr_get_tuning",
    "spr_set_volume", "spr_set_volume_device", "spr_volume_minmax",
    "spr_set_params", "spr_render", "spr_render_device", "spr_synchronize", "spr_get_stream",
    "spr_count_samples", "spr_volume_add", "spr_volume_add_dilated", "spr_volume_add_pooled",
    "spr_volume_add_pooled_pair", "spr_volume_read",
    "spr_volume_release", "spr_volume_info",
    "spr_render_batch", "spr_render_batch_noalloc", "spr_render_frames",
    "spr_render_frames_split", "spr_render_slot",
    "spr_render_slot_if", "spr_render_pair", "spr_render_coarse_pair",
    "spr_batch_graph_create",
    "spr_batch_graph_update",
    "spr_batch_graph_launch", "spr_batch_graph_destroy",
)


class SprParams(ctypes.Structure):
    """struct spr_params (include/spimrender.h)."""
    _fields_ = [
        ("inv_p", ctypes.c_float * 16),
        ("inv_m", ctypes.c_float * 16),
        ("box", ctypes.c_float * 6),
        ("min_val", ctypes.c_float),
        ("max_val", ctypes.c_float),
        ("gamma", ctypes.c_float),
        ("alpha_pow", ctypes.c_float),
        ("occ_strength", ctypes.c_float),
        ("occ_radius", ctypes.c_int),
        ("occ_n_points", ctypes.c_int),
        ("max_steps", ctypes.c_int),
        ("filter", ctypes.c_int),
    ]


def make_params(inv_p, inv_m, box=(-1, 1, -1, 1, -1, 1), min_val=0., max_val=0., gamma=1.,
                alpha_pow=0., occ_strength=.1, occ_radius=21, occ_n_points=30, max_steps=200,
                filter=FILTER_LINEAR):
    """An SprParams from Python values (every float rounded to float32)."""
    p = SprParams()
    p.inv_p[:] = [float(v) for v in np.asarray(inv_p, np.float32).ravel()]
    p.inv_m[:] = [float(v) for v in np.asarray(inv_m, np.float32).ravel()]
    p.box[:] = [float(v) for v in np.asarray(box, np.float32).ravel()]
    p.min_val = float(np.float32(min_val))
    p.max_val = float(np.float32(max_val))
    p.gamma = float(np.float32(gamma))
    p.alpha_pow = float(np.float32(alpha_pow))
    p.occ_strength = float(np.float32(occ_strength))
    p.occ_radius = int(occ_radius)
    p.occ_n_points = int(occ_n_points)
    p.max_steps = int(max_steps)
    p.filter = int(filter)
    return p


class SprJob(ctypes.Structure):
    """struct spr_job (include/spimrender.h)."""
    _fields_ = [
        ("slot", ctypes.c_int),
        ("inv_p", ctypes.c_float * 16),
        ("inv_m", ctypes.c_float * 16),
        ("d_out", ctypes.c_void_p),
        ("d_alpha", ctypes.c_void_p),
    ]


class SprFrame(ctypes.Structure):
    """struct spr_frame (include/spimrender.h)."""
    _fields_ = [
        ("slot", ctypes.c_int),
        ("w", ctypes.c_int),
        ("h", ctypes.c_int),
        ("max_steps", ctypes.c_int),
        ("inv_p", ctypes.c_float * 16),
        ("inv_m", ctypes.c_float * 16),
        ("box", ctypes.c_float * 6),
        ("min_val", ctypes.c_float),
        ("max_val", ctypes.c_float),
        ("gamma", ctypes.c_float),
        ("d_out", ctypes.c_void_p),
        ("d_alpha", ctypes.c_void_p),
    ]


def make_frame(slot, params, w, h, d_out, d_alpha):
    """An SprFrame of `slot` rendered with SprParams `params` (its matrices, box, transfer
    range, gamma and max_steps) into w x h device planes."""
    f = SprFrame()
    f.slot = int(slot)
    f.w, f.h = int(w), int(h)
    f.max_steps = params.max_steps
    f.inv_p[:] = list(params.inv_p)
    f.inv_m[:] = list(params.inv_m)
    f.box[:] = list(params.box)
    f.min_val, f.max_val, f.gamma = params.min_val, params.max_val, params.gamma
    f.d_out = d_out
    f.d_alpha = d_alpha
    return f


# numpy view of struct spr_job, for building large job tables without a Python loop
JOB_DTYPE = np.dtype({"names": ["slot", "inv_p", "inv_m", "d_out", "d_alpha"],
                      "formats": [np.int32, (np.float32, 16), (np.float32, 16), np.uint64,
                                  np.uint64],
                      "offsets": [SprJob.slot.offset, SprJob.inv_p.offset, SprJob.inv_m.offset,
                                  SprJob.d_out.offset, SprJob.d_alpha.offset],
                      "itemsize": ctypes.sizeof(SprJob)})


def jobs_from_arrays(slots, inv_p, inv_m, d_out, d_alpha):
    """A ctypes array of n spr_jobs from arrays: slots (n,), inv_p and inv_m (n, 16) or
    (n, 4, 4), d_out and d_alpha (n,) device addresses."""
    slots = np.asarray(slots)
    n = slots.shape[0]
    buf = np.zeros(n, JOB_DTYPE)
    buf["slot"] = slots
    buf["inv_p"] = np.asarray(inv_p, np.float32).reshape(n, 16)
    buf["inv_m"] = np.asarray(inv_m, np.float32).reshape(n, 16)
    buf["d_out"] = np.asarray(d_out, np.uint64)
    buf["d_alpha"] = np.asarray(d_alpha, np.uint64)
    arr = (SprJob * n)()
    ctypes.memmove(arr, buf.ctypes.data, buf.nbytes)
    return arr


# numpy views of struct spr_params and struct spr_frame (vectorised frame tables)
PARAMS_DTYPE = np.dtype({"names": [f[0] for f in SprParams._fields_],
                         "formats": [np.int32 if f[1] is ctypes.c_int else
                                     np.float32 if f[1] is ctypes.c_float else
                                     (np.float32, f[1]._length_) for f in SprParams._fields_],
                         "offsets": [getattr(SprParams, f[0]).offset for f in SprParams._fields_],
                         "itemsize": ctypes.sizeof(SprParams)})
FRAME_DTYPE = np.dtype({"names": [f[0] for f in SprFrame._fields_],
                        "formats": [np.int32 if f[1] is ctypes.c_int else
                                    np.float32 if f[1] is ctypes.c_float else
                                    np.uint64 if f[1] is ctypes.c_void_p else
                                    (np.float32, f[1]._length_) for f in SprFrame._fields_],
                        "offsets": [getattr(SprFrame, f[0]).offset for f in SprFrame._fields_],
                        "itemsize": ctypes.sizeof(SprFrame)})


def params_array(plist):
    """The SprParams of plist as one PARAMS_DTYPE array."""
    return np.frombuffer(b"".join(bytes(p) for p in plist), PARAMS_DTYPE)


class SprError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("spimrender error %d: %s" % (code, msg))
        self.code = code


_lib = None


def load():
    """Load libspimrender.so and declare argument types (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("%s is not built; run `python -m spimagine_amd.build`" % LIB_PATH)
    # One HIP runtime per process: torch ships its own libamdhip64 (SONAME libamdhip64.so.7)
    # and its extensions depend on it by file name, so a runtime this library pulled in
    # from /opt/rocm first would be a second, separate runtime and torch would find no
    # GPU.  Loading torch first makes this library's libamdhip64.so.7 dependency resolve
    # to torch's already loaded runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_void_p, c_float = ctypes.c_int, ctypes.c_void_p, ctypes.c_float
    fp = ctypes.POINTER(ctypes.c_float)
    L.spr_abi_version.argtypes = []
    L.spr_abi_version.restype = c_int
    L.spr_device_count.argtypes = [ctypes.POINTER(c_int)]
    L.spr_device_memory.argtypes = [c_int, ctypes.POINTER(ctypes.c_size_t),
                                    ctypes.POINTER(ctypes.c_size_t)]
    L.spr_create.argtypes = [c_int, ctypes.POINTER(c_void_p)]
    L.spr_destroy.argtypes = [c_void_p]
    L.spr_last_error.argtypes = [c_void_p]
    L.spr_last_error.restype = ctypes.c_char_p
    L.spr_set_layout.argtypes = [c_void_p, c_int]
    L.spr_get_layout.argtypes = [c_void_p, ctypes.POINTER(c_int)]
    L.spr_set_tuning.argtypes = [c_void_p, c_int, c_int]
    L.spr_get_tuning.argtypes = [c_void_p, c_int, ctypes.POINTER(c_int)]
    L.spr_set_volume.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_int]
    L.spr_set_volume_device.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_int]
    L.spr_volume_minmax.argtypes = [c_void_p, fp, fp]
    L.spr_set_params.argtypes = [c_void_p, ctypes.POINTER(SprParams)]
    L.spr_render.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p]
    L.spr_render_device.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    L.spr_synchronize.argtypes = [c_void_p]
    L.spr_get_stream.argtypes = [c_void_p, ctypes.POINTER(c_void_p)]
    L.spr_count_samples.argtypes = [c_void_p, c_int, c_int, c_int, ctypes.POINTER(ctypes.c_double)]
    L.spr_volume_add.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int,
                                 ctypes.POINTER(c_int)]
    L.spr_volume_add_dilated.argtypes = [c_void_p, c_int, c_int, ctypes.POINTER(c_int)]
    L.spr_volume_add_pooled.argtypes = [c_void_p, c_int, c_int, ctypes.POINTER(c_int)]
    L.spr_volume_add_pooled_pair.argtypes = [c_void_p, c_int, ctypes.POINTER(c_int)]
    L.spr_volume_read.argtypes = [c_void_p, c_int, c_void_p, ctypes.c_size_t]
    L.spr_volume_release.argtypes = [c_void_p, c_int]
    L.spr_volume_info.argtypes = [c_void_p, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                  ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                  ctypes.POINTER(ctypes.c_size_t)]
    jp = ctypes.POINTER(SprJob)
    L.spr_render_batch.argtypes = [c_void_p, ctypes.POINTER(SprParams), jp, c_int, c_int, c_int,
                                   c_int, c_int, c_void_p]
    L.spr_render_batch_noalloc.argtypes = [c_void_p, ctypes.POINTER(SprParams), jp, c_int, c_int,
                                           c_int, c_int, c_int, c_void_p]
    L.spr_render_frames_split.argtypes = [c_void_p, ctypes.POINTER(SprFrame), c_int, c_int,
                                          c_void_p, c_void_p]
    L.spr_render_frames.argtypes = [c_void_p, ctypes.POINTER(SprFrame), c_int, c_float, c_int,
                                    c_void_p]
    L.spr_render_slot.argtypes = [c_void_p, c_int, ctypes.POINTER(SprParams), c_int, c_int, c_int,
                                  c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p]
    L.spr_render_slot_if.argtypes = [c_void_p, c_int, ctypes.POINTER(SprParams), c_int, c_int,
                                     c_void_p, c_void_p, c_void_p, c_void_p]
    L.spr_render_pair.argtypes = [c_void_p, c_int, ctypes.POINTER(SprParams), c_int, c_int, c_int,
                                  c_int, c_void_p, c_void_p, c_void_p]
    L.spr_render_coarse_pair.argtypes = [c_void_p, c_int, c_int, ctypes.POINTER(SprParams), c_int,
                                         c_int, c_void_p, c_void_p, c_void_p]
    L.spr_batch_graph_create.argtypes = [c_void_p, jp, c_int, c_int, c_int,
                                         ctypes.POINTER(c_void_p)]
    L.spr_batch_graph_update.argtypes = [c_void_p, c_void_p, jp, c_int]
    L.spr_batch_graph_launch.argtypes = [c_void_p, c_void_p, c_void_p]
    L.spr_batch_graph_destroy.argtypes = [c_void_p, c_void_p]
    _lib = L
    return L


def check(rc, ctx=None):
    if rc != SPR_OK:
        msg = load().spr_last_error(ctx)
        raise SprError(rc, msg.decode() if msg else "")
    return rc


def device_count():
    n = ctypes.c_int(0)
    check(load().spr_device_count(ctypes.byref(n)))
    return n.value


def device_memory(device=0):
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(load().spr_device_memory(device, ctypes.byref(free), ctypes.byref(total)))
    return free.value, total.value


def ptr(a):
    """Host pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class Context:
    """Owns one spr_ctx (one device, one stream, a resident volume, output buffers)."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.spr_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if getattr(self, "handle", None):
            self.lib.spr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        return check(rc, self.handle)

    def set_layout(self, layout):
        """Layout of the next upload (LAYOUT_AUTO / LAYOUT_LINEAR / LAYOUT_PADDED)."""
        self._check(self.lib.spr_set_layout(self.handle, int(layout)))

    def layout(self):
        v = ctypes.c_int()
        self._check(self.lib.spr_get_layout(self.handle, ctypes.byref(v)))
        return v.value

    def set_tuning(self, key, value):
        """Launch tuning (TUNING_KEYS name or SPR_TUNE_* code); results do not change."""
        code = TUNING_KEYS[key] if isinstance(key, str) else int(key)
        self._check(self.lib.spr_set_tuning(self.handle, code, int(value)))

    def get_tuning(self, key):
        code = TUNING_KEYS[key] if isinstance(key, str) else int(key)
        v = ctypes.c_int()
        self._check(self.lib.spr_get_tuning(self.handle, code, ctypes.byref(v)))
        return v.value

    def set_volume(self, data):
        data = np.ascontiguousarray(data)
        if data.dtype not in DTYPE_CODES:
            raise NotImplementedError("unsupported voxel type %s" % data.dtype)
        nz, ny, nx = data.shape
        self._check(self.lib.spr_set_volume(self.handle, data.ctypes.data, DTYPE_CODES[data.dtype],
                                            nx, ny, nz))

    def set_volume_device(self, dev_ptr, dtype, shape):
        nz, ny, nx = shape
        self._check(self.lib.spr_set_volume_device(self.handle, ctypes.c_void_p(dev_ptr),
                                                   DTYPE_CODES[np.dtype(dtype)], nx, ny, nz))

    def volume_minmax(self):
        mn, mx = ctypes.c_float(), ctypes.c_float()
        self._check(self.lib.spr_volume_minmax(self.handle, ctypes.byref(mn), ctypes.byref(mx)))
        return mn.value, mx.value

    def set_params(self, inv_p, inv_m, box, min_val, max_val, gamma, alpha_pow, occ_strength,
                   occ_radius, occ_n_points, max_steps, filter):
        self._check(self.lib.spr_set_params(self.handle, ctypes.byref(make_params(
            inv_p, inv_m, box, min_val, max_val, gamma, alpha_pow, occ_strength, occ_radius,
            occ_n_points, max_steps, filter))))

    def render(self, method, num_parts, current_part, w, h, out=None, alpha=None, depth=None,
               normals=None, occlusion=None):
        self._check(self.lib.spr_render(self.handle, method, num_parts, current_part, w, h,
                                        ptr(out), ptr(alpha), ptr(depth), ptr(normals),
                                        ptr(occlusion)))

    def render_device(self, method, num_parts, current_part, w, h, d_out, d_alpha, d_depth=None,
                      d_normals=None, d_occlusion=None, stream=None):
        self._check(self.lib.spr_render_device(self.handle, method, num_parts, current_part, w, h,
                                               d_out, d_alpha, d_depth, d_normals, d_occlusion,
                                               stream))

    def synchronize(self):
        self._check(self.lib.spr_synchronize(self.handle))

    def stream(self):
        s = ctypes.c_void_p()
        self._check(self.lib.spr_get_stream(self.handle, ctypes.byref(s)))
        return s.value

    # ---- resident volume table and batches
    def volume_add(self, data=None, dev_ptr=None, dtype=None, shape=None):
        """Upload a host array (or a device buffer of `dtype`/`shape`) into a new slot."""
        slot = ctypes.c_int()
        if data is not None:
            data = np.ascontiguousarray(data)
            nz, ny, nx = data.shape
            self._check(self.lib.spr_volume_add(self.handle, data.ctypes.data, 0,
                                                DTYPE_CODES[data.dtype], nx, ny, nz,
                                                ctypes.byref(slot)))
        else:
            nz, ny, nx = shape
            self._check(self.lib.spr_volume_add(self.handle, ctypes.c_void_p(dev_ptr), 1,
                                                DTYPE_CODES[np.dtype(dtype)], nx, ny, nz,
                                                ctypes.byref(slot)))
        return slot.value

    def volume_add_dilated(self, slot, radius):
        """A new slot: the (2 radius + 1)^3 max filter of `slot`, computed on the device."""
        out = ctypes.c_int()
        self._check(self.lib.spr_volume_add_dilated(self.handle, int(slot), int(radius),
                                                    ctypes.byref(out)))
        return out.value

    def volume_add_pooled(self, slot, mode):
        """A new slot: `slot` pooled 2x2x2 by mean (mode 0) or max (mode 1) on the device."""
        out = ctypes.c_int()
        self._check(self.lib.spr_volume_add_pooled(self.handle, int(slot), int(mode),
                                                   ctypes.byref(out)))
        return out.value

    def volume_add_pooled_pair(self, slot):
        """A new channel-pair slot: the 2x mean (channel 0) and max (channel 1) poolings of
        the uint16 volume in `slot`, interleaved (spr_volume_add_pooled_pair)."""
        out = ctypes.c_int()
        self._check(self.lib.spr_volume_add_pooled_pair(self.handle, int(slot), ctypes.byref(out)))
        return out.value

    def volume_read(self, slot):
        """The slot's volume as a host (nz, ny, nx) array (uint32 for a channel pair: channel 0
        in the low 16 bits)."""
        info = self.volume_info(slot)
        dt = {DTYPE_F32: np.float32, DTYPE_U16: np.uint16, DTYPE_U8: np.uint8,
              DTYPE_U16X2: np.uint32}[info["dtype"]]
        out = np.empty(info["shape"], dt)
        self._check(self.lib.spr_volume_read(self.handle, int(slot), out.ctypes.data, out.nbytes))
        return out

    def volume_release(self, slot):
        self._check(self.lib.spr_volume_release(self.handle, int(slot)))

    def volume_info(self, slot):
        d, x, y, z = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        b = ctypes.c_size_t()
        self._check(self.lib.spr_volume_info(self.handle, int(slot), ctypes.byref(d), ctypes.byref(x),
                                             ctypes.byref(y), ctypes.byref(z), ctypes.byref(b)))
        return {"dtype": d.value, "shape": (z.value, y.value, x.value), "bytes": b.value}

    @staticmethod
    def make_jobs(jobs):
        """jobs: iterable of (slot, inv_p, inv_m, d_out, d_alpha) -> ctypes array."""
        jobs = list(jobs)
        arr = (SprJob * len(jobs))()
        for k, (slot, ip, im, o, a) in enumerate(jobs):
            arr[k].slot = int(slot)
            arr[k].inv_p[:] = [float(v) for v in np.asarray(ip, np.float32).ravel()]
            arr[k].inv_m[:] = [float(v) for v in np.asarray(im, np.float32).ravel()]
            arr[k].d_out = o
            arr[k].d_alpha = a
        return arr

    def render_batch(self, jobs, w, h, num_parts=1, current_part=0, stream=None, params=None):
        """params: an SprParams (make_params) or None for the spr_set_params state."""
        arr = jobs if isinstance(jobs, ctypes.Array) else self.make_jobs(jobs)
        pp = ctypes.byref(params) if params is not None else None
        self._check(self.lib.spr_render_batch(self.handle, pp, arr, len(arr), w, h, num_parts,
                                              current_part, stream))

    def render_batch_noalloc(self, jobs, w, h, num_parts=1, current_part=0, stream=None,
                             params=None):
        """render_batch through spr_render_batch_noalloc (job table and queue counters in the
        library's fixed device scratch; at most NOALLOC_MAX_JOBS jobs)."""
        arr = jobs if isinstance(jobs, ctypes.Array) else self.make_jobs(jobs)
        pp = ctypes.byref(params) if params is not None else None
        self._check(self.lib.spr_render_batch_noalloc(self.handle, pp, arr, len(arr), w, h,
                                                      num_parts, current_part, stream))

    def render_frames(self, frames, alpha_pow=0., filter=FILTER_LINEAR, stream=None):
        """frames: a list of SprFrame (make_frame); one persistent launch (spr_render_frames)."""
        arr = (SprFrame * len(frames))(*frames)
        self._check(self.lib.spr_render_frames(self.handle, arr, len(frames), float(alpha_pow),
                                               int(filter), stream))

    def render_frames_split(self, frames, seg=0, stream=None, d_gates=None):
        """Max projections of frames of scattered rays (spr_render_frames_split): each ray's
        samples split into segments of seg (32, 64, 128; 0 = 64) samples; frames: a list of
        SprFrame or a FRAME_DTYPE array; d_gates: None or a device pointer to one uint32
        flag per frame (0: leave the frame untouched)."""
        if isinstance(frames, np.ndarray):
            assert frames.dtype == FRAME_DTYPE and frames.flags.c_contiguous
            arr = frames.ctypes.data_as(ctypes.POINTER(SprFrame))
        else:
            arr = (SprFrame * len(frames))(*frames)
        self._check(self.lib.spr_render_frames_split(self.handle, arr, len(frames), int(seg),
                                                     d_gates, stream))

    def render_slot(self, slot, params, method, w, h, d_out, d_alpha, d_depth=None,
                    d_normals=None, d_occlusion=None, num_parts=1, current_part=0, stream=None):
        self._check(self.lib.spr_render_slot(self.handle, int(slot), ctypes.byref(params), method,
                                             num_parts, current_part, w, h, d_out, d_alpha, d_depth,
                                             d_normals, d_occlusion, stream))

    def render_slot_if(self, slot, params, w, h, d_out, d_alpha, d_flag, stream=None):
        """The max projection of render_slot, performed only if the device word at d_flag
        is nonzero when the kernels run (spr_render_slot_if)."""
        self._check(self.lib.spr_render_slot_if(self.handle, int(slot), ctypes.byref(params), w, h,
                                                d_out, d_alpha, d_flag, stream))

    def render_pair(self, slot, params, w, h, d_out, d_alpha, num_parts=1, current_part=0,
                    stream=None):
        """Both channels of a channel-pair slot into d_out (2, h, w); see spr_render_pair."""
        self._check(self.lib.spr_render_pair(self.handle, int(slot), ctypes.byref(params),
                                             num_parts, current_part, w, h, d_out, d_alpha, stream))

    def render_coarse_pair(self, slot_a, slot_b, params, w, h, d_out, d_alpha, stream=None):
        """Two coarse planes (slot_a, slot_b) into d_out (2, h, w); see spr_render_coarse_pair."""
        self._check(self.lib.spr_render_coarse_pair(self.handle, int(slot_a), int(slot_b),
                                                    ctypes.byref(params), w, h, d_out, d_alpha,
                                                    stream))

    def batch_graph(self, jobs, w, h):
        return BatchGraph(self, jobs, w, h)

    def count_samples(self, num_parts, w, h):
        n = ctypes.c_double()
        self._check(self.lib.spr_count_samples(self.handle, num_parts, w, h, ctypes.byref(n)))
        return int(n.value)


class BatchGraph:
    """A captured batch (spr_batch_graph_*): replay with launch(), retarget with update()."""

    def __init__(self, ctx, jobs, w, h):
        self.ctx = ctx
        arr = jobs if isinstance(jobs, ctypes.Array) else Context.make_jobs(jobs)
        self.n = len(arr)
        g = ctypes.c_void_p()
        ctx._check(ctx.lib.spr_batch_graph_create(ctx.handle, arr, self.n, w, h, ctypes.byref(g)))
        self.handle = g

    def update(self, jobs):
        arr = jobs if isinstance(jobs, ctypes.Array) else Context.make_jobs(jobs)
        self.ctx._check(self.ctx.lib.spr_batch_graph_update(self.ctx.handle, self.handle, arr,
                                                            len(arr)))

    def launch(self, stream=None):
        self.ctx._check(self.ctx.lib.spr_batch_graph_launch(self.ctx.handle, self.handle, stream))

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.spr_batch_graph_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
