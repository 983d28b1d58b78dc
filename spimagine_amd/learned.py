"""Learned reconstruction of a max projection (libspimlearn.so, include/spimlearn.h).

A frame is first rendered coarse -- half the resolution and a quarter of the ray samples
(max_steps / 4) -- twice: from the volume and from its max-filtered ("dilated") copy, which
bounds the samples the coarse step skips.  A small convolutional network on the matrix
cores turns those two planes into the full-resolution frame.  The exact renderer is the
reference: the network is trained on coarse/fine pairs the renderer produces, and
``LearnedMIP.render_judged`` checks every frame against the exact renderer on a sparse
pixel grid and falls back to the exact frame below the PSNR threshold (``validate``
measures full-frame PSNR on a set of poses).  No reference counterpart: the nearest analogue is
progressive multipass rendering (spimagine/volumerender/kernels/volume_kernel.cl:100-106,
driven from spimagine/gui/glwidget.py:633-635).
"""
import ctypes
import math
import os
import time

import numpy as np

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspimlearn.so")
WEIGHTS_PATH = os.path.join(HERE, "weights", "mip_x2_q4.npz")
EXPORTS = ("spl_abi_version", "spl_param_count", "spl_create", "spl_destroy", "spl_last_error",
           "spl_set_params", "spl_get_params", "spl_forward", "spl_gradients", "spl_train_step",
           "spl_judge_sse", "spl_judge_sse_frames", "spl_judge_flags", "spl_judge_flags_bound",
           "spl_upsample2")
C = 32
LAYER_NAMES = ("e0", "e0b", "e1", "e1b", "e2a", "e2", "u1", "u0", "u0b")
PSNR_THRESHOLD = 40.0

_lib_learn = None


def load():
    global _lib_learn
    if _lib_learn is not None:
        return _lib_learn
    if not os.path.exists(LIB_PATH):
        raise ImportError("%s is not built; run `python -m spimagine_amd.build`" % LIB_PATH)
    try:  # one HIP runtime per process (see _lib.load)
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_void_p, c_float = ctypes.c_int, ctypes.c_void_p, ctypes.c_float
    L.spl_abi_version.argtypes = []
    L.spl_param_count.argtypes = []
    L.spl_create.argtypes = [c_int, ctypes.POINTER(c_void_p)]
    L.spl_destroy.argtypes = [c_void_p]
    L.spl_last_error.argtypes = [c_void_p]
    L.spl_last_error.restype = ctypes.c_char_p
    L.spl_set_params.argtypes = [c_void_p, c_void_p, c_int]
    L.spl_get_params.argtypes = [c_void_p, c_void_p, c_int]
    L.spl_forward.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]
    L.spl_gradients.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p,
                                ctypes.POINTER(ctypes.c_double), c_void_p]
    L.spl_train_step.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_float,
                                 ctypes.POINTER(ctypes.c_double), c_void_p]
    L.spl_judge_sse.argtypes = [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_int, c_int,
                                c_int, c_int, c_void_p, c_int, c_void_p]
    L.spl_judge_sse_frames.argtypes = [c_void_p, ctypes.POINTER(c_void_p), c_int, c_int, c_int,
                                       c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                       c_void_p]
    L.spl_judge_flags_bound.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int, c_float,
                                        c_float, c_void_p, c_void_p]
    L.spl_judge_flags.argtypes = [c_void_p, c_void_p, c_int, c_int, c_float, c_void_p, c_void_p]
    L.spl_upsample2.argtypes = [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p]
    _lib_learn = L
    return L


def param_layout():
    """(name, shape) of each parameter block, in the flat order of spl_set_params."""
    out = [("in.w", (9, 2, C)), ("in.b", (C,))]
    for n in LAYER_NAMES:
        out += [(n + ".w", (9, C, C)), (n + ".b", (C,))]
    out += [("head.w", (9, C, 4)), ("head.b", (4,))]
    return out


def param_count():
    return sum(int(np.prod(s)) for _, s in param_layout())


def split_params(flat):
    """flat float32 array -> {name: array view of its shape}."""
    out, k = {}, 0
    for name, shape in param_layout():
        n = int(np.prod(shape))
        out[name] = flat[k:k + n].reshape(shape)
        k += n
    return out


def init_params(seed=0):
    """He-normal weights, zero biases, zero head: the untrained network returns the
    bilinear upsampling of the coarse projection."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(param_count(), np.float32)
    p = split_params(flat)
    p["in.w"][...] = rng.normal(0, math.sqrt(2. / (9 * 2)), p["in.w"].shape)
    for n in LAYER_NAMES:
        p[n + ".w"][...] = rng.normal(0, math.sqrt(2. / (9 * C)), p[n + ".w"].shape)
    return flat


class SplError(RuntimeError):
    def __init__(self, code, msg):
        RuntimeError.__init__(self, "libspimlearn error %d: %s" % (code, msg))
        self.code = code


class Net(object):
    """One network on a device (spl_create)."""

    def __init__(self, device=0, params=None):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.spl_create(int(device), ctypes.byref(h))
        if rc:
            raise SplError(rc, self.lib.spl_last_error(None).decode())
        self.handle = h
        self.n_params = self.lib.spl_param_count()
        if params is not None:
            self.set_params(params)

    def _check(self, rc):
        if rc:
            raise SplError(rc, self.lib.spl_last_error(self.handle).decode())

    def close(self):
        if getattr(self, "handle", None):
            self.lib.spl_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        self._check(self.lib.spl_set_params(self.handle, flat.ctypes.data, flat.size))

    def get_params(self):
        out = np.empty(self.n_params, np.float32)
        self._check(self.lib.spl_get_params(self.handle, out.ctypes.data, out.size))
        return out

    def forward(self, d_in, n, h, w, d_out, stream=None):
        self._check(self.lib.spl_forward(self.handle, d_in, n, h, w, d_out, stream))

    def gradients(self, d_in, d_target, n, h, w, stream=None):
        g = np.empty(self.n_params, np.float32)
        sse = ctypes.c_double()
        self._check(self.lib.spl_gradients(self.handle, d_in, d_target, n, h, w, g.ctypes.data,
                                           ctypes.byref(sse), stream))
        return g, sse.value

    def judge_sse(self, d_frame, w, h, d_patches, patch, gx, gy, ox, oy, d_sse,
                  accumulate=False, stream=None):
        """Squared error of a frame against gx x gy exact patches (spl_judge_sse)."""
        self._check(self.lib.spl_judge_sse(self.handle, d_frame, w, h, d_patches, patch, gx, gy,
                                           ox, oy, d_sse, int(bool(accumulate)), stream))

    def judge_sse_frames(self, d_frames, w, h, d_sparse, gx, gy, ox, oy, d_sse, d_sse4=None,
                         stream=None):
        """Summed squared (and, with d_sse4, fourth-power) errors of n frames (device
        pointers d_frames) against their sparse exact frames in one pass
        (spl_judge_sse_frames)."""
        arr = (ctypes.c_void_p * len(d_frames))(*d_frames)
        self._check(self.lib.spl_judge_sse_frames(self.handle, arr, len(d_frames), w, h, d_sparse,
                                                  gx, gy, ox, oy, d_sse, d_sse4, stream))

    def judge_flags_bound(self, d_sse, d_sse4, n, npix, mse_max, z, d_flags, stream=None):
        """flags[k] = 1 unless the judged MSE + z standard errors <= mse_max, on the device
        (spl_judge_flags_bound)."""
        self._check(self.lib.spl_judge_flags_bound(self.handle, d_sse, d_sse4, int(n), int(npix),
                                                   float(mse_max), float(z), d_flags, stream))

    def judge_flags(self, d_sse, n, npix, mse_max, d_flags, stream=None):
        """flags[k] = sse[k] / npix > mse_max, on the device (spl_judge_flags)."""
        self._check(self.lib.spl_judge_flags(self.handle, d_sse, int(n), int(npix), float(mse_max),
                                             d_flags, stream))

    def upsample2(self, d_src, w, h, d_dst, stream=None):
        self._check(self.lib.spl_upsample2(self.handle, d_src, w, h, d_dst, stream))

    def train_step(self, d_in, d_target, n, h, w, lr, stream=None, want_loss=True):
        sse = ctypes.c_double()
        self._check(self.lib.spl_train_step(self.handle, d_in, d_target, n, h, w, float(lr),
                                            ctypes.byref(sse) if want_loss else None, stream))
        return sse.value if want_loss else None


# coarse input planes: "dilated" = the volume and its 5^3 max filter; "pooled" = the volume
# pooled 2x2x2 by mean and by max (a quarter-size pyramid level: a coarse ray step of 4
# voxels is 2 pooled voxels, so the coarse render keeps the locality of a fine one)
INPUTS = ("dilated", "pooled")
INPUTS_DEFAULT = "pooled"


def load_weights(path):
    """(params, inputs) stored by `python -m spimagine_amd.learned`."""
    w = np.load(path)
    inputs = str(w["inputs"]) if "inputs" in w else "dilated"
    return w["params"], inputs


def psnr(a, b):
    """PSNR (dB) of a against b for images in [0, 1]."""
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10. * math.log10(1. / mse)


JUDGE_STRIDE = 16  # the judge traces one pixel in 16 x 16 of every frame
JUDGE_OFFSET = 1   # judge pixels (S i + 1, S j + 1): off the coarse rays, where errors are largest
JUDGE_SEG = 64     # samples per segment of the judge's split rays (spr_render_frames_split)
JUDGE_Z = 3.       # standard errors in the judge's upper confidence bound on a frame's MSE
LEARN_BATCH = 8    # learned frames per network pass (render_many)


def _shifted_inv_p(inv_p, w, h, ox, oy):
    """inv_p (n, 16) float32 times the clip-space translation by (2 ox / w, 2 oy / h),
    computed in float64 and rounded to float32 (n, 16)."""
    T = np.eye(4)
    T[0, 3] = 2. * ox / w
    T[1, 3] = 2. * oy / h
    iP = np.asarray(inv_p, np.float64).reshape(-1, 4, 4)
    return np.matmul(iP, T).astype(np.float32).reshape(-1, 16)


def sparse_frames(slot, plist, w, h, stride, offsets, d_out, d_alpha):
    """A FRAME_DTYPE table of the sparse frames of every frame of plist for each (ox, oy) in
    offsets (offset-major: entry j n + k is frame k at offset j; see sparse_params), each
    (w/stride) x (h/stride); the planes of entry i at d_out / d_alpha + 4 i (w/stride)
    (h/stride)."""
    P = _lib.params_array(plist)
    n, m = len(plist), len(offsets)
    sw, sh = w // stride, h // stride
    F = np.zeros(n * m, _lib.FRAME_DTYPE)
    F["slot"] = slot
    F["w"], F["h"] = sw, sh
    for name in ("max_steps", "inv_m", "box", "min_val", "max_val", "gamma"):
        F[name] = np.tile(P[name], (m,) + (1,) * (P[name].ndim - 1))
    for j, (ox, oy) in enumerate(offsets):
        F["inv_p"][j * n:(j + 1) * n] = (_shifted_inv_p(P["inv_p"], w, h, ox, oy)
                                         if ox or oy else P["inv_p"])
    step = 4 * sw * sh
    F["d_out"] = d_out + step * np.arange(n * m, dtype=np.uint64)
    F["d_alpha"] = d_alpha + step * np.arange(n * m, dtype=np.uint64)
    return F


def sparse_params(p, w, h, stride=JUDGE_STRIDE, offset=JUDGE_OFFSET, offset_y=None):
    """Parameters of a (w/stride) x (h/stride) exact render whose pixel (i, j) traces the ray
    of pixel (stride i + offset, stride j + offset_y) of the w x h frame (offset_y None: the
    same as offset): the pixel mapping u = x / W * 2 - 1 of the small frame is shifted by
    2 offset / w in clip space, i.e. inv_p is multiplied by a clip-space translation
    (offset 0: the rays of pixels (stride i, stride j) exactly, bit for bit)."""
    q = _lib.SprParams()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(q))
    oy = offset if offset_y is None else offset_y
    if offset or oy:
        iP = _shifted_inv_p(np.asarray(list(p.inv_p), np.float32), w, h, offset, oy)
        q.inv_p[:] = [float(v) for v in iP.ravel()]
    return q


def coarse_params(p):
    """The coarse pass's parameters: the same pose and transfer, a quarter of the steps."""
    q = _lib.SprParams()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(q))
    q.max_steps = max(16, p.max_steps // 4)
    return q


def _db(mse):
    return float("inf") if mse <= 0 else 10. * math.log10(1. / float(mse))


class JudgedBatch(object):
    """Device results of one render_judged_many call (read them once the stream has passed
    the call): the first stage's error sums sums1 = (squared, fourth-power) over npix
    pixels per frame and its flags; with a second stage, the sums over all 4 npix pixels
    (sums2, meaningful for the frames the first stage did not accept) and the final flags;
    z: the confidence bound's standard errors."""

    def __init__(self, sums1, flags1, npix, z, sums2=None, flags=None):
        self.sums1, self.flags1, self.npix, self.z = sums1, flags1, npix, z
        self.sums2 = sums2
        self.flags = flags1 if flags is None else flags

    def kept(self):
        """Per frame: True when the learned frame was kept (no exact fallback)."""
        return [int(f) == 0 for f in self.flags.cpu().numpy()]

    def refined(self):
        """Per frame: True when the second stage judged the frame."""
        if self.sums2 is None:
            return [False] * len(self.flags1)
        return [int(f) != 0 for f in self.flags1.cpu().numpy()]

    def _moments(self):
        s1 = self.sums1.cpu().numpy().astype(np.float64) / self.npix
        if self.sums2 is None:
            return s1
        s2 = self.sums2.cpu().numpy().astype(np.float64) / (4 * self.npix)
        return np.where(np.asarray(self.refined())[None, :], s2, s1)

    def _n(self):
        return np.where(np.asarray(self.refined()), 4 * self.npix, self.npix)

    def estimates(self):
        """Per frame: the PSNR estimated on the judge's pixels (dB; the last stage's)."""
        return [_db(v) for v in self._moments()[0]]

    def bounds(self):
        """Per frame: the PSNR of the judge's upper confidence bound on the MSE (dB), the
        value the last stage compared with the acceptance level."""
        m2, m4 = self._moments()
        ucb = m2 + self.z * np.sqrt(np.maximum(m4 - m2 * m2, 0.) / self._n())
        return [_db(v) for v in ucb]


class LearnedMIP(object):
    """Learned max projection over the volumes of one renderer context."""

    def __init__(self, ctx, params=None, device=0, radius=2, threshold=PSNR_THRESHOLD, net=None,
                 inputs=None):
        import torch
        self.torch = torch
        self.ctx = ctx
        if net is None:
            if params is None:
                params, stored = load_weights(WEIGHTS_PATH)
                inputs = inputs or stored
            net = Net(device, params)
        self.net = net
        self.radius = radius
        self.threshold = threshold
        if (inputs or INPUTS_DEFAULT) not in INPUTS:
            raise ValueError("unknown coarse inputs %r" % inputs)
        self.inputs = inputs or INPUTS_DEFAULT
        self.dilated = {}
        self.enabled = {}   # slot -> validated (min PSNR >= threshold)
        self._bufs = {}
        self._side = None   # the judge's second stream (sparse exact frames)
        # the judge: one pixel in judge_stride x judge_stride; a frame is kept when the
        # PSNR of its judged MSE plus judge_z standard errors reaches threshold + margin;
        # judge_slab_kb: the sparse frames' staging (their rays run judge_stride pixels
        # apart; None: the context's); judge_seg: samples per segment of the sparse
        # frames' split rays (spr_render_frames_split; None: one persistent batch launch
        # per set of non-pose parameters, spr_render_batch)
        self.judge_stride = JUDGE_STRIDE
        self.judge_z = JUDGE_Z
        self.margin = 0.
        self.judge_slab_kb = None
        self.judge_seg = JUDGE_SEG
        # the sparse frames on a second stream, concurrently with the learned frames (False:
        # on the caller's stream before them).  With one network pass per 8 frames the
        # concurrent form is faster (0.393 against 0.408 ms a frame); with a network pass
        # per frame it was not (0.555 against 0.551 ms): the two then thrashed each other's
        # caches and LDS without gaps to fill
        self.judge_concurrent = True
        # a second judge stage on the stride-(S/2) grid for the frames the first stage does
        # not accept (split rays only)
        self.judge_refine = True
        # render_many: the coarse frames of a group on this many side streams, round robin,
        # so that one frame's launch fills the GPU while another's last waves finish (a
        # 512^2 coarse frame is two rounds of one-wave workgroups over the CUs, and its tail
        # left the GPU half idle): 0.2241 against 0.2857 ms a learned frame in groups of 8,
        # judged 0.328 against 0.391 ms (4 streams: the same as 2).  0: on the caller's
        # stream one after another
        self.coarse_streams = 2
        self._coarse_side = []

    def attach(self, slot):
        """The slots whose coarse projections feed the network (created once): for pooled
        inputs of a uint16 volume one channel-pair slot (both poolings interleaved, rendered
        by spr_render_pair), else two one-channel slots."""
        if slot not in self.dilated:
            if self.inputs == "pooled":
                if self.ctx.volume_info(slot)["dtype"] == _lib.DTYPE_U16:
                    pair = self.ctx.volume_add_pooled_pair(slot)
                    self.dilated[slot] = (pair, pair)
                else:
                    self.dilated[slot] = (self.ctx.volume_add_pooled(slot, 0),
                                          self.ctx.volume_add_pooled(slot, 1))
            else:
                self.dilated[slot] = (slot, self.ctx.volume_add_dilated(slot, self.radius))
        return self.dilated[slot]

    def _buffer(self, key, shape):
        if key not in self._bufs:
            self._bufs[key] = self.torch.empty(shape, device="cuda")
        return self._bufs[key]

    def coarse(self, slot, p, w, h, planes, k=0, stream=None, alpha_key=None):
        """Renders the two coarse planes of frame k of `planes` (n, 2, h/2, w/2) in one
        launch (spr_render_pair on a channel-pair slot, else spr_render_coarse_pair; the
        coarse alpha lands in the "alpha" buffer, or in alpha buffer alpha_key)."""
        a, b = self.attach(slot)
        key = ("alpha", w, h) if alpha_key is None else ("alpha", w, h, alpha_key)
        scratch = self._buffer(key, (h // 2, w // 2))
        if a == b:
            self.ctx.render_pair(a, coarse_params(p), w // 2, h // 2, planes[k].data_ptr(),
                                 scratch.data_ptr(), stream=stream)
        else:
            self.ctx.render_coarse_pair(a, b, coarse_params(p), w // 2, h // 2,
                                        planes[k].data_ptr(), scratch.data_ptr(), stream=stream)

    def render(self, slot, p, w, h, d_out, stream):
        """The learned w x h frame of `slot` with parameters p into device buffer d_out."""
        if w % 8 or h % 8:
            raise ValueError("learned frames need w and h multiples of 8")
        planes = self._buffer(("planes", w, h), (1, 2, h // 2, w // 2))
        self.coarse(slot, p, w, h, planes, 0, stream)
        self.net.forward(planes.data_ptr(), 1, h // 2, w // 2, d_out, stream)

    def render_many(self, slot, plist, w, h, outs, stream, alphas=None):
        """The learned frames of plist into the (h, w) float32 tensors outs, the network run
        once per group of up to LEARN_BATCH frames (its launches then serve the whole group:
        at 512^2 each layer's time is mostly fixed latency); with alphas, the coarse alpha
        of each frame nearest-upsampled into alphas[k].  Frames land in place when outs are
        consecutive views of one tensor, else through a group buffer.  stream: a
        torch.cuda.Stream."""
        torch = self.torch
        if w % 8 or h % 8:
            raise ValueError("learned frames need w and h multiples of 8")
        n, sh = len(plist), stream.cuda_stream
        G = min(n, LEARN_BATCH)
        planes = self._buffer(("planes", w, h, G), (G, 2, h // 2, w // 2))
        coarse_alpha = self._buffer(("alpha", w, h), (h // 2, w // 2))
        base, fb = outs[0].data_ptr(), 4 * h * w
        inplace = all(o.is_contiguous() and o.data_ptr() == base + k * fb
                      for k, o in enumerate(outs))
        ns = self.coarse_streams
        while len(self._coarse_side) < ns:
            self._coarse_side.append(torch.cuda.Stream())
        for g0 in range(0, n, G):
            m = min(G, n - g0)
            if ns:
                # each side stream after the caller's earlier work (the previous group's
                # network pass reads the planes), the network after every side stream; a
                # coarse alpha scratch plane per side stream
                for st in self._coarse_side[:ns]:
                    st.wait_stream(stream)
                for k in range(m):
                    st = self._coarse_side[k % ns]
                    self.coarse(slot, plist[g0 + k], w, h, planes, k, st.cuda_stream,
                                alpha_key=k % ns)
                    if alphas is not None:
                        a = self._buffer(("alpha", w, h, k % ns), (h // 2, w // 2))
                        self.net.upsample2(a.data_ptr(), w // 2, h // 2,
                                           alphas[g0 + k].data_ptr(), st.cuda_stream)
                for st in self._coarse_side[:ns]:
                    stream.wait_stream(st)
            else:
                for k in range(m):
                    self.coarse(slot, plist[g0 + k], w, h, planes, k, sh)
                    if alphas is not None:
                        self.net.upsample2(coarse_alpha.data_ptr(), w // 2, h // 2,
                                           alphas[g0 + k].data_ptr(), sh)
            if inplace:
                self.net.forward(planes.data_ptr(), m, h // 2, w // 2, base + g0 * fb, sh)
            else:
                y = self._buffer(("y", w, h, G), (G, h, w))
                self.net.forward(planes.data_ptr(), m, h // 2, w // 2, y.data_ptr(), sh)
                with torch.cuda.stream(stream):
                    for k in range(m):
                        outs[g0 + k].copy_(y[k])

    def render_judged(self, slot, p, out, alpha, stream, margin=None):
        """One frame with a per-frame judge (render_judged_many for one frame)."""
        kept, est = self.render_judged_many(slot, [p], [out], [alpha], stream, margin)
        return kept[0], est[0]

    def render_judged_many(self, slot, plist, outs, alphas, stream, margin=None, wait=True):
        """Frames with a per-frame judge: the learned frames, the exact renderer on one pixel
        in S x S of every frame (S = judge_stride; a (w/S) x (h/S) render, sparse_params;
        the sparse frames of all frames in one launch of split rays,
        spr_render_frames_split, or with judge_seg None one batched launch per set of
        non-pose parameters), each learned frame's mean squared error on those pixels and
        its standard error (from the fourth-power sum), and the exact full frame for every
        frame whose MSE + judge_z standard errors is above the MSE of threshold + margin
        dB.  A scattered grid: the error of a learned frame sits in few pixels, and a
        sample of the pixels estimates it well only when spread over the whole frame
        (pixel blocks of the same count miss or over-weight the few bright structures);
        the bound widens where the sampled errors are concentrated.
        Nothing waits on the host: the sparse exact frames render on a second stream while
        the learned frames render on `stream` (judge_concurrent; else on `stream` before
        them), the learned frames one network pass per 8 (render_many); then, on `stream`,
        the judge's error sums
        (spl_judge_sse_frames), its decisions on the device (spl_judge_flags_bound) and
        the exact fallbacks, each rendered only if its frame's flag is set
        (spr_render_slot_if).
        outs / alphas: (h, w) float32 torch tensors; stream: a torch.cuda.Stream.  Learned
        frames take alpha from the coarse pass (exact at even pixels, nearest-upsampled).
        wait=True: synchronises `stream` and returns (kept flags, estimated PSNRs);
        wait=False: returns a JudgedBatch whose results are read later.  The slot must be
        batchable (padded layout, below 4 GiB)."""
        torch = self.torch
        h, w = outs[0].shape
        S = self.judge_stride
        margin = self.margin if margin is None else margin
        if w % S or h % S:
            raise ValueError("judged frames need w and h multiples of the judge stride %d" % S)
        sh = stream.cuda_stream
        n = len(plist)
        sw, sh_ = w // S, h // S
        if self.judge_concurrent:
            if self._side is None:
                self._side = torch.cuda.Stream()
            side = self._side
        else:
            side = stream
        with torch.cuda.stream(stream):
            sparse = torch.empty((n, sh_, sw), device="cuda")
            salpha = torch.empty_like(sparse)
            sse = torch.empty((2, n), device="cuda")  # squared, fourth-power error sums
            flags = torch.empty((n,), dtype=torch.int32, device="cuda")
        # the sparse exact frames: on `stream` before the learned frames, or on a second
        # stream concurrently with them (judge_concurrent)
        if side is not stream:
            side.wait_stream(stream)
        split = self.judge_seg is not None and all(
            p.alpha_pow == 0 and p.filter == _lib.FILTER_LINEAR for p in plist)
        keep = None
        if self.judge_slab_kb is not None:  # the sparse frames' launch tuning
            keep = self.ctx.get_tuning("slab_kb")
            self.ctx.set_tuning("slab_kb", self.judge_slab_kb)
        try:
            if split:  # every ray's samples split over several waves (few, scattered rays)
                frames = sparse_frames(slot, plist, w, h, S, [(JUDGE_OFFSET, JUDGE_OFFSET)],
                                       sparse.data_ptr(), salpha.data_ptr())
                self.ctx.render_frames_split(frames, seg=self.judge_seg, stream=side.cuda_stream)
            else:
                groups = {}
                for k, p in enumerate(plist):
                    groups.setdefault(bytes(p)[_lib.SprParams.box.offset:], []).append(k)
                for ks in groups.values():
                    qs = [sparse_params(plist[k], w, h, stride=S) for k in ks]
                    jobs = [(slot, list(q.inv_p), list(q.inv_m), sparse[k].data_ptr(),
                             salpha[k].data_ptr()) for k, q in zip(ks, qs)]
                    self.ctx.render_batch(jobs, sw, sh_, params=qs[0], stream=side.cuda_stream)
        finally:
            if keep is not None:
                self.ctx.set_tuning("slab_kb", keep)
        if side is not stream:
            sparse.record_stream(side)
            salpha.record_stream(side)
            done = torch.cuda.Event()
            done.record(side)
        self.render_many(slot, plist, w, h, outs, stream, alphas)
        if side is not stream:
            stream.wait_event(done)
        fptrs = [o.data_ptr() for o in outs]
        self.net.judge_sse_frames(fptrs, w, h, sparse.data_ptr(), sw, sh_, JUDGE_OFFSET,
                                  JUDGE_OFFSET, sse[0].data_ptr(), sse[1].data_ptr(), stream=sh)
        mse_max = 10. ** (-(self.threshold + margin) / 10.)
        self.net.judge_flags_bound(sse[0].data_ptr(), sse[1].data_ptr(), n, sw * sh_, mse_max,
                                   self.judge_z, flags.data_ptr(), stream=sh)
        sums2 = flags1 = None
        if split and self.judge_refine and S % 2 == 0:
            # second stage, for the frames the first did not accept: the other three pixels
            # of every (S/2) x (S/2) cell, offsets (S/2, 0), (0, S/2), (S/2, S/2) from the
            # first grid, rendered only for those frames (device gates), and the bound over
            # all four grids' pixels
            flags1 = flags
            offs = [(JUDGE_OFFSET + S // 2, JUDGE_OFFSET), (JUDGE_OFFSET, JUDGE_OFFSET + S // 2),
                    (JUDGE_OFFSET + S // 2, JUDGE_OFFSET + S // 2)]
            with torch.cuda.stream(stream):
                extra = torch.empty((3, n, sh_, sw), device="cuda")
                xalpha = torch.empty_like(extra)
                xsums = torch.empty((3, 2, n), device="cuda")
                gates = flags1.repeat(3)
            frames = sparse_frames(slot, plist, w, h, S, offs, extra.data_ptr(),
                                   xalpha.data_ptr())
            self.ctx.render_frames_split(frames, seg=self.judge_seg, stream=sh,
                                         d_gates=gates.data_ptr())
            for j, (ox, oy) in enumerate(offs):
                self.net.judge_sse_frames(fptrs, w, h, extra[j].data_ptr(), sw, sh_, ox, oy,
                                          xsums[j, 0].data_ptr(), xsums[j, 1].data_ptr(),
                                          stream=sh)
            with torch.cuda.stream(stream):
                sums2 = sse + xsums.sum(0)
                flags2 = torch.empty_like(flags1)
            self.net.judge_flags_bound(sums2[0].data_ptr(), sums2[1].data_ptr(), n, 4 * sw * sh_,
                                       mse_max, self.judge_z, flags2.data_ptr(), stream=sh)
            with torch.cuda.stream(stream):
                flags = flags1 & flags2
        for k, (p, out, alpha) in enumerate(zip(plist, outs, alphas)):
            self.ctx.render_slot_if(slot, p, w, h, out.data_ptr(), alpha.data_ptr(),
                                    flags.data_ptr() + 4 * k, stream=sh)
        if flags1 is None:
            res = JudgedBatch(sse, flags, sw * sh_, self.judge_z)
        else:
            res = JudgedBatch(sse, flags1, sw * sh_, self.judge_z, sums2, flags)
        if not wait:
            return res
        stream.synchronize()
        return res.kept(), res.estimates()

    def render_or_exact(self, slot, p, w, h, d_out, d_alpha, stream):
        """Learned frame when the slot validated, else the exact render."""
        if self.enabled.get(slot):
            self.render(slot, p, w, h, d_out, stream)
            return True
        self.ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, w, h, d_out, d_alpha, stream=stream)
        return False

    def validate(self, slot, param_list, w, h):
        """PSNR and max |error| of the learned frames against the exact ones; enables the
        learned path for `slot` iff every frame reaches the threshold."""
        torch = self.torch
        s = torch.cuda.Stream()
        exact = torch.empty((h, w), device="cuda")
        alpha = torch.empty_like(exact)
        learned = torch.empty_like(exact)
        rep = []
        for p in param_list:
            self.ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, w, h, exact.data_ptr(),
                                 alpha.data_ptr(), stream=s.cuda_stream)
            self.render(slot, p, w, h, learned.data_ptr(), s.cuda_stream)
            s.synchronize()
            e, l = exact.cpu().numpy(), learned.cpu().numpy()
            rep.append({"psnr": psnr(l, e), "max_abs": float(np.abs(l - e).max())})
        self.enabled[slot] = bool(rep) and min(r["psnr"] for r in rep) >= self.threshold
        return rep


# ------------------------------------------------------------------------------ training
def training_volumes(n, shape=(512, 512, 512), seed0=1):
    """Seeded uint16 blob volumes (2000 blobs per 512^3 recipe), as device int16 tensors."""
    from . import synth
    vols = []
    for k in range(n):
        nb = int(round(2000 * np.prod(shape) / 512 ** 3))
        vols.append(next(synth.blob_series_device(shape, nb, (3., 10.), (500., 3000.), 100., True,
                                                  5., 1, seed=seed0 + k)))
    return vols


def train(ctx, slots, ranges, steps=2500, batch=8, frame=1024, max_steps=896, lr=2e-3,
          seed=0, params=None, log=None, radius=2, inputs=INPUTS_DEFAULT):
    """Trains the network on coarse/fine pairs rendered from random poses of `slots`
    (ranges: per-slot (min, max) used to normalise).  Returns (params, history)."""
    import torch
    from . import synth
    from .utils import mat4_perspective
    from .volumerender.volumerender import render_matrices
    net = Net(0, init_params(seed) if params is None else params)
    lm = LearnedMIP(ctx, net=net, radius=radius, inputs=inputs)
    for s in slots:
        lm.attach(s)
    rng = np.random.default_rng(seed)
    P = mat4_perspective(45, 1., .1, 10)
    stream = torch.cuda.Stream()
    W = H = frame
    planes = torch.empty((batch, 2, H // 2, W // 2), device="cuda")
    target = torch.empty((batch, H, W), device="cuda")
    alpha = torch.empty((H, W), device="cuda")
    hist = []
    t0 = time.time()
    for it in range(steps):
        for k in range(batch):
            j = int(rng.integers(len(slots)))
            info = ctx.volume_info(slots[j])
            nz, ny, nx = info["shape"]
            M = synth.random_poses(1, seed=int(rng.integers(1 << 30)),
                                   distance=float(rng.uniform(3., 5.)))[0]
            iP, iM = render_matrices(M, P, (nx, ny, nz))
            lo, hi = ranges[j]
            p = _lib.make_params(iP, iM, (-1, 1, -1, 1, -1, 1), lo, hi, 1., 0., .1, 21, 30,
                                 max_steps, _lib.FILTER_LINEAR)
            ctx.render_slot(slots[j], p, _lib.METHOD_MAX_PROJECT, W, H, target[k].data_ptr(),
                            alpha.data_ptr(), stream=stream.cuda_stream)
            lm.coarse(slots[j], p, W, H, planes, k, stream.cuda_stream)
        frac = it / max(1, steps - 1)
        lr_t = lr * (.05 + .95 * .5 * (1 + math.cos(math.pi * frac)))
        want = it % 50 == 0 or it == steps - 1
        sse = net.train_step(planes.data_ptr(), target.data_ptr(), batch, H // 2, W // 2, lr_t,
                             stream.cuda_stream, want_loss=want)
        if want:
            mse = sse / (batch * H * W)
            hist.append((it, mse))
            if log:
                log("train step %d/%d  mse %.3e  (%.1f dB)  lr %.2e  %.0f s" % (
                    it, steps, mse, 10 * math.log10(1. / max(mse, 1e-30)), lr_t, time.time() - t0))
    stream.synchronize()
    return net.get_params(), hist


def main():
    """python -m spimagine_amd.learned [--steps N] [--volumes N] [--lr LR] [--init PATH]
    [--seed S] [--out PATH]: train on seeded 512^3 uint16 blob volumes (seeds 1..N) and
    save the parameters; --init continues from saved parameters (their input kind)."""
    import argparse
    import sys
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--volumes", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None)
    ap.add_argument("--out", default=WEIGHTS_PATH)
    ap.add_argument("--inputs", choices=INPUTS, default=INPUTS_DEFAULT)
    args = ap.parse_args()

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    params0, inputs, pairs0 = None, args.inputs, 0
    if args.init:
        params0, inputs = load_weights(args.init)
        with np.load(args.init) as w:
            pairs0 = int(w["pairs"]) if "pairs" in w.files else 0
    ctx = _lib.Context(0)
    slots, ranges = [], []
    for v in training_volumes(args.volumes):
        u = v.to(torch.int32) & 0xFFFF
        ranges.append((float(u.min()), float(u.max())))
        del u
        slots.append(ctx.volume_add(dev_ptr=v.data_ptr(), dtype=np.uint16, shape=tuple(v.shape)))
        del v
    log("training on %d volumes, %d steps x %d pairs" % (len(slots), args.steps, args.batch))
    params, hist = train(ctx, slots, ranges, steps=args.steps, batch=args.batch, lr=args.lr, log=log,
                         inputs=inputs, params=params0, seed=args.seed)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.savez(args.out, params=params, history=np.array(hist, np.float64),
             pairs=pairs0 + args.steps * args.batch, inputs=inputs)
    log("saved %s" % args.out)


if __name__ == "__main__":
    main()
