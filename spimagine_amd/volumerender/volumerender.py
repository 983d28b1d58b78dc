"""VolumeRenderer: spimagine's ray-casting renderer on MI355X.

Drop-in for spimagine/volumerender/volumerender.py:59-547.  The public surface is kept:
the class attribute ``dtypes``, the constructor, ``rebuild_program``, the setters,
``set_data``/``update_data``, ``render(...)`` (which returns None, like the reference)
and the ``output`` / ``output_alpha`` / ``output_depth`` (+ ``output_normals`` /
``output_occlusion`` after an iso render) arrays the viewer reads.

All pixels come from libspimrender.so (hand-written gfx950 kernels behind a C ABI);
this class keeps the host-side state exactly as the reference does (matrix inversion in
the same precision, downsampling rule, setter semantics) and hands it to the library.

Differences from the reference, all deliberate:
  - max_steps is a runtime argument (``set_max_steps``; default from ~/.spimagine),
    not a compile-time define, so changing it needs no rebuild;
  - ``numParts`` larger than max_steps/16 raises ValueError (the reference divides by
    zero inside the kernel, volume_kernel.cl:102);
  - the volume budget for downsampling is 0.7 x the device's total HBM (the reference
    uses 0.7 x CL_DEVICE_MAX_MEM_ALLOC_SIZE, volumerender.py:102).
"""
from __future__ import absolute_import, print_function

import logging
from time import time

import numpy as np
from scipy.linalg import inv

from .. import _lib
from .. import config
from ..utils.transform_matrices import mat4_identity, mat4_perspective, mat4_scale

logger = logging.getLogger(__name__)


def stack_scale_mat(shape_xyz, stackUnits):
    """mScale = diag(dx*Nx, dy*Ny, dz*Nz) / max(d*N) as float32 (volumerender.py:318-325)."""
    Nx, Ny, Nz = shape_xyz
    dx, dy, dz = stackUnits
    maxDim = max(d * N for d, N in zip([dx, dy, dz], [Nx, Ny, Nz]))
    return mat4_scale(1. * dx * Nx / maxDim, 1. * dy * Ny / maxDim, 1. * dz * Nz / maxDim)


def render_matrices(modelView, projection, shape_xyz, stackUnits=(1., 1., 1.)):
    """(invP, invM) as flat float32 arrays, exactly as update_matrices computes them:
    invM = inv(modelView . mScale), invP = inv(projection), each inverted in the precision
    numpy promotion gives the inputs (float64 for mat4_translate/rotation, float32 for
    mat4_perspective), then cast to float32 (volumerender.py:310-316)."""
    mScale = stack_scale_mat(shape_xyz, np.array(stackUnits))
    invM = inv(np.dot(1. * modelView, mScale))
    invP = inv(projection)
    return invP.flatten().astype(np.float32), invM.flatten().astype(np.float32)


def host_plane(shape):
    """A fresh float32 host array for an output plane.  Pinned host memory when torch can
    provide it (its caching host allocator recycles a block once the array is garbage), so
    the library's device-to-host copies run at DMA rate: a 1024^2 max projection's three
    planes (12 MB) took 2.04 ms render + copy into pinned arrays against 2.52 ms into
    pageable ones.  Pageable numpy memory otherwise; the pixels are the same either way."""
    try:
        import torch
        return torch.empty(shape, dtype=torch.float32, pin_memory=True).numpy()
    except Exception:  # no torch, or no device to pin for: pageable memory
        return np.empty(shape, np.float32)


class DeviceVolume(object):
    """The resident volume (the reference's ``dataImg`` OCLImage): ``shape`` is
    (Nx, Ny, Nz) like an OpenCL image, ``dtype`` the voxel type."""

    def __init__(self, ctx, shape, dtype):
        self._ctx = ctx
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)

    def min_max(self):
        """GPU min/max of the volume (replaces GLWidget._get_min_max, glwidget.py:328-344)."""
        return self._ctx.volume_minmax()


class VolumeRenderer(object):
    """renders a data volume by ray casting (max projection, opacity transfer, iso surface)

    usage:
        rend = VolumeRenderer((400, 400))
        rend.set_data(d)
        rend.set_units([1., 1., 2.])
        rend.set_modelView(mat4_rotation(.7, 0, 1, 0))
        rend.render()
        out = rend.output
    """
    dtypes = [np.float32, np.uint16, np.uint8]
    interpolation_defines = {"linear": _lib.FILTER_LINEAR, "nearest": _lib.FILTER_NEAREST}

    def __init__(self, size=None, interpolation='linear', device=None):
        """e.g. size = (300, 300)"""
        if device is None:
            device = config.__ID_DEVICE__
        self._ctx = _lib.Context(device)
        self.isGPU = True
        _, total = _lib.device_memory(device)
        self.memMax = .7 * total
        self.max_steps = config.__DEFAULTMAXSTEPS__

        self.rebuild_program(interpolation=interpolation)

        self.projection = np.zeros((4, 4))
        self.modelView = np.zeros((4, 4))
        self._invM = np.zeros(16, np.float32)
        self._invP = np.zeros(16, np.float32)

        if size:
            self.resize(size)
        else:
            self.resize((200, 200))

        self.set_dtype()
        self.set_gamma()
        self.set_max_val()
        self.set_min_val()

        self.set_occ_strength(.1)
        self.set_occ_radius(21)
        self.set_occ_n_points(30)

        self.set_alpha_pow()
        self.set_box_boundaries()
        self.set_units()

        self.set_modelView()
        self.set_projection()

    # ---------------------------------------------------------------- program
    def rebuild_program(self, interpolation="linear"):
        """Select the sampler filter (the reference recompiles with -D SAMPLER_FILTER)."""
        if interpolation not in VolumeRenderer.interpolation_defines:
            raise KeyError("interpolation = '%s' not defined ,valid: %s"
                           % (interpolation, list(VolumeRenderer.interpolation_defines.keys())))
        self.interpolation = interpolation
        self._filter = VolumeRenderer.interpolation_defines[interpolation]

    def set_max_steps(self, max_steps):
        """Samples per ray (config key max_steps)."""
        if int(max_steps) < 1:
            raise ValueError("max_steps must be positive")
        self.max_steps = int(max_steps)

    # ---------------------------------------------------------------- buffers
    def set_dtype(self, dtype=None):
        if hasattr(self, "dtype") and dtype is self.dtype:
            return
        if dtype is None:
            dtype = self.dtypes[0]
        if dtype in self.dtypes:
            self.dtype = dtype
        else:
            raise NotImplementedError("data type should be either %s not %s" % (self.dtypes, dtype))
        self.reset_buffer()

    def resize(self, size):
        self.width, self.height = size
        self.reset_buffer()

    def reset_buffer(self):
        self.output = np.zeros((self.height, self.width), dtype=np.float32)
        self.output_alpha = np.zeros((self.height, self.width), dtype=np.float32)
        self.output_depth = np.zeros((self.height, self.width), dtype=np.float32)

    # ---------------------------------------------------------------- data
    def _get_downsampled_data_slices(self, data):
        """Integer-stride slices when the volume exceeds the memory budget, else None
        (volumerender.py:199-212)."""
        Nstep = int(np.ceil((1. * data.nbytes / self.memMax) ** (1. / 3)))
        slices = tuple(slice(0, d, Nstep) for d in data.shape)
        if Nstep > 1:
            logger.info("downsample image by factor of  %s" % Nstep)
            return slices
        return None

    def set_max_val(self, maxVal=0.):
        self.maxVal = maxVal

    def set_min_val(self, minVal=0.):
        self.minVal = minVal

    def set_gamma(self, gamma=1.):
        self.gamma = gamma

    def set_occ_strength(self, occ=.2):
        self.occ_strength = occ

    def set_occ_radius(self, rad=21):
        self.occ_radius = rad

    def set_occ_n_points(self, n_points=31):
        self.occ_n_points = n_points

    def set_alpha_pow(self, alphaPow=0.):
        self.alphaPow = alphaPow

    def set_data(self, data, autoConvert=True, copyData=False):
        logger.debug("set_data")
        if not autoConvert and data.dtype not in self.dtypes:
            raise NotImplementedError("data type should be either %s not %s" % (self.dtypes, data.dtype))

        if data.dtype.type in self.dtypes:
            self.set_dtype(data.dtype.type)
            _data = data
        else:
            print("converting type from %s to %s" % (data.dtype.type, self.dtype))
            _data = data.astype(self.dtype, copy=False)

        self.dataSlices = self._get_downsampled_data_slices(_data)
        if self.dataSlices is not None:
            self.set_shape(_data[self.dataSlices].shape[::-1])
        else:
            self.set_shape(_data.shape[::-1])

        t = time()
        self.update_data(_data, copyData=copyData)
        logger.debug("update data: %s ms" % (1000. * (time() - t)))
        self.update_matrices()

    def set_shape(self, dataShape):
        """dataShape = (Nx, Ny, Nz)."""
        self.dataImg = DeviceVolume(self._ctx, dataShape, self.dtype)

    def update_data(self, data, copyData=False):
        if self.dataSlices is not None:
            self._data = data[self.dataSlices].copy()
        else:
            self._data = data.copy() if copyData else data
        if self._data.dtype != self.dtype:
            self._data = self._data.astype(self.dtype, copy=False)
        self._ctx.set_volume(self._data)

    # ---------------------------------------------------------------- pose
    def set_box_boundaries(self, boxBounds=[-1, 1, -1, 1, -1, 1]):
        self.boxBounds = np.array(boxBounds)

    def set_units(self, stackUnits=np.ones(3)):
        self.stackUnits = np.array(stackUnits)

    def set_projection(self, projection=mat4_perspective()):
        self.projection = projection
        self.update_matrices()

    def set_modelView(self, modelView=mat4_identity()):
        self.modelView = 1. * modelView
        self.update_matrices()

    def update_matrices(self):
        """invM = inv(modelView . mScale), invP = inv(projection), flattened to float32
        (volumerender.py:310-316; the inversion precision follows the input dtypes)."""
        if hasattr(self, "dataImg"):
            self._invP, self._invM = render_matrices(self.modelView, self.projection,
                                                     self.dataImg.shape, self.stackUnits)

    def _stack_scale_mat(self):
        return stack_scale_mat(self.dataImg.shape, self.stackUnits)

    # ---------------------------------------------------------------- render
    def _push_params(self):
        self._ctx.set_params(self._invP, self._invM, np.asarray(self.boxBounds, np.float32),
                             self.minVal, self.maxVal, self.gamma, self.alphaPow,
                             self.occ_strength, self.occ_radius, self.occ_n_points,
                             self.max_steps, self._filter)

    def _render_max_project(self, dtype=np.float32, numParts=1, currentPart=0):
        if dtype not in (np.float32, np.uint16, np.uint8):
            raise NotImplementedError("wrong dtype: %s", dtype)
        if numParts < 1 or (self.max_steps // numParts) // 16 < 1:
            raise ValueError("numParts=%s leaves no 16-sample block of max_steps=%s"
                             % (numParts, self.max_steps))
        self._push_params()
        out, alpha, depth = (host_plane((self.height, self.width)) for _ in range(3))
        self._ctx.render(_lib.METHOD_MAX_PROJECT, int(numParts), int(currentPart), self.width,
                         self.height, out, alpha, depth)
        self.output, self.output_alpha, self.output_depth = out, alpha, depth

    def _render_isosurface(self):
        """iso_surface -> conv_vec(7) -> occlusion -> conv(5) -> shading (volumerender.py:446-506)."""
        self._push_params()
        out, alpha, depth, occ = (host_plane((self.height, self.width)) for _ in range(4))
        normals = host_plane((self.height, self.width, 3))
        self._ctx.render(_lib.METHOD_ISO_SURFACE, 1, 0, self.width, self.height, out, alpha,
                         depth, normals, occ)
        self.output, self.output_alpha, self.output_depth = out, alpha, depth
        self.output_normals, self.output_occlusion = normals, occ

    def render(self, data=None, stackUnits=None,
               minVal=None, maxVal=None, gamma=None,
               modelView=None, projection=None,
               boxBounds=None, return_alpha=False, method="max_project",
               numParts=1, currentPart=0):
        """Render into self.output (and friends); returns None like the reference.
        boxBounds and return_alpha are accepted and ignored (volumerender.py:508-547)."""
        if data is not None:
            self.set_data(data)
        if maxVal is not None:
            self.set_max_val(maxVal)
        if minVal is not None:
            self.set_min_val(minVal)
        if gamma is not None:
            self.set_gamma(gamma)
        if stackUnits is not None:
            self.set_units(stackUnits)
        if modelView is not None:
            self.set_modelView(modelView)
        if projection is not None:
            self.set_projection(projection)

        if not hasattr(self, 'dataImg'):
            print("no data provided, set_data(data) before")
            return

        if modelView is None and not hasattr(self, 'modelView'):
            print("no modelView provided and set_modelView() not called before!")
            return

        if method == "max_project":
            self._render_max_project(self.dtype, numParts, currentPart)

        if method == "iso_surface":
            self._render_isosurface()
