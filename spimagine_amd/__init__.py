"""spimagine_amd: spimagine's volume ray caster (max projection, opacity transfer,
iso surface) as hand-written gfx950 HIP kernels behind a C ABI (libspimrender.so,
include/spimrender.h), driven by a drop-in ``VolumeRenderer``.

    from spimagine_amd.volumerender import VolumeRenderer
    from spimagine_amd.utils import mat4_perspective, mat4_translate, mat4_rotation
"""
__version__ = "0.1.0"
