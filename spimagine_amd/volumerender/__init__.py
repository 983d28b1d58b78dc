"""spimagine.volumerender equivalent: ``VolumeRenderer`` re-exported at package level
(the reference's volumerender/__init__.py is empty, so only the
``volumerender.volumerender`` path resolves there; both resolve here)."""
from .volumerender import VolumeRenderer, DeviceVolume  # noqa: F401
