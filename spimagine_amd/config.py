"""Renderer settings read from ~/.spimagine.

Same file and keys as spimagine/config/config.py:14-31 and myconfigparser.py:19-60:
"key = value" lines without a section header.  Only the keys the ray caster uses are
read here:
  max_steps       samples per ray (default 200); a runtime kernel argument here, the
                  reference compiles it in as -D maxSteps (volumerender.py:138-139)
  interpolation   default sampler filter for VolumeRenderer ("linear")
  id_device       HIP device index (the reference's OpenCL device id)
"""
import configparser
import itertools
import logging
import os

logger = logging.getLogger(__name__)

__CONFIGFILE__ = os.path.expanduser("~/.spimagine")

defaults = {
    "id_device": 0,
    "max_steps": 200,
    "interpolation": "linear",
}


def read_config(path=__CONFIGFILE__):
    """Parse a section-less key=value file; a missing or unreadable file gives {}."""
    parser = configparser.ConfigParser()
    try:
        with open(path) as f:
            parser.read_file(itertools.chain(("[dummy]",), f))
    except Exception as e:  # missing file, parse error: defaults apply
        logger.debug("config %s not read: %s", path, e)
        return {}
    return dict(parser.items("dummy")) if parser.has_section("dummy") else {}


def get_param(name, type_, path=__CONFIGFILE__):
    vals = read_config(path)
    try:
        return type_(vals.get(name, defaults[name]))
    except (TypeError, ValueError):
        return type_(defaults[name])


__DEFAULTMAXSTEPS__ = get_param("max_steps", int)
__DEFAULT_INTERP__ = get_param("interpolation", str)
__ID_DEVICE__ = get_param("id_device", int)
