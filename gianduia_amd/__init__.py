"""gianduia_amd — MI355X-native ray-tracing hot path of the Gianduia renderer.

The compute path is libgnd_hip.so (hand-written HIP for gfx950 plus the host
scene shell), reached through the C-ABI in include/gnd_hip.h and
include/gnd_host.h.  This package is the Python driver of that library.
"""
from ._abi import (GND_CAMERA_PERSPECTIVE, GND_CAMERA_THINLENS, GND_FILTER_BOX, GND_FILTER_GAUSSIAN,
                   GND_INTEGRATOR_PATH, GND_INTEGRATOR_VOLPATH, load_library)
from .api import (Device, DeviceScene, GndError, HostScene, Integrator, Job, SceneObjects, camera_lookat,
                  integrator_names, load_exr, resolve, resolve_aovs, sampler_floats)

__all__ = ["Device", "DeviceScene", "GndError", "HostScene", "Integrator", "Job", "SceneObjects", "camera_lookat",
           "integrator_names",
           "load_exr", "resolve", "resolve_aovs", "sampler_floats", "load_library",
           "GND_CAMERA_PERSPECTIVE", "GND_CAMERA_THINLENS", "GND_FILTER_BOX", "GND_FILTER_GAUSSIAN",
           "GND_INTEGRATOR_PATH", "GND_INTEGRATOR_VOLPATH"]
