"""MI355X-native improved W-stacking imager (drop-in for cuGridder's curig).

``from cugridder_amd import ms2dirty, dirty2ms`` gives the NIFTY-compatible
functions of the reference's ``curig`` package (python/curagridder/
__init__.py:1-3); ``vis2dirty`` / ``dirty2vis`` the plain complex ones; and
``Context`` runs the same pipeline on device-resident torch tensors.
"""
from .cursl import dirty2ms, ms2dirty
from .curagridder import (dirty2vis, dirty2vis_mixed, nufft2d1, nufft2d2, vis2dirty,
                          vis2dirty_mixed)
from ._lib import LIB_PATH, lib

__all__ = ["ms2dirty", "dirty2ms", "vis2dirty", "dirty2vis", "vis2dirty_mixed",
           "dirty2vis_mixed", "nufft2d1", "nufft2d2", "Context", "lib", "LIB_PATH"]


def __getattr__(name):
    if name == "Context":
        from .context import Context
        return Context
    raise AttributeError(name)
