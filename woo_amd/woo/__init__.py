"""Woo's Python surface (``woo.core``, ``woo.dem``, ``woo.utils``) over the MI355X engine.

Scenes are scripted as with the reference (py/__init__.py, py/utils.py)::

    import woo_amd.woo as woo            # or woo_amd.woo.install(), then ``import woo``
    from woo_amd.woo.dem import *
    S = woo.core.Scene(fields=[DemField(gravity=(0, 0, -9.81))])
    S.dem.par.add([woo.utils.sphere((0, 0, .1), .01), woo.utils.wall(0, axis=2, sense=1)])
    S.engines = woo.utils.defaultEngines(damping=.2)
    S.run(1000)

The classes come from the compiled host module ``woo_amd._woo`` (C++), whose engines
drive the HIP kernels through the C ABI in include/woo_amd.h.
"""
import sys

from . import core, dem, triangulated, utils  # noqa: F401


def install():
    """Register this package as ``woo`` (and ``woo.core``, ``woo.dem``, ``woo.utils``, ``woo.triangulated``),
    so that scripts written for the reference import it unchanged."""
    me = sys.modules[__name__]
    sys.modules.setdefault("woo", me)
    for sub in ("core", "dem", "utils", "triangulated"):
        sys.modules.setdefault("woo." + sub, sys.modules[__name__ + "." + sub])
    return me
