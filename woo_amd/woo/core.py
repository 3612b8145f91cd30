"""woo.core: Scene, engines, nodes (src/core/Scene.hpp, src/core/Engine.hpp)."""
from woo_amd._woo import Cell, Engine, Node, Object, PeriodicEngine, PyRunner, Scene  # noqa: F401
