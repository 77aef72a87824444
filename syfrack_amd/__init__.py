"""MI355X (gfx950) sphere-traced distance-estimator path of SyFrack, behind a C ABI.

``syfrack_amd.abi`` loads the in-tree HIP library; ``syfrack_amd.scene`` lowers classic
fractal scenes to job descriptions; ``syfrack_amd.render`` mirrors the reference Renderer.
"""
from . import abi, scene  # noqa: F401

__all__ = ["abi", "scene"]
