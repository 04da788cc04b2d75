"""MI355X-native RPMD force-and-propagate path (DCM-UPB/RPMD_Mainz semantics).

* ``capi``    -- the reference's force-routine interface over the HIP C-ABI (librpmd_mi355x.so)
* ``configs`` -- water models, seeded lattices and the reference's parameter set-up
* ``build``   -- compiles the HIP library for gfx950 in-tree
"""
from . import configs  # noqa: F401

__all__ = ["configs", "capi", "build"]
