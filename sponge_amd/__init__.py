"""MI355X-native (gfx950) nonbonded hot path of the SPONGE MD engine.

The product is libsponge_nb.so (hand-written HIP, C ABI in include/sponge_nb.h);
`sponge_amd.nb` mirrors SPONGE's NEIGHBOR_LIST / LENNARD_JONES_INFORMATION /
Particle_Mesh_Ewald over that ABI, `sponge_amd.system` builds the seeded
synthetic water boxes.
"""
from . import system  # noqa: F401

__all__ = ["system", "nb", "build"]
