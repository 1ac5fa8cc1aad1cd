"""magace-mi355x: MI355X-native evaluator of magnetic-ACE energies, forces and torques.

The package directory is ``mag-ace_amd/`` (not an importable identifier); it is loaded as
``magace_amd`` by ``__graft_entry__.load_package()``.  Contents:

* ``libmagace.so``  -- HIP kernels + the libace-compatible C-ABI (include/magace.h)
* ``capi``          -- ctypes binding of that C-ABI
* ``systems``       -- synthetic bcc-Fe workloads and the pair_spin_ace array packing
* ``potgen``        -- synthetic potential files in the reference text format
* ``fit``           -- energies/forces of configuration batches as a torch.autograd.Function
                       of the coefficients, and the parameter-file writer
* ``md``            -- device-resident spin-lattice dynamics (graph-captured steps)
* ``neighbors``     -- GPU neighbour lists with the reference neilist's pair sets, single
                       structures or batches packed as packstructures packs them
"""
from . import capi, fit, md, neighbors, potgen, systems  # noqa: F401

__all__ = ["capi", "fit", "md", "neighbors", "potgen", "systems"]
