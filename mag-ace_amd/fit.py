"""Fitting on the GPU: energies and forces of a batch of configurations as a
``torch.autograd.Function`` of the expansion coefficients.

The reference ships the pieces a fit needs -- the coefficient contraction
(src/forces.F90:93-141), the structure packing (src/neiutil.F90:5-264) and the parameter
writer (src/util.F90:1110-1362) -- but not the fitting driver.  Here:

* ``Batch`` packs configurations with the GPU neighbour build (the packstructures layout);
* ``EnergyForces.apply(c, batch)`` returns per-configuration energies and per-atom forces
  (FP64); its backward gives dL/dc for any loss L(E, F) through ``magace_fit_backward``
  (forward-mode differentiation of the density along dL/dF, one accumulation over the
  monomials), so a loss like sum w_E (E - E_ref)^2 + w_F |F - F_ref|^2 trains with any
  torch optimiser;
* ``write_potential(path)`` writes the current coefficients in the reference's file format.

Coefficients are ordered as the file's B functions (order 1, then 2, 3, 4).
"""
from __future__ import annotations

import numpy as np

from . import capi
from .neighbors import DeviceNeighbors


class Batch:
    """Configurations packed one after another with device neighbour lists (acelib_
    conventions).  structures: dicts with pos (n, 3), cell (3, 3 rows = lattice vectors),
    sdir (n, 3), smag (n,) and optional periodic."""

    def __init__(self, structures, rcut, device="cuda:0"):
        import torch
        self.torch = torch
        dev = torch.device(device)
        self.natoms = np.array([len(s["pos"]) for s in structures], np.int32)
        self.nstruc = len(structures)
        self.natom = int(self.natoms.sum())
        pos = np.concatenate([np.asarray(s["pos"], np.float64) for s in structures])
        mag = np.concatenate([np.asarray(s["sdir"], np.float64) * np.asarray(s["smag"], np.float64)[:, None]
                              for s in structures])
        per = [int(bool(s.get("periodic", True))) for s in structures]
        cells = np.stack([np.asarray(s["cell"], np.float64) for s in structures])
        self.nb = DeviceNeighbors(self.natoms, per, cells, rcut, device=device)
        L = self.nb.build(torch.tensor(pos, device=dev), torch.tensor(mag, device=dev))
        self.L = L
        self.npairs = L["npairs"]
        self.atom_struct = torch.tensor(np.repeat(np.arange(self.nstruc, dtype=np.int32), self.natoms), device=dev)
        self.device = dev

    def ptrs(self):
        L = self.L
        return (self.natom, self.npairs, L["offsets"].data_ptr(), L["nlist"].data_ptr(), L["dlist"].data_ptr(),
                L["elist"].data_ptr())

    def forward(self, energies, forces):
        L = self.L
        capi.fit_forward(*self.ptrs(), L["dlist_momi"].data_ptr(), L["elist_momi"].data_ptr(),
                         L["dlist_mom"].data_ptr(), L["elist_mom"].data_ptr(), self.atom_struct.data_ptr(),
                         self.nstruc, energies.data_ptr(), forces.data_ptr(),
                         self.torch.cuda.current_stream(self.device).cuda_stream)

    def backward(self, gE, gF, grad):
        L = self.L
        capi.fit_backward(*self.ptrs(), L["dlist_momi"].data_ptr(), L["dlist_mom"].data_ptr(),
                          L["elist_mom"].data_ptr(), self.atom_struct.data_ptr(),
                          gE.data_ptr() if gE is not None else 0, gF.data_ptr(), grad.data_ptr(),
                          self.torch.cuda.current_stream(self.device).cuda_stream)


try:
    import torch as _t

    class EnergyForces(_t.autograd.Function):  # noqa: F811
        @staticmethod
        def forward(ctx, coeffs, batch):
            capi.fit_set_coefficients(coeffs.detach().to("cpu", _t.float64).numpy())
            E = _t.zeros(batch.nstruc, dtype=_t.float64, device=batch.device)
            F = _t.zeros(batch.natom, 3, dtype=_t.float64, device=batch.device)
            batch.forward(E, F)
            ctx.batch = batch
            ctx.save_for_backward(coeffs.detach().clone())
            return E, F

        @staticmethod
        def backward(ctx, gE, gF):
            (coeffs,) = ctx.saved_tensors
            batch = ctx.batch
            # re-evaluate so that the device state (Z, W) belongs to these coefficients and atoms
            capi.fit_set_coefficients(coeffs.to("cpu", _t.float64).numpy())
            E = _t.zeros(batch.nstruc, dtype=_t.float64, device=batch.device)
            F = _t.zeros(batch.natom, 3, dtype=_t.float64, device=batch.device)
            batch.forward(E, F)
            gF = (gF if gF is not None else _t.zeros_like(F)).contiguous().to(_t.float64)
            gE = gE.contiguous().to(_t.float64) if gE is not None else None
            grad = _t.zeros(coeffs.numel(), dtype=_t.float64, device=batch.device)
            batch.backward(gE, gF, grad)
            return grad.to(coeffs.device), None
except ImportError:  # pragma: no cover
    pass


def write_potential(path):
    """Current coefficients in the reference file format (src/util.F90:1110-1362)."""
    capi.write_potential(path)
