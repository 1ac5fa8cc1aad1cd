"""Spin-lattice dynamics on the device (libmagace ``magace_md_*``).

Host-side mirror of the loop LAMMPS runs around pair_spin_ace
(lammps/pair_spin_ace.cpp:340-669 force/field assembly, :677-986 per-spin sweep), kept on
the GPU: velocity Verlet for positions (metal units: A, ps, amu, eV), rotations of the spins
about Omega_i = -(2/hbar) dE/dm_i with the field extrapolated to mid-step (3/2 G_n - 1/2
G_{n-1}), one evaluation per step, neighbour lists rebuilt
every ``rebuild_every`` steps and the steps in between replayed from a captured HIP graph.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi

HBAR = 6.582119569e-4  # eV ps
MVV2E = 1.0364269e-4   # amu A^2 / ps^2 -> eV
KB = 8.617333262e-5    # eV / K
MASS_FE = 55.845


class SpinLatticeMD:
    def __init__(self, pos, cell, sdir, smag, vel=None, mass=MASS_FE, dt=1e-3, rcut=5.0, skin=0.5,
                 rebuild_every=10):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        self.n = len(pos)
        vel = np.zeros_like(pos) if vel is None else np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        cellf = np.ascontiguousarray(np.asarray(cell, np.float64).reshape(3, 3)).reshape(-1)
        h = C.c_void_p()
        rc = capi.lib().magace_md_create(self.n, cellf, float(rcut + skin), float(mass), float(dt),
                                         int(rebuild_every), pos.reshape(-1), vel.reshape(-1),
                                         np.ascontiguousarray(sdir, np.float64).reshape(-1),
                                         np.ascontiguousarray(smag, np.float64).reshape(-1), C.byref(h))
        capi._check(rc)
        self.h = h

    def run(self, nsteps, use_graph=True):
        capi._check(capi.lib().magace_md_run(self.h, int(nsteps), int(bool(use_graph))))

    def state(self):
        n = self.n
        x, v, s, f, g = (np.zeros(3 * n) for _ in range(5))
        e = np.zeros(n)
        vir = np.zeros(9)
        ep, ek = C.c_double(), C.c_double()
        cnt = np.zeros(4, np.int64)
        rc = capi.lib().magace_md_read(self.h, x.ctypes.data, v.ctypes.data, s.ctypes.data, f.ctypes.data,
                                       g.ctypes.data, e.ctypes.data, C.addressof(ep), C.addressof(ek),
                                       vir.ctypes.data, cnt.ctypes.data)
        capi._check(rc)
        return dict(x=x.reshape(n, 3), v=v.reshape(n, 3), sdir=s.reshape(n, 3), f=f.reshape(n, 3),
                    dEdm=g.reshape(n, 3), eatom=e, epot=ep.value, ekin=ek.value, virial=vir.reshape(3, 3),
                    steps=int(cnt[0]), rebuilds=int(cnt[1]), captures=int(cnt[2]), pairs=int(cnt[3]))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            capi.lib().magace_md_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def maxwell_velocities(n, temperature, rng, mass=MASS_FE):
    """Velocities (A/ps) from the Maxwell distribution at `temperature` K, zero total momentum."""
    sigma = np.sqrt(KB * temperature / (mass * MVV2E))
    v = rng.normal(0.0, sigma, size=(n, 3))
    return v - v.mean(axis=0)
