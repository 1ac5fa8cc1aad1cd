"""Synthetic workloads and the caller-side packing of the libace arrays.

The packing reproduces what LAMMPS ``PairSpinACE::compute`` hands to ``acelib_``
(lammps/pair_spin_ace.cpp:419-543): a full neighbour list flattened per central atom,
1-based ``nstart``/``nstop`` ranges, ``elist`` = (x_i - x_j)/r, neighbour moment magnitude
``dlist_mom`` = sp[j][3] and unit direction ``elist_mom`` = sp[j][0:3], and the same for
the central atom (``dlist_momi``/``elist_momi``).  Neighbour lists here are built on the
host with numpy/scipy; they are input preparation (the caller's job in the reference
path, where LAMMPS builds them), not part of the timed evaluation.
"""
from __future__ import annotations

import numpy as np

A_FE = 2.83


def bcc(ncell: int, a: float = A_FE):
    """bcc lattice of ncell^3 conventional cells (2 atoms each) in a cubic periodic box."""
    g = np.arange(ncell, dtype=np.float64)
    cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
    corners = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    pos = np.concatenate([corners, corners + 0.5], axis=0)
    # interleave the two basis atoms per cell so nearby indices are nearby in space
    pos = pos.reshape(2, -1, 3).transpose(1, 0, 2).reshape(-1, 3) * a
    cell = np.eye(3) * (a * ncell)
    return pos, cell


def bcc_box(n1: int, n2: int, n3: int, a: float = A_FE):
    """bcc lattice of n1 x n2 x n3 conventional cells in an orthorhombic periodic box."""
    g = np.stack(np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij"), -1)
    corners = g.reshape(-1, 3).astype(np.float64)
    pos = np.stack([corners, corners + 0.5], axis=1).reshape(-1, 3) * a
    return pos, np.diag([n1 * a, n2 * a, n3 * a])


def config_batch(nconf: int = 256, seed: int = 2029, lo: int = 16, hi: int = 1024, a: float = A_FE):
    """Batched-fit workload: nconf periodic bcc Fe configurations with sizes log-uniform in
    [lo, hi] atoms (rounded to 2*n1*n2*n3), isotropic strain U[-3%, +3%], 0-2 random
    vacancies, jitter N(0, 0.05 A), directions uniform on the sphere with magnitudes
    N(2.2, 0.1) muB.  Returns a list of dicts: pos, cell (rows = lattice vectors), sdir,
    smag, mag = sdir * smag, periodic."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nconf):
        target = np.exp(rng.uniform(np.log(lo), np.log(hi)))
        nc = max(1, int(round(target / 2)))
        n1 = max(1, int(round(nc ** (1.0 / 3.0))))
        n2 = max(1, int(round(np.sqrt(nc / n1))))
        n3 = max(1, int(round(nc / (n1 * n2))))
        strain = rng.uniform(-0.03, 0.03)
        pos, cell = bcc_box(n1, n2, n3, a * (1.0 + strain))
        nvac = int(rng.integers(0, 3))
        if nvac and len(pos) > nvac + 1:
            pos = np.delete(pos, rng.choice(len(pos), nvac, replace=False), axis=0)
        pos = jitter(pos, rng, 0.05)
        sdir, smag = spins_uniform(len(pos), rng)
        out.append(dict(pos=pos, cell=cell, sdir=sdir, smag=smag, mag=sdir * smag[:, None], periodic=True))
    return out


def jitter(pos, rng, sigma=0.05):
    return pos + rng.normal(0.0, sigma, size=pos.shape)


def spins_uniform(n, rng, mean=2.2, std=0.1):
    """Unit directions uniform on the sphere, magnitudes N(mean, std) (muB)."""
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    mag = rng.normal(mean, std, size=n)
    return v, mag


def spins_cone(n, rng, max_deg=30.0, mean=2.2, std=0.1):
    """Ferromagnetic along +z with polar angle uniform in [0, max_deg], uniform azimuth."""
    th = np.deg2rad(rng.uniform(0.0, max_deg, size=n))
    ph = rng.uniform(0.0, 2 * np.pi, size=n)
    v = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    mag = rng.normal(mean, std, size=n)
    return v, mag


def neighbor_pairs(pos, cell, rcut):
    """Full neighbour list (both i->j and j->i) of a periodic orthorhombic box.

    Returns (i, j, d) with d = x_j + shift - x_i for every pair with |d| < rcut, sorted by
    (i, j, shift).  Small boxes (edge < 2 rcut) enumerate the periodic images explicitly.
    """
    L = np.diag(cell).astype(np.float64)
    n = len(pos)
    if np.all(L > 2.0 * rcut + 1e-9) and n > 64:
        from scipy.spatial import cKDTree
        wrapped = np.mod(pos, L)
        tree = cKDTree(wrapped, boxsize=L)
        pairs = tree.query_pairs(rcut, output_type="ndarray")
        i = np.concatenate([pairs[:, 0], pairs[:, 1]])
        j = np.concatenate([pairs[:, 1], pairs[:, 0]])
        d = wrapped[j] - wrapped[i]
        d -= L * np.round(d / L)
    else:
        reps = np.ceil(rcut / L).astype(int)
        shifts = np.array([[a, b, c] for a in range(-reps[0], reps[0] + 1)
                           for b in range(-reps[1], reps[1] + 1)
                           for c in range(-reps[2], reps[2] + 1)], dtype=np.float64) * L
        ii, jj, dd = [], [], []
        for s in shifts:
            d = pos[None, :, :] + s[None, None, :] - pos[:, None, :]
            r2 = np.einsum("ijk,ijk->ij", d, d)
            m = r2 < rcut * rcut
            if not np.any(s):
                np.fill_diagonal(m, False)
            a, b = np.nonzero(m)
            ii.append(a); jj.append(b); dd.append(d[a, b])
        i, j, d = np.concatenate(ii), np.concatenate(jj), np.concatenate(dd)
    r = np.linalg.norm(d, axis=1)
    keep = r < rcut
    i, j, d = i[keep], j[keep], d[keep]
    order = np.lexsort((d[:, 2], d[:, 1], d[:, 0], j, i))
    return i[order].astype(np.int64), j[order].astype(np.int64), d[order]


def pack_acelib(pos, cell, sdir, smag, rcut):
    """Arrays of acelib_ (src/ace.F90:79-107) in the pair_spin_ace convention.

    Also returns 'offsets' (0-based CSR, natom+1) and 'pair_i'/'pair_j' for assembly.
    """
    n = len(pos)
    i, j, d = neighbor_pairs(pos, cell, rcut)
    r = np.linalg.norm(d, axis=1)
    if np.any(r < 1e-6):
        raise ValueError("Atoms with identical positions.")  # pair_spin_ace.cpp:502-506
    counts = np.bincount(i, minlength=n)
    offsets = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=offsets[1:])
    nstart = (offsets[:-1] + 1).astype(np.int32)
    nstop = offsets[1:].astype(np.int32)   # inclusive 1-based: offsets[i+1]-1+1
    out = dict(
        natom=n, nall=n,
        nstart=nstart, nstop=nstop,
        nlist=(j + 1).astype(np.int32),
        occ=np.ones(n, np.int32), occlist=np.ones(len(j), np.int32),
        dlist=np.ascontiguousarray(r),
        elist=np.ascontiguousarray((-d / r[:, None]).ravel()),
        dlist_momi=np.ascontiguousarray(smag, dtype=np.float64),
        elist_momi=np.ascontiguousarray(sdir, dtype=np.float64).ravel(),
        dlist_mom=np.ascontiguousarray(smag[j], dtype=np.float64),
        elist_mom=np.ascontiguousarray(sdir[j], dtype=np.float64).ravel(),
        offsets=offsets, pair_i=i, pair_j=j,
    )
    return out


def assemble(a, eout, flist, tlist, tlist0):
    """Caller-side assembly of pair_spin_ace.cpp:571-641.

    f_i -= F_p, f_j += F_p ; fm_i -= T0_i ; fm_j -= T_p ; E = sum eout ;
    virial W_ab = sum_p u_a F_b with u = x_i - x_j (the tally of pair_spin_ace.cpp:635-636).
    Returns (energy, forces (N,3), fm (N,3), virial (3,3)).
    """
    n = a["natom"]
    F = flist.reshape(-1, 3)
    T = tlist.reshape(-1, 3)
    f = np.zeros((n, 3))
    fm = -np.asarray(tlist0, dtype=np.float64).reshape(-1, 3).copy()
    for k in range(3):
        f[:, k] += np.bincount(a["pair_j"], weights=F[:, k], minlength=n)
        f[:, k] -= np.bincount(a["pair_i"], weights=F[:, k], minlength=n)
        fm[:, k] -= np.bincount(a["pair_j"], weights=T[:, k], minlength=n)
    u = a["elist"].reshape(-1, 3) * a["dlist"][:, None]
    vir = u.T @ F
    return float(np.sum(eout)), f, fm, vir


def make_system(ncell, seed, spins="uniform", rcut=5.0, jitter_sigma=0.05, cone_deg=30.0):
    rng = np.random.default_rng(seed)
    pos, cell = bcc(ncell)
    pos = jitter(pos, rng, jitter_sigma)
    if spins == "uniform":
        sdir, smag = spins_uniform(len(pos), rng)
    elif spins == "cone":
        sdir, smag = spins_cone(len(pos), rng, cone_deg)
    else:
        raise ValueError(spins)
    return dict(pos=pos, cell=cell, sdir=sdir, smag=smag, rcut=rcut, seed=seed)
