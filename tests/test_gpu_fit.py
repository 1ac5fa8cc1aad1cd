"""Fitting gradients on the GPU (magace_fit_*, fit.EnergyForces).

Pins: the energy part of dL/dc is sum_i gE_s(i) fc F'_d B_k,i, checked against the per-atom
B vectors of the oracle (the reference's prepB restated), with one density and with two
(linear embedding: the same B vectors for both densities' coefficients); the full gradient
(energy and force terms) against central differences of L(c) = sum_s gE_s E_s +
sum_j gF_j . F_j, which are exact up to rounding when E is linear in c (nfemb 0) and,
Richardson-extrapolated, checked to 1e-7 for nonlinear embeddings (one or two densities) with
core repulsion and the inner cutoff active; body orders 5 and 6 (products of A entries in the
monomials) by central differences with the linear embedding.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import HEADERS, PKG

capi, potgen, systems, fit, nbr = PKG.capi, PKG.potgen, PKG.systems, PKG.fit, PKG.neighbors
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU visible", allow_module_level=True)

DEV = torch.device("cuda:0")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def structures(seed, sizes=((2, 2, 2), (3, 2, 2), (2, 2, 1))):
    rng = np.random.default_rng(seed)
    out = []
    for n1, n2, n3 in sizes:
        pos, cell = systems.bcc_box(n1, n2, n3, 2.83 * rng.uniform(0.98, 1.02))
        pos = systems.jitter(pos, rng, 0.08)
        sdir, smag = systems.spins_uniform(len(pos), rng)
        out.append(dict(pos=pos, cell=cell, sdir=sdir, smag=smag))
    return out


# Two densities.  Linear embedding: dE_i/dc_{k,d} = B_k,i for both densities.  For central
# differences the embedding is the shifted one (nfemb 2, s = 1 + rho), nonlinear on the second
# density (exponent 0.6): around s = 0 the reference's embedding has an unbounded second
# derivative for exponents below 2/3 and drops the derivative of its first term where that
# term is below 1e-10 (src/forces.F90:1134-1233), so F' is not exactly dF/drho there, and
# finite differences over atoms with rho near 0 miss the analytic gradient by up to 1e-6.
_ND2S = dict(HEADERS["H500ND2"], nfemb=2)
FIT_HEADERS = dict(HEADERS, H500ND2LIN=dict(HEADERS["H500ND2"], nfemb=0), H500ND2S=_ND2S,
                   H500ND2CORE=dict(_ND2S, core_pre=25.0, core_lambda=0.35, ecut=14.0, decut=8.0))


def load(tmp_path, header, seed, c_scale=1e-2):
    pot = str(tmp_path / "Fe.pbe.in")
    h = FIT_HEADERS[header]
    with open(pot, "w") as f:
        f.write(potgen.render(h, capi.enumerate_basis(h), seed, c_scale=c_scale))
    capi.init_file(pot)
    return pot


def loss(batch, c, gE, gF):
    capi.fit_set_coefficients(c)
    E = torch.zeros(batch.nstruc, dtype=torch.float64, device=DEV)
    F = torch.zeros(batch.natom, 3, dtype=torch.float64, device=DEV)
    batch.forward(E, F)
    torch.cuda.synchronize()
    return float((E * gE).sum() + (F * gF).sum())


def gradient(batch, c, gE, gF):
    capi.fit_set_coefficients(c)
    E = torch.zeros(batch.nstruc, dtype=torch.float64, device=DEV)
    F = torch.zeros(batch.natom, 3, dtype=torch.float64, device=DEV)
    batch.forward(E, F)
    g = torch.zeros(len(c), dtype=torch.float64, device=DEV)
    batch.backward(gE, gF.contiguous(), g)
    return g.cpu().numpy()


@pytest.mark.parametrize("header", ["H500", "H500ND2LIN", "H56"])
def test_energy_gradient_equals_oracle_b_vectors(tmp_path, header):
    say("fit: energy gradient vs oracle B vectors,", header)
    pot = load(tmp_path, header, 2040)
    nd = int(FIT_HEADERS[header].get("ndensity", 1))
    structs = structures(1)
    batch = fit.Batch(structs, 5.0)
    c = capi.fit_get_coefficients()
    gE = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64, device=DEV)
    gF = torch.zeros(batch.natom, 3, dtype=torch.float64, device=DEV)
    g = gradient(batch, c, gE, gF)
    o = O.Oracle.load(pot)
    want1 = np.zeros(sum(o.counts()))
    for s, w in zip(structs, gE.cpu().numpy()):
        a = nbr.neilist(s["pos"], s["sdir"] * s["smag"][:, None], s["cell"], 5.0)
        want1 += w * o.bvectors(a).sum(axis=0)
    # coefficient order: per body order, density-major
    cuts = np.cumsum([0] + list(o.counts()))
    want = np.concatenate([np.tile(want1[cuts[k]:cuts[k + 1]], nd) for k in range(6)])
    assert len(want) == len(g)
    err = np.abs(g - want).max() / np.abs(want).max()
    say("  rel err %.2e" % err)
    assert err < 1e-10


@pytest.mark.parametrize("header,seed,tol", [("H500", 2041, 1e-9), ("H500CORE", 2042, 1e-7), ("HFE", 2043, 1e-7),
                                             ("H6W", 2046, 1e-9), ("H500ND2S", 2047, 1e-7),
                                             ("H500ND2CORE", 2048, 1e-7), ("H5", 2049, 1e-9),
                                             ("H56", 2050, 1e-9)])
def test_full_gradient_matches_central_differences(tmp_path, header, seed, tol):
    say("fit: dL/dc vs central differences,", header)
    load(tmp_path, header, seed)
    batch = fit.Batch(structures(seed), 5.0)
    rng = np.random.default_rng(seed)
    c0 = capi.fit_get_coefficients()
    gE = torch.tensor(rng.normal(size=batch.nstruc), dtype=torch.float64, device=DEV)
    gF = torch.tensor(rng.normal(size=(batch.natom, 3)), dtype=torch.float64, device=DEV)
    g = gradient(batch, c0, gE, gF)

    def central(d):
        # linear embedding: L is linear in c and one central difference is exact for any step;
        # nonlinear embeddings: Richardson extrapolation of steps h and h/2 (error O(h^4))
        if int(FIT_HEADERS[header].get("nfemb", 0)) == 0:
            h = 1e-2
            return (loss(batch, c0 + h * d, gE, gF) - loss(batch, c0 - h * d, gE, gF)) / (2 * h)
        h = 1e-5
        f1 = (loss(batch, c0 + h * d, gE, gF) - loss(batch, c0 - h * d, gE, gF)) / (2 * h)
        f2 = (loss(batch, c0 + 0.5 * h * d, gE, gF) - loss(batch, c0 - 0.5 * h * d, gE, gF)) / h
        return (4.0 * f2 - f1) / 3.0

    for trial in range(3):
        d = rng.normal(size=len(c0))
        fd = central(d)
        an = float(g @ d)
        say("  trial %d: analytic %.10e  central difference %.10e" % (trial, an, fd))
        assert abs(an - fd) <= tol * max(1.0, abs(fd)), (an, fd)
    # single coefficients (first B1, a middle one, the last order-4 function)
    for k in (0, len(c0) // 3, len(c0) - 1):
        d = np.zeros(len(c0))
        d[k] = 1.0
        fd = central(d)
        assert abs(g[k] - fd) <= tol * max(1.0, abs(fd)), (k, g[k], fd)
    capi.fit_set_coefficients(c0)


def test_torch_training_step_reduces_loss(tmp_path):
    """Teacher/student: forces and energies of a teacher potential are the targets; the
    autograd gradient of the student's loss matches a central difference and Adam steps
    reduce the loss."""
    say("fit: torch autograd loss and optimiser steps")
    load(tmp_path, "H500", 2044)
    batch = fit.Batch(structures(7), 5.0)
    c_teacher = capi.fit_get_coefficients()
    E_ref = torch.zeros(batch.nstruc, dtype=torch.float64, device=DEV)
    F_ref = torch.zeros(batch.natom, 3, dtype=torch.float64, device=DEV)
    capi.fit_set_coefficients(c_teacher)
    batch.forward(E_ref, F_ref)
    rng = np.random.default_rng(3)
    c = torch.tensor(c_teacher + 0.3 * np.abs(c_teacher).mean() * rng.normal(size=len(c_teacher)),
                     dtype=torch.float64, device=DEV, requires_grad=True)

    def L(cc):
        E, F = fit.EnergyForces.apply(cc, batch)
        return ((E - E_ref) ** 2).sum() / batch.nstruc + 10.0 * ((F - F_ref) ** 2).sum() / batch.natom

    l0 = L(c)
    l0.backward()
    g = c.grad.detach().cpu().numpy()
    d = rng.normal(size=len(g))
    h = 1e-4 * np.abs(c_teacher).mean()
    with torch.no_grad():
        cd = torch.tensor(d, dtype=torch.float64, device=DEV)
        fd = (float(L(c + h * cd)) - float(L(c - h * cd))) / (2 * h)
    say("  analytic %.8e  central %.8e" % (g @ d, fd))
    assert abs(g @ d - fd) <= 1e-6 * max(1e-12, abs(fd))
    opt = torch.optim.Adam([c], lr=2e-3 * np.abs(c_teacher).mean())
    losses = [float(l0)]
    for _ in range(20):
        opt.zero_grad()
        l = L(c)
        l.backward()
        opt.step()
        losses.append(float(l))
    say("  loss %.4e -> %.4e" % (losses[0], losses[-1]))
    assert losses[-1] < 0.5 * losses[0]


def test_config_batch_fit_step(tmp_path):
    """The batched-fit workload (256 configurations, lmax-4 header): one forward + backward."""
    say("fit: 256-configuration batch")
    load(tmp_path, "H500", 2029)
    structs = systems.config_batch()
    batch = fit.Batch(structs, 5.0)
    c = torch.tensor(capi.fit_get_coefficients(), dtype=torch.float64, device=DEV, requires_grad=True)
    E, F = fit.EnergyForces.apply(c, batch)
    (E.sum() + (F ** 2).sum()).backward()
    assert torch.isfinite(c.grad).all()
    say("  atoms", batch.natom, "pairs", batch.npairs, "|grad|", float(c.grad.norm()))


def test_split_halves_leave_the_fit_unchanged(tmp_path):
    """Batches of 2048+ atoms run the forward evaluation as two halves of the atom range on two
    streams (MAGACE_SPLIT_STREAMS=0 keeps one sequence); the backward pass reads the per-atom
    buffers both halves wrote.  Energies, forces and dL/dc equal one sequence to 1e-12."""
    load(tmp_path, "H500", 2034)
    batch = fit.Batch(structures(11, sizes=((10, 10, 10), (4, 4, 4))), 5.0)
    assert batch.natom >= 2048
    say("fit: two halves,", batch.natom, "atoms")
    rng = np.random.default_rng(11)
    c0 = capi.fit_get_coefficients()
    gE = torch.tensor(rng.normal(size=batch.nstruc), dtype=torch.float64, device=DEV)
    gF = torch.tensor(rng.normal(size=(batch.natom, 3)), dtype=torch.float64, device=DEV)
    old = os.environ.get("MAGACE_SPLIT_STREAMS")
    out = {}
    try:
        for flag in ("0", "1"):
            os.environ["MAGACE_SPLIT_STREAMS"] = flag
            capi.fit_set_coefficients(c0)
            E = torch.zeros(batch.nstruc, dtype=torch.float64, device=DEV)
            F = torch.zeros(batch.natom, 3, dtype=torch.float64, device=DEV)
            batch.forward(E, F)
            g = torch.zeros(len(c0), dtype=torch.float64, device=DEV)
            batch.backward(gE, gF.contiguous(), g)
            out[flag] = (E.cpu().numpy(), F.cpu().numpy(), g.cpu().numpy())
    finally:
        if old is None:
            os.environ.pop("MAGACE_SPLIT_STREAMS", None)
        else:
            os.environ["MAGACE_SPLIT_STREAMS"] = old
    for a, b, what in zip(out["1"], out["0"], ("E", "F", "dL/dc")):
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-12 * scale, what



def test_backward_refused_after_another_evaluation(tmp_path):
    """The backward pass reads the Z and W its forward left on the device.  Any other
    evaluation in between (here a mixed-precision magace_compute_device on the same atoms)
    overwrites them, so magace_fit_backward must refuse instead of returning a wrong dL/dc."""
    say("fit: backward after an intervening evaluation")
    load(tmp_path, "H500", 2045)
    batch = fit.Batch(structures(2), 5.0)
    E = torch.zeros(batch.nstruc, dtype=torch.float64, device=DEV)
    F = torch.zeros(batch.natom, 3, dtype=torch.float64, device=DEV)
    batch.forward(E, F)
    L = batch.L
    N, P = batch.natom, batch.npairs
    outs = [torch.zeros(n, dtype=torch.float64, device=DEV) for n in (N, 3 * P, 3 * P, 3 * N)]
    capi.set_precision(capi.MIXED)
    try:
        capi.compute_device(N, P, L["offsets"].data_ptr(), L["dlist"].data_ptr(), L["elist"].data_ptr(),
                            L["dlist_momi"].data_ptr(), L["elist_momi"].data_ptr(), L["dlist_mom"].data_ptr(),
                            L["elist_mom"].data_ptr(), *[o.data_ptr() for o in outs],
                            stream=torch.cuda.current_stream().cuda_stream)
    finally:
        capi.set_precision(capi.FP64)
    gF = torch.zeros(batch.natom, 3, dtype=torch.float64, device=DEV)
    g = torch.zeros(capi.fit_ncoef(), dtype=torch.float64, device=DEV)
    with pytest.raises(capi.MagaceError, match="must follow"):
        batch.backward(None, gF, g)
    # a new forward makes the backward valid again
    batch.forward(E, F)
    batch.backward(None, gF, g)


def test_full_batch_loss_gradient(tmp_path):
    """The batched-fit workload (256 configurations, lmax-4 header, linear embedding) with
    L = sum_s (E_s - E_s^ref)^2 + sum_j |F_j - F_j^ref|^2 against FP64 targets from a teacher
    potential (the same coefficients perturbed by 30%).  The energy part of dL/dc
    (sum_s 2 dE_s sum_{i in s} B_i) equals the oracle's B vectors on all 256 configurations
    (<= 1e-9 relative); the full gradient along a random direction equals a central
    difference of L (L is quadratic in c for a linear embedding, so one central difference is
    exact up to rounding; <= 1e-8 relative)."""
    import oracle as O

    say("fit: full 256-configuration batch, L = dE^2 + dF^2")
    pot = load(tmp_path, "H500", 2029)
    structs = systems.config_batch()
    batch = fit.Batch(structs, 5.0)
    c0 = capi.fit_get_coefficients()
    rng = np.random.default_rng(2029)
    c_t = c0 * (1.0 + 0.3 * rng.normal(size=len(c0)))
    E_ref = torch.zeros(batch.nstruc, dtype=torch.float64, device=DEV)
    F_ref = torch.zeros(batch.natom, 3, dtype=torch.float64, device=DEV)
    capi.fit_set_coefficients(c_t)
    batch.forward(E_ref, F_ref)
    c = torch.tensor(c0, dtype=torch.float64, device=DEV, requires_grad=True)

    def L(cc):
        E, F = fit.EnergyForces.apply(cc, batch)
        return ((E - E_ref) ** 2).sum() + ((F - F_ref) ** 2).sum(), E

    l0, E0 = L(c)
    l0.backward()
    g = c.grad.detach().cpu().numpy()
    say("  atoms %d pairs %d  L %.6e  |dL/dc| %.3e" % (batch.natom, batch.npairs, float(l0), np.linalg.norm(g)))
    # energy part alone, against the oracle's per-atom B vectors
    gE = (2.0 * (E0.detach() - E_ref)).contiguous()
    gF = torch.zeros(batch.natom, 3, dtype=torch.float64, device=DEV)
    ge = gradient(batch, c0, gE, gF)
    o = O.Oracle.load(pot)
    want = np.zeros_like(ge)
    w = gE.cpu().numpy()
    for s, ws in zip(structs, w):
        a = nbr.neilist(s["pos"], s["sdir"] * s["smag"][:, None], s["cell"], 5.0)
        want += ws * o.bvectors(a).sum(axis=0)
    err = np.abs(ge - want).max() / np.abs(want).max()
    say("  energy part vs oracle B vectors: rel err %.2e" % err)
    assert err < 1e-9
    d = rng.normal(size=len(c0))
    h = 1e-3 * np.abs(c0).mean()
    with torch.no_grad():
        dd = torch.tensor(d, dtype=torch.float64, device=DEV)
        fd = (float(L(c + h * dd)[0]) - float(L(c - h * dd)[0])) / (2 * h)
    an = float(g @ d)
    say("  dL/dc . d: analytic %.10e  central difference %.10e" % (an, fd))
    assert abs(an - fd) <= 1e-8 * max(1e-12, abs(fd))
    capi.fit_set_coefficients(c0)
