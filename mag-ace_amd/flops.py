"""Algorithmic FLOPs per atom-step of the reference algorithm (the roofline numerator).

Counted from the reference loop structure at minimal arithmetic, not from what the GPU
kernels execute (they recompute splines and Y_lm in the force pass, which is not counted):
complex multiply = 6, complex MAC = 8, real x complex MAC = 4, Horner cubic with
derivative = 11, Y_lm with gradient = 30 per (l, m >= 0).

  (i)   spline lookups          src/radial.F90:53-160      per pair and per atom
  (ii)  Y_lm + gradients        src/spherical.F90:159-282  2 per pair + 1 per atom
  (iii) A, A2 accumulation      src/a_basis.F90:193-291
  (iv)  B products              src/b_basis.F90:91-307     sum_B m-combinations x rank x 6
  (v)   dB/dA                   same count as (iv)
  (vi)  weights                 src/forces.F90:150-331     slots x m-range real x complex
  (vii) force pass              src/forces.F90:422-623
  (viii) torque pass            src/forces.F90:688-961
  (ix)  c.B                     2 nB

Grouped per GPU kernel: density = (i)+(ii)+(iii), adjoint = (iv)+(v)+(vi)+(ix),
forces = (vii)+(viii).
"""
from __future__ import annotations


def flops_model(h: dict, basis: dict, raw: dict, nneigh: float) -> dict:
    """h: header dict; basis: order -> index tuples (enumeration); raw: m-combination counts
    {2: n2, 3: n3, 4: n4}; nneigh: mean neighbours per atom inside the cutoff."""
    nfa = h["nr3_a"] * (h["l3a"] + 1)
    nfm = h["nr2_m"] * (h["l2m"] + 1)
    lha = (h["l3a"] + 1) * (h["l3a"] + 2) // 2
    lhm = (h["l2m"] + 1) * (h["l2m"] + 2) // 2
    # (i) + (ii)
    spl_pair = 11 * (h["nb_a"] + nfa + 1 + h["nb_m"] + nfm)
    spl_atom = 11 * (h["nb_m"] + nfm)
    ylm_pair = 30 * (lha + lhm)
    ylm_atom = 30 * lhm
    # (iii) entries the reference fills per pair (a_basis.F90:248-291, 193-201)
    cA = h["nr3_a"] * h["nr3_m"] * sum((2 * l2 + 1) + 2 * l1 * (l2 + 1)
                                       for l1 in range(h["l3a"] + 1) for l2 in range(h["l3m"] + 1))
    cA2 = h["nb_a"] * h["nr2_m"] * sum(l + 1 for l in range(h["l2m"] + 1))
    acc_pair = 8 * cA + 4 * cA2
    # (iv) + (v)
    bprod = 6 * (2 * raw[2] + 3 * raw[3] + 4 * raw[4])
    # (vi) forceweights slots x m-range (forces.F90:158-232)
    wts = 0
    for b in basis.get(2, []):
        wts += 2 * (b[3] + 1)
    for b in basis.get(3, []):
        l1, l0p, l1p, l2p = b[5], b[6], b[7], b[8]
        wts += (2 * l1 + 1) * (l1p + 1) + (2 * l1 + 1) * (l2p + 1) + (l0p + 1)
    for b in basis.get(4, []):
        l1, l2, l3, k0, k1, k2, k3 = b[7], b[8], b[9], b[10], b[11], b[12], b[13]
        wts += (2 * l1 + 1) * (k1 + 1) + (2 * l2 + 1) * (k2 + 1) + (2 * l3 + 1) * (k3 + 1) + (k0 + 1)
    wts *= 4
    nB = sum(len(basis.get(k, [])) for k in (1, 2, 3, 4))
    cdotb = 2 * nB
    # (vii) + (viii): folded loops of forcecompute / torquecompute
    fA = h["nr3_a"] * h["nr3_m"] * sum(1 + 2 * l1 * l2 + l1 + l2
                                       for l1 in range(h["l3a"] + 1) for l2 in range(h["l3m"] + 1))
    force_pair = 3 * 8 * fA + 3 * 4 * cA2 + 6
    torque_pair = 3 * 8 * fA + 3 * 8 * cA2
    torque_atom = 3 * 8 * h["nr2_m"] * sum(l + 1 for l in range(h["l2m"] + 1)) + 6 * h["nb_m"]
    density = nneigh * (spl_pair + ylm_pair + acc_pair) + spl_atom + ylm_atom
    adjoint = 2 * bprod + wts + cdotb
    forces = nneigh * (force_pair + torque_pair) + torque_atom
    return {k: float(v) for k, v in dict(density=density, adjoint=adjoint, forces=forces,
                                          total=density + adjoint + forces).items()}
