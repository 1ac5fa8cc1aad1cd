"""Synthetic magnetic-ACE potential files in the reference's text format.

The layout is the one the reference reads (src/util.F90:551-653 header, :670-1106 records)
and writes (src/util.F90:532-547, 1110-1362): a 14-line header followed by
``<type> <norder> <nd>`` / payload record pairs.  One ``clu`` record is written per
enumerated basis function, so every coefficient is set explicitly.

Random parameters (documented, deterministic from ``seed`` with numpy default_rng):
  * expansion coefficients c ~ N(0, 1e-2) (standard deviation 1e-2), one per B function;
  * atomic radial coefficients crad_atomic ~ N(0, 1)/sqrt(nradbase_atomic);
  * magnetic radial coefficients crad_magnetic ~ N(0, 1)/sqrt(nradbase_magnetic);
  * by default core repulsion off (prefactor 0), embedding linear (nfemb 0), ndensity 1;
    headers may override par (p1a, p1b), core (core_pre, core_lambda), ecut/decut.
"""
from __future__ import annotations

import numpy as np

BASE = dict(nr5_a=0, nr5_m=0, l5a=0, l5m=0, nr6_a=0, nr6_m=0, l6a=0, l6m=0,
            ndensity=1, npot=5, nfemb=0, radtype_a=3, radtype_m=7,
            lambda_a=5.25, lambda_m=5.25, cut_a=5.0, cut_m=4.0, dcut_a=0.01, dcut_m=0.01,
            ecut=10000.0, decut=250.0, p1a=1.0, p1b=1.0, core_pre=0.0, core_lambda=0.0)

#: "nradbase l2max 8 3 2 4 / nradial3 l3max 4 1 4 1 / nradial4 l4max 2 1 2 1" (489 B functions)
H500 = dict(BASE, nb_a=8, nb_m=3, nr2_m=2, l2m=4, nr3_a=4, nr3_m=1, l3a=4, l3m=1,
            nr4_a=2, nr4_m=1, l4a=2, l4m=1)
#: "10 3 2 6 / 7 1 6 1 / 3 1 3 1" (2043 B functions)
H2000 = dict(BASE, nb_a=10, nb_m=3, nr2_m=2, l2m=6, nr3_a=7, nr3_m=1, l3a=6, l3m=1,
             nr4_a=3, nr4_m=1, l4a=3, l4m=1)
#: the shape of the reference's example potential (example/Fe.pbe.in header lines 8-14:
#: 16 16 5 2 / 4 3 1 2 / 3 3 0 0, nfemb 1) with synthetic parameters and a nonlinear
#: embedding exponent
HFE = dict(BASE, nb_a=16, nb_m=16, nr2_m=5, l2m=2, nr3_a=4, nr3_m=3, l3a=1, l3m=2,
           nr4_a=3, nr4_m=3, l4a=0, l4m=0, nfemb=1, p1a=1.0, p1b=0.75)
#: lmax-4 header with two densities, E = F(rho_1) + F(rho_2) (src/forces.F90:1066-1085), a
#: nonlinear embedding on the second density
H500ND2 = dict(H500, ndensity=2, nfemb=1, p1a=1.0, p1b=1.0, p1a_2=0.7, p1b_2=0.6)
#: lmax 6 with l3max_magnetic 2 (nine magnetic columns of A): "12 4 2 6 / 4 1 6 2 / 2 1 2 2"
H6W = dict(BASE, nb_a=12, nb_m=4, nr2_m=2, l2m=6, nr3_a=4, nr3_m=1, l3a=6, l3m=2,
           nr4_a=2, nr4_m=1, l4a=2, l4m=2)
#: sixteen atomic and six magnetic radial base functions, five magnetic radial functions,
#: l3max_magnetic 2 at lmax 4: "16 6 5 4 / 6 1 4 2 / 2 1 2 2"
H4W = dict(BASE, nb_a=16, nb_m=6, nr2_m=5, l2m=4, nr3_a=6, nr3_m=1, l3a=4, l3m=2,
           nr4_a=2, nr4_m=1, l4a=2, l4m=2)
#: body order 5, radial only (l5max 0): two atomic, one magnetic radial function
H5 = dict(H500, nr5_a=2, nr5_m=1)
#: body orders 5 and 6, radial only, on two magnetic radial functions of A (medium class)
H56 = dict(H500, nr3_m=2, nr4_m=2, nr5_a=2, nr5_m=2, nr6_a=2, nr6_m=2)
#: tiny header for fast unit tests
HTINY = dict(BASE, nb_a=4, nb_m=2, nr2_m=2, l2m=2, nr3_a=2, nr3_m=1, l3a=2, l3m=1,
             nr4_a=2, nr4_m=1, l4a=2, l4m=1)


def _f(x: float) -> str:
    return repr(float(x))


def render(header: dict, basis: dict, seed: int, element: str = "Fe", xc: str = "pbe",
           smear: float = 0.2, c_scale: float = 1e-2, values: dict | None = None) -> str:
    """Text of a potential file.  ``basis[k]`` = (nvar_k, width) int array of index tuples
    in the enumeration order of the reference (b1index/b2index/b3index/b4index columns).

    ``values`` (one density only) replaces the random draws with given parameters:
    ``crad_a`` [nr3_a][l3a+1][nb_a], ``crad_m`` [nr2_m][l2m+1][nb_m] and ``c`` {order: array
    in enumeration order}; the header's p1a/p1b/core_pre/core_lambda give par and core."""
    h = dict(BASE, **header)
    rng = np.random.default_rng(seed)
    if values is not None:
        if int(h.get("ndensity", 1)) != 1:
            raise ValueError("explicit values are supported for one density")
        ca, cm = np.asarray(values["crad_a"]), np.asarray(values["crad_m"])
        cv = {k: np.asarray(values["c"].get(k, np.zeros(0))).ravel() for k in (1, 2, 3, 4, 5, 6)}
        for k in (1, 2, 3, 4, 5, 6):
            if len(cv[k]) != len(np.asarray(basis.get(k, np.zeros((0, 1))))):
                raise ValueError("coefficient count of order %d does not match the basis" % k)
    L = []
    L.append(" Parameters (last)")
    L.append(" created        2026           1           1           0           0           0")
    L.append(" Elements %s" % element)
    L.append(" XC %s" % xc)
    L.append(" gauss-smearing   %s" % _f(smear))
    L.append(" lambda    %s        %s" % (_f(h["lambda_a"]), _f(h["lambda_m"])))
    L.append(" rcut dcut ecut decut    %s %s %s %s %s %s" % tuple(
        _f(h[k]) for k in ("cut_a", "cut_m", "dcut_a", "dcut_m", "ecut", "decut")))
    L.append(" nradbase l2max %d %d %d %d" % (h["nb_a"], h["nb_m"], h["nr2_m"], h["l2m"]))
    L.append(" nradial3 l3max %d %d %d %d" % (h["nr3_a"], h["nr3_m"], h["l3a"], h["l3m"]))
    L.append(" nradial4 l4max %d %d %d %d" % (h["nr4_a"], h["nr4_m"], h["l4a"], h["l4m"]))
    L.append(" nradial5 l5max %d %d %d %d" % (h["nr5_a"], h["nr5_m"], h["l5a"], h["l5m"]))
    L.append(" nradial6 l6max %d %d %d %d" % (h["nr6_a"], h["nr6_m"], h["l6a"], h["l6m"]))
    L.append(" ndensity %d" % h["ndensity"])
    L.append(" npot-nfemb-radtype %d %d %d %d" % (h["npot"], h["nfemb"], h["radtype_a"], h["radtype_m"]))
    for k, v in ((1, h["p1a"]), (2, h["p1b"])):
        L.append(" par %d 1" % k)
        L.append(" %s %s F" % (element, _f(v)))
    for k, v in ((1, h["core_pre"]), (2, h["core_lambda"])):
        L.append(" core %d 0" % k)
        L.append(" %s %s %s F" % (element, element, _f(v)))
    nb_a, nb_m = h["nb_a"], h["nb_m"]
    for nr in range(1, h["nr3_a"] + 1):
        for l in range(0, h["l3a"] + 1):
            for nb in range(1, nb_a + 1):
                v = ca[nr - 1, l, nb - 1] if values is not None else rng.normal() / np.sqrt(nb_a)
                L.append(" rad_atomic 0 0")
                L.append(" %s %s %d %d %d %s T" % (element, element, nb, nr, l, _f(v)))
    for nr in range(1, h["nr2_m"] + 1):
        for l in range(0, h["l2m"] + 1):
            for nb in range(1, nb_m + 1):
                v = cm[nr - 1, l, nb - 1] if values is not None else rng.normal() / np.sqrt(nb_m)
                L.append(" rad_magnetic 0 0")
                L.append(" %s %s %d %d %d %s T" % (element, element, nb, nr, l, _f(v)))
    for order in (1, 2, 3, 4, 5, 6):
        idx = np.asarray(basis.get(order, np.zeros((0, 1), np.int32)))
        els = " ".join([element] * order)
        for k, row in enumerate(idx):
            v = cv[order][k] if values is not None else rng.normal(0.0, c_scale)
            L.append(" clu %d 1" % order)
            L.append(" %s %s %s T" % (els, " ".join(str(int(x)) for x in row), _f(v)))
    # further densities (ndensity 2): their par and clu records follow, so that files with one
    # density are unchanged
    for nd in range(2, int(h.get("ndensity", 1)) + 1):
        for k, key in ((1, "p1a_%d" % nd), (2, "p1b_%d" % nd)):
            L.append(" par %d %d" % (k, nd))
            L.append(" %s %s F" % (element, _f(h.get(key, 1.0))))
        for order in (1, 2, 3, 4, 5, 6):
            idx = np.asarray(basis.get(order, np.zeros((0, 1), np.int32)))
            els = " ".join([element] * order)
            for row in idx:
                L.append(" clu %d %d" % (order, nd))
                L.append(" %s %s %s T" % (els, " ".join(str(int(v)) for v in row),
                                          _f(rng.normal(0.0, c_scale))))
    return "\n".join(L) + "\n"


def write(path, header, basis, seed, **kw):
    with open(path, "w") as f:
        f.write(render(header, basis, seed, **kw))
    return path
