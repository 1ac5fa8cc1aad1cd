"""Shared test helpers: golden cases, potentials, error metrics."""
import hashlib
import os

import numpy as np

import __graft_entry__

PKG = __graft_entry__.load_package()
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
CASES = ("tiny_16", "h500_128", "h2000_16", "hfe_16", "edge_h500core", "nd2_16", "radt46_16", "radt58_16",
         "m0_16", "h6w_16", "h4w_16", "o5_16", "o56_16", "l4m_16")
# l4max_magnetic above l3max_magnetic: pinned to the reference's energies and its energy's
# central differences only (the C restatement covers A's l' range l3max_magnetic)
REFERENCE_ONLY = ("l4m_16",)
ORACLE_CASES = tuple(c for c in CASES if c not in REFERENCE_ONLY)

HEADERS = {"HTINY": PKG.potgen.HTINY, "H500": PKG.potgen.H500, "H2000": PKG.potgen.H2000,
           "HFE": PKG.potgen.HFE, "H500ND2": PKG.potgen.H500ND2,
           "H500R46": dict(PKG.potgen.H500, radtype_a=4, radtype_m=6),
           "H500R58": dict(PKG.potgen.H500, radtype_a=5, radtype_m=8),
           "H500CORE": dict(PKG.potgen.H500, core_pre=25.0, core_lambda=0.35, ecut=14.0, decut=8.0,
                            nfemb=2, p1a=1.0, p1b=0.9),
           "H500M0": dict(PKG.potgen.H500, nr3_m=2, l3m=0, l4m=0),
           "H6W": PKG.potgen.H6W, "H4W": PKG.potgen.H4W, "H5": PKG.potgen.H5, "H56": PKG.potgen.H56,
           "HL4": dict(PKG.potgen.H500, l4m=2)}
HIGHER_HEADERS = ("HL4",)


def potential_text(header_name, seed, enum="product"):
    h = HEADERS[header_name]
    if enum == "product":
        basis = PKG.capi.enumerate_basis(h)
    else:
        import oracle as O
        o = O.Oracle.from_header(h)
        basis = {k: o.basis(k) for k in (1, 2, 3, 4, 5, 6)}
    return PKG.potgen.render(h, basis, seed)


def write_potential(path, header_name, seed, enum="product", sha=None):
    text = potential_text(header_name, seed, enum)
    if sha is not None:
        assert hashlib.sha256(text.encode()).hexdigest() == sha, "regenerated potential differs"
    with open(path, "w") as f:
        f.write(text)
    return path


def load_case(name, tmpdir, enum="product"):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    a = PKG.systems.pack_acelib(z["pos"], z["cell"], z["sdir"], z["smag"], float(z["rlist"]))
    pot = write_potential(os.path.join(str(tmpdir), name + ".pbe.in"), str(z["header"]), int(z["pot_seed"]),
                          enum, str(z["pot_sha256"]))
    ref = {k: z[k] for k in ("eout", "flist", "tlist", "tlist0")}
    return dict(a=a, ref=ref, pot=pot, header=str(z["header"]), z=z)


def relerr(got, ref):
    """Norm-wise relative error ||got-ref||_inf / ||ref||_inf."""
    s = np.abs(ref).max() if np.size(ref) else 0.0
    d = np.abs(np.asarray(got) - np.asarray(ref)).max() if np.size(ref) else 0.0
    return d / s if s > 0 else d


def per_atom_errors(a, got, ref):
    """Max |dE_i| (eV), max |dF_i| (eV/A), max |d fm_i| (eV/muB) after caller assembly."""
    s = PKG.systems
    e1, f1, m1, _ = s.assemble(a, got["eout"], got["flist"], got["tlist"], got["tlist0"])
    e0, f0, m0, _ = s.assemble(a, ref["eout"], ref["flist"], ref["tlist"], ref["tlist0"])
    return (np.abs(got["eout"] - ref["eout"]).max(), np.abs(f1 - f0).max(), np.abs(m1 - m0).max())


FLOAT_KEYS = ("lambda_a", "lambda_m", "cut_a", "cut_m", "dcut_a", "dcut_m", "ecut", "decut", "p1a", "p1b",
              "core_pre", "core_lambda")


def fe_example(tmpdir):
    """The reference's fitted example potential (example/Fe.pbe.in), regenerated from the
    parameters stored in golden/fe_example.npz (tests/golden/make_fe_example.py), and the
    fixture itself.  Returns (path, fixture)."""
    z = np.load(os.path.join(GOLDEN, "fe_example.npz"))
    h = {str(k): (float(v) if str(k) in FLOAT_KEYS else int(round(v)))
         for k, v in zip(z["header_keys"], z["header_vals"])}
    basis = PKG.capi.enumerate_basis(h)
    vals = dict(crad_a=z["crad_a"], crad_m=z["crad_m"], c={k: z["c%d" % k] for k in (1, 2, 3, 4)})
    text = PKG.potgen.render(h, basis, 0, values=vals)
    assert hashlib.sha256(text.encode()).hexdigest() == str(z["pot_sha256"]), "regenerated potential differs"
    path = os.path.join(str(tmpdir), "Fe.pbe.in")
    with open(path, "w") as f:
        f.write(text)
    return path, z


FE_CASES = ("notebook", "nc16", "nc128")


def fe_case(z, name, rcut=4.5):
    """acelib_ arrays and reference outputs of one system of the example-potential fixture."""
    a = PKG.systems.pack_acelib(z[name + "_pos"], z[name + "_cell"], z[name + "_sdir"], z[name + "_smag"], rcut)
    ref = {k: z["%s_%s" % (name, k)] for k in ("eout", "flist", "tlist", "tlist0")}
    return a, ref
