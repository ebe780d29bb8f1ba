"""Seeded test systems shared by the CPU and GPU tests (cached per process)."""
import functools

import numpy as np

from alma import synth


@functools.lru_cache(maxsize=None)
def chain(n_res, seed):
    return synth.make_chain(n_res, seed)


@functools.lru_cache(maxsize=None)
def two_chains_ss():
    s = synth.concat([synth.make_chain(18, 11, seq=["CYS"] + ["ALA"] * 8 + ["GLY"] * 9),
                      synth.make_chain(14, 12, seq=["GLY"] * 6 + ["CYS"] + ["SER"] * 7)], gap=2)
    return s


_SS_SEQ = ["ALA", "CYS", "LEU", "GLY", "SER", "VAL", "LYS", "GLU", "ASP", "CYS", "THR", "ILE",
           "ASN", "GLY", "PHE", "CYS", "ARG", "GLN", "TYR", "CYS", "PRO", "HIS", "MET", "TRP"]


@functools.lru_cache(maxsize=None)
def chain_ss():
    """24-residue compact chain with two disulfides (SG moved to 2.04 A)."""
    return synth._add_disulfides(synth.make_chain(len(_SS_SEQ), 21, seq=list(_SS_SEQ)), 2)


@functools.lru_cache(maxsize=None)
def copies(n_copies, n_ss=0, step=52.0, row=0, step_y=66.0):
    """n_copies of the 300-residue chain (4,965 atoms each), copy k shifted by
    step A per grid cell so neighbouring copies share an interface: along x,
    or on a grid of `row` copies per row (x, step) stacked in y (step_y);
    n_ss SG-SG pairs moved to 2.04 A across the complex.  Closest atoms of
    neighbouring copies: 1.7 A at step 52, 4.9 A at 58 (x); 4.1 A at 66 (y)."""
    base = chain(300, 2)
    parts = []
    for k in range(n_copies):
        c = dict(base)
        gx, gy = (k % row, k // row) if row else (k, 0)
        c["xyz"] = base["xyz"] + np.array([gx * step, gy * step_y, 0.0])
        parts.append(c)
    sysd = synth.concat(parts, gap=2)
    if n_ss:
        sysd = synth._add_disulfides(sysd, n_ss)
    return sysd


def bridge():
    """tests/bridge_test.py:83-85 of the reference: three particles."""
    return {"xyz": np.array([[0.0, 0, 0], [5.0, 0, 0], [10.0, 0, 0]]),
            "charge": np.array([1.0, -1.0, 0.5]), "rstar": np.array([1.9, 1.9, 1.9]),
            "eps": np.array([0.1, 0.1, 0.1])}


def topology(sysd, ss_pairs=()):
    """Product BondTopology (pybind drop-in) for a system dict."""
    import protein_physics as PP
    t = PP.BondTopology()
    t.build(list(sysd["resnames"]), list(sysd["atomnames"]), [int(v) for v in sysd["res_idx"]])
    for i, j in ss_pairs:
        t.add_disulfide(int(i), int(j))
    return t


def particles(sysd, mod=None):
    if mod is None:
        import protein_physics as mod
    x = sysd["xyz"]
    return [mod.Particle(float(x[k, 0]), float(x[k, 1]), float(x[k, 2]), float(sysd["charge"][k]),
                         float(sysd["rstar"][k]), float(sysd["eps"][k])) for k in range(len(x))]


def csr(t):
    from oracle.oracle import topo_to_csr
    return topo_to_csr(t)
