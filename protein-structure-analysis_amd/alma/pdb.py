"""PDB files in: particles, residue metadata and disulfides for the energy and
sampling core, the input side of the reference's GUI (``_parse_pdb``,
python/gui_main.py:281-605).

The reference reads files through Bio.PDB.PDBParser (biopython 1.87,
permissive mode); this reader restates the parts of that parser the physics
path sees:

  * ATOM / HETATM records of the first MODEL (a file without MODEL records is
    one model); fixed columns: name 13-16, altloc 17, resname 18-20, chain 22,
    resseq 23-26, icode 27, x/y/z 31-54, occupancy 55-60;
  * coordinates are single precision, as Bio.PDB stores them (the particles
    carry the float32 value widened to double);
  * atoms are visited chain by chain in order of each chain's first record,
    residues in order of first appearance within their chain (a chain or a
    residue id that comes back later is continued, as Bio.PDB does), atoms in
    order of first appearance within the residue;
  * alternate locations: one atom per name, the location with the highest
    occupancy (the first on ties); a repeated name without an altloc is
    ignored;
  * residue id = (hetero flag, resseq, icode), flag " " for ATOM, "W" for
    HETATM water (HOH / WAT), "H_<resname>" for other HETATM records.

On top of that, the rules of _parse_pdb (gui_main.py:383-560):

  * HETATM water (any name in WATER_RESNAMES) and nucleic-acid residues (any
    record) are dropped, as are atoms with non-finite coordinates;
  * residue index: standard (ATOM) residues are numbered 0, 1, 2, ... in
    order of appearance, HETATM residues 100000 + that count, so no peptide
    bond is inferred to them;
  * parameters from ff14sb.get_atom_params, then every hydrogen of a heavy
    atom's template that is absent from that residue instance has its charge
    folded onto the heavy atom (united atom, amber_params.py:1004-1071);
  * disulfides: every pair of CYS SG atoms of ATOM records closer than 2.5 A
    (d^2 < 6.25 on the float32 coordinates), in order of the SG atoms.

Use::

    s = pdb.load("1XQ8.pdb")            # or .pdb.gz
    topo = pdb.topology(s, protein_physics)
    particles = pdb.particles(s, protein_physics)
"""
from __future__ import annotations

import gzip
import math

import numpy as np

from .ff14sb import NUCLEOTIDE_RESNAMES, WATER_RESNAMES, get_atom_params, missing_hydrogen_charge

SS_CUTOFF2 = 6.25   # (2.5 A)^2, gui_main.py:585-591
HET_RES_BASE = 100000


def _open(path):
    if str(path).endswith(".gz"):
        return gzip.open(path, "rt")
    return open(path)


def _f32(v: float) -> float:
    return float(np.float32(v))


class _Residue:
    __slots__ = ("resname", "het", "resseq", "icode", "atoms", "order")

    def __init__(self, resname, het, resseq, icode):
        self.resname = resname
        self.het = het
        self.resseq = resseq
        self.icode = icode
        self.atoms = {}     # name -> [x, y, z, occupancy]
        self.order = []     # names in order of first appearance


def read_records(lines):
    """Chains -> residues -> atoms of the first model, as Bio.PDB's permissive
    parser builds them.  Returns a list of (chain_id, _Residue)."""
    chains = {}          # chain id -> {residue id -> _Residue} (insertion ordered)
    models = 0
    for line in lines:
        rec = line[0:6]
        if rec == "MODEL ":
            models += 1
            if models > 1:
                break
            continue
        if rec == "ENDMDL":
            if models == 0:
                models = 1
            break
        if rec not in ("ATOM  ", "HETATM"):
            continue
        fullname = line[12:16]
        parts = fullname.split()
        name = parts[0] if len(parts) == 1 else fullname
        altloc = line[16] if len(line) > 16 else " "
        resname = line[17:20].strip()
        chain = line[21] if len(line) > 21 else " "
        resseq = int(line[22:26].split()[0])
        icode = line[26] if len(line) > 26 else " "
        if rec == "HETATM":
            het = "W" if resname in ("HOH", "WAT") else "H_" + resname
        else:
            het = " "
        try:
            x = _f32(float(line[30:38]))
            y = _f32(float(line[38:46]))
            z = _f32(float(line[46:54]))
        except ValueError as e:
            raise ValueError("invalid or missing coordinates: %r" % line.rstrip()) from e
        try:
            occ = float(line[54:60])
        except ValueError:
            occ = 0.0
        residues = chains.setdefault(chain, {})
        rid = (het, resseq, icode)
        res = residues.get(rid)
        if res is None:
            res = residues[rid] = _Residue(resname, het, resseq, icode)
        elif res.resname != resname:
            raise ValueError("residue %s%d%s of chain %r appears with two names (%s, %s): point-mutation "
                             "disorder is not supported" % (het, resseq, icode.strip(), chain, res.resname, resname))
        have = res.atoms.get(name)
        if have is None:
            res.atoms[name] = [x, y, z, occ]
            res.order.append(name)
        elif altloc != " " and occ > have[3]:
            have[:] = [x, y, z, occ]
        # else: a later location with no higher occupancy, or a duplicate
    return [(cid, r) for cid, residues in chains.items() for r in residues.values()]


def load(path):
    """A structure as a system dict (float64 arrays; atom order as the
    reference's particle list):

      xyz (N,3), charge, rstar, eps, resnames, atomnames, res_idx (int),
      ss_pairs (K,2) SG-SG disulfides, het (bool, HETATM atom),
      chain (str per atom), resseq (int per atom),
      ca_indices (standard-residue CA atoms), heavy_indices (standard-residue
      non-hydrogen atoms), skipped (atoms dropped for non-finite coordinates).
    """
    with _open(path) as f:
        records = read_records(f)
    xyz, q, r, e = [], [], [], []
    resnames, atomnames, res_idx, het, chain_of, resseq = [], [], [], [], [], []
    ca, heavy, sg = [], [], []
    rmap = {}
    skipped = 0
    for cid, res in records:
        if res.het != " " and res.resname in WATER_RESNAMES:
            continue
        if res.resname in NUCLEOTIDE_RESNAMES:
            continue
        for name in res.order:
            x, y, z, _ = res.atoms[name]
            if not (math.isfinite(x) and math.isfinite(y) and math.isfinite(z)):
                skipped += 1
                continue
            k = len(xyz)
            std = res.het == " "
            if std and name == "CA":
                ca.append(k)
            if std and not name.startswith("H"):
                heavy.append(k)
            if std and res.resname == "CYS" and name == "SG":
                sg.append(k)
            key = (cid, res.resseq, res.icode)
            if key not in rmap:
                rmap[key] = len(rmap) if std else HET_RES_BASE + len(rmap)
            c, rs, ep = get_atom_params(res.resname, name)
            xyz.append((x, y, z))
            q.append(c)
            r.append(rs)
            e.append(ep)
            resnames.append(res.resname)
            atomnames.append(name)
            res_idx.append(rmap[key])
            het.append(not std)
            chain_of.append(cid)
            resseq.append(res.resseq)
    # united-atom charge fold over the atoms actually present per residue
    present = {}
    for ri, nm in zip(res_idx, atomnames):
        present.setdefault(ri, set()).add(nm)
    for k in range(len(q)):
        extra = missing_hydrogen_charge(resnames[k], atomnames[k], present[res_idx[k]])
        if extra:
            q[k] += extra
    xyz = np.array(xyz, dtype=np.float64).reshape(-1, 3)
    ss = []
    for a in range(len(sg)):
        for b in range(a + 1, len(sg)):
            d = xyz[sg[a]] - xyz[sg[b]]
            if float(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) < SS_CUTOFF2:
                ss.append((sg[a], sg[b]))
    return {
        "xyz": xyz, "charge": np.array(q), "rstar": np.array(r), "eps": np.array(e),
        "resnames": resnames, "atomnames": atomnames, "res_idx": np.array(res_idx, dtype=np.int64),
        "ss_pairs": np.array(ss, dtype=np.int32).reshape(-1, 2), "het": np.array(het, dtype=bool),
        "chain": chain_of, "resseq": np.array(resseq, dtype=np.int64),
        "ca_indices": np.array(ca, dtype=np.int64), "heavy_indices": np.array(heavy, dtype=np.int64),
        "skipped": skipped,
    }


def particles(sysd, physics_mod):
    """Particle list of the drop-in module (is_water False, as the reference)."""
    x, q, r, e = sysd["xyz"], sysd["charge"], sysd["rstar"], sysd["eps"]
    return [physics_mod.Particle(float(x[k, 0]), float(x[k, 1]), float(x[k, 2]), float(q[k]), float(r[k]),
                                 float(e[k]), False) for k in range(len(x))]


def topology(sysd, physics_mod, parts=None):
    """BondTopology in the reference's order (gui_main.py:562-591): build,
    concerted side-chain pairs from the coordinates, then the disulfides."""
    t = physics_mod.BondTopology()
    t.build(list(sysd["resnames"]), list(sysd["atomnames"]), [int(v) for v in sysd["res_idx"]])
    t.identify_concerted_sidechain_pairs(parts if parts is not None else particles(sysd, physics_mod))
    for i, j in np.asarray(sysd["ss_pairs"]).reshape(-1, 2):
        t.add_disulfide(int(i), int(j))
    return t
