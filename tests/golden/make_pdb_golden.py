"""Writes pdb_golden.npz: parity vectors for the PDB input path.

Sources, all in this container where the reference is present:
  * the reference's parameter module python/amber_params.py (pure Python, no
    third-party imports), imported read-only from /root/reference/python:
    get_atom_params over every (residue, atom) key of its tables plus probes
    (ions, unknown residues, element fallbacks), its hydrogen-parent table,
    and for each fixture structure the per-atom (charge, R*, epsilon) that
    _parse_pdb assigns (get_atom_params plus missing_hydrogen_charge over the
    atoms present in each residue, python/gui_main.py:534-560), on the atom
    list of this package's reader;
  * the reference module built from its own sources (oracle/_ref/alma_ref,
    see make_golden.py): BondTopology (build, concerted side-chain pairs,
    disulfides in _parse_pdb's order), the per-term energies / total of
    calculate_potential for those particles, seeded run_mc_diagnostic
    records on two of them, and the GUI's replica-exchange loop
    (gui_main.py:1960-2013, restated here over the module's
    run_landscape_segment with seeded engines) on one.

The structures are the reference's own data files (data/1XQ8.pdb: NMR with
explicit hydrogens; data/SM_P01308.pdb: three disulfides; data/SM_P21543.pdb:
a Ca2+ HETATM and one disulfide), committed gzipped under pdb/.

Run here:  make -C oracle && python tests/golden/make_pdb_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_PY = "/root/reference/python"
sys.path.insert(0, os.path.join(ROOT, "protein-structure-analysis_amd"))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import make_golden as G  # noqa: E402
from alma import pdb  # noqa: E402

STRUCTURES = ["1XQ8", "SM_P01308", "SM_P21543"]
KT = 0.0019872041 * 300.0
DIAG = {"1XQ8": (120, 31), "SM_P01308": (200, 57)}   # run_mc_diagnostic: steps, seed
# replica exchange over landscape segments: replicas, snapshots, steps per
# snapshot, swap interval, ladder ratio, max angle, engine seed, exchange seed
PT = {"SM_P01308": (4, 6, 15, 3, 1.6, 0.12, 21, 4)}
DIAG_KEYS = ("move_type", "proposed_delta", "accepted", "d_hardcore", "d_nonbonded_soft", "d_sasa", "d_dih",
             "d_ss")
PROBES = [("XYZ", n) for n in ("CL1", "BR", "1HB", "OXT", "P1", "FE", "ZN1", "SE", "CA", "C7", "N3", "O1P",
                               "S1", "H12", "F2", "I1", "MG", "NA", "K", "MN1", "Q9", "2HG1")]
PROBES += [(r, "CA") for r in ("MG", "CA", "ZN", "FE", "FE2", "FE3", "MN", "NI", "CU", "CO", "NA", "CL", "K", "MO")]
PROBES += [("GLY", "HA"), ("PRO", "H"), ("PRO", "OXT"), ("ALA", "OXT"), ("HIS", "HE2"), ("MET", "H1"),
           ("ALA", "H2"), ("LYS", "H3"), ("CYS", "HG"), ("ATP", "H61"), ("HEM", "FE"), ("NAD", "C4N")]


def gui_pt(R, ps, t, n_rep, n_snap, sps, swap_interval, ladder_ratio, max_angle, seed, exchange_seed):
    """The GUI's replica-exchange loop (python/gui_main.py:1960-2013) over
    the reference module's run_landscape_segment, slot k's engine seeded
    with seed + k: slot 0's pooled snapshots and energies, exchange counts."""
    rng = np.random.default_rng(exchange_seed)
    ladder = [KT * ladder_ratio ** k for k in range(n_rep)]
    engines = [R.PhysicsEngine() for _ in range(n_rep)]
    for k, e in enumerate(engines):
        R.seed(e, seed + k)
    states, energies, cur = [ps] * n_rep, [None] * n_rep, [max_angle] * n_rep
    snaps, ens, att, acc = [], [], 0, 0
    for seg in range(max(1, n_snap // swap_interval)):
        for k in range(n_rep):
            sn, en, cm = engines[k].run_landscape_segment(states[k], t, swap_interval, sps, ladder[k], cur[k])
            states[k], energies[k], cur[k] = sn[-1], en[-1], cm
            if k == 0:
                snaps.extend(sn)
                ens.extend(en)
        for a in range(seg % 2, n_rep - 1, 2):
            b = a + 1
            x = (energies[a] - energies[b]) * ((1.0 / ladder[a]) - (1.0 / ladder[b]))
            att += 1
            if rng.random() < (1.0 if x >= 0 else np.exp(x)):
                states[a], states[b] = states[b], states[a]
                energies[a], energies[b] = energies[b], energies[a]
                cur[a], cur[b] = cur[b], cur[a]
                acc += 1
    xyz = np.array([[[p.x, p.y, p.z] for p in sn] for sn in snaps])
    return xyz, np.array(ens), np.array([att, acc])


def main():
    sys.path.insert(0, REF_PY)
    import amber_params as AP  # the reference's parameter tables
    R = G.load_ref()
    out = {}
    keys = sorted(set(AP.PARTIAL_CHARGES) | set(AP.ATOM_TYPE) | set(PROBES))
    out["tab_res"] = np.array([k[0] for k in keys])
    out["tab_atom"] = np.array([k[1] for k in keys])
    out["tab_params"] = np.array([AP.get_atom_params(r, a) for r, a in keys])
    kids = [(r, h, ",".join(v)) for (r, h), v in AP._HEAVY_CHILDREN.items()]
    out["hp_res"] = np.array([k[0] for k in kids])
    out["hp_heavy"] = np.array([k[1] for k in kids])
    out["hp_children"] = np.array([k[2] for k in kids])
    out["hp_all_missing"] = np.array([AP.missing_hydrogen_charge(r, h, set()) for r, h, _ in kids])
    for name in STRUCTURES:
        s = pdb.load(os.path.join(HERE, "pdb", name + ".pdb.gz"))
        present = {}
        for ri, a in zip(s["res_idx"], s["atomnames"]):
            present.setdefault(int(ri), set()).add(a)
        prm = []
        for res, a, ri in zip(s["resnames"], s["atomnames"], s["res_idx"]):
            q, rr, e = AP.get_atom_params(res, a)
            extra = AP.missing_hydrogen_charge(res, a, present[int(ri)])
            if extra:
                q += extra
            prm.append((q, rr, e))
        prm = np.array(prm)
        out[name + "/n"] = np.array(len(s["xyz"]))
        out[name + "/params"] = prm
        out[name + "/ss"] = s["ss_pairs"]
        ps = [R.Particle(float(x[0]), float(x[1]), float(x[2]), float(p[0]), float(p[1]), float(p[2]))
              for x, p in zip(s["xyz"], prm)]
        t = R.BondTopology()
        t.build(list(s["resnames"]), list(s["atomnames"]), [int(v) for v in s["res_idx"]])
        t.identify_concerted_sidechain_pairs(ps)
        for i, j in s["ss_pairs"]:
            t.add_disulfide(int(i), int(j))
        for k, v in G.topo_arrays(t).items():
            out[name + "/" + k] = v
        out[name + "/terms"] = np.array(R.terms(ps, t))
        out[name + "/potential"] = np.array(R.PhysicsEngine().calculate_potential(ps, t))
        if name in DIAG:
            # run_mc_diagnostic with the member generator seeded
            eng = R.PhysicsEngine()
            R.seed(eng, DIAG[name][1])
            rec = eng.run_mc_diagnostic(ps, t, DIAG[name][0], KT, 0.12)
            for key, v in zip(DIAG_KEYS, rec):
                out[name + "/diag_" + key] = np.array(v)
            out[name + "/diag_steps"] = np.array(DIAG[name][0])
            out[name + "/diag_seed"] = np.array(DIAG[name][1])
        if name in PT:
            xyz, e, sw = gui_pt(R, ps, t, *PT[name])
            out[name + "/pt_args"] = np.array(PT[name], dtype=np.float64)
            out[name + "/pt_xyz"] = xyz
            out[name + "/pt_e"] = e
            out[name + "/pt_swaps"] = sw
        print(name, len(ps), "atoms, total", out[name + "/potential"])
    np.savez_compressed(os.path.join(HERE, "pdb_golden.npz"), **out)


if __name__ == "__main__":
    main()
