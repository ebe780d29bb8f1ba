"""PDB input path (alma.pdb, alma.ff14sb.get_atom_params) against vectors from
the reference's own parameter module and its compiled physics module
(tests/golden/make_pdb_golden.py), on three of the reference's data files:
  * get_atom_params over every (residue, atom) key of the reference tables
    plus probes (ions, unknown residues, element fallbacks): exact;
  * hydrogen-parent table (children and their order) and the united-atom
    charge sums: exact;
  * per-atom charge / R* / epsilon of each structure, as _parse_pdb assigns
    them: exact; disulfides detected: equal;
  * the reader itself against the PDB text: atom counts of the first model,
    float32 coordinates, residue numbering, CA list;
  * topology of the drop-in module on the parsed structures: bit-exact;
  * the FP64 oracle's energy terms against the reference's: 1e-11 relative.
CPU only."""
import gzip
import os

import numpy as np
import pytest

from alma import ff14sb as F
from alma import pdb
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STRUCTURES = ("1XQ8", "SM_P01308", "SM_P21543")


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "pdb_golden.npz"), allow_pickle=False)


def _path(name):
    return os.path.join(GOLDEN, "pdb", name + ".pdb.gz")


def test_parameter_table_matches_reference(golden):
    bad = []
    for r, a, ref in zip(golden["tab_res"], golden["tab_atom"], golden["tab_params"]):
        got = F.get_atom_params(str(r), str(a))
        if tuple(got) != tuple(ref):
            bad.append((str(r), str(a), got, tuple(ref)))
    assert not bad, bad[:10]
    assert len(golden["tab_res"]) > 600


def test_hydrogen_parents_match_reference(golden):
    ref = {(str(r), str(h)): tuple(str(c).split(",")) for r, h, c in
           zip(golden["hp_res"], golden["hp_heavy"], golden["hp_children"])}
    assert F.HEAVY_CHILDREN == ref
    for (r, h), v in zip(ref, golden["hp_all_missing"]):
        assert F.missing_hydrogen_charge(r, h, set()) == v, (r, h)
        assert F.missing_hydrogen_charge(r, h, set(ref[(r, h)])) == 0.0


def _records(name):
    """ATOM / HETATM lines of the first model, straight from the text."""
    out = []
    with gzip.open(_path(name), "rt") as f:
        for line in f:
            if line.startswith("ENDMDL"):
                break
            if line.startswith(("ATOM  ", "HETATM")):
                out.append(line)
    return out


@pytest.mark.parametrize("name", STRUCTURES)
def test_structure_parameters_match_reference(golden, name):
    s = pdb.load(_path(name))
    assert len(s["xyz"]) == int(golden[name + "/n"])
    ref = golden[name + "/params"]
    got = np.stack([s["charge"], s["rstar"], s["eps"]], axis=1)
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]
    assert np.array_equal(s["ss_pairs"], golden[name + "/ss"])


@pytest.mark.parametrize("name", STRUCTURES)
def test_reader_follows_the_text(name):
    s = pdb.load(_path(name))
    lines = _records(name)
    assert len(s["xyz"]) == len(lines)     # no water, nucleotides, altlocs or NaN in these files
    for k, line in enumerate(lines):
        want = [float(np.float32(float(line[c:c + 8]))) for c in (30, 38, 46)]
        assert s["xyz"][k].tolist() == want, k
        assert s["atomnames"][k] == line[12:16].strip() and s["resnames"][k] == line[17:20].strip(), k
    std = ~s["het"]
    assert len(s["ca_indices"]) == len(set(s["res_idx"][std].tolist()))
    # standard residues numbered 0.. in order of appearance, HETATM from 100000
    first = np.unique(s["res_idx"][std], return_index=True)[1]
    assert np.array_equal(np.sort(s["res_idx"][std][np.sort(first)]), np.arange(len(first)))
    assert np.all(s["res_idx"][s["het"]] >= pdb.HET_RES_BASE)


def test_ion_and_disulfides():
    s = pdb.load(_path("SM_P21543"))
    ion = np.flatnonzero(s["het"])
    assert len(ion) == 1 and s["resnames"][ion[0]] == "CA"
    assert (s["charge"][ion[0]], s["rstar"][ion[0]], s["eps"][ion[0]]) == F.ION_PARAMS["CA"]
    ss = pdb.load(_path("SM_P01308"))["ss_pairs"]
    assert len(ss) == 3
    x = pdb.load(_path("SM_P01308"))["xyz"]
    for i, j in ss:
        assert np.sum((x[i] - x[j]) ** 2) < 6.25


def test_united_atom_fold_keeps_each_residue_charge():
    """With hydrogens folded onto their heavy atoms, every complete standard
    residue carries its template's total charge (hydrogens included), both
    for a heavy-atom-only model and for an explicit-hydrogen structure
    (1XQ8), where only absent hydrogens are folded.  A hydrogen not folded
    would be off by 0.03-0.45 e."""
    for name in ("SM_P01308", "SM_P21543", "1XQ8"):
        s = pdb.load(_path(name))
        res = {}
        for k, ri in enumerate(s["res_idx"]):
            if not s["het"][k]:
                res.setdefault(int(ri), []).append(k)
        checked = 0
        for ri, atoms in res.items():
            rn = s["resnames"][atoms[0]]
            rec = F.RESIDUES[rn]
            names = {s["atomnames"][k] for k in atoms}
            if any(a not in names for a, _, _ in rec if not a.startswith("H")):
                continue    # incomplete residue
            q = sum(s["charge"][k] for k in atoms)
            assert abs(q - sum(c for _, _, c in rec)) < 1e-9, (name, ri, rn, q)
            checked += 1
        assert checked > 0.9 * len(res), (name, checked, len(res))


@pytest.fixture(scope="module", params=STRUCTURES)
def parsed(request):
    import protein_physics as PP
    s = pdb.load(_path(request.param))
    t = pdb.topology(s, PP)
    return request.param, s, t, O.topo_to_csr(t)


def test_topology_matches_reference(golden, parsed):
    name, s, t, c = parsed
    g = {k.split("/", 1)[1]: golden[k] for k in golden.files if k.startswith(name + "/")}
    excl = t.excl
    assert np.array_equal(np.concatenate([[0], np.cumsum([len(v) for v in excl])]), g["t_excl_off"])
    assert np.array_equal(np.array([j for v in excl for j in v]), g["t_excl_idx"])
    for key, got in (("t_rb_i", c["rb_i"]), ("t_rb_j", c["rb_j"]), ("t_rb_kind", c["rb_kind"]),
                     ("t_side_off", c["side_off"]), ("t_dih_a", c["dih_a"]), ("t_dih_d", c["dih_d"]),
                     ("t_dih_toff", c["dih_toff"]), ("t_dih_n", c["dih_n"])):
        assert np.array_equal(got, g[key]), (name, key)
    assert np.array_equal(c["rb_scale"], g["t_rb_scale"])
    assert np.array_equal(np.array(t.disulfide_pairs).reshape(-1, 2), g["t_ss"])
    assert np.array_equal(np.array(t.concerted_pairs).reshape(-1, 2), g["t_cp"])
    assert np.array_equal(np.array(t.concerted_sidechain_pairs).reshape(-1, 2), g["t_csp"])
    assert np.array_equal(np.array(t.bonds).reshape(-1, 2), g["t_bonds"])


def test_oracle_energy_matches_reference(golden, parsed):
    name, s, t, c = parsed
    got = O.energy_terms(s, c)
    ref = golden[name + "/terms"]
    for k, term in enumerate(O.TERMS):
        assert abs(got[term] - ref[k]) <= 1e-11 * max(abs(ref[k]), 1e-300) or got[term] == ref[k], \
            (name, term, got[term], ref[k])
    assert abs(got["TOTAL"] - float(golden[name + "/potential"])) <= 1e-11 * abs(float(golden[name + "/potential"]))


DIAG_KEYS = ("move_type", "proposed_delta", "accepted", "d_hardcore", "d_nonbonded_soft", "d_sasa", "d_dih",
             "d_ss")


@pytest.mark.parametrize("name", ("1XQ8", "SM_P01308"))
def test_oracle_mc_diagnostic_matches_reference(golden, name):
    """run_mc_diagnostic records (seeded member generator) from the
    reference module against the oracle: move types and accept decisions
    exact, |delta| exact, energy changes within 1e-9 of the move's scale."""
    import protein_physics as PP
    s = pdb.load(_path(name))
    c = O.topo_to_csr(pdb.topology(s, PP))
    steps, seed = int(golden[name + "/diag_steps"]), int(golden[name + "/diag_seed"])
    got = O.mc_diagnostic(s, c, steps, 0.0019872041 * 300.0, 0.12, seed=seed)
    ref = {k: golden[name + "/diag_" + k] for k in DIAG_KEYS}
    assert np.array_equal(got["move_type"], ref["move_type"])
    assert np.array_equal(got["accepted"], ref["accepted"])
    assert np.array_equal(got["proposed_delta"], ref["proposed_delta"])
    scale = sum(np.abs(ref[k]) for k in DIAG_KEYS[3:]) + 1e-12
    for k in DIAG_KEYS[3:]:
        assert np.all(np.abs(got[k] - ref[k]) <= 1e-9 * scale), (name, k)


def test_exchange_generator_is_numpys():
    """The replica-exchange draws come from numpy.random.default_rng(seed)
    .random() (python/gui_main.py:1961, 2009): the restated SeedSequence +
    PCG64 stream equals numpy's bit for bit, one- and two-word seeds."""
    for seed in (0, 1, 3, 12345, 2**32 - 1, 2**32 + 7, 2**63 + 11):
        assert np.array_equal(O.pcg64_random(seed, 64), np.random.default_rng(seed).random(64)), seed


def test_oracle_landscape_pt_matches_reference(golden):
    """The oracle's PT driver against the GUI's loop over the reference
    module's run_landscape_segment (seeded engines; tests/golden): pooled
    snapshots bit for bit, energies to 1e-12, exchange counts equal."""
    import protein_physics as PP
    s = pdb.load(_path("SM_P01308"))
    c = O.topo_to_csr(pdb.topology(s, PP))
    R, ns, sps, si, lr, ma, seed, ex = golden["SM_P01308/pt_args"].tolist()
    o = O.landscape_pt(s, c, int(R), int(ns), int(sps), 0.0019872041 * 300.0, lr, int(si), ma, seed=int(seed),
                       exchange_seed=int(ex))
    assert np.array_equal(o["swaps"], golden["SM_P01308/pt_swaps"])
    assert np.array_equal(o["snapshots"], golden["SM_P01308/pt_xyz"])
    ref = golden["SM_P01308/pt_e"]
    assert np.all(np.abs(o["energies"] - ref) <= 1e-12 * np.abs(ref))


def test_null_move_rule():
    """ALMA_NULL_DE (include/alma.h) and the oracle's null_floor: with the
    floor, a move with |dE| <= 1e-10 takes the Metropolis draw.  Against the
    reference's rule (floor 0) the two runs agree step for step up to the
    first null move whose rounding-noise dE is negative (there the
    reference accepts without a draw): same move, both accepted, and the
    generator streams part from the next step on."""
    import re
    from alma import _native as N
    hdr = open(os.path.join(os.path.dirname(GOLDEN), "..", "include", "alma.h")).read()
    assert float(re.search(r"#define ALMA_NULL_DE (\S+)", hdr).group(1)) == N.NULL_DE
    import protein_physics as PP
    s = pdb.load(_path("SM_P01308"))
    c = O.topo_to_csr(pdb.topology(s, PP))
    kt = 0.0019872041 * 300.0
    ref = O.mc_run(s, c, 1500, kt, 0.12, seed=7, mode=O.MODE_LANDSCAPE, trace=True)["trace"]
    flo = O.mc_run(s, c, 1500, kt, 0.12, seed=7, mode=O.MODE_LANDSCAPE, trace=True, null_floor=N.NULL_DE)["trace"]
    null = np.abs(ref["dE"]) <= N.NULL_DE
    neg = np.flatnonzero(null & (ref["dE"] < 0))
    assert len(neg) > 0 and np.sum(null & (ref["dE"] >= 0)) > 0
    k = int(neg[0])
    for key in ("kind", "bond", "delta", "accepted", "words_before"):
        assert np.array_equal(ref[key][:k + 1], flo[key][:k + 1]), key
    assert ref["accepted"][k] == 1
    assert flo["words_before"][k + 1] == ref["words_before"][k + 1] + 2   # the floor's draw: 2 words
    # every null move is accepted under either rule
    assert np.all(ref["accepted"][null] == 1)


def test_oracle_landscape_pt_rules():
    """The oracle's PT driver: one replica equals a plain landscape
    trajectory of the same length (same generator, a fresh build per
    segment, step threaded); exchange counts follow the alternating-pair
    schedule; pooled energies are fresh calculate_potential values."""
    import protein_physics as PP
    s = pdb.load(_path("SM_P01308"))
    c = O.topo_to_csr(pdb.topology(s, PP))
    kt = 0.0019872041 * 300.0
    r = O.landscape_pt(s, c, 3, 6, 15, kt, 1.6, 3, 0.12, seed=5, exchange_seed=2)
    assert r["swaps"][0] == 1 + 1          # segment 0: pair (0,1); segment 1: pair (1,2)
    assert len(r["energies"]) == 6
    last = dict(s, xyz=r["snapshots"][-1])
    assert abs(O.energy_terms(last, c)["TOTAL"] - r["energies"][-1]) <= 1e-12 * abs(r["energies"][-1])
    one = O.landscape_pt(s, c, 1, 2, 15, kt, 1.6, 2, 0.12, seed=9, exchange_seed=0)
    traj = O.mc_run(s, c, 30, kt, 0.12, seed=9, mode=O.MODE_LANDSCAPE, snapshot_every=15)
    assert np.array_equal(one["snapshots"], traj["snapshots"])
    # (the oracle's energy sums run over OpenMP threads: order not fixed)
    assert np.allclose(one["energies"], traj["energies"], rtol=1e-12, atol=0)
