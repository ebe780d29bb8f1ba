"""Golden vectors written by the reference module built from its own
sources (tests/golden/make_golden.py) against
  * the host topology builder of the drop-in module (bit-exact arrays and
    orders, incl. the hash-map order of concerted_pairs),
  * the FP64 oracle: neighbour list (offsets and SHA-256 of the entries),
    Born radii, each energy term, calculate_potential, seeded
    generate_ensemble candidates and a seeded landscape trajectory,
  * the libstdc++ distribution rules the oracle restates, and the
    reference's dihedral convention.
CPU only."""
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KT = 0.59616
SYSTEMS = ("chain56", "ss24")


def _load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.int32).tobytes()).hexdigest()


def _sys(g):
    return {"xyz": g["xyz"], "charge": g["charge"], "rstar": g["rstar"], "eps": g["eps"]}


def _particles(g):
    import protein_physics as PP
    x = g["xyz"]
    return [PP.Particle(float(x[k, 0]), float(x[k, 1]), float(x[k, 2]), float(g["charge"][k]),
                        float(g["rstar"][k]), float(g["eps"][k])) for k in range(len(x))]


def _topology(g):
    import protein_physics as PP
    t = PP.BondTopology()
    t.build([str(v) for v in g["resnames"]], [str(v) for v in g["atomnames"]],
            [int(v) for v in g["res_idx"]])
    for i, j in g["ss_pairs"]:
        t.add_disulfide(int(i), int(j))
    t.identify_concerted_sidechain_pairs(_particles(g))
    return t


@pytest.fixture(scope="module", params=SYSTEMS)
def case(request):
    g = _load(request.param)
    t = _topology(g)
    return request.param, g, t, O.topo_to_csr(t)


def test_topology_bitexact(case):
    name, g, t, c = case
    excl = t.excl
    assert np.array_equal(np.concatenate([[0], np.cumsum([len(v) for v in excl])]), g["t_excl_off"])
    assert np.array_equal(np.array([j for v in excl for j in v]), g["t_excl_idx"])
    for key, got in (("t_rb_i", c["rb_i"]), ("t_rb_j", c["rb_j"]), ("t_rb_kind", c["rb_kind"]),
                     ("t_side_off", c["side_off"]), ("t_dih_a", c["dih_a"]), ("t_dih_b", c["dih_b"]),
                     ("t_dih_c", c["dih_c"]), ("t_dih_d", c["dih_d"]), ("t_dih_toff", c["dih_toff"]),
                     ("t_dih_n", c["dih_n"])):
        assert np.array_equal(got, g[key]), (name, key)
    assert _sha(c["side_idx"]) == str(g["t_side_sha"])
    # FP64 tables: exact
    assert np.array_equal(c["rb_scale"], g["t_rb_scale"])
    assert np.array_equal(c["dih_v2"], g["t_dih_v2"])
    assert np.array_equal(c["dih_gamma"], g["t_dih_gamma"])
    assert np.array_equal(np.array(t.disulfide_pairs).reshape(-1, 2), g["t_ss"])
    assert np.array_equal(np.array(t.concerted_pairs).reshape(-1, 2), g["t_cp"])
    assert np.array_equal(np.array(t.concerted_sidechain_pairs).reshape(-1, 2), g["t_csp"])
    assert np.array_equal(np.array(t.bonds).reshape(-1, 2), g["t_bonds"])


def test_nlist_and_born(case):
    name, g, _, _ = case
    offs, js = O.nlist(g["xyz"])
    assert np.array_equal(offs, g["nl_off"])
    assert _sha(js) == str(g["nl_sha"])
    a, _ = O.born(g["xyz"], g["rstar"], offs, js)
    assert np.max(np.abs(a - g["born"]) / g["born"]) <= 1e-13


def test_energy_terms(case):
    name, g, _, c = case
    s = _sys(g)
    for got, ref in ((O.energy_terms(s, c), g["terms_topo"]), (O.energy_terms(s), g["terms_notopo"])):
        for k, term in enumerate(O.TERMS):
            assert abs(got[term] - ref[k]) <= 1e-11 * max(abs(ref[k]), 1e-300) or got[term] == ref[k], \
                (name, term, got[term], ref[k])
    assert abs(O.energy_terms(s, c)["TOTAL"] - g["potential_topo"]) <= 1e-11 * abs(g["potential_topo"])
    assert abs(O.energy_terms(s)["TOTAL"] - g["potential_notopo"]) <= 1e-11 * abs(g["potential_notopo"])


def test_seeded_generate_ensemble(case):
    """The serial candidate loop draws every candidate from one generator:
    candidate 1 starts where candidate 0 stopped."""
    name, g, _, c = case
    s = _sys(g)
    seed, steps = int(g["ens_seed"]), int(g["ens_steps"])
    r0 = O.mc_run(s, c, steps, KT, 0.12, seed=seed)
    assert np.max(np.abs(r0["xyz"] - g["ens_xyz"][0])) <= 1e-9, name
    r1 = O.mc_run(s, c, steps, KT, 0.12, seed=seed, skip=r0["stats"]["words"])
    assert np.max(np.abs(r1["xyz"] - g["ens_xyz"][1])) <= 1e-9, name


def test_seeded_landscape_trajectory(case):
    name, g, _, c = case
    s = _sys(g)
    n_snap, every = int(g["land_snaps"]), int(g["land_every"])
    r = O.mc_run(s, c, n_snap * every, KT, 0.12, seed=int(g["land_seed"]), mode=O.MODE_LANDSCAPE,
                 snapshot_every=every)
    assert np.max(np.abs(r["energies"] - g["land_energies"]) / np.abs(g["land_energies"])) <= 1e-11
    assert np.max(np.abs(r["snapshots"][-1] - g["land_last_xyz"])) <= 1e-9


def test_bridge_three_particles():
    g = _load("bridge")
    got = O.energy_terms({k: g[k] for k in ("xyz", "charge", "rstar", "eps")})
    for k, term in enumerate(O.TERMS):
        assert abs(got[term] - g["terms"][k]) <= 1e-12 * max(abs(g["terms"][k]), 1e-300), term
    assert abs(got["TOTAL"] - g["potential"]) <= 1e-12 * abs(g["potential"])


def test_rng_rules():
    g = _load("rng")
    # [rand.predef]: the 10000th output of a default-constructed mt19937
    assert int(g["mt_5489_10000"]) == 4123659995
    assert O.mt_words(5489, 10000)[-1] == 4123659995
    for s in (0, 1, 1234):
        assert np.array_equal(O.mt_words(s, 16), g["mt_%d_first" % s])
    # uniform_int_distribution<int>(0, hi) and uniform_real_distribution<double>
    # as the oracle's generator class draws them
    for hi in (1, 2, 6, 99, 1199, 65535, 2147483646):
        assert np.array_equal(O.uniform_int(1234, hi, 64), g["uint_%d" % hi]), hi
    assert np.array_equal(O.uniform_real(99, 0.0, 1.0, 64), g["ureal_u01"])
    assert np.array_equal(O.uniform_real(99, -0.5235987755982988, 0.5235987755982988, 64), g["ureal_pm30"])


def test_dihedral_convention():
    g = _load("dihedral")
    got = np.array([O.dihedral_angle(q) for q in g["xyz"]])
    assert np.max(np.abs(got - g["angle"])) <= 1e-13
    # closed form: phi_ref = atan2(|b2| sin(phi), -cos(phi)) with the IUPAC angle
    for q, ref in zip(g["xyz"], g["angle"]):
        b1, b2, b3 = q[0] - q[1], q[2] - q[1], q[3] - q[2]
        n1, n2 = np.cross(-b1, b2), np.cross(b2, b3)
        iupac = np.arctan2(np.dot(np.cross(n1, n2), b2 / np.linalg.norm(b2)), np.dot(n1, n2))
        want = np.arctan2(np.linalg.norm(b2) * np.sin(iupac) * np.linalg.norm(n1) * np.linalg.norm(n2),
                          -np.cos(iupac) * np.linalg.norm(n1) * np.linalg.norm(n2))
        assert abs(np.angle(np.exp(1j * (want - ref)))) <= 1e-9
