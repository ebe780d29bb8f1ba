"""Structures read from PDB files (the reference's own data files) through the
GPU path: per-term energies against the reference's calculate_potential on
the same particles (tests/golden/pdb_golden.npz), the drop-in module's
calculate_potential, forces and MC trajectories against the FP64 oracle.
Tolerances as for the synthetic systems: 1e-5 relative per term; forces
1e-5 * (|g_i| + rms g); trajectories with the same accept decisions,
coordinates within 1e-8 A."""
import os

import numpy as np
import pytest

from alma import _native as N
from alma import pdb
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STRUCTURES = ("1XQ8", "SM_P01308", "SM_P21543")
KT = 0.0019872041 * 300.0


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "pdb_golden.npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def _parsed(name):
    import protein_physics as PP
    s = pdb.load(os.path.join(GOLDEN, "pdb", name + ".pdb.gz"))
    t = pdb.topology(s, PP)
    return s, t, O.topo_to_csr(t)


@pytest.mark.parametrize("name", STRUCTURES)
def test_energy_terms_match_reference(ctx, golden, name):
    s, t, c = _parsed(name)
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    got = ctx.energy()
    ref = golden[name + "/terms"]
    for k, term in enumerate(O.TERMS):
        if ref[k] == 0.0:
            assert got[term] == 0.0, (name, term)
        else:
            assert abs(got[term] - ref[k]) <= 1e-5 * abs(ref[k]), (name, term, got[term], ref[k])


@pytest.mark.parametrize("name", ("1XQ8", "SM_P21543"))
def test_dropin_calculate_potential(golden, name):
    import protein_physics as PP
    s, t, _ = _parsed(name)
    e = PP.PhysicsEngine().calculate_potential(pdb.particles(s, PP), t)
    ref = float(golden[name + "/potential"])
    assert abs(e - ref) <= 1e-5 * abs(ref), (name, e, ref)


def test_forces_match_oracle(ctx):
    s, t, c = _parsed("SM_P01308")
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    e, g, gt = ctx.forces()
    E, G, GT = O.forces(s, c)
    assert abs(e - E) <= 1e-5 * abs(E)
    err = np.linalg.norm(g - G, axis=1)
    mag = np.linalg.norm(G, axis=1)
    assert np.all(err <= 1e-5 * (mag + np.sqrt(np.mean(mag ** 2))))
    assert np.all(np.abs(gt - GT) <= 1e-5 * (np.abs(GT) + np.sqrt(np.mean(GT ** 2))))


@pytest.mark.parametrize("name,mode", [("1XQ8", N.MC_ENSEMBLE), ("SM_P01308", N.MC_LANDSCAPE),
                                       ("SM_P21543", N.MC_ENSEMBLE)])
def test_mc_trajectory_matches_oracle(ctx, name, mode):
    s, t, c = _parsed(name)
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    steps = 60
    gpu = ctx.mc_run(mode, steps, n_chains=2, temperature=KT, max_angle=0.12, seed=2024, trace=True)
    orc = O.mc_run(s, c, steps, KT, 0.12, seed=2024, mode=mode, trace=True, l1=True, null_floor=N.NULL_DE)
    gt, ot = gpu["trace"], orc["trace"]
    assert np.array_equal(gt["kind"], ot["kind"]), name
    assert np.array_equal(gt["accepted"], ot["accepted"]), name
    assert np.all(np.abs(gt["dE"] - ot["dE"]) <= 1e-5 * ot["dE_l1"] + 1e-9), name
    assert np.max(np.abs(gpu["xyz"][0] - orc["xyz"])) < 1e-8, name
    o2 = O.mc_run(s, c, steps, KT, 0.12, seed=2025, mode=mode, null_floor=N.NULL_DE)
    assert np.max(np.abs(gpu["xyz"][1] - o2["xyz"])) < 1e-8, name


DIAG_KEYS = ("move_type", "proposed_delta", "accepted", "d_hardcore", "d_nonbonded_soft", "d_sasa", "d_dih",
             "d_ss")


@pytest.mark.parametrize("name,steps,seed", [("SM_P01308", 200, 57), ("1XQ8", 120, 31)])
def test_mc_diagnostic_matches_oracle(ctx, name, steps, seed):
    """run_mc_diagnostic on the device against the oracle's (itself pinned to
    the reference's records): same logged moves, |delta| and decisions,
    energy changes within 1e-5 of each move's per-item scale."""
    s, t, c = _parsed(name)
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    got = ctx.mc_diagnostic(steps, temperature=KT, max_angle=0.12, seed=seed)
    orc = O.mc_diagnostic(s, c, steps, KT, 0.12, seed=seed, l1=True, null_floor=N.NULL_DE)
    assert np.array_equal(got["move_type"], orc["move_type"]), name
    assert np.array_equal(got["accepted"], orc["accepted"]), name
    assert np.array_equal(got["proposed_delta"], orc["proposed_delta"]), name
    tol = 1e-5 * orc["dE_l1"] + 1e-9
    for k in DIAG_KEYS[3:]:
        assert np.all(np.abs(got[k] - orc[k]) <= tol), (name, k)


def test_dropin_run_mc_diagnostic(golden):
    """The drop-in method against the reference module's own records
    (seeded): the tuple of 8 lists."""
    import protein_physics as PP
    s, t, c = _parsed("SM_P01308")
    rec = PP.PhysicsEngine().run_mc_diagnostic(pdb.particles(s, PP), t, 200, KT, 0.12, seed=57)
    assert len(rec) == 8
    ref = [golden["SM_P01308/diag_" + k] for k in DIAG_KEYS]
    assert list(rec[0]) == ref[0].tolist() and list(rec[2]) == ref[2].tolist()
    assert np.array_equal(np.array(rec[1]), ref[1])
    # error scale of each move: sum of |per-item change| (from the oracle's
    # replay of the same moves)
    tol = 1e-5 * O.mc_diagnostic(s, c, 200, KT, 0.12, seed=57, l1=True, null_floor=N.NULL_DE)["dE_l1"] + 1e-9
    for got, want in zip(rec[3:], ref[3:]):
        assert np.all(np.abs(np.array(got) - want) <= tol)


def test_landscape_pt_matches_oracle(ctx):
    """The GUI's replica-exchange driver on the device against the oracle's:
    4 slots, 2 segments of 3 snapshots x 15 steps on insulin (concerted,
    crankshaft and torsion moves, disulfides): same pooled snapshots
    (1e-8 A), fresh energies (1e-5 relative) and exchange counts."""
    s, t, c = _parsed("SM_P01308")
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    g = ctx.landscape_pt(4, 6, 15, temperature=KT, ladder_ratio=1.6, swap_interval=3, max_angle=0.12, seed=21,
                         exchange_seed=4)
    o = O.landscape_pt(s, c, 4, 6, 15, KT, 1.6, 3, 0.12, seed=21, exchange_seed=4, null_floor=N.NULL_DE)
    assert np.array_equal(g["swaps"], o["swaps"])
    assert np.max(np.abs(g["snapshots"] - o["snapshots"])) < 1e-8
    assert np.all(np.abs(g["energies"] - o["energies"]) <= 1e-5 * np.abs(o["energies"]))


def test_dropin_run_parallel_tempering():
    import protein_physics as PP
    s, t, c = _parsed("SM_P01308")
    snaps, es = PP.PhysicsEngine().run_parallel_tempering(pdb.particles(s, PP), t, 6, 15, KT, 0.12, 4, 1.6, 3,
                                                          rng_seed=4, seed=21)
    o = O.landscape_pt(s, c, 4, 6, 15, KT, 1.6, 3, 0.12, seed=21, exchange_seed=4, null_floor=N.NULL_DE)
    assert len(snaps) == 6 and len(es) == 6
    x = np.array([[p.x, p.y, p.z] for p in snaps[-1]])
    assert np.max(np.abs(x - o["snapshots"][-1])) < 1e-8
    assert np.all(np.abs(np.array(es) - o["energies"]) <= 1e-5 * np.abs(o["energies"]))


def test_landscape_pt_pool_analysis_on_device(ctx):
    """The pool's Calpha mean-distance / contact-frequency matrices computed
    on the device while the snapshots are taken equal protein_analysis on the
    pooled snapshots (oracle loop, bit for bit)."""
    s, t, c = _parsed("SM_P01308")
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    ca = s["ca_indices"]
    g = ctx.landscape_pt(3, 4, 10, temperature=KT, ladder_ratio=1.6, swap_interval=2, max_angle=0.12, seed=3,
                         exchange_seed=1, analysis_atoms=ca, contact_cutoff=7.5)
    om, of = O.ensemble_matrices(g["snapshots"][:, ca, :], 7.5)
    assert np.array_equal(g["mean_dist"], om) and np.array_equal(g["contact_freq"], of)
    with pytest.raises(ValueError):
        ctx.landscape_pt(2, 2, 5, temperature=KT, swap_interval=1, analysis_atoms=[0, len(s["xyz"])])


@pytest.mark.parametrize("r,ns,sps,si", [(1, 4, 10, 2), (2, 3, 12, 5), (5, 6, 8, 2)])
def test_landscape_pt_shapes_match_oracle(ctx, r, ns, sps, si):
    """One slot (no exchanges), a swap interval longer than the run (one
    segment of swap_interval snapshots, as max(1, n // k) gives), five slots
    over three segments (both pair offsets)."""
    s, t, c = _parsed("SM_P01308")
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    print("landscape pt", r, ns, sps, si, flush=True)
    g = ctx.landscape_pt(r, ns, sps, temperature=KT, ladder_ratio=1.4, swap_interval=si, max_angle=0.12, seed=7,
                         exchange_seed=11)
    o = O.landscape_pt(s, c, r, ns, sps, KT, 1.4, si, 0.12, seed=7, exchange_seed=11, null_floor=N.NULL_DE)
    assert len(g["energies"]) == max(1, ns // si) * si
    assert np.array_equal(g["swaps"], o["swaps"])
    assert np.max(np.abs(g["snapshots"] - o["snapshots"])) < 1e-8
    assert np.all(np.abs(g["energies"] - o["energies"]) <= 1e-5 * np.abs(o["energies"]))


def test_landscape_pt_capacity_growth(ctx, monkeypatch):
    """Row capacities far too small at the start: the chains grow them and
    rebuild, and the run equals the oracle's as with ample capacities."""
    s, t, c = _parsed("SM_P01308")
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    monkeypatch.setenv("ALMA_MC_TEST_CAPS", "64,0")
    g = ctx.landscape_pt(3, 4, 10, temperature=KT, ladder_ratio=1.6, swap_interval=2, max_angle=0.12, seed=5,
                         exchange_seed=2)
    assert ctx.mc_timing()["regrows"] >= 1
    o = O.landscape_pt(s, c, 3, 4, 10, KT, 1.6, 2, 0.12, seed=5, exchange_seed=2, null_floor=N.NULL_DE)
    assert np.array_equal(g["swaps"], o["swaps"])
    assert np.max(np.abs(g["snapshots"] - o["snapshots"])) < 1e-8


def test_landscape_pt_bad_arguments(ctx):
    s, t, c = _parsed("SM_P01308")
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    for kw in ({"n_replicas": 0}, {"n_snapshots": 0}, {"steps_per_snap": 0}, {"swap_interval": 0}):
        args = dict(n_replicas=2, n_snapshots=2, steps_per_snap=5, swap_interval=1)
        args.update(kw)
        with pytest.raises(ValueError):
            ctx.landscape_pt(args["n_replicas"], args["n_snapshots"], args["steps_per_snap"], temperature=KT,
                             swap_interval=args["swap_interval"])
    with pytest.raises(ValueError):
        ctx.landscape_pt(2, 2, 5, temperature=-1.0, swap_interval=1)
