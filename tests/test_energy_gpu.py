"""GPU parity of the full evaluation (calculate_potential) against the FP64
oracle: neighbour list and Born radii, and each energy term within 1e-5
relative error (FP32 pair arithmetic, FP64 reductions)."""
import numpy as np
import pytest

import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    from alma import _native
    c = _native.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _check_terms(got, ref, tag):
    for k in O.TERMS:
        if ref[k] == 0.0:
            assert got[k] == 0.0, (tag, k, got[k])
            continue
        assert _rel(got[k], ref[k]) <= REL_TOL, (tag, k, got[k], ref[k], _rel(got[k], ref[k]))


CASES = [("chain56", lambda: S.chain(56, 1)), ("chain120", lambda: S.chain(120, 7)),
         ("chain300", lambda: S.chain(300, 2))]


@pytest.mark.parametrize("name,make", CASES)
def test_nlist_and_born_bitexact_order(ctx, name, make):
    s = make()
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"])
    ctx.energy()
    offs, js = ctx.nlist()
    ro, rj = O.nlist(s["xyz"])
    assert np.array_equal(offs, ro)
    assert np.array_equal(js, rj)
    a = ctx.born_radii()
    ra, _ = O.born(s["xyz"], s["rstar"], ro, rj)
    assert np.max(np.abs(a - ra) / ra) <= 1e-6


@pytest.mark.parametrize("name,make", CASES)
def test_energy_terms_with_topology(ctx, name, make):
    s = make()
    t = S.topology(s)
    c = S.csr(t)
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    _check_terms(ctx.energy(), O.energy_terms(s, c), name)


@pytest.mark.parametrize("name,make", CASES[:2])
def test_energy_terms_without_topology(ctx, name, make):
    s = make()
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"])
    _check_terms(ctx.energy(), O.energy_terms(s), name + "-notopo")


def test_disulfide_system(ctx):
    s = S.two_chains_ss()
    sg = [i for i, n in enumerate(s["atomnames"]) if n == "SG"]
    t = S.topology(s, [(sg[0], sg[1])])
    c = S.csr(t)
    assert len(c["ss_i"]) == 1
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    _check_terms(ctx.energy(), O.energy_terms(s, c), "ss")


def test_bridge_three_particles(ctx):
    s = S.bridge()
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"])
    _check_terms(ctx.energy(), O.energy_terms(s), "bridge")


def test_empty_and_single(ctx):
    ctx.set_system(np.zeros((0, 3)), np.zeros(0), np.zeros(0), np.zeros(0))
    assert ctx.energy()["TOTAL"] == 0.0
    s = S.bridge()
    ctx.set_system(s["xyz"][:1], s["charge"][:1], s["rstar"][:1], s["eps"][:1])
    _check_terms(ctx.energy(), O.energy_terms({k: v[:1] for k, v in s.items()}), "single")


def test_dropin_module_matches(ctx):
    import protein_physics as PP
    s = S.chain(56, 1)
    t = S.topology(s)
    eng = PP.PhysicsEngine()
    e = eng.calculate_potential(S.particles(s), t)
    ref = O.energy_terms(s, S.csr(t))["TOTAL"]
    assert _rel(e, ref) <= REL_TOL
    assert "MI" in PP.PhysicsEngine.device_name() or len(PP.PhysicsEngine.device_name()) > 0
