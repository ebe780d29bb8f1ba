"""GPU forces (dE/dx with the Born chain rule, dE/dtheta per rotatable bond)
against the FP64 oracle gradient, whose own correctness is pinned by central
differences in test_forces_cpu.py.

Tolerance: the pair derivatives are FP32 per pair with FP64 accumulation, so
a component's error is relative to the size of the forces acting on that
atom: |g - g64| <= 1e-5 * (|g64_i| + rms(g64)) per atom (norms over xyz), and
the same rule for the torsion gradients."""
import numpy as np
import pytest

import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    from alma import _native
    c = _native.Context(0)
    yield c
    c.close()


def _check_vec(got, ref, tag):
    got = np.asarray(got).reshape(len(ref), -1)
    ref = np.asarray(ref).reshape(len(ref), -1)
    err = np.linalg.norm(got - ref, axis=1)
    mag = np.linalg.norm(ref, axis=1)
    rms = np.sqrt(np.mean(mag ** 2)) if len(mag) else 0.0
    bad = err > TOL * (mag + rms)
    assert not bad.any(), (tag, np.flatnonzero(bad)[:8], err[bad][:8], mag[bad][:8], rms)


CASES = [("chain56", lambda: S.chain(56, 1)), ("chain300", lambda: S.chain(300, 2)),
         ("chain_ss", S.chain_ss)]


@pytest.mark.parametrize("name,make", CASES)
def test_forces_match_oracle(ctx, name, make):
    s = make()
    pairs = [tuple(int(v) for v in p) for p in s["ss_pairs"]]
    c = S.csr(S.topology(s, pairs))
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    e, g, gt = ctx.forces()
    E, G, GT = O.forces(s, c)
    assert abs(e - E) <= 1e-5 * abs(E), (name, e, E)
    _check_vec(g, G, name + "/dEdx")
    _check_vec(gt, GT, name + "/dEdtheta")


def test_forces_without_topology(ctx):
    s = S.chain(56, 1)
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"])
    e, g, gt = ctx.forces()
    E, G, _ = O.forces(s, None)
    assert len(gt) == 0
    assert abs(e - E) <= 1e-5 * abs(E)
    _check_vec(g, G, "notopo")


def test_forces_small_systems(ctx):
    s = S.bridge()
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"])
    e, g, _ = ctx.forces()
    E, G, _ = O.forces(s, None)
    assert abs(e - E) <= 1e-5 * abs(E)
    _check_vec(g, G, "bridge")
    ctx.set_system(np.zeros((0, 3)), np.zeros(0), np.zeros(0), np.zeros(0))
    e, g, gt = ctx.forces()
    assert e == 0.0 and g.shape == (0, 3)


def test_forces_consistent_after_mc(ctx):
    """The gradient at a state produced on the device equals the oracle's at
    the same coordinates (the context reloads the MC end state)."""
    s = S.chain(56, 1)
    c = S.csr(S.topology(s))
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], c)
    from alma import _native
    r = ctx.mc_run(_native.MC_TORSION30, 300, n_chains=1, temperature=0.59616, seed=9)
    x = r["xyz"][0]
    ctx.set_positions(x)
    e, g, gt = ctx.forces()
    d = dict(s)
    d["xyz"] = x
    E, G, GT = O.forces(d, c)
    assert abs(e - E) <= 1e-5 * abs(E)
    _check_vec(g, G, "after-mc")
    _check_vec(gt, GT, "after-mc/theta")


def test_dropin_forces_and_autograd(ctx):
    import torch

    import protein_physics as PP
    from alma import autograd as AG
    s = S.chain(56, 1)
    t = S.topology(s)
    c = S.csr(t)
    eng = PP.PhysicsEngine()
    e, g, gt = eng.forces(S.particles(s), t)
    E, G, GT = O.forces(s, c)
    assert abs(e - E) <= 1e-5 * abs(E)
    _check_vec(np.asarray(g).reshape(-1, 3), G, "dropin")
    _check_vec(np.asarray(gt), GT, "dropin/theta")
    # torch.autograd.Function over Cartesian coordinates on the GPU
    model = AG.EnergyModel(s["charge"], s["rstar"], s["eps"], c)
    x = torch.tensor(s["xyz"], dtype=torch.float64, device="cuda", requires_grad=True)
    en = model(x)
    en.backward()
    assert abs(en.item() - E) <= 1e-5 * abs(E)
    _check_vec(x.grad.cpu().numpy(), G, "autograd")
    # chain rule through a downstream function: d(2E + 3)/dx = 2 dE/dx
    x.grad = None
    (2.0 * model(x) + 3.0).backward()
    _check_vec(x.grad.cpu().numpy(), 2.0 * G, "autograd-scaled")
