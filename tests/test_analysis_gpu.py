"""Ensemble metrics of protein_analysis on the device (alma_ensemble_matrices):
bit-exact against the reference module's plain build (tests/golden) and
against the oracle's serial loop on shapes around the kernel's tile (4 rows x
64 columns) and unroll (4 snapshots) boundaries, at the GUI's largest case
(500 Calpha, 1,000 snapshots), with one output or both, and through the
device-pointer entry on torch tensors."""
import os

import numpy as np
import pytest

from alma import _native as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ENS = ("rw", "pt", "empty_snaps", "empty_atoms", "one_atom")


@pytest.fixture(scope="module")
def g():
    return np.load(os.path.join(GOLDEN, "analysis.npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", ENS)
def test_dropin_matrices_match_reference(g, name):
    import protein_analysis as pa
    x = g["in/ens/" + name]
    print("drop-in matrices", name, x.shape, flush=True)
    assert np.array_equal(pa.compute_mean_dist_matrix(x), g["plain/mean/" + name])
    assert np.array_equal(pa.compute_contact_freq_matrix(x), g["plain/freq/" + name])
    assert np.array_equal(pa.compute_contact_freq_matrix(x, 6.0), g["plain/freq6/" + name])
    assert np.array_equal(pa.compute_contact_freq_matrix(x, cutoff=6.0), g["plain/freq6/" + name])


SHAPES = [(1, 2), (16, 8), (17, 9), (5, 63), (33, 64), (15, 65), (64, 300), (3, 1000), (40, 129)]


@pytest.mark.parametrize("s,n", SHAPES)
def test_matrices_match_oracle(ctx, s, n):
    print("matrices", s, n, flush=True)
    rng = np.random.default_rng(s * 1000 + n)
    x = np.cumsum(rng.normal(size=(s, n, 3)) * 2.3, axis=1)
    m, f = ctx.ensemble_matrices(x, 7.0)
    om, of = O.ensemble_matrices(x, 7.0)
    assert np.array_equal(m, om) and np.array_equal(f, of)
    m1, f1 = ctx.ensemble_matrices(x, 7.0, freq=False)
    m2, f2 = ctx.ensemble_matrices(x, 7.0, mean=False)
    assert f1 is None and m2 is None
    assert np.array_equal(m1, om) and np.array_equal(f2, of)


def test_gui_largest_case(ctx):
    """500 Calpha x 1,000 snapshots (the GUI's 1YPI-sized ensemble): equal to
    the oracle bit for bit; symmetric, zero diagonal, frequencies in k / S."""
    rng = np.random.default_rng(11)
    x = np.cumsum(rng.normal(size=(1000, 500, 3)) * 2.3, axis=1)
    print("largest case", x.shape, flush=True)
    m, f = ctx.ensemble_matrices(x)
    om, of = O.ensemble_matrices(x)
    assert np.array_equal(m, om) and np.array_equal(f, of)
    assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0) and np.all(np.diag(f) == 0)
    k = f * 1000
    assert np.all(np.abs(k - np.round(k)) < 1e-9)
    t = ctx.ensemble_timing()
    assert t["kernel_ms"] > 0 and t["tiles"] > 0


def test_device_pointer_entry(ctx):
    import torch
    rng = np.random.default_rng(5)
    x = np.cumsum(rng.normal(size=(40, 150, 3)) * 2.3, axis=1)
    d = torch.tensor(x, device="cuda:0")
    dm = torch.empty((150, 150), dtype=torch.float64, device="cuda:0")
    df = torch.empty((150, 150), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    ctx.ensemble_matrices_device(d.data_ptr(), 40, 150, 8.0, dm.data_ptr(), df.data_ptr())
    m, f = ctx.ensemble_matrices(x, 8.0)
    assert np.array_equal(dm.cpu().numpy(), m) and np.array_equal(df.cpu().numpy(), f)
    ctx.ensemble_matrices_device(d.data_ptr(), 40, 150, 8.0, 0, df.data_ptr())
    assert np.array_equal(df.cpu().numpy(), f)
    # no snapshot: zero matrices
    dm.fill_(1.0)
    torch.cuda.synchronize()
    ctx.ensemble_matrices_device(d.data_ptr(), 0, 150, 8.0, dm.data_ptr(), 0)
    assert float(dm.abs().sum()) == 0.0


def test_bad_arguments(ctx):
    with pytest.raises(ValueError):
        ctx.ensemble_matrices(np.zeros((3, 4)))
    with pytest.raises(ValueError):
        ctx.ensemble_matrices(np.zeros((3, 4, 3)), mean=False, freq=False)
    with pytest.raises(ValueError):
        ctx.ensemble_matrices_device(0, 2, 4, 8.0, 0, 0)


def test_special_distances(ctx):
    """Coincident atoms (distance 0), subnormal and near-overflow squared
    distances: the square root's fallback path, bit-equal to the oracle."""
    x = np.zeros((3, 6, 3))
    x[:, 1] = x[:, 0]                       # coincident
    x[:, 2] = [1e-160, 0.0, 0.0]            # d^2 ~ 1e-320 (subnormal)
    x[:, 3] = [3e-155, 2e-155, 0.0]         # d^2 below 2^-767
    x[:, 4] = [1e150, -1e150, 5e149]        # d^2 ~ 2e300
    x[:, 5] = [0.7, 1.3, -2.2]
    x[1] *= 1.0000001
    m, f = ctx.ensemble_matrices(x)
    om, of = O.ensemble_matrices(x)
    assert np.array_equal(m, om) and np.array_equal(f, of)
