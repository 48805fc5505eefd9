"""The learned-compensation network on the GPU: forward and gradients
against the numpy oracle (BF16-rounded at the MFMA inputs, FP32/FP64
accumulation), Adam training lowers the loss, the compensated
factorization matches the oracle's, and the learned solve falls back to
plain ILU when it misses its step budget."""
from __future__ import annotations

import numpy as np
import pytest

import par_ilu0_cpp_amd as P
from oracle import gnn_oracle as G
from par_ilu0_cpp_amd import generators as gen
from par_ilu0_cpp_amd.learned import DeviceSystem, Gnn, transpose_map

pytestmark = pytest.mark.gpu


def dev(*arrs):
    import torch

    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


def grf(s, seed):
    return gen.grf_diffusion3d(s, sigma=1.5, corr_len=4.0, seed=seed)


@pytest.mark.parametrize("s", [5, 11])
def test_forward_matches_oracle(s):
    import torch

    a = grf(s, seed=s)
    p = np.random.default_rng(s).normal(0, 0.25, G.N_PARAMS).astype(np.float32)
    net = Gnn(p)
    rp, ci, va = dev(a.row_ptr, a.col_idx, a.value)
    om = torch.zeros(a.n, dtype=torch.float32, device="cuda")
    net.predict(a.n, rp.data_ptr(), ci.data_ptr(), va.data_ptr(), om.data_ptr())
    feat, w = G.features(a.row_ptr, a.col_idx, a.value)
    ref, _, _ = G.forward(p.astype(np.float64), a.row_ptr, a.col_idx, feat, w, use_bf16=True)
    got = om.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 5e-3, np.max(np.abs(got - ref))
    assert np.allclose(net.params(), p)


def test_gradients_match_oracle():
    a = grf(9, seed=4)
    rng = np.random.default_rng(4)
    p = rng.normal(0, 0.25, G.N_PARAMS).astype(np.float32)
    t = rng.uniform(0.3, 0.95, a.n).astype(np.float32)
    net = Gnn(p)
    rpt, cit, tmap = transpose_map(a.row_ptr, a.col_idx)
    d = dev(a.row_ptr, a.col_idx, a.value, rpt, cit, tmap, t)
    loss, g = net.gradients(a.n, *[x.data_ptr() for x in d])
    feat, w = G.features(a.row_ptr, a.col_idx, a.value)
    lref, gref = G.loss_grad(p.astype(np.float64), a.row_ptr, a.col_idx, feat, w, t.astype(np.float64), True)
    assert loss == pytest.approx(lref, rel=2e-2)
    blocks = [(G.OFF_W[l], G.OFF_W[l] + 2 * G.IN_DIM[l] * G.HID) for l in range(3)]
    blocks += [(G.OFF_B[l], G.OFF_B[l] + G.HID) for l in range(3)] + [(G.OFF_WO, G.N_PARAMS)]
    for lo, hi in blocks:
        err = np.linalg.norm(g[lo:hi] - gref[lo:hi]) / max(np.linalg.norm(gref[lo:hi]), 1e-12)
        assert err <= 3e-2, (lo, hi, err)


def test_training_lowers_loss():
    a = grf(10, seed=5)
    t = np.full(a.n, 0.9, dtype=np.float32)
    rpt, cit, tmap = transpose_map(a.row_ptr, a.col_idx)
    d = dev(a.row_ptr, a.col_idx, a.value, rpt, cit, tmap, t)
    net = Gnn()
    losses = [net.train_step(a.n, *[x.data_ptr() for x in d], lr=3e-3) for _ in range(60)]
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])


def test_compensated_factorization_matches_oracle(oracle):
    import torch

    a = grf(8, seed=6)
    omega = np.random.default_rng(6).uniform(0.0, 1.0, a.n).astype(np.float32)
    rp, ci, va, om = dev(a.row_ptr, a.col_idx, a.value, omega)
    ilu = P.IluSolver(0)
    assert ilu.SymbolicDevice(a.n, rp.data_ptr(), ci.data_ptr(), 0) == a.nnz
    ilu.SetDiagonalCompensation(om.data_ptr())
    assert ilu.FactorizeDevice(va.data_ptr(), True) == 0
    fac = ilu.DeviceFactors()
    ref, _ = oracle.ilu_factorize_comp(a.row_ptr, a.col_idx, a.value, omega)
    assert np.max(np.abs(fac - ref) / np.maximum(np.abs(ref), 1e-300)) <= 1e-12
    ilu.SetDiagonalCompensation(0)
    assert ilu.FactorizeDevice(va.data_ptr(), True) == 0
    plain, _, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, a.value)
    assert np.array_equal(ilu.DeviceFactors(), plain)


@pytest.mark.parametrize("lof", [1, 2])
def test_compensated_ilu1_factorization_matches_oracle(oracle, lof):
    """The per-row compensation on an ILU(k) pattern with fill (the learned
    form runs compensated ILU(1)): the dropped updates are those falling
    outside F's pattern, and the factors match the oracle's to 1e-12."""
    a = gen.convdiff3d(9)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, lof)
    frp, fci, fval = o[0], o[1], o[3]
    omega = np.random.default_rng(lof).uniform(0.0, 1.0, a.n).astype(np.float32)
    rp, ci, va, om = dev(a.row_ptr, a.col_idx, a.value, omega)
    ilu = P.IluSolver(0)
    assert ilu.SymbolicDevice(a.n, rp.data_ptr(), ci.data_ptr(), lof) == fci.size
    ilu.SetDiagonalCompensation(om.data_ptr())
    assert ilu.FactorizeDevice(va.data_ptr(), True) == 0
    fac = ilu.DeviceFactors()
    ref, _ = oracle.ilu_factorize_comp(frp, fci, fval, omega)
    plain, _, _ = oracle.ilu_factorize(frp, fci, fval)
    assert np.max(np.abs(fac - ref) / np.maximum(np.abs(ref), 1e-300)) <= 1e-12
    assert not np.allclose(ref, plain)  # the fill dropped at this level changes the pivots
    ilu.SetDiagonalCompensation(0)


def test_learned_solve_fallback():
    a = grf(12, seed=7)
    b = gen.rhs_uniform(a.n, seed=7)
    sys_ = DeviceSystem(a, b, lof=0, m=50)
    ok0, its0, res0 = sys_.solve_with_omega(None)
    net = Gnn()
    ok, its, res, fb = sys_.solve_learned(net, budget=0)  # any step count misses a zero budget
    assert fb and ok and its == its0 and res <= 1.01e-8
    ok2, its2, res2, fb2 = sys_.solve_learned(net, budget=10**6)
    assert not fb2 and ok2 and res2 <= 1.01e-8


@pytest.mark.parametrize("seed", [910_000, 910_001])
def test_trained_network_beats_plain_ilu(seed):
    """Ablation with the shipped weights (trained on 64^3 GRF diffusion) on
    unseen seeds of the same family and size: the learned compensation needs
    far fewer ILU(1)-GMRES(50) steps than plain ILU(1) (measured on MI355X
    for seeds 910000-910005: 0.39-0.47 of the plain count), without falling
    back, and the FP64 true residual meets the tolerance.  A constant
    omega = 0.98 (the most frequent training label) is the non-learned
    reference point: the network must not do worse than it."""
    from par_ilu0_cpp_amd.learned import WEIGHTS

    a = gen.grf_diffusion3d(64, sigma=1.5, corr_len=8.0, seed=seed)
    b = gen.rhs_uniform(a.n, seed=seed)
    sys_ = DeviceSystem(a, b, lof=1, m=50)
    ok0, its0, res0 = sys_.solve_with_omega(None)
    okc, itsc, resc = sys_.solve_with_omega(0.98)
    net = Gnn.load(WEIGHTS)
    ok, its, res, fb = sys_.solve_learned(net, budget=its0)
    assert ok0 and okc and ok and not fb
    assert res <= 1.01e-8 and res0 <= 1.01e-8
    assert its <= 0.55 * its0, (its, its0)
    assert its <= itsc + 2, (its, itsc)
