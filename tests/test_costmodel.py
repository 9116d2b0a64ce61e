"""The learned cost model: the float64 oracle is pinned by finite differences and a
closed-form Adam step (CPU); the HIP forward/backward/Adam kernels match it and training
reduces the loss (GPU)."""

import os

import numpy as np
import pytest

from oracle import mlp_oracle
from tensororder_amd import _lib, costmodel, network, planners, schedule
from tensororder_amd.formula import parse_dimacs_text
from tensororder_amd.instances import random_3cnf


def _params(seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(_lib.MLP_NPARAMS) * 0.3


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(1)
    x = rng.standard_normal((40, 16))
    t = rng.standard_normal((40, 2))
    _, g = mlp_oracle.loss_and_grad(p, x, t)
    for i in rng.choice(p.size, 60, replace=False):
        e = np.zeros_like(p)
        e[i] = 1e-6
        lp, _ = mlp_oracle.loss_and_grad(p + e, x, t)
        lm, _ = mlp_oracle.loss_and_grad(p - e, x, t)
        assert g[i] == pytest.approx((lp - lm) / 2e-6, rel=1e-4, abs=1e-8)


def test_oracle_adam_first_step_is_signed_lr():
    # after one step from zero moments the bias-corrected update is lr * g / (|g| + eps)
    g = np.array([0.5, -2.0, 1e-3])
    p, m, v = mlp_oracle.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 0.01, 0.9, 0.999, 1e-12, 1)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-6)


def test_features_are_finite_and_shaped():
    f = parse_dimacs_text(random_3cnf(24, 50, 3), "cachet")
    plan = planners.best_plan(network.cnf_count(f), "factor", solver=None)
    import emulator

    tree, tables, G, total = emulator.plan_tables(plan, emulator.HostOps)
    nodes = schedule.analyse(tree, tables)
    schedule.peak_bytes(nodes, tables, 1, _lib.TO_F64)
    X, mult = costmodel.plan_launch_features(nodes, 1, 8, 1)
    assert X.shape[1] == len(costmodel.FEATURES) == 16
    assert np.isfinite(X).all() and (mult == 1).all()


@pytest.mark.gpu
def test_hip_mlp_matches_oracle_and_trains():
    from tensororder_amd.hip_api import HipAPI

    api = HipAPI()
    model = costmodel.CostModel(api, seed=3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((3000, 16)).astype(np.float32)
    w = rng.standard_normal((16, 2))
    t = (np.tanh(x @ w) + 0.1 * x[:, :2] ** 2).astype(np.float32)
    p = model.params.cpu().numpy().astype(np.float64)
    # forward
    xd = api._torch.from_numpy(x).to(api._device)
    td = api._torch.from_numpy(t).to(api._device)
    _, _, y = model.forward_raw(xd)
    api.synchronize()
    _, _, y_ref = mlp_oracle.forward(p, x)
    np.testing.assert_allclose(y.cpu().numpy(), y_ref, rtol=1e-4, atol=1e-4)
    # backward
    loss, grad = model.gradient(xd, td)
    loss_ref, grad_ref = mlp_oracle.loss_and_grad(p, x, t)
    assert loss == pytest.approx(loss_ref, rel=1e-4)
    np.testing.assert_allclose(grad.cpu().numpy(), grad_ref, rtol=2e-3, atol=2e-5)
    # Adam
    m0 = np.zeros_like(p)
    p_ref, _, _ = mlp_oracle.adam(p, grad.cpu().numpy().astype(np.float64), m0, m0, 1e-3, 0.9, 0.999, 1e-8, 1)
    model.adam_step(1e-3)
    api.synchronize()
    np.testing.assert_allclose(model.params.cpu().numpy(), p_ref, rtol=1e-5, atol=1e-6)
    # training reduces the loss
    hist = model.fit(x, t, epochs=60, batch=512, lr=3e-3)
    assert hist[-1] < 0.5 * hist[0]


CANDIDATES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cnf70_candidates.json")


def test_candidate_plans_share_one_count():
    # every candidate (decomposition x slice spec) is an exact plan of the same network:
    # the oracle's float64 counts agree, whichever one a rule picks
    from oracle import numpy_oracle
    from tensororder_amd import plan_io

    f = parse_dimacs_text(random_3cnf(22, 44, 5, True), "cachet")
    net = network.cnf_count(f)
    tds = planners.decompose(net.structure_text(), None, tie_seeds=[0, 1])
    plans, labels = planners.plan_candidates(net, "factor", tds, 3)
    assert len(plans) >= 4 and all(len(p.groups_to_slice) == 3 for p in plans)
    ref = planners.reference_candidate(labels)
    assert labels[ref]["slicer"] == "greedy_mem"
    assert labels[ref]["unsliced_maxrank"] == min(l["unsliced_maxrank"] for l in labels)
    counts = [numpy_oracle.contract_sliced(plan_io.plan_to_dict(p)) for p in plans]
    np.testing.assert_allclose(counts, counts[ref], rtol=1e-12)


def test_candidate_fixture_rebuilds():
    from tensororder_amd import train_costmodel

    plans, labels, ref, doc = train_costmodel.load_candidates(CANDIDATES)
    assert labels == doc["candidates"] and ref == doc["reference"]
    assert all(len(p.groups_to_slice) == doc["slices"] for p in plans)


@pytest.mark.gpu
def test_trained_model_choice_beats_reference_rule():
    # the committed model picks among the 24 candidate plans of the 4096-slice formula;
    # both picks run in full with the FP32-scaled executor: same count, less time
    from tensororder_amd import train_costmodel
    from tensororder_amd.hip_api import HipF32API

    api = HipF32API()
    model = costmodel.load_default(api)
    plans, labels, ref, doc = train_costmodel.load_candidates(CANDIDATES)
    choice, preds = costmodel.learned_choice(model, api, plans)
    print("reference #%d predicted %.3f s; learned #%d predicted %.3f s" % (ref, preds[ref], choice, preds[choice]),
          flush=True)
    assert preds[choice] <= preds[ref]
    t_ref, v_ref = train_costmodel.measure(api, plans[ref], warmup=0)
    t_new, v_new = train_costmodel.measure(api, plans[choice], warmup=0)
    print("measured: reference %.3f s, learned %.3f s" % (t_ref, t_new), flush=True)
    assert abs(v_new - v_ref) <= 1e-6 * abs(v_ref)
    assert t_new < t_ref
