"""Cost model on the GPU: gradient and forward kernels against the float64
reference, training, and cost-ordered scheduling (the output must equal the
FIFO run and the oracle; only the order of work may change)."""
import sys

import numpy as np
import pytest

import cuqc_amd as qc
import oracle
from cuqc_amd import costmodel as cm
from tests.graphs import SMALL_CASES

pytestmark = pytest.mark.gpu

# fp32 accumulation against float64 with identical bf16 weights:
# |g_gpu - g_ref| <= GRAD_RTOL * max|g_ref| + GRAD_ATOL
GRAD_RTOL, GRAD_ATOL = 2e-3, 1e-6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _synthetic(n, seed):
    rng = np.random.default_rng(seed)
    X = np.exp(rng.normal(2.0, 1.0, (n, cm.NF))).astype(np.float32)
    Z = np.log1p(X.astype(np.float64))
    y = (0.8 * Z[:, 1] - 0.5 * Z[:, 0] + 0.3 * Z[:, 9] * Z[:, 4] / 3 + 0.1 * rng.normal(size=n)).astype(np.float32)
    return X, y


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 5000])
def test_gradient_kernel_matches_reference(n):
    _say(f"[gpu] mlp grad n={n}")
    X, y = _synthetic(n, n)
    m = qc.CostModel.create(seed=n)
    p, nrm = m.params()
    g, loss = m.grad(X, y)
    z = cm.standardize(X, nrm)
    t = (y.astype(np.float64) - nrm[24]) / nrm[25]
    ref_loss, ref_g = cm.loss_and_grad(z, t, cm.bf16_round(p).astype(np.float64))
    tol = GRAD_RTOL * np.max(np.abs(ref_g)) + GRAD_ATOL
    assert np.max(np.abs(g - ref_g)) <= tol
    assert loss == pytest.approx(ref_loss, rel=1e-3)


def test_forward_kernel_matches_reference():
    _say("[gpu] mlp forward")
    X, _ = _synthetic(3000, 7)
    m = qc.CostModel.create(seed=2)
    p, nrm = m.params()
    nrm[:12] = np.log1p(X).mean(0)
    nrm[12:24] = np.log1p(X).std(0)
    nrm[24], nrm[25] = 1.5, 2.0
    m.set_params(norm=nrm)
    out = m.predict(X)
    y, _ = cm.forward(cm.standardize(X, nrm), cm.bf16_round(p).astype(np.float64))
    ref = y * 2.0 + 1.5
    assert np.max(np.abs(out - ref)) <= 1e-3 * max(1.0, np.max(np.abs(ref)))


def test_training_fits_a_smooth_target():
    _say("[gpu] mlp train")
    X, y = _synthetic(4000, 11)
    m = qc.CostModel.train(X[:3000], y[:3000], steps=400, lr=3e-3, seed=1)
    pred = m.predict(X[3000:])
    resid = np.mean((pred - y[3000:]) ** 2)
    assert resid < 0.25 * np.var(y[3000:])
    # the trained fp32 master weights agree with their bf16 copy in use
    p, _ = m.params()
    _, nrm = m.params()
    ref, _ = cm.forward(cm.standardize(X[3000:], nrm), cm.bf16_round(p).astype(np.float64))
    assert np.max(np.abs(pred - (ref * nrm[25] + nrm[24]))) <= 1e-3 * max(1.0, np.max(np.abs(pred)))


@pytest.fixture(scope="module")
def cfg0_model():
    g = qc.Graph.generate(0, 1)
    off, adj = g.csr()
    ref = oracle.mine(off, adj, g.gamma, g.min_size, trace=True)
    sub = cm.subtree_sizes(np.asarray(ref["trace_parent"]))
    model = qc.CostModel.train(ref["trace_feats"], np.log(sub.astype(np.float32)), steps=300, seed=0)
    return g, oracle.mine(off, adj, g.gamma, g.min_size), model, ref, sub


def test_cost_model_ranks_subtrees(cfg0_model):
    g, _, model, ref, sub = cfg0_model
    pred = model.predict(ref["trace_feats"])
    rank = lambda a: np.argsort(np.argsort(a, kind="stable"), kind="stable")
    rho = np.corrcoef(rank(pred), rank(np.log(sub)))[0, 1]
    _say(f"[gpu] cost model spearman {rho:.3f} over {len(sub)} tasks")
    assert rho > 0.3


def _compare(res, ref):
    for k in ("tasks", "branches", "children", "raw_cliques", "w_adj", "sum_t", "sum_t_written"):
        assert res.stats[k] == ref["stats"][k], k
    assert oracle.canon(res.raw) == oracle.canon(ref["raw"])
    assert res.maximal_sorted_tuples() == oracle.nonmax(ref["raw"])


@pytest.mark.parametrize("chunk", [-1, 4])
def test_cost_order_matches_oracle(cfg0_model, chunk):
    g, ref, model, _, _ = cfg0_model
    _say(f"[gpu] cost order config0 chunk={chunk}")
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", order="cost", cost_model=model, split_chunk=chunk)
    _compare(res, ref)


@pytest.mark.parametrize("name,build,gamma,ms", SMALL_CASES, ids=[c[0] for c in SMALL_CASES])
def test_cost_order_small(name, build, gamma, ms):
    off, adj = build()
    ref = oracle.mine(off, adj, gamma, ms)
    g = qc.Graph.from_csr(off, adj)
    model = qc.CostModel.create(seed=len(name))
    _say(f"[gpu] cost order {name}")
    res = qc.mine(g, gamma, ms, 0, mode="persistent", order="cost", cost_model=model, lds_tile=64)
    _compare(res, ref)


def test_cost_order_requires_model_and_fifo_trace():
    g = qc.Graph.generate(0, 1)
    with pytest.raises(qc.QCError):
        qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", order="cost")
    m = qc.CostModel.create(0)
    with pytest.raises(qc.QCError):
        qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", order="cost", cost_model=m, trace=True)


@pytest.mark.parametrize("kind", ["nan", "huge", "negated"])
def test_cost_order_with_corrupted_model(cfg0_model, kind):
    """The model only orders work: a corrupted one changes nothing."""
    g, ref, model, _, _ = cfg0_model
    p, nrm = model.params()
    bad = qc.CostModel.create(seed=9)
    if kind == "nan":
        bad.set_params(np.full_like(p, np.nan), nrm)
    elif kind == "huge":
        bad.set_params(np.full_like(p, 3.0e38), nrm)
    else:
        bad.set_params(-p, nrm)
    _say(f"[gpu] cost order corrupted model {kind}")
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", order="cost", cost_model=bad, split_chunk=4)
    _compare(res, ref)
