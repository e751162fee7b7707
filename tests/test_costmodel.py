"""Cost-model math on the CPU: the float64 reference gradient against finite
differences, bf16 rounding, subtree sizes, and the model file round trip
(no call here reaches the GPU)."""
import numpy as np
import pytest

import cuqc_amd as qc
from cuqc_amd import costmodel as cm


def test_layout_matches_library():
    assert cm.NP == qc.NPARAMS == 12 * 64 + 64 + 64 * 64 + 64 + 64 + 1
    assert qc.NNORM == 2 * 12 + 2 + 3


def test_bf16_round_known_values():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e-3], np.float32)
    y = cm.bf16_round(x)
    # halfway cases round to the even mantissa; 1 + 2^-7 is representable
    assert list(y[:4]) == [1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7]
    assert y[4] == -2.5
    assert abs(y[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8
    assert np.all(cm.bf16_round(y) == y)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, cm.NP)
    z = rng.normal(0, 1, (40, cm.NF))
    t = rng.normal(0, 1, 40)
    loss, g = cm.loss_and_grad(z, t, p)
    idx = rng.choice(cm.NP, 200, replace=False)
    idx = np.concatenate([idx, [cm.OB3, cm.OW3, cm.OB2, cm.OB1, cm.OW1]])
    h = 1e-6
    for i in idx:
        e = np.zeros(cm.NP)
        e[i] = h
        fd = (cm.loss_and_grad(z, t, p + e)[0] - cm.loss_and_grad(z, t, p - e)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i
    # the loss is the mean squared error of the forward pass
    y, _ = cm.forward(z, p)
    assert loss == pytest.approx(float(np.mean((y - t) ** 2)), rel=1e-12)


def test_gradient_of_output_bias_is_mean_residual():
    # closed form: dL/db3 = 2 mean(y - t); a planted error term must fail it
    rng = np.random.default_rng(4)
    p = rng.normal(0, 0.2, cm.NP)
    z = rng.normal(0, 1, (17, cm.NF))
    t = rng.normal(0, 1, 17)
    y, _ = cm.forward(z, p)
    _, g = cm.loss_and_grad(z, t, p)
    assert g[cm.OB3] == pytest.approx(2 * np.mean(y - t), rel=1e-12)


def test_subtree_sizes():
    # 0 -> (1 -> 3, 2), 4 is a second root
    parent = np.array([-1, 0, 0, 1, -1])
    assert list(cm.subtree_sizes(parent)) == [4, 2, 1, 1, 1]
    # children past a depth window count as leaves of their parent
    assert list(cm.subtree_sizes(parent, [0, 2, 5, 0, 1])) == [11, 4, 6, 1, 2]


def test_oracle_trace_counts_children_past_the_window():
    import cuqc_amd as qc
    import oracle
    g = qc.Graph.generate(0, 1)
    off, adj = g.csr()
    r = oracle.mine(off, adj, g.gamma, g.min_size, trace=True, max_depth=2)
    assert int(np.sum(r["trace_beyond"])) == r["stats"]["beyond"] > 0
    depth = r["trace_feats"][:, 8]
    assert np.all(r["trace_beyond"][depth < 2] == 0)
    sub = cm.subtree_sizes(r["trace_parent"], r["trace_beyond"])
    assert sub[0] == r["stats"]["tasks"] + r["stats"]["beyond"]


def test_model_file_round_trip(tmp_path):
    m = qc.CostModel.create(seed=5)
    p, nrm = m.params()
    assert p.shape == (qc.NPARAMS,) and nrm.shape == (qc.NNORM,)
    # He initialisation: hidden-layer scale sqrt(2 / fan_in)
    assert 0.3 < np.std(p[cm.OW1:cm.OB1]) / np.sqrt(2 / 12) < 1.7
    nrm2 = nrm.copy()
    nrm2[:12] = np.arange(12)
    m.set_params(norm=nrm2)
    m.save(tmp_path / "m.bin")
    m2 = qc.CostModel.load(tmp_path / "m.bin")
    p2, n2 = m2.params()
    assert np.array_equal(p2, p) and np.array_equal(n2, nrm2)
    (tmp_path / "bad.bin").write_bytes(b"XXXX")
    with pytest.raises(qc.QCError):
        qc.CostModel.load(tmp_path / "bad.bin")
