"""CPU pinning of the oracle's learned-stopping restatement.

The FP64 gradient is checked against central finite differences of the
oracle's own loss; the logit against a network whose output is known in
closed form; the probe-search traces against invariants and against the
oracle's plain search run at the probe's parameters.
"""
import numpy as np

import oracle
from efanna_amd import learned
from tests import helpers

NP = 5313
W1, B1, W2, B2, W3, B3 = 0, 1024, 1088, 5184, 5248, 5312


def test_param_count():
    assert oracle.lib().orc_mlp_param_count() == NP
    assert learned.param_count() == NP


def test_logit_closed_form():
    # z = b3 + W3[0] * relu(W2[0][0] * relu(W1[0][0] * x0)) with BF16-exact values
    w = np.zeros(NP, np.float32)
    w[W1 + 0 * 64 + 0] = 1.0
    w[W2 + 0 * 64 + 0] = 2.0
    w[W3 + 0] = 3.0
    w[B3] = 0.5
    norm = np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    f = np.zeros((3, 16), np.float32)
    f[:, 0] = [1.5, -1.0, 0.25]
    np.testing.assert_array_equal(oracle.mlp_logits(w, norm, f), np.float32([9.5, 0.5, 2.0]))
    # standardisation: (f - mu) * inv_sd
    norm2 = norm.copy()
    norm2[0], norm2[16] = 1.0, 4.0
    np.testing.assert_array_equal(oracle.mlp_logits(w, norm2, f), np.float32([0.5 + 6 * 2.0, 0.5, 0.5]))


def test_logit_rounds_parameters_to_bf16():
    w = np.zeros(NP, np.float32)
    w[W1] = 1.0
    w[W2] = 1.0
    w[W3] = 1.0 + 2.0 ** -10  # not a BF16 value: rounds to 1.0
    norm = np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    f = np.zeros((1, 16), np.float32)
    f[0, 0] = 3.0
    assert oracle.mlp_logits(w, norm, f)[0] == np.float32(3.0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = learned.init_params(1) + rng.normal(0, 0.05, NP).astype(np.float32)
    f = rng.normal(0, 1, size=(6, 16)).astype(np.float32)
    y = np.array([1, 0, 1, 1, 0, 0], np.uint8)
    norm = np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    loss, g = oracle.mlp_grad64(w, norm, f, y, round_params=False)
    assert loss > 0
    idx = np.concatenate([rng.choice(NP, 300, replace=False), [B3, W3, W3 + 63, B2, B1]])
    eps = np.float32(2.0 ** -11)
    ok = 0
    for i in idx:
        a, b = w.copy(), w.copy()
        a[i] += eps
        b[i] -= eps
        la, _ = oracle.mlp_grad64(a, norm, f, y, round_params=False)
        lb, _ = oracle.mlp_grad64(b, norm, f, y, round_params=False)
        fd = (la - lb) / (float(a[i]) - float(b[i]))
        ok += abs(fd - g[i]) <= 1e-3 * abs(g[i]) + 1e-7
    # a step may cross a ReLU kink for a few parameters; every other one agrees
    assert ok >= 0.97 * len(idx)
    # W3 and b3 enter linearly: no kinks, all must agree
    for i in [B3, W3, W3 + 63]:
        a, b = w.copy(), w.copy()
        a[i] += eps
        b[i] -= eps
        fd = (oracle.mlp_grad64(a, norm, f, y, False)[0] - oracle.mlp_grad64(b, norm, f, y, False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-4 * abs(g[i]) + 1e-8


def test_probe_traces_invariants():
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=200, centres=50)
    full, probe, k = (4, 300, 50, 8), (100, 20, 3), 10
    plain, _, _ = oix.search(q, k, *full, 0)
    feat, label = oix.search_probe(q, k, *full, *probe, mode=2, plain=plain)
    assert (label[:, 0] != 255).all()
    reached = label != 255
    # once a checkpoint is missed, the later ones are too
    assert (np.diff(reached.astype(int), axis=1) <= 0).all()
    assert (label[reached] <= k).all()
    e = feat[:, :, 0]
    assert (e[reached] == np.broadcast_to(np.arange(probe[2] + 1), e.shape)[reached]).all()
    assert (feat[:, :, 3][reached] <= 1.0).all() and (feat[:, :, 1][reached] <= feat[:, :, 2][reached]).all()
    assert (feat[:, 0, 12] == probe[0]).all()          # ids measured after the seeds
    assert (feat[:, 0, 14] == k).all()                 # every top-k id is new at checkpoint 0
    assert (feat[:, 0, 13] == 0).all()
    # the last reached checkpoint's label = overlap of the probe search's result with the plain result
    pids, _, _ = oix.search(q, k, probe[2], probe[0], probe[1], full[3], 0)
    last = np.array([row[row != 255][-1] for row in label])
    overlap = np.array([len(set(a.tolist()) & set(b.tolist())) for a, b in zip(pids, plain)])
    np.testing.assert_array_equal(last, overlap)


def test_learned_search_extremes():
    """A network that always stops returns the probe search's results; one
    that never stops returns the full plain search's."""
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=200, centres=50)
    full, probe, k = (4, 300, 50, 8), (100, 20, 3), 10
    norm = np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    w = np.zeros(NP, np.float32)
    w[B3] = 4.0  # logit 4 everywhere
    ids, _, stopped = oix.search_probe(q, k, *full, *probe, mode=1, params=w, norm=norm, zthr=2.0)
    assert stopped.all()
    # stopped at the seeds checkpoint: the probe's seeds-only result
    pids, _, _ = oix.search(q, k, 0, probe[0], probe[1], full[3], 0)
    np.testing.assert_array_equal(ids, pids)
    w[B3] = -4.0
    ids, _, stopped = oix.search_probe(q, k, *full, *probe, mode=1, params=w, norm=norm, zthr=2.0)
    assert not stopped.any()
    fids, _, _ = oix.search(q, k, *full, 0)
    np.testing.assert_array_equal(ids, fids)


def test_continuation_pins():
    """The continuation form (mode 3) against what its definition fixes.

    (a) A network that always stops returns the probe's seeds-only result.
    (b) With the probe's extend_to equal to the search's and no probe epoch,
        a network that never stops returns the plain search's result: the
        continuation takes the next ids of the leaf sequence, so the seeds
        are the plain search's first init_num ids, and the best R of them
        that lie among the probe's seeds are within the probe's best R.
    (c) With probe epochs, a network that never stops continues every query:
        the result is a valid top k (distinct ids, sorted true distances),
        its recall stays within 0.01 of the plain search's, and it differs
        from the restart form (mode 1) on some query, so the continuation
        path is the one taken.
    A planted mistake in the oracle (the continuation skipping one sequence
    position, or dropping the probe's pool) fails (b)."""
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=200, centres=50)
    full, k = (4, 300, 50, 8), 10
    norm = np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    w = np.zeros(NP, np.float32)
    w[B3] = 4.0
    probe = (100, 20, 3)
    ids, _, stopped = oix.search_probe(q, k, *full, *probe, mode=3, params=w, norm=norm, zthr=2.0)
    assert stopped.all()
    pids, _, _ = oix.search(q, k, 0, probe[0], probe[1], full[3], 0)
    np.testing.assert_array_equal(ids, pids)
    w[B3] = -4.0
    fids, fd, _ = oix.search(q, k, *full, 0)
    for pp in (60, 150, 300):
        ids, d, stopped = oix.search_probe(q, k, *full, pp, full[2], 0, mode=3, params=w, norm=norm, zthr=2.0)
        assert not stopped.any()
        np.testing.assert_array_equal(ids, fids)
        np.testing.assert_array_equal(d.view(np.uint32), fd.view(np.uint32))
    # no epoch and k = extend_to: the whole pool, the best 50 of the first 300 seeds
    seeds0 = (0, 300, 50, 8)
    sids, _, _ = oix.search(q, 50, *seeds0, 0)
    for pp in (60, 150):
        ids, _, _ = oix.search_probe(q, 50, *seeds0, pp, 50, 0, mode=3, params=w, norm=norm, zthr=2.0)
        np.testing.assert_array_equal(ids, sids)
    ids, d, stopped = oix.search_probe(q, k, *full, *probe, mode=3, params=w, norm=norm, zthr=2.0)
    assert not stopped.any()
    for row, dr, qq in zip(ids, d, q):
        assert len(set(row.tolist())) == k
        assert (np.diff(dr) >= 0).all()
        true = np.array([oracle.l2(qq, base[i]) for i in row], np.float32)
        np.testing.assert_array_equal(dr.view(np.uint32), true.view(np.uint32))
    gt, _ = oracle.exact_knn(base, q, k)
    assert helpers.recall(ids, gt, k) >= helpers.recall(fids, gt, k) - 0.01
    rids, _, _ = oix.search_probe(q, k, *full, *probe, mode=1, params=w, norm=norm, zthr=2.0)
    np.testing.assert_array_equal(rids, fids)
    assert (ids != rids).any(axis=1).sum() > 0


def test_continuation_refuses_probe_outside_search():
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=10, centres=50)
    w = np.zeros(NP, np.float32)
    norm = np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    for probe in ((400, 20, 2), (100, 60, 2), (5, 5, 2)):
        try:
            oix.search_probe(q, 10, 4, 300, 50, 8, *probe, mode=3, params=w, norm=norm, zthr=0.0)
        except oracle.OracleError:
            continue
        raise AssertionError(f"probe {probe} accepted")
