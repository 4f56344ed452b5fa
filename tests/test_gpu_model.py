"""The learned configuration model on the GPU.

* The FP32 MLP's hand-written inference and gradient kernels against a plain
  PyTorch float64 restatement of the same 8-64-64-1 network (tolerances are
  written in each test; the kernels accumulate in FP32, the gradient with
  float atomics in an unspecified order).
* Training reduces the loss.
* The calibration sweep: measured lookup times are positive, measured
  false-positive rates fall with bits per key and agree with the reference's
  FPR formula (fpr.hpp:11-60, restated in the oracle) for plain configurations.
* The exhaustive choice is the argmin of the table over every candidate; the
  learned choice is a valid configuration whose table cost is within a few
  percent of it; a saved model reloads to the same choices.
"""
import os

import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dtl_hip as D  # noqa: E402

KIN, KH = 8, 64


def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    D.load()
    return "cuda"


def unpack(p):
    p = torch.from_numpy(p.astype(np.float64))
    o = 0
    W1 = p[o:o + KH * KIN].reshape(KH, KIN); o += KH * KIN
    b1 = p[o:o + KH]; o += KH
    W2 = p[o:o + KH * KH].reshape(KH, KH); o += KH * KH
    b2 = p[o:o + KH]; o += KH
    W3 = p[o:o + KH]; o += KH
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def torch_forward(params, x):
    W1, b1, W2, b2, W3, b3 = params
    h1 = torch.relu(x @ W1.T + b1)
    h2 = torch.relu(h1 @ W2.T + b2)
    return h2 @ W3 + b3


def test_mlp_forward_matches_torch():
    d = dev()
    mlp = D.MLP(seed=3)
    assert mlp.n_params == KH * KIN + KH + KH * KH + KH + KH + 1
    p = mlp.params()
    rng = np.random.default_rng(5)
    p[KH * KIN:KH * KIN + KH] = rng.normal(0, 0.1, KH)          # nonzero biases
    p[-1] = 0.25
    mlp.set_params(p)
    assert np.array_equal(mlp.params(), p)
    x = rng.normal(0, 1, (1000, KIN)).astype(np.float32)
    y = mlp.predict(torch.from_numpy(x).to(d)).cpu().numpy()
    ref = torch_forward(unpack(p), torch.from_numpy(x.astype(np.float64))).numpy()
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-5)
    mlp.close()


@pytest.mark.parametrize("n,weighted", [(1, False), (63, False), (300, False), (4096, False), (300, True),
                                        (4096, True)])
def test_mlp_gradient_matches_autograd(n, weighted):
    d = dev()
    mlp = D.MLP(seed=11)
    p = mlp.params()
    rng = np.random.default_rng(n)
    p[KH * KIN:KH * KIN + KH] = rng.normal(0, 0.1, KH)
    mlp.set_params(p)
    x = rng.normal(0, 1, (n, KIN)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    w = rng.uniform(0.02, 3.0, n).astype(np.float32) if weighted else np.ones(n, np.float32)
    wd = torch.from_numpy(w).to(d) if weighted else None
    g = mlp.gradient(torch.from_numpy(x).to(d), torch.from_numpy(t).to(d), wd).cpu().numpy()
    pt = torch.from_numpy(p.astype(np.float64)).requires_grad_(True)
    o = 0
    parts = []
    for size in (KH * KIN, KH, KH * KH, KH, KH, 1):
        parts.append(pt[o:o + size]); o += size
    params = (parts[0].reshape(KH, KIN), parts[1], parts[2].reshape(KH, KH), parts[3], parts[4], parts[5][0])
    y = torch_forward(params, torch.from_numpy(x.astype(np.float64)))
    loss = 0.5 * (torch.from_numpy(w.astype(np.float64)) * (y - torch.from_numpy(t.astype(np.float64))) ** 2).mean()
    loss.backward()
    ref = pt.grad.numpy()
    # FP32 accumulation with atomics: absolute error relative to the gradient's scale
    scale = np.abs(ref).max()
    np.testing.assert_allclose(g, ref, rtol=0, atol=2e-4 * scale + 1e-7)
    assert np.array_equal(mlp.params(), p)  # the gradient call leaves the weights alone
    mlp.close()


def test_mlp_training_reduces_loss():
    d = dev()
    rng = np.random.default_rng(2)
    n = 8192
    x = rng.uniform(-1, 1, (n, KIN)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2] - 0.5 * np.abs(x[:, 3])).astype(np.float32)
    xd, td = torch.from_numpy(x).to(d), torch.from_numpy(t).to(d)
    mlp = D.MLP(seed=1)
    before = float(((mlp.predict(xd) - td) ** 2).mean())
    mlp.train(xd, td, epochs=30, batch=256, lr=2e-3, seed=4)
    after = float(((mlp.predict(xd) - td) ** 2).mean())
    assert after < 0.05 * before, (before, after)
    mlp.close()


@pytest.fixture(scope="module")
def small_model():
    dev()
    p = D.CalibParams(log2_m_min=13, log2_m_max=20, probe_keys_log2=24, reps=3, fpr_keys_log2=20, flags=0)
    m = D.CostModel.calibrate(p)
    yield m
    m.close()


def test_model_configs_are_the_reference_space(small_model):
    M = small_model
    cfgs = [M.config(i).key() for i in range(M.config_count())]
    assert len(cfgs) == len(set(cfgs)) == 279
    for wb, w, s, z, k in cfgs:
        assert wb // 8 * w <= 64 and (z == 1 or s == w)   # model/util.hpp:16-63
        D.bbf_desc(wb, 1 << 20, k, w, s, z)                 # every candidate is compiled


def test_calibration_tables(small_model):
    M = small_model
    for i in range(M.config_count()):
        c = M.config(i)
        for lm in range(13, 21):
            assert 0 < M.lookup_ns(i, 2.0 ** lm) < 50
        f4, f24 = M.fpr(i, 4.0), M.fpr(i, 24.0)
        assert 0 < f24 < f4 <= 1, c.as_dict()
        if c.z == 1 and c.k <= 8:
            # the reference's formula for word/sector-blocked filters; the measured
            # rate over 2^20 absent keys must agree within sampling + model error
            for bpk in (8.0, 12.0):
                n = 1 << 16
                m = int(bpk * n) + 1
                desc = D.bbf_desc(c.word_bits, m, c.k, c.w, c.s, c.z)
                model = O.fpr_config(desc.m_actual, n, c.k, c.word_bits // 8, c.w, c.s, c.z)
                meas = M.fpr(i, bpk)
                if model * (1 << 20) > 200:
                    assert 0.5 * model < meas < 2.0 * model, (c.as_dict(), bpk, model, meas)


def candidate_ms(n, c, lo=2.0 ** 13, hi=2.0 ** 20):
    out = []
    for b in range(4, 21):
        m = min(hi, max(lo, 2.0 * c.word_bits * c.w + 1, float(n) * b))
        if not out or out[-1] != m:
            out.append(m)
    return out


@pytest.mark.parametrize("n,tw", [(1 << 10, 1.0), (1 << 12, 30.0), (1 << 14, 1000.0), (5000, 7.0)])
def test_exhaustive_choice_is_table_argmin(small_model, n, tw):
    M = small_model
    ch = M.choose(n, tw, use_mlp=False)
    best = np.inf
    for i in range(M.config_count()):
        c = M.config(i)
        for m in candidate_ms(n, c):
            c.m = int(m)
            best = min(best, M.cost(c, n, tw))
    assert ch.cost_ns == pytest.approx(best, rel=1e-12)
    assert M.cost(ch, n, tw) == pytest.approx(ch.cost_ns, rel=1e-12)
    D.bbf_desc(ch.word_bits, ch.m, ch.k, ch.w, ch.s, ch.z)


def test_learned_choice_and_persistence(small_model, tmp_path):
    M = small_model
    M.train(train_fraction=1.0, samples=400000, epochs=16, seed=3)
    regrets, rand = [], []
    rng = np.random.default_rng(0)
    for ln in range(10, 17):
        for tw in (1.0, 10.0, 100.0, 1000.0):
            ml = M.choose(1 << ln, tw, use_mlp=True)
            ex = M.choose(1 << ln, tw, use_mlp=False)
            D.bbf_desc(ml.word_bits, ml.m, ml.k, ml.w, ml.s, ml.z)
            assert ml.cost_ns >= ex.cost_ns * (1 - 1e-12)
            regrets.append(ml.cost_ns / ex.cost_ns - 1)
            c = M.config(int(rng.integers(0, M.config_count())))
            c.m = ex.m
            rand.append(M.cost(c, 1 << ln, tw) / ex.cost_ns - 1)
    # per-query regret of the learned choice, measured against the calibrated table:
    # far below that of a random configuration (this small calibration grid,
    # 2^13..2^20 bits, is noisier than the full one the benchmark reports)
    assert np.mean(regrets) < 0.25 * np.mean(rand), (np.mean(regrets), np.mean(rand))
    assert np.median(regrets) < 0.10, regrets
    path = str(tmp_path / "model.bin")
    M.save(path)
    L = D.CostModel.load_file(path)
    assert L.calibration_seconds() == M.calibration_seconds()
    for ln, tw in ((11, 3.0), (15, 300.0)):
        for use in (True, False):
            a, b = M.choose(1 << ln, tw, use), L.choose(1 << ln, tw, use)
            assert a.key() == b.key() and a.m == b.m and a.cost_ns == b.cost_ns
    L.close()
    os.remove(path)


def regret_stats(M, lns, tws_cycles, clock_ghz=2.4):
    r = []
    for ln in lns:
        for tc in tws_cycles:
            tw = tc / clock_ghz
            ml = M.choose(1 << ln, tw, use_mlp=True)
            ex = M.choose(1 << ln, tw, use_mlp=False)
            assert ml.cost_ns >= ex.cost_ns * (1 - 1e-12)
            r.append(ml.cost_ns / ex.cost_ns - 1)
    r = np.array(r)
    return float(r.mean()), float(np.percentile(r, 95)), float((r <= 0.05).mean())


def test_learned_model_full_calibration_regret():
    """The full calibration sweep (279 configurations x 2^13..2^33 bits) and the
    default training run, then the learned choice against the exhaustive
    optimum of the same calibrated table: on the query grid the benchmark
    reports (n = 2^10..2^28 in steps of 4x, tw = 1..1000 GPU cycles at 2.4 GHz)
    and on a grid between its points (odd powers of two, tw between the
    benchmark's values), where neither n nor tw was a query of the model's
    selection.  The learned choice is priced by the table (the reference's
    cost function, cost_fn.hpp:143-166: lookup time + tw x false positives),
    so the regret is a property of the model, not of timing noise at query
    time."""
    dev()
    M = D.CostModel.calibrate()
    M.train()
    on = regret_stats(M, range(10, 29, 2), (1, 2, 5, 10, 20, 50, 100, 200, 500, 1000))
    off = regret_stats(M, range(11, 28, 2), (1.5, 3, 7, 15, 30, 70, 150, 300, 700))
    print("regret on grid: mean %.4f p95 %.4f within5 %.3f; between: mean %.4f p95 %.4f within5 %.3f"
          % (on + off), flush=True)
    assert on[0] <= 0.02 and on[1] <= 0.06 and on[2] >= 0.9, on
    # between the grid points the calibration is interpolated: a looser bound
    assert off[0] <= 0.03 and off[1] <= 0.10 and off[2] >= 0.85, off
    M.close()
