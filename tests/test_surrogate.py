"""The learned SCHEFF field surrogate.

* the float64 reference network (oracle/nn_ref.py) is pinned by finite
  differences;
* the GPU network (BF16 weights, FP32 accumulation, the library's own
  forward and backward kernels) matches it, forward and gradients;
* trained on (charge, field) pairs the engine's grouped SCHEFF solve emits
  while tracking, it reproduces held-out fields."""
import numpy as np
import pytest

from oracle import nn_ref as R


def _table(cin, cout, widths, nbot):
    """The layer table of hps_surrogate.hip's architecture."""
    C0, C1, C2, C3 = widths
    rows, off = [], [0]

    def add(mode, ci, co, dil=1, skip=-1):
        t = R.TAPS[mode]
        w = off[0]
        a = w + co * t * ci
        off[0] = a + co * 3
        rows.append((mode, ci, co, t, dil, skip, w, a))
        return len(rows) - 1

    add(R.POINT, cin, C0)
    e0 = add(R.SAME, C0, C0)
    add(R.DOWN, C0, C1)
    e1 = add(R.SAME, C1, C1)
    add(R.DOWN, C1, C2)
    e2 = add(R.SAME, C2, C2)
    add(R.DOWN, C2, C3)
    for k in range(nbot):
        add(R.SAME, C3, C3, dil=1 << k)
    add(R.UP, C3, C2, skip=e2)
    add(R.SAME, C2, C2)
    add(R.UP, C2, C1, skip=e1)
    add(R.SAME, C1, C1)
    add(R.UP, C1, C0, skip=e0)
    add(R.SAME, C0, C0)
    add(R.POINT, C0, cout)
    return rows, off[0]


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    table, n = _table(3, 5, (4, 5, 6, 7), 2)
    p = rng.normal(0, 0.4, n)
    x = rng.normal(0, 1, (2, 16, 3))
    u = rng.normal(0, 0.5, 2)
    y = rng.normal(0, 1, (2, 13, 5))
    _, g = R.loss_and_grad(p, table, x, u, y, 13, 5)
    for i in rng.choice(n, 60, replace=False):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.loss_and_grad(pp, table, x, u, y, 13, 5)[0].mean() -
              R.loss_and_grad(pm, table, x, u, y, 13, 5)[0].mean()) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_reference_is_linear_in_the_charge():
    """No bias, no activation: the network's output is linear in its input
    for a fixed conditioning, as the SCHEFF field is in the charge."""
    rng = np.random.default_rng(4)
    table, n = _table(3, 5, (4, 5, 6, 7), 2)
    p = rng.normal(0, 0.4, n)
    x1, x2 = rng.normal(0, 1, (2, 2, 16, 3))
    u = np.array([0.3, -0.7])
    f = lambda x: R.forward(p, table, x, u)
    assert np.allclose(f(2.0 * x1 - 3.0 * x2), 2.0 * f(x1) - 3.0 * f(x2), rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
def test_gpu_network_matches_reference_forward_and_gradient():
    from hpsim_amd.surrogate import Network
    widths, nbot = (16, 32, 32, 48), 2
    net = Network(nr=16, nz=64, widths=widths, nbot=nbot, max_batch=4, seed=5)
    table = net.layers()
    ref_table, n = _table(16, net.cpad, widths, nbot)
    assert table == [tuple(r) for r in ref_table] and net.num_params == n
    net.set_norm(1.0, 0.0, 1.0)
    rng = np.random.default_rng(8)
    p = net.params()
    film = np.zeros(n, bool)
    for r in table:
        film[r[7]:r[7] + 3 * r[2]] = True
    p[film] = rng.normal(0, 0.2, film.sum()).astype(np.float32)
    net.set_params(p)
    B = 3
    x = rng.normal(0, 1, (B, 64, 16)).astype(np.float32)
    u = rng.normal(0, 0.6, B).astype(np.float32)
    y = rng.normal(0, 1, (B, 65, 34)).astype(np.float32)
    # the reference with the GPU's BF16 weights (FiLM coefficients stay FP32)
    pr = p.astype(np.float64)
    pr[~film] = R.bf16(p[~film])
    xr = np.zeros((B, net.Z0, 16))
    xr[:, :64] = x
    out = R.forward(pr, table, xr, u.astype(np.float64))
    got = net.predict(x, u)
    want = out[:, :65, :34]
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-4
    loss, g = net.loss_grad(x, u, y)
    rl, rg = R.loss_and_grad(pr, table, xr, u.astype(np.float64), y.astype(np.float64), 65, 34)
    assert np.allclose(loss, rl, rtol=1e-4)
    cos = float(np.dot(g, rg) / (np.linalg.norm(g) * np.linalg.norm(rg)))
    assert cos > 0.9999, cos
    assert np.linalg.norm(g - rg) / np.linalg.norm(rg) < 1e-2
    for r in table:   # every layer's gradient block, weights and FiLM
        w0, a0, co, taps, ci = r[6], r[7], r[2], r[3], r[1]
        for lo, hi in ((w0, a0), (a0, a0 + 3 * co)):
            e = np.linalg.norm(g[lo:hi] - rg[lo:hi]) / max(np.linalg.norm(rg[lo:hi]), 1e-30)
            assert e < 2e-2, (r, e)


def _train_small(steps=20000):
    """A small network trained on the grouped SCHEFF solves of two short
    tracking runs (scan members and beams of the accelerating lattice)."""
    from hpsim_amd import configs as cfg
    from hpsim_amd import surrogate as S
    wl = S.training_workloads(seeds=(201,), nbeams=24, ramp_beams=4)
    wl[0] = cfg.batched_scan(nbeams=24, nmin=16384, nmax=131072, periods=12, seed=201)
    net = S.Network(nr=32, nz=128, widths=(32, 48, 64, 64), nbot=4, max_batch=32, seed=3)
    net.reserve(40000)
    S.collect(net, wl, log=print)
    net.shuffle(1)
    n = net.size()
    hold = n // 10
    net.fit_norm(0, n - hold)
    before = net.eval(n - hold, hold)
    losses = net.train(steps, batch=32, lr=1e-3, lr_end=1e-5, seed=1, first=0, count=n - hold)
    after = net.eval(n - hold, hold)
    return net, n, before, after, losses


@pytest.mark.gpu
def test_training_on_the_engine_fields_reproduces_held_out_fields():
    """Pairs captured from the engine's own grouped SCHEFF solve while
    tracking; Adam with the library's kernels; the held-out field error
    falls from the untrained network's to a few per cent."""
    net, n, before, after, losses = _train_small()
    print(f"{n} pairs, held-out relative L2 error: untrained median {np.median(before):.3f}, "
          f"trained median {np.median(after):.4f}, 95th pct {np.percentile(after, 95):.4f}",
          flush=True)
    assert n > 2000
    assert np.median(before) > 0.5
    assert losses[-50:].mean() < 0.1 * losses[:50].mean()
    assert np.median(after) < 0.05
