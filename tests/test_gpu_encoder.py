"""Amortization encoder on the GPU (gp_enc_*, through the C ABI) against the
CPU oracle oracle/encoder.py.

The oracle rounds to BF16 where the GPU does; what remains is the order and
width of accumulation (FP32 MFMA sums vs float64) and, rarely, a hidden
activation that lands on the other side of a BF16 rounding boundary.  The
tolerances below are set for that: about 1e-6 relative on FP32-accumulated
values, a few 1e-4 on norms of quantities that pass through BF16 roundings.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

F, Q, H1, H2, BATCH = 200, 5, 128, 64, 128


def _data(N=300, seed=0):
    from gaspacho2_amd import synth
    return synth.make_dataset(N, F, Q, 8, seed=seed, model_seed=7)


def _nrel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _encoder(ds, X, seed=1):
    from gaspacho2_amd import lib
    enc = lib.Encoder(F, Q, H1=H1, H2=H2, batch=BATCH, capacity=ds.N, seed=seed)
    enc.add(ds.N, ds.Yp, ds.Yi, ds.Yx, ds.ls, X)
    return enc


def test_encoder_loss_and_gradient_match_oracle():
    from gaspacho2_amd.synth import dense_counts
    from oracle import encoder as E
    ds = _data()
    X = np.random.default_rng(5).normal(0, 1, (ds.N, Q))
    enc = _encoder(ds, X)
    for n in (100, BATCH):
        loss = enc.loss_grad(n)
        params, grad = enc.get("params"), enc.get("grad")
        p = E.unpack(params, F, H1, H2, Q)
        A0 = E.features(dense_counts(ds)[:n], ds.ls[:n])
        lo, go = E.loss_grad(p, A0, X[:n])
        print(f"n={n}: loss gpu {loss:.8f} oracle {lo:.8f}")
        assert abs(loss - lo) <= 1e-5 * abs(lo)
        gg = E.unpack(grad, F, H1, H2, Q)
        full = enc.unpack(grad)
        # padded rows and columns of the parameter layout get exactly zero gradient
        assert not np.any(full["W1"][:, F:]) and not np.any(full["W3"][2 * Q:]) and not np.any(full["b3"][2 * Q:])
        for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
            r = _nrel(gg[k], go[k])
            print(f"  grad {k}: norm rel err {r:.2e}")
            assert r < 2e-3, (k, r)


def test_encoder_predict_matches_oracle_forward():
    from gaspacho2_amd.synth import dense_counts
    from oracle import encoder as E
    ds = _data(N=333, seed=2)            # not a multiple of the batch
    enc = _encoder(ds, np.zeros((ds.N, Q)))
    mu, lv = enc.predict(ds.N, ds.Yp, ds.Yi, ds.Yx, ds.ls)
    p = E.unpack(enc.get("params"), F, H1, H2, Q)
    z3 = E.forward(p, E.features(dense_counts(ds), ds.ls))["z3"]
    r_mu, r_lv = _nrel(mu, z3[:, :Q]), _nrel(lv, z3[:, Q:])
    print(f"predict: mu norm rel err {r_mu:.2e}, logvar {r_lv:.2e}, max {rel(mu, z3[:, :Q]):.2e}")
    assert r_mu < 1e-3 and r_lv < 1e-3
    assert rel(mu, z3[:, :Q]) < 1e-2


def test_encoder_adam_step_matches_oracle():
    from oracle import encoder as E
    ds = _data()
    X = np.random.default_rng(6).normal(0, 1, (ds.N, Q))
    enc = _encoder(ds, X)
    w = enc.get("params").astype(np.float64)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    for t in (1, 2, 3):
        enc.loss_grad(BATCH)
        g = enc.get("grad").astype(np.float64)
        enc.adam_step(1e-3)
        w, m, v = E.adam(w, g, m, v, t, 1e-3)
        wg = enc.get("params")
        err = np.max(np.abs(wg - w))
        print(f"adam step {t}: max abs err {err:.2e}")
        assert err < 2e-6 * max(1.0, np.max(np.abs(w)))
        assert np.max(np.abs(enc.get("adam_m") - m)) <= 1e-6 * max(np.max(np.abs(m)), 1e-30)
        w = wg.astype(np.float64)      # continue from the GPU's FP32 state


def test_encoder_training_learns_the_latent_map():
    """Trained on the true latents of 4,000 cells (Q = 3, gene functions on 32
    centres), the encoder predicts the latents of 1,000 fresh cells (same gene
    functions, new seed) much better than their mean does (a float32 torch MLP
    of the same shape reaches R^2 0.73 with a squared loss here), and its
    training loss falls."""
    from gaspacho2_amd import lib, synth
    q = 3
    tr = synth.make_dataset(4000, F, q, 8, seed=11, model_seed=7, gen_M=32, pca=False)
    te = synth.make_dataset(1000, F, q, 8, seed=12, model_seed=7, gen_M=32, pca=False)
    enc = lib.Encoder(F, q, H1=H1, H2=H2, batch=BATCH, capacity=tr.N, seed=3)
    enc.add(tr.N, tr.Yp, tr.Yi, tr.Yx, tr.ls, tr.X_true)
    losses = enc.train(20, lr=1e-3)
    mu, _ = enc.predict(te.N, te.Yp, te.Yi, te.Yx, te.ls)
    r2 = 1.0 - np.sum((mu - te.X_true) ** 2) / np.sum((te.X_true - te.X_true.mean(0)) ** 2)
    print(f"epoch losses {losses[0]:.3f} -> {losses[-1]:.3f}; held-out R^2 {r2:.3f}")
    assert losses[-1] < losses[0] - 1.0
    assert r2 > 0.5


def test_encoder_warm_start_gain_over_pca_start():
    """What the trained encoder gains over the PCA start (small data sets:
    N = 2,000, G = 200, Q = 5, M = 32; 8 training seeds sharing the gene
    functions, one fresh seed; mixed-precision fits, one outer iteration of at
    most 100 L-BFGS steps).

    Asserted (a property of the starting point, stable under rounding-level
    changes): with the "aligned" target the warm start has a higher objective
    after the quasi-likelihood start than the PCA latent (measured +1.8e-3
    relative), and the warm-started fit does not end materially below the
    PCA-started one.  Printed, not asserted (the L-BFGS path amplifies
    rounding differences): the final objectives and the step at which the
    warm-started fit passes the PCA-started final objective (measured: 47 or
    95 of 100 steps depending on last-bit differences elsewhere in the fit).
    At N = 100,000, G = 2,000, M = 256 the residual target is the one that
    gains (30 of 100 steps saved), which `bench.py --workload warmstart`
    measures."""
    from gaspacho2_amd import lib, synth, warmstart
    data = [synth.make_dataset(2000, 200, 5, 32, seed=s, model_seed=1000) for s in range(9)]
    T0 = lib.init_ta_with_per(32, 5)
    train = [(d, warmstart.fit_from(d, d.X0, T0, 1, 100)[0]) for d in data[:8]]
    ds = data[8]
    _, obj_p, _, tr_p = warmstart.fit_from(ds, ds.X0, T0, 1, 100)
    start = lambda tr: float(tr[0][list(tr[2]).index(1)])
    res = {}
    for target in warmstart.TARGETS:
        enc, losses = warmstart.train_encoder(train, target, epochs=100)
        X0 = warmstart.warm_start(enc, ds, target)
        enc.close()
        _, obj_e, _, tr_e = warmstart.fit_from(ds, X0, T0, 1, 100)
        n_hit, _ = warmstart.steps_to_reach(tr_e, obj_p)
        res[target] = (obj_e, start(tr_e), losses)
        print(f"{target}: loss {losses[0]:.3f} -> {losses[-1]:.3f}; start {start(tr_e):.8e} vs PCA "
              f"{start(tr_p):.8e}; final {obj_e:.8e} vs PCA {obj_p:.8e} ({(obj_e - obj_p) / abs(obj_p):+.2e}); "
              f"passes the PCA final after {n_hit} steps")
    obj_e, st_e, losses = res["aligned"]
    assert losses[-1] < losses[0] - 3.0
    assert st_e > start(tr_p)
    assert obj_e > obj_p - 1e-3 * abs(obj_p)
