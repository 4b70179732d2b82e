"""The workloads at their stated sizes on the GPU: the N=100,000 / Q=10 /
M=256 ELBO + gradient against the FP64 oracle on a gene sample and the
mixed path against the FP64 path on all 2,000 genes, the SE-ARD kernel at
the N=2,000 / G=200 / M=32 size, the batched uneven-subset fits at
N=200,000 / G=20,000 / M=128, and the 2000-512-256-20 encoder."""
import numpy as np
import pytest

from conftest import make_oracle, rel

pytestmark = pytest.mark.gpu


def _ctx(ds, T0, ard_only=True, precision=0):
    from gaspacho2_amd import lib
    return lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0,
                         T0, ard_only=ard_only, precision=precision)


def _blocks(N, p, Q, M):
    return {"gOmega": slice(0, N), "gdelta": slice(N, N + p), "gX": slice(N + p, N + p + N * Q),
            "gT": slice(N + p + N * Q, N + p + N * Q + M * Q),
            "grho": slice(N + p + N * Q + M * Q, N + p + N * Q + M * Q + Q)}


def test_n100k_m256_gene_sample_matches_oracle():
    """N=100,000 cells, Q=10, M=256 (P=257), SE-ARD, 8 genes: two IRLS steps,
    the quasi-likelihood start and one gradient on the FP64 path against the
    oracle (objectives 1e-11, arrays 1e-9; the tau / rho gradient blocks and
    C2, sums over all 100,000 cells, 2e-8), then the mixed ELBO within 1e-6
    and the mixed gradient within 1e-4 of the oracle's."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(100000, 8, 10, 256, seed=0, pca=False)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    o = make_oracle(ds, T0=T0, ard_only=True)
    g = _ctx(ds, T0)
    o.init_fit()
    g.init_fit()
    for name in ("K", "KinvKmn", "titsias", "Beta"):
        assert rel(g.get(name), o.get(name)) < 1e-9, name
    for jtr in range(2):
        a, b = o.irls_step(jtr), g.irls_step(jtr)
        assert abs(a - b) <= 1e-11 * abs(a), (jtr, a, b)
    for name in ("Beta", "phi"):
        assert rel(g.get(name), o.get(name)) < 1e-9, name
    a, b = o.ql_start(), g.ql_start()
    assert abs(a - b) <= 1e-11 * abs(a), (a, b)
    o.grad()
    g.grad()
    go, gg = o.get("grad"), g.get("grad")
    errs = {k: rel(gg[s], go[s]) for k, s in _blocks(ds.N, len(ds.nh) - 1, ds.Q, ds.M).items()}
    print(f"N=100K M=256, 8 genes, FP64 gradient blocks: {errs}")
    for k, r in errs.items():
        # gT and grho sum C1-weighted terms over 100,000 cells: the two
        # summation orders part by ~6e-9 of the largest entry at this N
        assert r < (2e-8 if k in ("gT", "grho") else 1e-9), k
    assert rel(g.get("C2"), o.get("C2")) < 2e-8            # also a sum over the 100,000 cells
    eo = o.elbo(1)
    g.set_precision(lib.PREC_MIXED)
    em = g.elbo(1)
    o.grad()
    g.grad()
    gm = g.get("grad")
    print(f"N=100K M=256, 8 genes: mixed ELBO rel err {abs(em - eo) / abs(eo):.2e}, "
          f"gradient {rel(gm, o.get('grad')):.2e}")
    assert abs(em - eo) <= 1e-6 * abs(eo)
    assert rel(gm, o.get("grad")) < 1e-4
    g.close()


def test_n100k_g2k_m256_mixed_close_to_fp64():
    """All 2,000 genes at N=100,000, Q=10, M=256, SE-ARD: the mixed ELBO within
    1e-6 of the FP64 path's (itself pinned to the oracle above), the gradient
    finite and within 1e-4 of the FP64 gradient's largest entry."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(100000, 2000, 10, 256, seed=0)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = _ctx(ds, T0)
    g.init_fit()
    for jtr in range(2):
        g.irls_step(jtr)
    g.ql_start()
    e64 = g.elbo(1)
    g.grad()
    g64 = g.get("grad")
    g.set_precision(lib.PREC_MIXED)
    em = g.elbo(1)
    g.grad()
    gm = g.get("grad")
    print(f"N=100K G=2K M=256: mixed ELBO rel err {abs(em - e64) / abs(e64):.2e}, gradient {rel(gm, g64):.2e}")
    assert np.all(np.isfinite(gm))
    assert abs(em - e64) <= 1e-6 * abs(e64)
    assert rel(gm, g64) < 1e-4
    g.close()


def test_se_ard_n2k_g200_m32_elbo_and_gradient():
    """The SE-ARD kernel (kernel_funcs.cu:221-255) at N=2,000, G=200, Q=5,
    M=32: the FP64 ELBO and every gradient block against the oracle, and the
    mixed path within 1e-7 (ELBO) and 1e-4 (gradient)."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(2000, 200, 5, 32, seed=0)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    o = make_oracle(ds, T0=T0, ard_only=True)
    g = _ctx(ds, T0)
    o.init_fit()
    g.init_fit()
    for jtr in range(2):
        o.irls_step(jtr)
        g.irls_step(jtr)
    a, b = o.ql_start(), g.ql_start()
    assert abs(a - b) <= 1e-11 * abs(a)
    o.grad()
    g.grad()
    go, gg = o.get("grad"), g.get("grad")
    for k, s in _blocks(ds.N, len(ds.nh) - 1, ds.Q, ds.M).items():
        assert rel(gg[s], go[s]) < 1e-9, k
    eo = o.elbo(1)
    g.set_precision(lib.PREC_MIXED)
    em = g.elbo(1)
    o.grad()
    g.grad()
    assert abs(em - eo) <= 1e-7 * abs(eo)
    assert rel(g.get("grad"), o.get("grad")) < 1e-4
    g.close()


def _gene_oracle(ds, bd, covoff, k, T0):
    """The oracle on gene k's own cells and design (a one-gene state)."""
    from oracle.oracle import OracleState
    c = bd.cell[bd.off[k]:bd.off[k + 1]]
    n, pg = c.size, int(bd.ncov[k])
    Zcov = np.ones((n, pg))
    Zcov[:, 1:] = bd.cov[covoff[k]:covoff[k] + n * (pg - 1)].reshape(n, pg - 1, order="F")
    nh = np.array([0, 1, pg, pg + ds.M], np.int32)
    lo, hi = bd.yoff[k], bd.yoff[k + 1]
    return OracleState(n, 1, ds.M, ds.Q, nh, np.array([0, hi - lo], np.int32), bd.ypos[lo:hi],
                       bd.yval[lo:hi], Zcov, ds.ls[c], np.asarray(ds.X0)[c], T0)


def test_batched_n200k_g20k_m128_sampled_genes_match_oracle():
    """G=20,000 genes, each on its own subset (sizes log-uniform in [500,
    50,000]) of N=200,000 cells, M=128, Q=10, p_g ~ U{2..16}: 4 IRLS steps of
    the persistent launch replayed as a graph; every gene's beta, phi and
    objective finite, the summed objective rising at every step; 24 sampled
    genes (the smallest, the largest, the widest design and 21 at random)
    against the oracle on their own cells: beta and phi within 1e-8, the
    per-gene objectives within 1e-10."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(200000, 1, 10, 128, seed=0, pca=False)
    bd = synth.make_batched(ds, 20000, 500, 50000, seed=0)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = _ctx(ds, T0, ard_only=False)
    g.init_fit()
    bf = lib.BatchedFit(g, bd.off, bd.cell, bd.ncov, bd.cov, bd.yoff, bd.ypos, bd.yval)
    bf.init()
    objs = []
    for _ in range(4):
        _, ob = bf.run(1, 0.0, graph=True)
        objs.append(ob)
    assert np.all(np.isfinite(objs)) and np.all(np.diff(objs) > 0), objs
    PB = bf.PB
    B = bf.get("Beta").reshape(bf.J, PB)
    phi, lk = bf.get("phi"), bf.get("lkhd")
    assert np.all(np.isfinite(B)) and np.all(np.isfinite(phi)) and np.all(np.isfinite(lk))
    sizes = np.diff(bd.off)
    covoff = np.zeros(bf.J + 1, np.int64)
    covoff[1:] = np.cumsum(sizes.astype(np.int64) * (bd.ncov.astype(np.int64) - 1))
    rng = np.random.default_rng(1)
    pick = sorted(set([int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argmax(bd.ncov))]
                      + list(rng.choice(bf.J, 21, replace=False))))
    worst = 0.0
    for k in pick:
        o = _gene_oracle(ds, bd, covoff, k, T0)
        o.init_fit()
        for j in range(4):
            lo = o.irls_step(j)
        P = int(bd.ncov[k]) + ds.M
        rb = rel(B[k, :P], o.get("Beta")[:P])
        worst = max(worst, rb)
        assert rb < 1e-8, (k, sizes[k], rb)
        assert abs(phi[k] - o.get("phi")[0]) <= 1e-8 * abs(phi[k]), k
        assert abs(lk[k] - o.get("lkhds")[0]) <= 1e-10 * abs(lk[k]), k
        o.close()
    print(f"batched N=200K G=20K: objectives {objs}, {len(pick)} genes vs oracle, worst beta rel err {worst:.2e}")
    bf.close()
    g.close()


def test_encoder_2000_512_256_20_loss_and_gradient_match_oracle():
    """The encoder at its stated shape (2,000 log1p features, 512 and 256
    hidden, 2 x Q = 20 outputs, batch 1,024): loss and every weight gradient
    of one batch against the oracle (same BF16 rounding points)."""
    from gaspacho2_amd import lib, synth
    from gaspacho2_amd.synth import dense_counts
    from oracle import encoder as E
    F, Q, H1, H2, B = 2000, 10, 512, 256, 1024
    ds = synth.make_dataset(1100, F, Q, 8, seed=3, model_seed=7, pca=False)
    X = np.random.default_rng(4).normal(0, 1, (ds.N, Q))
    enc = lib.Encoder(F, Q, H1=H1, H2=H2, batch=B, capacity=ds.N, seed=2)
    enc.add(ds.N, ds.Yp, ds.Yi, ds.Yx, ds.ls, X)
    loss = enc.loss_grad(B)
    p = E.unpack(enc.get("params"), F, H1, H2, Q)
    A0 = E.features(dense_counts(ds)[:B], ds.ls[:B])
    lo, go = E.loss_grad(p, A0, X[:B])
    gg = E.unpack(enc.get("grad"), F, H1, H2, Q)
    errs = {k: float(np.linalg.norm(gg[k] - go[k]) / np.linalg.norm(go[k])) for k in go}
    print(f"encoder 2000-512-256-20: loss {loss:.8f} vs {lo:.8f}; gradient norm rel errs {errs}")
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    for k, r in errs.items():
        assert r < 2e-3, (k, r)
    enc.close()
