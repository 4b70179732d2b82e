"""GPU (HIP, through the C ABI) vs the FP64 CPU oracle, launch group by launch
group, on the same seeded inputs.  FP64 path: tolerances are a few hundred
ulps relative to the largest entry (summation order differs: tree reductions
and blocked factorizations on the GPU, serial loops in the oracle)."""
import numpy as np
import pytest

from conftest import make_oracle, rel

pytestmark = pytest.mark.gpu

TOL = 1e-9          # arrays, FP64 path
TOL_OBJ = 1e-11     # objectives, FP64 path


def _pair(ds, ard_only=False, precision=0):
    from gaspacho2_amd import lib
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    o = make_oracle(ds, T0=T0, ard_only=ard_only)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0,
                      T0, ard_only=ard_only, precision=precision)
    return o, g


def _cmp(o, g, name, tol=TOL, sl=None):
    a, b = g.get(name), o.get(name)
    if sl is not None:
        a, b = a[sl], b[sl]
    r = rel(a, b)
    assert r < tol, f"{name}: rel err {r:.3e}"
    return r


def _psi_lower(x, J, P):
    x = x[:J * P * P].reshape(J, P, P).transpose(0, 2, 1)
    return np.tril(x)


def test_init_fit(small_ds):
    o, g = _pair(small_ds)
    o.init_fit()
    g.init_fit()
    for name in ("Z", "K", "Kinv", "KinvKmn", "titsias", "DeltaInv", "Beta"):
        _cmp(o, g, name)


@pytest.mark.parametrize("Q", [40, 64])
def test_init_fit_wide_latent(Q):
    """Kuf, Kmm and the basis for a latent of Q = 40 and 64 dimensions (the
    block's X rows in LDS take 80 and 128 KB there) against the oracle."""
    from gaspacho2_amd import synth
    ds = synth.make_dataset(700, 6, Q, 24, seed=4, pca=False)
    for ard in (False, True):
        o, g = _pair(ds, ard_only=ard)
        o.init_fit()
        g.init_fit()
        for name in ("Z", "K", "Kinv", "KinvKmn", "titsias"):
            _cmp(o, g, name)
        g.close()


@pytest.mark.parametrize("ard", [False, True])
def test_irls_and_quasi_likelihood_start(small_ds, ard):
    o, g = _pair(small_ds, ard_only=ard)
    o.init_fit()
    g.init_fit()
    J, P = small_ds.J, small_ds.P
    for jtr in range(4):
        lo, lg = o.irls_step(jtr), g.irls_step(jtr)
        assert abs(lo - lg) <= TOL_OBJ * abs(lo), (jtr, lo, lg)
    for name in ("W", "Nu", "Beta", "phi", "lkhds"):
        _cmp(o, g, name, sl=slice(0, small_ds.N * small_ds.J) if name == "W" else None)
    assert rel(_psi_lower(g.get("Psi"), J, P), _psi_lower(o.get("Psi"), J, P)) < TOL
    lo, lg = o.ql_start(), g.ql_start()
    assert abs(lo - lg) <= TOL_OBJ * abs(lo)
    _cmp(o, g, "lkhds")


@pytest.mark.parametrize("ard", [False, True])
def test_elbo_and_gradient(small_ds, ard):
    """ELBO + gradient unit vs the oracle with the periodic-dim-0 kernel
    (kernel_funcs.cu:61-97) and with SE-ARD on every dimension
    (kernel_funcs.cu:221-255): the cell gradient, the C1 T product, and the
    tau / rho gradients take their non-periodic branches for dim 0."""
    o, g = _pair(small_ds, ard_only=ard)
    o.init_fit()
    g.init_fit()
    for jtr in range(3):
        o.irls_step(jtr)
        g.irls_step(jtr)
    o.ql_start()
    g.ql_start()
    o.grad()
    g.grad()
    N, p, Q, M, J, P = small_ds.N, len(small_ds.nh) - 1, small_ds.Q, small_ds.M, small_ds.J, small_ds.P
    _cmp(o, g, "tBeta")
    _cmp(o, g, "gradDeltaTmp")
    _cmp(o, g, "C1")
    _cmp(o, g, "C2")
    go, gg = o.get("grad"), g.get("grad")
    blocks = {"gOmega": slice(0, N), "gdelta": slice(N, N + p), "gX": slice(N + p, N + p + N * Q),
              "gT": slice(N + p + N * Q, N + p + N * Q + M * Q),
              "grho": slice(N + p + N * Q + M * Q, N + p + N * Q + M * Q + Q)}
    for k, s in blocks.items():
        assert rel(gg[s], go[s]) < TOL, k
    inv_g = g.get("Psi")[J * P * P:]
    inv_o = o.get("Psi")[J * P * P:]
    assert rel(inv_g, inv_o) < 1e-8
    # ELBO at a moved point
    X = o.get("X")
    X[: N * Q] += 0.01 * np.sin(np.arange(N * Q))
    o.set("X", X)
    g.set("X", X)
    lo, lg = o.elbo(1), g.elbo(1)
    assert abs(lo - lg) <= TOL_OBJ * abs(lo)


def test_lbfgs_line_search_sequence(small_ds):
    o, g = _pair(small_ds)
    o.init_fit()
    g.init_fit()
    for jtr in range(3):
        o.irls_step(jtr)
        g.irls_step(jtr)
    lk_o = [o.ql_start()]
    lk_g = [g.ql_start()]
    for jtr in range(6):
        o.grad()
        g.grad()
        o.lbfgs_step(jtr, 0)
        g.lbfgs_step(jtr, 0)
        sr = 1.0
        for ktr in range(1, 8):
            a, b = o.elbo(jtr + 1), g.elbo(jtr + 1)
            assert abs(a - b) <= 1e-9 * abs(a), (jtr, ktr, a, b)
            if a > lk_o[-1] or ktr == 7:
                break
            sr *= 0.1
            o.downdate(jtr, sr, 0)
            g.downdate(jtr, sr, 0)
        lk_o.append(a)
        lk_g.append(b)
    for name in ("Omega", "delta", "X", "T", "rho"):
        _cmp(o, g, name, tol=1e-8)


def test_fit_short_schedule(tiny_ds):
    o, g = _pair(tiny_ds)
    lo = o.fit(2, 6)
    lg = g.fit(2, 6)
    assert abs(lo - lg) <= 1e-8 * abs(lo)
    _cmp(o, g, "X", tol=1e-7)


def test_config0_size_elbo_and_gradient():
    """N=2,000, G=200, Q=5, M=32 (the small CPU-runnable workload)."""
    from gaspacho2_amd import synth
    ds = synth.make_dataset(2000, 200, 5, 32, seed=0)
    o, g = _pair(ds)
    o.init_fit()
    g.init_fit()
    for jtr in range(2):
        o.irls_step(jtr)
        g.irls_step(jtr)
    a, b = o.ql_start(), g.ql_start()
    assert abs(a - b) <= TOL_OBJ * abs(a)
    o.grad()
    g.grad()
    assert rel(g.get("grad"), o.get("grad")) < TOL


def test_non_pd_factor_is_zero_filled(tiny_ds):
    o, g = _pair(tiny_ds)
    o.init_fit()
    g.init_fit()
    o.irls_step(0)
    g.irls_step(0)
    N, J, P = tiny_ds.N, tiny_ds.J, tiny_ds.P
    W = o.get("W")
    W[N * 2: N * 3] *= -50.0          # gene 2 gets negative weights -> indefinite Gram
    o.set("W", W)
    g.set("W", W)
    o.call("orc_get_psi")
    bad_o = o.call("orc_chol_psi")
    r = g.elbo(0)
    assert bad_o == 1 and g.last_nonpd == 1
    Pg = g.get("Psi")[:J * P * P].reshape(J, P * P)
    assert not Pg[2].any()
    assert not np.isfinite(r)


def test_mixed_precision_gram_error_bounds():
    """FP16 hi/lo MFMA Gram (FP32 accumulate inside a cell segment, FP64 across
    segments) against the FP64 Gram of the same whitened basis at the same W
    (IRLS run in FP64) at the N=2,000 / G=200 / M=32 workload: entries within
    5e-7 of the largest entry of their gene (measured 1e-7: FP32 accumulation
    dominates; the hi/lo operand split alone gives ~2e-8), ELBO within 1e-7 of
    the oracle."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(2000, 200, 5, 32, seed=0, n_numeric=2)
    o, g = _pair(ds, precision=lib.PREC_FP64)
    o.init_fit()
    g.init_fit()
    for jtr in range(2):
        o.irls_step(jtr)
        g.irls_step(jtr)
    o.ql_start()
    g.ql_start()
    J, P = ds.J, ds.P
    iu = np.triu_indices(P, 1)
    g.set_precision(lib.PREC_FP64_WHITENED)
    g.elbo(1)
    Pw = g.get("Psi")[:J * P * P].reshape(J, P, P)[:, iu[1], iu[0]]
    g.set_precision(lib.PREC_MIXED)
    f = g.elbo(1)
    e = o.elbo(1)
    Pm = g.get("Psi")[:J * P * P].reshape(J, P, P)[:, iu[1], iu[0]]
    err = np.max(np.abs(Pm - Pw), axis=1) / np.max(np.abs(Pw), axis=1)
    print(f"mixed Gram max rel err per gene: {err.max():.3e}; ELBO rel err {abs(e - f) / abs(e):.3e}")
    assert err.max() < 5e-7
    assert abs(e - f) <= 1e-7 * abs(e), (e, f, abs(e - f) / abs(e))


def test_whitened_basis_fp64_matches_oracle():
    """The whitened basis [covariates | Knm L^-T] with FP64 arithmetic gives the
    reference ELBO and gradient (the basis change is exact algebra: the log|L|
    terms cancel and Psi^-1, beta-tilde, U map back through L^-1)."""
    from gaspacho2_amd import lib
    from gaspacho2_amd import synth
    ds = synth.make_dataset(600, 30, 4, 16, seed=5, n_numeric=2, n_levels=2)
    o, g = _pair(ds)
    o.init_fit()
    g.init_fit()
    for jtr in range(3):
        o.irls_step(jtr)
        g.irls_step(jtr)
    o.ql_start()
    g.ql_start()
    g.set_precision(lib.PREC_FP64_WHITENED)
    a, b = o.elbo(1), g.elbo(1)
    assert abs(a - b) <= 1e-10 * abs(a), (a, b)
    o.grad()
    g.grad()
    go, gg = o.get("grad"), g.get("grad")
    N, p = ds.N, len(ds.nh) - 1
    assert rel(gg, go) < 1e-8
    assert rel(gg[N:N + p], go[N:N + p]) < 1e-8
    _cmp(o, g, "tBeta", tol=1e-8)
    _cmp(o, g, "C2", tol=1e-8)


def test_mixed_precision_gradient_close_to_fp64():
    """Mixed path (whitened basis, FP16 hi/lo Gram) vs the FP64 oracle: ELBO and
    the whole gradient vector at the N=2,000 / G=200 / M=32 workload."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(2000, 200, 5, 32, seed=0)
    o, g = _pair(ds)
    o.init_fit()
    g.init_fit()
    for jtr in range(2):
        o.irls_step(jtr)
        g.irls_step(jtr)
    o.ql_start()
    g.ql_start()
    g.set_precision(lib.PREC_MIXED)
    a, b = o.elbo(1), g.elbo(1)
    o.grad()
    g.grad()
    gr = rel(g.get("grad"), o.get("grad"))
    print(f"mixed: ELBO rel err {abs(a - b) / abs(a):.3e}, gradient rel err {gr:.3e}")
    assert abs(a - b) <= 1e-7 * abs(a)
    assert gr < 1e-4


@pytest.mark.parametrize("shape", [(2000, 200, 5, 32, 0, 0), (3000, 48, 10, 256, 0, 0)])
def test_fp32_tn_products_keep_the_mixed_gradient(shape):
    """The mixed path's z^T beta-tilde and C1a products in FP32 (option
    tn_f32, the default; ggplvm.cu:335-345 tn and C1 pieces) against the same
    products in FP64: the mixed gradient's distance from the FP64 unit's
    changes by no more than 2e-6 of its largest entry (at N = 100K, M = 256:
    1.61e-6 in FP64 and 1.59e-6 in FP32), and the two gradients agree within
    1e-5 of it; the ELBO is unchanged (the products are gradient-only)."""
    from gaspacho2_amd import lib, synth
    N, G, Q, M, nnum, nlev = shape
    ds = synth.make_dataset(N, G, Q, M, seed=5, n_numeric=nnum, n_levels=nlev)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0,
                      T0, precision=lib.PREC_FP64)
    g.init_fit()
    for jtr in range(2):
        g.irls_step(jtr)
    g.ql_start()
    g.elbo(1)
    g.grad()
    g64 = g.get("grad")
    g.set_precision(lib.PREC_MIXED)
    out = {}
    for v in (0, 1):
        g.set_option("tn_f32", v)
        e = g.elbo(1)
        g.grad()
        out[v] = (e, g.get("grad"))
    d0, d1 = rel(out[0][1], g64), rel(out[1][1], g64)
    d01 = rel(out[1][1], out[0][1])
    print(f"tn products FP64 / FP32: gradient vs FP64 unit {d0:.2e} / {d1:.2e}, between them {d01:.2e}")
    assert out[0][0] == out[1][0]
    assert d1 <= d0 + 2e-6 and d01 < 1e-5
    g.close()


@pytest.mark.parametrize("shape", [(300, 20, 3, 8, 2, 3), (2000, 200, 5, 32, 0, 0)])
def test_mixed_irls_steps_close_to_fp64(shape):
    """IRLS steps on the whitened FP16 hi/lo Gram (option irls_mixed, the
    default of a mixed context) against the FP64 steps of the reference basis
    from the same start, with covariates (first shape) and without: the GLM
    objectives within 1e-8, beta and phi within 1e-5 of their largest entry."""
    from gaspacho2_amd import lib, synth
    N, G, Q, M, nnum, nlev = shape
    ds = synth.make_dataset(N, G, Q, M, seed=1, n_numeric=nnum, n_levels=nlev)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    ctx = {}
    for prec in (lib.PREC_FP64, lib.PREC_MIXED):
        ctx[prec] = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls,
                                  ds.X0, T0, precision=prec)
        ctx[prec].init_fit()
    assert ctx[lib.PREC_MIXED].get_option("irls_mixed") == 1
    worst = 0.0
    for jtr in range(6):
        a, b = ctx[lib.PREC_FP64].irls_step(jtr), ctx[lib.PREC_MIXED].irls_step(jtr)
        worst = max(worst, abs(a - b) / abs(a))
    rb = rel(ctx[lib.PREC_MIXED].get("Beta"), ctx[lib.PREC_FP64].get("Beta"))
    rp = rel(ctx[lib.PREC_MIXED].get("phi"), ctx[lib.PREC_FP64].get("phi"))
    print(f"mixed IRLS: objective rel err {worst:.2e}, beta {rb:.2e}, phi {rp:.2e}")
    assert worst < 1e-8
    assert rb < 1e-5 and rp < 1e-5
    for g in ctx.values():
        g.close()


@pytest.mark.parametrize("shape", [(300, 20, 3, 8, 2, 3), (2000, 200, 5, 32, 0, 0),
                                   (3000, 24, 5, 256, 0, 0)])
def test_fused_irls_step_matches_separate_kernels(shape):
    """The fused IRLS step of a mixed context (option irls_fused, the default:
    W / Omega, nu, R and the sums built inside the Gram kernel from
    eta = z^T beta, ggplvm.cu:468-514 with 621-680) and the same step on
    separate kernels (W and nu written, then the FP32 weights, the Gram, R and
    the sums, the GLM objective), both against the FP64 steps from the same
    start.  After the first step the fused form's W and nu, written out only
    when read, equal the separate kernels' to rounding (1e-13: it computes W
    as exp(eta - d) Omega).  Over six steps both mixed forms keep the
    objectives and the quasi-likelihood start that follows within 1e-7 of
    FP64 (at M = 256 both measure ~1.2e-8; at M = 8 and 32 they are within the
    1e-8 of the test above), beta and phi within 1e-5 of their largest entry
    plus twice the FP64 path's own spread (its beta from the Cholesky on 8-
    and on 16-column panels: at M = 256 on 3,000 cells the reference-basis
    factors are conditioned so that this alone moves beta by 5.5e-6; 1e-11 at
    M = 8 and 32), and the fused objectives no further from FP64 than twice
    the separate form's (the fused Gram scales its FP16 operands per cell
    segment rather than per gene, so the two round differently).  Covariates
    (the whole-design kernel), M = 32 and M = 256 (the GP-block kernel at the
    main workload's width); N = 300 leaves cell segments that hold only
    padding."""
    from gaspacho2_amd import lib, synth
    N, G, Q, M, nnum, nlev = shape
    ds = synth.make_dataset(N, G, Q, M, seed=2, n_numeric=nnum, n_levels=nlev)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    ctx = {}
    for key, prec, fused in (("f64", lib.PREC_FP64, 0), ("f64b", lib.PREC_FP64, 0),
                             ("sep", lib.PREC_MIXED, 0), ("fused", lib.PREC_MIXED, 1)):
        g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls,
                          ds.X0, T0, precision=prec)
        g.set_option("irls_fused", fused)
        if key == "f64b":
            g.set_option("chol_nb", 16)
        g.init_fit()
        ctx[key] = g
    assert ctx["fused"].get_option("irls_fused") == 1
    err = {"sep": 0.0, "fused": 0.0}
    for jtr in range(6):
        ref = ctx["f64"].irls_step(jtr)
        ctx["f64b"].irls_step(jtr)
        for k in err:
            err[k] = max(err[k], abs(ctx[k].irls_step(jtr) - ref) / abs(ref))
        if jtr == 0:
            rw = rel(ctx["fused"].get("W"), ctx["sep"].get("W"))
            rn = rel(ctx["fused"].get("Nu"), ctx["sep"].get("Nu"))
    rb = {k: rel(ctx[k].get("Beta"), ctx["f64"].get("Beta")) for k in err}
    rp = {k: rel(ctx[k].get("phi"), ctx["f64"].get("phi")) for k in err}
    sb = rel(ctx["f64b"].get("Beta"), ctx["f64"].get("Beta"))
    sp = rel(ctx["f64b"].get("phi"), ctx["f64"].get("phi"))
    q64 = ctx["f64"].ql_start()
    rq = {k: abs(ctx[k].ql_start() - q64) / abs(q64) for k in err}
    print(f"first step W {rw:.2e}, Nu {rn:.2e}; vs FP64 (fused / separate): objectives "
          f"{err['fused']:.2e} / {err['sep']:.2e}, beta {rb['fused']:.2e} / {rb['sep']:.2e}, "
          f"phi {rp['fused']:.2e} / {rp['sep']:.2e}, quasi-likelihood start "
          f"{rq['fused']:.2e} / {rq['sep']:.2e}; FP64 spread: beta {sb:.2e}, phi {sp:.2e}")
    assert rw < 1e-13 and rn < 1e-13
    for k in err:
        assert err[k] < 1e-7 and rq[k] < 1e-7, k
        assert rb[k] < 1e-5 + 2 * sb and rp[k] < 1e-5 + 2 * sp, k
    assert err["fused"] <= 2.0 * err["sep"] + 1e-9
    for g in ctx.values():
        g.close()


@pytest.mark.parametrize("slo", [1, 0])
def test_mixed_gradient_contraction_at_m256(slo):
    """Mixed path at P = 257 (M = 256, the widest column-tile variant of both
    FP16 MFMA kernels).  The FP16 contraction against the FP64 contraction of
    the same whitened inverses: with S_j as hi/lo (option gc_slo 1, the
    default) Uw within 1e-6 of its largest entry (measured 1.4e-7), with S_j
    in FP16 alone and zw as hi/lo (gc_slo 0: two MFMAs per tile instead of
    three) within 5e-4, FP16's unit roundoff (measured 1.2e-4).  The whole
    mixed path against the FP64 path on the GPU (itself pinned to the oracle
    above): ELBO within 1e-7, gradient within 1e-4 of its largest entry
    (measured 7.6e-7 and 5.4e-6)."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(3000, 48, 10, 256, seed=3)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0,
                      T0, precision=lib.PREC_FP64)
    g.init_fit()
    for jtr in range(2):
        g.irls_step(jtr)
    g.ql_start()
    e64 = g.elbo(1)
    g.grad()
    g64 = g.get("grad")
    g.set_precision(lib.PREC_MIXED)
    g.set_option("gc_slo", slo)
    em = g.elbo(1)
    g.set_option("gc_mixed", 0)
    g.grad()
    uw64 = g.get("Uw")
    g.set_option("gc_mixed", 1)
    g.grad()
    gm, uwm = g.get("grad"), g.get("Uw")
    ue, ge = rel(uwm, uw64), rel(gm, g64)
    print(f"M=256 mixed, gc_slo={slo}: ELBO rel err {abs(em - e64) / abs(e64):.3e}, contraction rel err "
          f"{ue:.3e}, gradient rel err {ge:.3e}")
    assert abs(em - e64) <= 1e-7 * abs(e64)
    assert ue < (1e-6 if slo else 5e-4)
    assert ge < 1e-4


@pytest.mark.parametrize("M", [32, 160])
def test_inverse_vtv_forms_agree(M):
    """The per-gene inverses S_j = V_j^T V_j (solvePsi, ggplvm.cu:46-53) with
    the product on FP64 MFMA (option vtv_mfma, the default) against the FP64
    VALU form: the inverses within 1e-12 of their largest entry and the
    gradient within the FP64 path's array tolerance (1e-9; both measured
    equal: the MFMA adds the rows' products in the loop's order); P = 33 (one
    64-column tile) and P = 161 (three tiles, the last partial)."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(1500, 24, 5, M, seed=17)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    out = {}
    for form in (1, 0):
        g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0,
                          T0, precision=lib.PREC_FP64)
        g.set_option("vtv_mfma", form)
        g.init_fit()
        for jtr in range(2):
            g.irls_step(jtr)
        g.ql_start()
        g.grad()
        J, P = ds.J, ds.P
        out[form] = (g.get("Psi")[J * P * P:].copy(), g.get("grad").copy())
        g.close()
    ri, rg = rel(out[1][0], out[0][0]), rel(out[1][1], out[0][1])
    print(f"V^T V on MFMA vs VALU, P={ds.P}: inverses {ri:.2e}, gradient {rg:.2e}")
    assert ri < 1e-12 and rg < TOL


@pytest.mark.parametrize("M", [32, 96])
def test_fp16_s_contraction_other_widths(M):
    """The peeled contraction with S_j in FP16 alone (option gc_slo 0) at two
    more column-tile counts (M / 16 = 2 and 6): Uw within FP16's unit
    roundoff (5e-4 of its largest entry) of the FP64 contraction of the same
    whitened inverses, and the ELBO untouched (the contraction is
    gradient-only)."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(2000, 40, 5, M, seed=13)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0,
                      T0, precision=lib.PREC_FP64)
    g.init_fit()
    for jtr in range(2):
        g.irls_step(jtr)
    g.ql_start()
    g.set_precision(lib.PREC_MIXED)
    e1 = g.elbo(1)
    g.set_option("gc_mixed", 0)
    g.grad()
    uw64 = g.get("Uw")
    g.set_option("gc_mixed", 1)
    g.set_option("gc_slo", 0)
    e0 = g.elbo(1)
    g.grad()
    ue = rel(g.get("Uw"), uw64)
    g.close()
    print(f"M={M} FP16-S contraction: Uw rel err {ue:.3e}")
    assert e0 == e1
    assert ue < 5e-4


@pytest.mark.parametrize("M", [32, 40, 256])
def test_single_matrix_cholesky_matches_batched_kernel(M):
    """The one-matrix Cholesky of the Kmm factor (option chol1, the default;
    cholKmm / choldc chol.cu:19-43) against the batched kernel on the same
    matrix: the factor within 1e-10 of its largest entry (right-looking
    against left-looking order; M = 40 ends on a partial 32-column panel) and
    the ELBO after two IRLS steps within 1e-9, the FP64 path's array
    tolerance (256 inducing points in 5 latent dimensions make Kmm
    ill-conditioned enough that the two orders differ by 8.9e-12 and
    1.7e-10 there).  The M = 256 FP64 oracle comparison at N = 100K runs with
    it as well."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(1500, 24, 5, M, seed=6)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    out = {}
    for one in (0, 1):
        g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls,
                          ds.X0, T0, precision=lib.PREC_FP64)
        g.set_option("chol1", one)
        g.init_fit()
        out[one] = {"Kinv": g.get("Kinv")}
        for jtr in range(2):
            g.irls_step(jtr)
        g.ql_start()
        out[one]["elbo"] = g.elbo(1)
        g.close()
    rk = rel(out[1]["Kinv"], out[0]["Kinv"])
    re = abs(out[1]["elbo"] - out[0]["elbo"]) / abs(out[0]["elbo"])
    print(f"one-matrix Cholesky vs batched, M={M}: Kmm factor {rk:.2e}, ELBO {re:.2e}")
    assert rk < 1e-10 and re < TOL


@pytest.mark.parametrize("shape", [(300, 20, 3, 8, 2, 3), (2000, 200, 5, 32, 0, 0)])
def test_blocked_solve_matches_row_solve(shape):
    """The per-gene solves x_j = Psi_j^-1 b_j (cholsl, chol.cu:45-56) blocked
    by 16 rows (option cholsl_blk, the default) against the row-by-row
    kernel: IRLS objectives and the quasi-likelihood start within 1e-11,
    beta within the FP64 path's array tolerance (1e-9); P = 17 with
    covariates and P = 33, both ending on a partial block.  (At M = 256 on
    1,500 cells the reference-basis factors are conditioned so badly that any
    change of summation order moves beta by 1e-4; the N = 100K, M = 256
    oracle comparison runs with the blocked solve.)"""
    from gaspacho2_amd import lib, synth
    N, G, Q, M, nnum, nlev = shape
    ds = synth.make_dataset(N, G, Q, M, seed=8, n_numeric=nnum, n_levels=nlev)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    out = {}
    for blk in (0, 1):
        g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls,
                          ds.X0, T0, precision=lib.PREC_FP64)
        g.set_option("cholsl_blk", blk)
        g.init_fit()
        obj = [g.irls_step(jtr) for jtr in range(3)]
        obj.append(g.ql_start())
        out[blk] = (np.array(obj), g.get("Beta"))
        g.close()
    ro = float(np.max(np.abs(out[1][0] - out[0][0]) / np.abs(out[0][0])))
    rb = rel(out[1][1], out[0][1])
    print(f"blocked vs row solve: objectives {ro:.2e}, beta {rb:.2e}")
    assert ro < TOL_OBJ and rb < TOL


@pytest.mark.parametrize("shape", [(300, 20, 3, 8, 2, 3, "fp64"), (2000, 200, 5, 32, 0, 0, "fp64"),
                                   (3000, 60, 5, 64, 0, 0, "whitened")])
def test_cholesky_panel_widths_agree(shape):
    """The batched Cholesky of the Psi_j (choldc, chol.cu:19-43) on 8-column
    panels (option chol_nb 8: eight workgroups per CU) against 16-column
    panels: the factors and matrices within 1e-10 of their largest entry, the
    IRLS objectives and the quasi-likelihood start within 1e-11, beta within
    the FP64 path's array tolerance (1e-9; 1e-8 on the whitened basis at
    P = 65, whose conditioning carries a 4e-13 difference of the factors to
    1.8e-9 in beta).  P = 14 with covariates, P = 33, and P = 65 on the
    whitened basis (three k-chunks, a partial last panel)."""
    from gaspacho2_amd import lib, synth
    N, G, Q, M, nnum, nlev, basis = shape
    ds = synth.make_dataset(N, G, Q, M, seed=8, n_numeric=nnum, n_levels=nlev)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    prec = lib.PREC_FP64 if basis == "fp64" else lib.PREC_FP64_WHITENED
    out = {}
    for nb in (16, 8):
        g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls,
                          ds.X0, T0, precision=prec)
        g.set_option("chol_nb", nb)
        assert g.get_option("chol_nb") == nb
        g.init_fit()
        obj = [g.irls_step(jtr) for jtr in range(3)]
        obj.append(g.ql_start())
        out[nb] = (np.array(obj), g.get("Psi"), g.get("Beta"))
        g.close()
    ro = float(np.max(np.abs(out[8][0] - out[16][0]) / np.abs(out[16][0])))
    rp = rel(out[8][1], out[16][1])
    rb = rel(out[8][2], out[16][2])
    print(f"8- vs 16-column panels, P={ds.P} {basis}: objectives {ro:.2e}, Psi and factors {rp:.2e}, "
          f"beta {rb:.2e}")
    assert ro < TOL_OBJ and rp < 1e-10 and rb < (TOL if basis == "fp64" else 1e-8)


def test_options_reject_bad_values(tiny_ds):
    from gaspacho2_amd import lib
    _, g = _pair(tiny_ds)
    assert g.get_option("gc_groups") >= 1
    with pytest.raises(lib.GPError):
        g.set_option("gram_splits", 0)
    with pytest.raises(lib.GPError):
        g.set_option("no_such_option", 1)
    g.set_option("gc_groups", 7)
    assert g.get_option("gc_groups") == 7
    with pytest.raises(lib.GPError):
        g.set_option("gram_waves", 6)
    with pytest.raises(lib.GPError):
        g.set_option("chol_nb", 12)
    assert g.get_option("gram_waves") == 8
    with pytest.raises(lib.GPError):
        g.set_option("gc_form", 4)
    g.set_option("gc_form", 3)
    assert g.get_option("gc_form") == 3


@pytest.mark.parametrize("shape", [(2000, 40, 5, 32), (3000, 24, 10, 256), (3000, 24, 10, 240)])
def test_gram_four_wave_form_matches_eight_wave_form(shape):
    """The four-wave mixed Gram (fragments shared across a wave's blocks) sums
    every tile in the same order as the eight-wave form, so the factors and
    the ELBO agree to the last bit; P = 33, 257 and 241 (last column group
    narrower than 4 tiles)."""
    from gaspacho2_amd import lib, synth
    N, G, Q, M = shape
    ds = synth.make_dataset(N, G, Q, M, seed=4)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0,
                      precision=lib.PREC_FP64)
    g.init_fit()
    g.irls_step(0)
    g.ql_start()
    g.set_precision(lib.PREC_MIXED)
    out = {}
    for w in (8, 4):
        g.set_option("gram_waves", w)
        e = g.elbo(1)
        out[w] = (e, g.get("Psi").copy())
    assert out[4][0] == out[8][0]
    assert np.array_equal(out[4][1], out[8][1])
    g.close()


@pytest.mark.parametrize("shape", [(2000, 24, 5, 32), (3000, 24, 10, 256), (3000, 16, 10, 240),
                                   (3000, 16, 10, 224), (700, 8, 3, 8)])
def test_contraction_row_block_form_matches_16_cell_form(shape):
    """The 128-cell row-block contraction (gc_form 1) feeds every accumulator
    the same MFMA sequence as the 16-cell form (gc_form 0): the gradient
    agrees to the last bit; the factored form (gc_form 2) within 1e-6 on the
    contraction Uw and 1e-5 on the gradient, and its one-wave-per-SIMD tiling
    (gc_form 3) to the last bit of form 2; P = 33, 257, 241, 225 and 9
    (column-tile counts 3, 17, 16, 15 and 1: every remainder mod 4, N not a
    multiple of 128)."""
    from gaspacho2_amd import lib, synth
    N, G, Q, M = shape
    ds = synth.make_dataset(N, G, Q, M, seed=6)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0,
                      precision=lib.PREC_FP64)
    g.init_fit()
    g.irls_step(0)
    g.ql_start()
    g.set_precision(lib.PREC_MIXED)
    g.set_option("gc_peel", 0)              # form 2 on the whole design, as forms 0, 1, 3
    out = {}
    for form in (0, 1, 2, 3):
        g.set_option("gc_form", form)
        g.elbo(1)
        g.grad()
        out[form] = (g.get("grad").copy(), g.get("C1").copy(), g.get("Uw").copy())
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    # form 2 rounds differently (the weight multiplies each gene's FP32 sum)
    ru, rg = rel(out[2][2], out[0][2]), rel(out[2][0], out[0][0])
    print(f"gc_form 2 vs 0: Uw rel err {ru:.2e}, gradient rel err {rg:.2e}")
    assert ru < 1e-6 and rg < 1e-5
    assert np.array_equal(out[3][2], out[2][2])
    assert np.array_equal(out[3][0], out[2][0])
    g.close()


@pytest.mark.parametrize("shape", [(2000, 24, 5, 32), (3000, 24, 10, 256), (3000, 16, 4, 64),
                                   (2500, 20, 6, 160)])
def test_contraction_intercept_peeled_form(shape):
    """The intercept-peeled contraction (MFMAs over the GP block only, the
    intercept's k term in each gene's update, the intercept output column
    from one FP32 GEMM and a row dot) against the factored form on the whole
    design and against the FP64 contraction of the same inverses: Uw as close
    to FP64 as the whole-design form is (measured 3e-7 at M = 256), the
    gradient within 1e-5 of the factored form's and within 1e-4 of the FP64
    one's; M = 32, 256, 64 and 160 (2, 16, 4 and 10 GP column tiles).  Its
    wave tilings and prefetch depths feed every accumulator the same MFMA
    sequence: same bits."""
    from gaspacho2_amd import lib, synth
    N, G, Q, M = shape
    ds = synth.make_dataset(N, G, Q, M, seed=8)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0,
                      ard_only=True, precision=lib.PREC_FP64)
    g.init_fit()
    g.irls_step(0)
    g.ql_start()
    g.set_precision(lib.PREC_MIXED)
    g.elbo(1)
    out = {}
    for name, opts in (("peel", {"gc_peel": 1, "gc_mixed": 1}), ("full", {"gc_peel": 0, "gc_mixed": 1}),
                       ("f64", {"gc_mixed": 0})):
        for k, v in opts.items():
            g.set_option(k, v)
        g.grad()
        out[name] = (g.get("Uw").copy(), g.get("grad").copy())
    ru, rg = rel(out["peel"][0], out["full"][0]), rel(out["peel"][1], out["full"][1])
    ru64, rg64 = rel(out["peel"][0], out["f64"][0]), rel(out["peel"][1], out["f64"][1])
    rf64 = rel(out["full"][0], out["f64"][0])
    Np = g.N
    u64 = out["f64"][0].reshape(-1, Np)
    col0 = float(np.max(np.abs(out["peel"][0].reshape(-1, Np)[0] - u64[0])) / np.max(np.abs(u64)))
    colg = float(np.max(np.abs(out["peel"][0].reshape(-1, Np)[1:] - u64[1:])) / np.max(np.abs(u64)))
    print(f"M={M} peeled vs full: Uw {ru:.2e}, grad {rg:.2e}; vs FP64: Uw {ru64:.2e} (intercept column "
          f"{col0:.2e}, GP columns {colg:.2e}), grad {rg64:.2e}; full vs FP64: Uw {rf64:.2e}")
    # about as close to FP64 as the factored form on the whole design (both
    # are at the FP32-accumulation level: 2e-7 to 2e-6 here), and within
    # 5e-6 of FP64 everywhere
    assert ru64 < 5e-6 and ru64 <= 3.0 * rf64 + 1e-7
    assert ru < 5e-6
    assert rg < 1e-5 and rg64 < 1e-4
    # the wave tilings and prefetch depths of the peeled kernel give the same bits
    g.set_option("gc_mixed", 1)
    g.set_option("gc_peel", 1)
    for opts in ({"gc_prefetch": 0}, {"gc_prefetch": 1}, {"gc_wave_rows": 4, "gc_prefetch": 3},
                 {"gc_wave_rows": 8, "gc_waves": 4, "gc_prefetch": 1}, {"gc_prefetch": 0}):
        for k, v in opts.items():
            g.set_option(k, v)
        g.grad()
        assert np.array_equal(g.get("Uw"), out["peel"][0]), opts
    g.close()


@pytest.mark.parametrize("shape", [(2000, 24, 5, 32), (3000, 24, 10, 256), (3000, 16, 4, 64),
                                   (2500, 20, 6, 160), (2500, 12, 6, 320)])
def test_gram_intercept_peeled_form(shape):
    """The intercept-peeled mixed Gram (symmetric FP16 hi/lo Gram over the GP
    block only, the intercept row and column from split-K FP32 products
    Zw^T (W / Omega) summed in FP64) against the Gram over the whole design
    and against the FP64 Gram of the same whitened basis: entries within
    5e-7 of each gene's largest (as the whole-design form), the ELBO within
    1e-7 of FP64; the eight-, four- and (16 GP tiles) twelve-wave
    producer/consumer forms of the peeled Gram give the same bits.  M = 320 (21 GP tiles) is wider than the symmetric kernel's
    tiles: both forms run the tiled general kernel there, the peeled one over
    the GP block."""
    from gaspacho2_amd import lib, synth
    N, G, Q, M = shape
    ds = synth.make_dataset(N, G, Q, M, seed=11)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0,
                      ard_only=True, precision=lib.PREC_FP64)
    g.init_fit()
    g.irls_step(0)
    g.ql_start()
    J, P = ds.J, ds.P
    iu = np.triu_indices(P)
    out = {}
    for name, prec, opts in (("f64", lib.PREC_FP64_WHITENED, {}),
                             ("full", lib.PREC_MIXED, {"gram_peel": 0}),
                             ("peel", lib.PREC_MIXED, {"gram_peel": 1}),
                             ("peel4", lib.PREC_MIXED, {"gram_peel": 1, "gram_waves": 4}),
                             ("peel12", lib.PREC_MIXED, {"gram_peel": 1, "gram_waves": 12})):
        g.set_precision(prec)
        for k, v in opts.items():
            g.set_option(k, v)
        e = g.elbo(1)
        out[name] = (e, g.get("Psi")[:J * P * P].reshape(J, P, P)[:, iu[1], iu[0]].copy())
    ref = out["f64"][1]
    scale = np.max(np.abs(ref), axis=1)
    errs = {k: float(np.max(np.max(np.abs(out[k][1] - ref), axis=1) / scale)) for k in ("full", "peel")}
    e64 = out["f64"][0]
    print(f"M={M} Gram rel err vs FP64: {errs}; ELBO rel err peeled {abs(out['peel'][0] - e64) / abs(e64):.2e}, "
          f"whole {abs(out['full'][0] - e64) / abs(e64):.2e}")
    assert errs["peel"] < 5e-7 and errs["full"] < 5e-7
    assert abs(out["peel"][0] - e64) <= 1e-7 * abs(e64)
    assert np.array_equal(out["peel4"][1], out["peel"][1]) and out["peel4"][0] == out["peel"][0]
    assert np.array_equal(out["peel12"][1], out["peel"][1]) and out["peel12"][0] == out["peel"][0]
    g.close()


def test_autotune_picks_valid_settings_and_keeps_results():
    """gp_autotune picks gram_splits / gc_groups among its candidates and the
    mixed ELBO and gradient after tuning stay within the mixed-path error of
    the FP64 values.  The pick depends on timings, so every Gram split count it
    can choose is also checked: on this dataset the gradient error (all in the
    inducing-point block) is 8.3e-5, 1.26e-4 and 8.5e-5 of max |grad| at 8, 16
    and 32 splits -- where the FP32 partial sums are cut changes their
    rounding -- and the contraction settings move it by < 1e-7."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(3000, 40, 5, 64, seed=4)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0,
                      T0, precision=lib.PREC_FP64)
    g.init_fit()
    for jtr in range(2):
        g.irls_step(jtr)
    g.ql_start()
    e64 = g.elbo(1)
    g.grad()
    g64 = g.get("grad")
    g.set_precision(lib.PREC_MIXED)
    g.set_precision(lib.PREC_FP64)
    with pytest.raises(lib.GPError):
        g.autotune()                      # FP64 path has nothing to tune
    g.set_precision(lib.PREC_MIXED)
    choice = g.autotune()
    assert choice["gram_splits"] in (8, 16, 32) and choice["gc_groups"] in (16, 32, 64)
    assert choice["gram_waves"] in (4, 8, 12) and choice["gc_form"] in (0, 1, 2, 3)
    em = g.elbo(1)
    g.grad()
    assert abs(em - e64) <= 1e-7 * abs(e64)
    assert rel(g.get("grad"), g64) < 2e-4
    for splits in (8, 16, 32):
        g.set_option("gram_splits", splits)
        em = g.elbo(1)
        g.grad()
        assert abs(em - e64) <= 1e-7 * abs(e64), splits
        assert rel(g.get("grad"), g64) < 2e-4, splits


@pytest.mark.parametrize("precision", [0, 1])
def test_gradient_reuses_likelihood_solve(precision):
    """gp_grad right after gp_elbo takes beta-tilde from the likelihood's own
    solve; after any state-changing call (here gp_set_option) it solves again.
    Same inputs, same kernels: the two gradients agree to the last bit."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(1500, 30, 4, 48, seed=7, n_numeric=1)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0,
                      T0, precision=lib.PREC_FP64)
    g.init_fit()
    for jtr in range(2):
        g.irls_step(jtr)
    g.ql_start()
    g.set_precision(precision)
    g.elbo(2)
    g.grad()
    reused = g.get("grad").copy()
    tb = g.get("tBeta").copy()
    g.set_option("gc_groups", g.get_option("gc_groups"))   # no-op value, clears the reuse
    g.grad()
    assert np.array_equal(g.get("grad"), reused)
    assert np.array_equal(g.get("tBeta"), tb)
    g.close()


@pytest.mark.parametrize("precision,M,nnum", [(0, 48, 1), (1, 48, 1), (1, 64, 0)])
def test_graph_replay_matches_eager(precision, M, nnum):
    """gp_eval_graph (the ELBO + gradient unit recorded as a HIP graph) gives
    the eager results, replay after replay, with eager work and parameter
    changes between the replays; M = 64 with no covariates takes the
    intercept-peeled contraction and its FP32 GEMM inside the graph."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(1500, 30, 4, M, seed=6, n_numeric=nnum)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0,
                      T0, precision=lib.PREC_FP64)
    g.init_fit()
    for jtr in range(2):
        g.irls_step(jtr)
    g.ql_start()
    g.set_precision(precision)
    X0 = g.get("X")
    NQ = ds.N * ds.Q
    for rep in range(4):
        X = X0.copy()
        X[:NQ] += 0.003 * rep * np.cos(np.arange(NQ))
        g.set("X", X)
        e_eager = g.elbo(3)
        g.grad()
        gr_eager = g.get("grad")
        e_graph = g.eval_graph(3)
        gr_graph = g.get("grad")
        assert abs(e_graph - e_eager) <= 1e-13 * abs(e_eager), (rep, e_eager, e_graph)
        assert rel(gr_graph, gr_eager) <= 1e-12, rep
        # a second replay at the same parameters reproduces the first
        assert g.eval_graph(3) == e_graph


def test_score_tests_match_oracle(small_ds):
    """getScoreStats (ggplvm_de.cu:491-532): GPU statistics against the oracle
    after the IRLS phase (FP64 factors, reference basis), for doubles and for
    0/1/2 dosage bytes, including a test list with repeated genes / SNPs."""
    o, g = _pair(small_ds)
    o.init_fit()
    g.init_fit()
    for jtr in range(3):
        o.irls_step(jtr)
        g.irls_step(jtr)
    N, J = small_ds.N, small_ds.J
    rng = np.random.default_rng(8)
    S = 7
    D = rng.integers(0, 3, (N, S)).astype(np.uint8)
    gene = rng.integers(0, J, 60).astype(np.int32)
    snp = rng.integers(0, S, 60).astype(np.int32)
    so = o.score_stats(D.astype(np.float64), gene, snp)
    sg = g.score_tests(D, gene, snp)
    sd = g.score_tests(D.astype(np.float64), gene, snp)
    assert np.all(np.isfinite(so))
    assert rel(sg, so) < 1e-9
    assert np.array_equal(sg, sd)


@pytest.mark.parametrize("ard", [False, True])
def test_interaction_score_tests_and_p_values_match_oracle(small_ds, ard):
    """Interaction tests (ggplvm_de.cu:989-994): the test column g * d centred
    over cells, built here on the host as the reference does and given to the
    oracle's getScoreStats, against the GPU's on-device construction from the
    dosage bytes and the per-cell factor; two-sided normal p-values against
    scipy's normal survival function; an absent factor gives the plain
    dosage statistics."""
    import scipy.stats
    o, g = _pair(small_ds, ard_only=ard)
    o.init_fit()
    g.init_fit()
    for jtr in range(3):
        o.irls_step(jtr)
        g.irls_step(jtr)
    N, J = small_ds.N, small_ds.J
    rng = np.random.default_rng(12)
    S = 9
    D = rng.integers(0, 3, (N, S)).astype(np.uint8)
    d = rng.uniform(0.0, 3.0, N)
    Gi = D.astype(np.float64) * d[:, None]
    Gi -= Gi.mean(axis=0)
    gene = np.repeat(np.arange(J), S).astype(np.int32)
    snp = np.tile(np.arange(S), J).astype(np.int32)
    so = o.score_stats(Gi, gene, snp)
    sg, pg = g.score_tests_interaction(D, d, gene, snp)
    print(f"interaction score statistics: max rel err {rel(sg, so):.2e}")
    assert rel(sg, so) < 1e-9
    pref = 2.0 * scipy.stats.norm.sf(np.abs(sg))
    assert np.max(np.abs(pg - pref) / pref) < 1e-12
    s0, _ = g.score_tests_interaction(D, None, gene, snp)
    assert np.array_equal(s0, g.score_tests(D, gene, snp))


def test_donor_genotype_score_tests_match_expanded_genotypes(small_ds):
    """Genotypes given per donor (cell i carries donor[i]'s dosages) give the
    statistics and p-values of the same genotypes expanded to one row per
    cell (the reference's host-side layout) to the last bit, plain and
    with an interaction factor; an out-of-range donor index is an error."""
    from gaspacho2_amd import lib
    _, g = _pair(small_ds)
    g.init_fit()
    for jtr in range(3):
        g.irls_step(jtr)
    N, J = small_ds.N, small_ds.J
    rng = np.random.default_rng(13)
    ND, S = 7, 5
    donor = rng.integers(0, ND, N).astype(np.int32)
    Dd = rng.integers(0, 3, (ND, S)).astype(np.uint8)
    Dc = Dd[donor]
    gene = np.repeat(np.arange(J), S).astype(np.int32)
    snp = np.tile(np.arange(S), J).astype(np.int32)
    d = rng.normal(0.0, 1.0, N)
    for f in (None, d):
        s1, p1 = g.score_tests_interaction(Dc, f, gene, snp)
        s2, p2 = g.score_tests_donor(donor, Dd, f, gene, snp)
        assert np.array_equal(s1, s2) and np.array_equal(p1, p2)
    bad = donor.copy()
    bad[3] = ND
    with pytest.raises(lib.GPError):
        g.score_tests_donor(bad, Dd, None, gene, snp)


def test_score_tests_mixed_basis_close_to_oracle():
    """Score statistics from the mixed path's whitened factors (after gp_elbo)
    against the oracle's reference-basis factors at the same W."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(2000, 40, 5, 32, seed=2, n_numeric=1)
    o, g = _pair(ds)
    o.init_fit()
    g.init_fit()
    for jtr in range(2):
        o.irls_step(jtr)
        g.irls_step(jtr)
    o.ql_start()
    g.ql_start()
    g.set_precision(lib.PREC_MIXED)
    o.elbo(1)
    g.elbo(1)
    rng = np.random.default_rng(9)
    D = rng.integers(0, 3, (ds.N, 10)).astype(np.uint8)
    gene = np.repeat(np.arange(ds.J), 10).astype(np.int32)
    snp = np.tile(np.arange(10), ds.J).astype(np.int32)
    so = o.score_stats(D.astype(np.float64), gene, snp)
    sg = g.score_tests(D, gene, snp)
    print(f"mixed-basis score statistics: max rel err {rel(sg, so):.3e}")
    assert rel(sg, so) < 1e-6


@pytest.mark.parametrize("prec", ["f64", "mixed"])
def test_score_mfma_form_matches_column_loop(prec):
    """The score-test partials on FP64 MFMA with every model-matrix column in
    one pass over the cells (option score_mfma, the default) against the
    column-per-thread loop (score_mfma 0), interaction tests with per-donor
    genotypes at P = 97: in an FP64 context both read the FP64 model matrix
    (statistics within 1e-10 of each other; measured equal); in a mixed
    context after gp_elbo the MFMA form reads the FP32 copy of Zw (within
    1e-8, FP32's rounding of the model matrix; measured 4.2e-10)."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(3000, 24, 5, 96, seed=4)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0,
                      precision=lib.PREC_FP64)
    g.init_fit()
    for jtr in range(2):
        g.irls_step(jtr)
    g.ql_start()
    if prec == "mixed":
        g.set_precision(lib.PREC_MIXED)
        g.elbo(1)
    rng = np.random.default_rng(21)
    D, S = 40, 6
    donor = rng.integers(0, D, ds.N).astype(np.int32)
    dos = rng.integers(0, 3, (D, ds.J * S)).astype(np.uint8)
    factor = rng.uniform(0.0, 2.0, ds.N)
    gene = np.repeat(np.arange(ds.J), S).astype(np.int32)
    snp = np.arange(ds.J * S).astype(np.int32)
    out = {}
    for form in (1, 0):
        g.set_option("score_mfma", form)
        out[form] = g.score_tests_donor(donor, dos, factor, gene, snp)
    g.close()
    rs, rp = rel(out[1][0], out[0][0]), rel(out[1][1], out[0][1])
    print(f"score partials, MFMA vs column loop ({prec}): statistics {rs:.2e}, p-values {rp:.2e}")
    assert np.all(np.isfinite(out[1][0]))
    assert rs < (1e-10 if prec == "f64" else 1e-8)


def _batched_oracle(ds, bd, g, T0):
    """The oracle on gene g's own cells and design (one-gene state)."""
    from oracle.oracle import OracleState
    c = bd.cell[bd.off[g]:bd.off[g + 1]]
    n, pg = c.size, int(bd.ncov[g])
    cov = bd.cov[sum((bd.off[k + 1] - bd.off[k]) * (bd.ncov[k] - 1) for k in range(g)):][: n * (pg - 1)]
    Zcov = np.ones((n, pg))
    Zcov[:, 1:] = cov.reshape(n, pg - 1, order="F").astype(np.float64)
    nh = np.array([0, 1, pg, pg + ds.M], np.int32)
    lo, hi = bd.yoff[g], bd.yoff[g + 1]
    X0 = np.asarray(ds.X0)[c]
    return OracleState(n, 1, ds.M, ds.Q, nh, np.array([0, hi - lo], np.int32), bd.ypos[lo:hi],
                       bd.yval[lo:hi], Zcov, ds.ls[c], X0, T0)


def test_batched_grid_autotune_keeps_results():
    """The batched pass's grid (persistent workgroups per CU) only changes how
    genes are dealt to workgroups: 4 IRLS steps (replayed as a graph) give
    the same beta, phi and objectives at every grid, and after
    gp_bfit_autotune (which keeps beta and phi and drops the recorded graph)
    the steps continue as they would have."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(3000, 4, 4, 24, seed=12)
    bd = synth.make_batched(ds, 60, 40, 2500, seed=3)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0)
    g.init_fit()
    ref = None
    for m in (2, 1, 3):
        bf = lib.BatchedFit(g, bd.off, bd.cell, bd.ncov, bd.cov, bd.yoff, bd.ypos, bd.yval)
        bf.set_grid(m)
        bf.init()
        _, o2 = bf.run(2, tol=0.0, graph=True)
        if m == 3:
            assert bf.autotune() in (1, 2, 3, 4)
        _, o4 = bf.run(2, tol=0.0, graph=True)
        out = (o2, o4, bf.get("Beta").copy(), bf.get("phi").copy())
        bf.close()
        if ref is None:
            ref = out
        assert out[0] == ref[0] and out[1] == ref[1], m
        assert np.array_equal(out[2], ref[2]) and np.array_equal(out[3], ref[3]), m
    with pytest.raises(lib.GPError):
        bf2 = lib.BatchedFit(g, bd.off, bd.cell, bd.ncov, bd.cov, bd.yoff, bd.ypos, bd.yval)
        try:
            bf2.set_grid(0)
        finally:
            bf2.close()
    g.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batched_subset_glm_matches_oracle(graph):
    """Batched per-gene IRLS on uneven cell subsets (persistent launch) against
    the oracle run gene by gene on each subset: beta and phi within 1e-8
    (the unweighted initial Gram with the Kmm/delta prior block is
    ill-conditioned: FP64 summation-order differences reach ~1e-9 in beta),
    the summed objective within 1e-10, over 6 IRLS steps."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(3000, 4, 4, 24, seed=12)
    bd = synth.make_batched(ds, 14, 40, 2500, seed=3)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0)
    g.init_fit()
    bf = lib.BatchedFit(g, bd.off, bd.cell, bd.ncov, bd.cov, bd.yoff, bd.ypos, bd.yval)
    bf.init()
    ors = [_batched_oracle(ds, bd, k, T0) for k in range(bf.J)]
    for o in ors:
        o.init_fit()
    PB = bf.PB
    B0 = bf.get("Beta").reshape(bf.J, PB)
    for k, o in enumerate(ors):
        P = int(bd.ncov[k]) + ds.M
        assert rel(B0[k, :P], o.get("Beta")[:P]) < 1e-8, k
    if graph:
        its, obj = bf.run(6, tol=0.0, graph=True)
        assert its == 6
        objs_o = [sum(o.irls_step(j) for o in ors) for j in range(6)]
        assert abs(obj - objs_o[-1]) <= 1e-10 * abs(objs_o[-1])
    else:
        for j in range(6):
            og = bf.step()
            oo = sum(o.irls_step(j) for o in ors)
            assert abs(og - oo) <= 1e-10 * abs(oo), (j, og, oo)
    B = bf.get("Beta").reshape(bf.J, PB)
    phi = bf.get("phi")
    for k, o in enumerate(ors):
        P = int(bd.ncov[k]) + ds.M
        assert rel(B[k, :P], o.get("Beta")[:P]) < 1e-8, k
        assert abs(phi[k] - o.get("phi")[0]) <= 1e-8 * abs(phi[k]), k
    bf.close()


@pytest.mark.parametrize("case", ["covariates", "intercept"])
def test_pca_init_matches_oracle(case, tmp_path):
    """gp_pca_init (dense nu0, covariate fit, SYRK, rocSOLVER syevd, projection)
    against the numpy restatement of init.cu:304-357: eigenvalues within 1e-10,
    X within 1e-8 of its largest entry.  The covariate case has a full set of
    one-hot levels next to the intercept (dependent columns).  The pcainit
    tool writes the same X as Xinit.bin."""
    import os
    import subprocess
    from conftest import ROOT
    from gaspacho2_amd import lib, synth
    from oracle import pca
    if case == "covariates":
        ds = synth.make_dataset(600, 40, 4, 8, seed=7, n_numeric=2, n_levels=3)
    else:
        ds = synth.make_dataset(800, 60, 5, 8, seed=8)
    N, Q = ds.N, ds.Q
    ncol = int(ds.nh[-2])
    Z = np.asarray(ds.Zcov).reshape(N, -1, order="F")[:, :ncol]
    Xo, lo = pca.pca_init(N, ds.Yp, ds.Yi, ds.Yx, ds.ls, Z, Q)
    Xg, lg = lib.pca_init(N, ds.Yp, ds.Yi, ds.Yx, ds.ls, Z, Q)
    gaps = np.diff(lo[::-1]) / lo[0]
    print(f"{case}: eigenvalues {lo}, smallest relative gap {gaps.min():.2e}; "
          f"X rel err {rel(Xg, Xo):.2e}, eigenvalue rel err {rel(lg, lo):.2e}")
    assert rel(lg, lo) < 1e-10
    assert rel(Xg, Xo) < 1e-8
    for f, name in ((synth.write_sparsemat, "Y.bin"), (synth.write_dataframe, "DF.bin")):
        f(str(tmp_path / name), ds)
    synth.write_doubles(str(tmp_path / "ls.bin"), ds.ls)
    tool = os.path.join(ROOT, "gaspacho2_amd", "pcainit")
    p = subprocess.run([tool, "Y.bin", "DF.bin", "ls.bin", str(Q), "Xinit.bin"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    Xt = np.frombuffer(open(tmp_path / "Xinit.bin", "rb").read(), dtype="<f8").reshape(Q, N).T
    assert np.array_equal(Xt, Xg)


@pytest.mark.parametrize("M", [288, 512, 600])
def test_wide_design_fp64_matches_oracle(M):
    """P = M + 1 beyond the 272-column tiles of the fast kernels: P16 = 304
    and 528 (FP64 MFMA Gram and contraction past their instantiated widths)
    and 608 (the generic FP64 Gram past its LDS bound), SE-ARD, Q = 20.  IRLS
    objectives, factors, ELBO and gradient against the oracle."""
    from gaspacho2_amd import synth
    N, J, Q = (500, 3, 20) if M == 600 else (600, 4, 20)
    ds = synth.make_dataset(N, J, Q, M, seed=5, pca=False)
    o, g = _pair(ds, ard_only=True)
    o.init_fit()
    g.init_fit()
    for jtr in range(2):
        a, b = o.irls_step(jtr), g.irls_step(jtr)
        assert abs(a - b) <= TOL_OBJ * abs(a), (jtr, a, b)
    P = ds.P
    assert rel(_psi_lower(g.get("Psi"), J, P), _psi_lower(o.get("Psi"), J, P)) < TOL
    a, b = o.ql_start(), g.ql_start()
    assert abs(a - b) <= TOL_OBJ * abs(a)
    o.grad()
    g.grad()
    assert rel(g.get("grad"), o.get("grad")) < TOL
    _cmp(o, g, "C2")


@pytest.mark.parametrize("M,wide", [(288, 1), (512, 1), (512, 0)])
def test_wide_design_mixed_close_to_fp64(M, wide):
    """The mixed path at P16 = 304 and 528 against the FP64 whitened path of
    the same state (pinned to the oracle by the whitened-basis test): Gram
    entries within 5e-7 of each gene's largest, ELBO within 1e-7, gradient
    within 1e-4 of its largest entry.  M = 512 runs both Gram forms: the GP
    block as two symmetric halves and four rectangles on the symmetric
    kernel's chunk loop (option gram_wide, the default) and the general tiled
    kernel."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(3000, 24, 20, M, seed=9, pca=False)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0,
                      T0, ard_only=True, precision=lib.PREC_FP64)
    g.init_fit()
    for jtr in range(2):
        g.irls_step(jtr)
    g.ql_start()
    J, P = ds.J, ds.P
    iu = np.triu_indices(P, 1)
    g.set_precision(lib.PREC_FP64_WHITENED)
    e64 = g.elbo(1)
    Pw = g.get("Psi")[:J * P * P].reshape(J, P, P)[:, iu[1], iu[0]]
    g.grad()
    g64 = g.get("grad")
    g.set_precision(lib.PREC_MIXED)
    g.set_option("gram_wide", wide)
    em = g.elbo(1)
    Pm = g.get("Psi")[:J * P * P].reshape(J, P, P)[:, iu[1], iu[0]]
    g.grad()
    gm = g.get("grad")
    err = np.max(np.abs(Pm - Pw), axis=1) / np.max(np.abs(Pw), axis=1)
    ge = rel(gm, g64)
    print(f"M={M} gram_wide={wide} mixed: Gram max rel err {err.max():.2e}, "
          f"ELBO {abs(em - e64) / abs(e64):.2e}, gradient {ge:.2e}")
    assert err.max() < 5e-7
    assert abs(em - e64) <= 1e-7 * abs(e64)
    assert ge < 1e-4
    g.close()
