"""Pins for the FP64 CPU oracle: every check here is independent of the
oracle's own formulas (LAPACK/numpy routines, closed forms, finite
differences, the Woodbury identity, fixed-point conditions, published KATs)."""
import ctypes as C

import numpy as np
import pytest
import scipy.stats

from conftest import make_oracle, rel
from oracle import oracle as O


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _spd(n, seed, cond=1e3):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0, np.log10(cond), n)
    return (Q * ev) @ Q.T


# ------------------------------------------------------------------ chol.cu
@pytest.mark.parametrize("n", [1, 5, 37, 64])
def test_choldc_matches_lapack(n):
    lib = O.load()
    A = _spd(n, n)
    a = np.asfortranarray(np.triu(A)).ravel(order="F").copy()
    assert lib.orc_choldc(_dp(a), n) == 0
    out = a.reshape(n, n, order="F")
    L = np.linalg.cholesky(A)
    assert rel(np.tril(out), L) < 1e-12
    # strict upper triangle keeps the input (chol.cu writes lower+diag only)
    iu = np.triu_indices(n, 1)
    assert np.array_equal(out[iu], A[iu])


def test_choldc_indefinite_zero_fills():
    lib = O.load()
    A = _spd(6, 1)
    A[3, 3] = -5.0
    a = np.asfortranarray(np.triu(A)).ravel(order="F").copy()
    assert lib.orc_choldc(_dp(a), 6) == 1
    assert not a.any()


def test_cholsl_and_getRb_match_numpy():
    lib = O.load()
    n = 23
    A = _spd(n, 5)
    a = np.asfortranarray(np.triu(A)).ravel(order="F").copy()
    lib.orc_choldc(_dp(a), n)
    b = np.random.default_rng(0).standard_normal(n)
    x = np.zeros(n)
    lib.orc_cholsl(_dp(a), n, _dp(b), _dp(x))
    assert rel(x, np.linalg.solve(A, b)) < 1e-11
    L = np.linalg.cholesky(A)
    bb = b.copy()
    lib.orc_getRb(_dp(a), _dp(bb), n)
    assert rel(bb, L.T @ b) < 1e-12
    K = _spd(n, 9)
    Ku = np.asfortranarray(np.triu(K)).ravel(order="F").copy()
    assert abs(lib.orc_quadform(_dp(Ku), _dp(b), n) - b @ K @ b) < 1e-10 * abs(b @ K @ b)
    v = np.array([2.0, 4.0, 9.0])
    lib.orc_mscale(3, _dp(v), 10.0)
    assert abs(v.mean() - 10.0) < 1e-12 and abs(v[1] / v[0] - 2.0) < 1e-14


# ------------------------------------------------------------------ kernel_funcs.cu
def _kmats(X, T, rho, ard=0):
    lib = O.load()
    N, Q = X.shape
    M = T.shape[0]
    Xf = np.asfortranarray(X).ravel(order="F").copy()
    Tf = np.asfortranarray(T).ravel(order="F").copy()
    Knm = np.zeros(N * M)
    K = np.zeros(M * M)
    Ki = np.zeros(M * M)
    lib.orc_kernel_mats(N, M, Q, _dp(Xf), _dp(Tf), _dp(np.asarray(rho, float)), ard, _dp(Knm), _dp(K), _dp(Ki))
    return Knm.reshape(N, M, order="F"), K.reshape(M, M, order="F")


def test_kernel_closed_forms():
    rho = np.array([5.0, 10.0, 2.0])
    T = np.array([[0.3, 0.1, -0.2], [1.0, 2.0, 0.5]])
    X = np.array([[0.3, 0.1, -0.2],                      # equal to T0 -> 1
                  [0.3 + 2 * np.pi, 0.1, -0.2],          # period 2 pi in dim 0 -> 1
                  [0.3 + np.pi, 0.1, -0.2],              # sin^2(pi/2) = 1 -> exp(-1/rho0)
                  [0.3, 0.1 + np.sqrt(2 * 10.0), -0.2],  # d^2/(2 rho) = 1 -> e^-1
                  [0.3, 0.1, -0.2 + 2.0]])               # 4/(2*2) = 1 -> e^-1
    Knm, K = _kmats(X, T, rho)
    assert abs(Knm[0, 0] - 1) < 1e-15
    assert abs(Knm[1, 0] - 1) < 1e-14
    assert abs(Knm[2, 0] - np.exp(-1 / 5.0)) < 1e-14
    assert abs(Knm[3, 0] - np.exp(-1)) < 1e-14
    assert abs(Knm[4, 0] - np.exp(-1)) < 1e-14
    # Kmm: unit diagonal, lower triangle zero, upper = k(T_a, T_b)
    assert K[0, 0] == 1 and K[1, 1] == 1 and K[1, 0] == 0
    assert abs(K[0, 1] - Knm[0, 1]) < 1e-15


def test_kernel_factorises_over_dims():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((7, 4))
    T = rng.standard_normal((3, 4))
    rho = rng.uniform(0.5, 3, 4)
    full, _ = _kmats(X, T, rho)
    prod = np.ones_like(full)
    for q in range(4):
        r = np.ones(4) * 1e300
        r[q] = rho[q]
        Xq = np.zeros_like(X); Tq = np.zeros_like(T)
        Xq[:, q] = X[:, q]; Tq[:, q] = T[:, q]
        kq, _ = _kmats(Xq, Tq, np.where(np.arange(4) == q, rho, 1.0))
        prod *= kq
    assert rel(full, prod) < 1e-13
    # SE-ARD-only variant: dim 0 is squared-exponential too
    ard, _ = _kmats(X, T, rho, ard=1)
    d = sum((X[:, [q]] - T[None, :, q]) ** 2 / (2 * rho[q]) for q in range(4))
    assert rel(ard, np.exp(-d)) < 1e-13


# ------------------------------------------------------------------ DataFrame.cu
def test_dataframe_expansion_matches_one_hot(small_ds):
    lib = O.load()
    ds = small_ds
    ncov = len(ds.c)
    nh = np.zeros(ncov + 3, np.int32)
    P = lib.orc_dataframe_layout(ncov, ds.c.ctypes.data_as(C.POINTER(C.c_int)),
                                 ds.v.ctypes.data_as(C.POINTER(C.c_int)), ds.M,
                                 nh.ctypes.data_as(C.POINTER(C.c_int)))
    assert P == ds.P and list(nh) == list(ds.nh)
    Z = np.zeros(ds.N * nh[-2])
    lib.orc_dataframe_expand(ds.N, ncov, ds.c.ctypes.data_as(C.POINTER(C.c_int)),
                             ds.v.ctypes.data_as(C.POINTER(C.c_int)), ds.body,
                             nh.ctypes.data_as(C.POINTER(C.c_int)), _dp(Z))
    assert np.array_equal(Z.reshape(ds.N, -1, order="F"), ds.Zcov)


# ------------------------------------------------------------------ Gram, likelihood
def _dense_delta(o):
    P = o.P
    D = o.get("DeltaInv").reshape(P, P, order="F")
    return np.triu(D) + np.triu(D, 1).T


def test_gram_matches_dense(small_ds):
    o = make_oracle(small_ds)
    o.init_fit()
    o.call("orc_get_w_for_beta")
    o.call("orc_get_psi")
    N, J, P = o.N, o.J, o.P
    Z = o.get("Z").reshape(N, P, order="F")
    W = o.get("W")[:N * J].reshape(N, J, order="F")
    Om = o.get("Omega")[:N]
    D = _dense_delta(o)
    Psi = o.get("Psi")[:J * P * P].reshape(J, P, P).transpose(0, 2, 1)
    for j in (0, 7, J - 1):
        G = Z.T @ (Z * (W[:, j] / Om)[:, None]) + D
        assert rel(np.triu(Psi[j]), np.triu(G)) < 1e-12
        assert not np.tril(Psi[j], -1).any()


def test_lkhd_equals_dense_gaussian_logpdf(tiny_ds):
    """getLkhdj == log N(nu_j; 0, phi_j (diag(Omega/W_j) + Z Delta Z^T)) + N/2 log 2pi."""
    o = make_oracle(tiny_ds)
    o.init_fit()
    for jtr in range(3):
        o.irls_step(jtr)
    o.ql_start()
    N, J, P = o.N, o.J, o.P
    Z = o.get("Z").reshape(N, P, order="F")
    W = o.get("W")[:N * J].reshape(N, J, order="F")
    Nu = o.get("Nu").reshape(N, J, order="F")
    Om = o.get("Omega")[:N]
    phi = o.get("phi")
    o.call("orc_expand_delta_k")
    Delta = np.linalg.inv(_dense_delta(o))
    lk = o.get("lkhds")
    for j in range(J):
        C_ = phi[j] * (np.diag(Om / W[:, j]) + Z @ Delta @ Z.T)
        ref = scipy.stats.multivariate_normal(np.zeros(N), C_).logpdf(Nu[:, j]) + N / 2 * np.log(2 * np.pi)
        assert abs(lk[j] - ref) < 1e-8 * abs(ref)


def test_irls_fixed_point(tiny_ds):
    """At the IRLS fixed point  Z^T ((y - W)/Omega) = DeltaInv beta."""
    o = make_oracle(tiny_ds)
    o.init_fit()
    for jtr in range(160):
        o.irls_step(jtr)
    o.call("orc_get_w_for_beta")
    N, J, P = o.N, o.J, o.P
    Z = o.get("Z").reshape(N, P, order="F")
    W = o.get("W")[:N * J].reshape(N, J, order="F")
    Om = o.get("Omega")[:N]
    B = o.get("Beta").reshape(P, J, order="F")
    D = _dense_delta(o)
    ds = tiny_ds
    from gaspacho2_amd.synth import dense_counts
    Y = dense_counts(ds)
    lhs = Z.T @ ((Y - W) / Om[:, None])
    rhs = D @ B
    assert rel(lhs, rhs) < 1e-6


def test_glm_objective_is_poisson_loglik_with_titsias_term(tiny_ds):
    """getLkhdPois (ggplvm.cu:703-724) per gene equals the Poisson log-pmf of
    the counts at rate W plus log(y!), minus wt*delta_GP*sum_i t_i y_ij: the
    titsias factor leaves the log term only (the -W term keeps it)."""
    import scipy.special
    import scipy.stats
    from gaspacho2_amd.synth import dense_counts
    o = make_oracle(tiny_ds)
    o.init_fit()
    o.irls_step(0)
    o.irls_step(1)
    N, J = o.N, o.J
    W = o.get("W")[:N * J].reshape(N, J, order="F")
    t = o.get("titsias")[:N]
    dgp = o.get("delta")[len(tiny_ds.nh) - 2]
    Y = dense_counts(tiny_ds)
    wt = 0.1                                  # ggplvm.cu:20
    ref = (scipy.stats.poisson.logpmf(Y, W) + scipy.special.gammaln(Y + 1)).sum(0) \
        - wt * dgp * (t[:, None] * Y).sum(0)
    lk = o.get("lkhds")[:J]
    assert np.max(np.abs(lk - ref)) < 1e-10 * np.max(np.abs(ref))
    # a plausible slip (titsias factor removed from -W too) fails the check
    alt = ref - (W * (np.exp(-wt * dgp * t)[:, None] - 1)).sum(0)
    assert np.max(np.abs(lk - alt)) > 1e-6 * np.max(np.abs(ref))


# ------------------------------------------------------------------ gradient
def _fd(o, name, idx, eps, log=False):
    v = o.get(name)
    v0 = v.copy()
    v[idx] = v0[idx] * np.exp(eps) if log else v0[idx] + eps
    o.set(name, v)
    ep = o.elbo(1)
    v[idx] = v0[idx] * np.exp(-eps) if log else v0[idx] - eps
    o.set(name, v)
    em = o.elbo(1)
    o.set(name, v0)
    return (ep - em) / (2 * eps)


@pytest.mark.parametrize("ard", [0, 1])
def test_gradient_matches_finite_differences(tiny_ds, ard):
    o = make_oracle(tiny_ds, ard_only=bool(ard))
    o.init_fit()
    for jtr in range(4):
        o.irls_step(jtr)
    o.ql_start()
    o.grad()
    g = o.get("grad")
    N, Q, M, p = o.N, o.Q, o.M, o.p
    o.elbo(1)
    checks = [("Omega", i, i, True) for i in (0, 11, N - 1)]
    checks += [("delta", l, N + l, True) for l in range(p)]
    checks += [("X", a, N + p + a, False) for a in (2, N + 5, 2 * N + 9)]
    checks += [("T", a, N + p + N * Q + a, False) for a in (1, M + 3, 2 * M + 4)]
    checks += [("rho", q, N + p + N * Q + M * Q + q, True) for q in range(Q)]
    for name, idx, gi, lg in checks:
        fd = _fd(o, name, idx, 1e-6, log=lg)
        assert abs(fd - g[gi]) <= 1e-5 * max(1.0, abs(g[gi])), (name, idx, fd, g[gi])


# ------------------------------------------------------------------ L-BFGS
def test_lbfgs_two_loop_equals_dense_bfgs():
    """Two-loop recursion == H_k g with the dense inverse-BFGS update
    H <- (I - r s y^T) H (I - r y s^T) + r s s^T from H0 = gamma I."""
    lib = O.load()
    n, hist = 9, 10
    rng = np.random.default_rng(2)
    A = _spd(n, 3, cond=20)
    for itr in (1, 4, 10, 13):
        S = np.zeros((n, hist)); Y = np.zeros((n, hist))
        pairs = []
        for k in range(itr):
            s = rng.standard_normal(n)
            y = A @ s
            S[:, k % hist] = s; Y[:, k % hist] = y
            pairs.append((s, y))
        g = 1e-3 * rng.standard_normal(n)
        work = np.zeros(n + hist)
        z = np.zeros(n)
        lib.orc_lbfgs_two_loop(_dp(np.asfortranarray(S).ravel(order="F").copy()),
                               _dp(np.asfortranarray(Y).ravel(order="F").copy()), _dp(g), n, hist,
                               itr, _dp(work), _dp(z))
        use = pairs[-hist:]
        s_l, y_l = use[-1]
        H = (s_l @ y_l) / (y_l @ y_l) * np.eye(n)
        for s, y in use:
            r = 1.0 / (y @ s)
            V = np.eye(n) - r * np.outer(y, s)
            H = V.T @ H @ V + r * np.outer(s, s)
        assert rel(z, -np.clip(H @ g, -1, 1)) < 1e-10


def _theta(o, rhoflag):
    """The L-BFGS coordinates of ggplvm.cu:865-934: [log Omega | log delta | X | T | log rho_0]."""
    N, p, Q, M = o.N, o.p, o.Q, o.M
    parts = [np.log(o.get("Omega")[:N]), np.log(o.get("delta")[:p]), o.get("X")[:N * Q],
             o.get("T")[:M * Q]]
    if rhoflag:
        parts.append(np.log(o.get("rho")[:1]))
    return np.concatenate(parts)


def _dense_bfgs_direction(pairs, g):
    """H g for the inverse-BFGS matrix built from the last 10 (s, y) pairs,
    H0 = (s'y / y'y) I of the newest pair."""
    use = pairs[-10:]
    s_l, y_l = use[-1]
    H = (s_l @ y_l) / (y_l @ y_l) * np.eye(g.size)
    for s, y in use:
        r = 1.0 / (y @ s)
        V = np.eye(g.size) - r * np.outer(y, s)
        H = V.T @ H @ V + r * np.outer(s, s)
    return H @ g


@pytest.mark.parametrize("rhoflag", [0, 1])
def test_lbfgs_update_and_downdate_follow_dense_bfgs_on_a_quadratic(tiny_ds, rhoflag):
    """orc_lbfgs + orc_downdate (ggplvm.cu:818-934, util_cuda.cu:187-232,
    546-555) driven by the exact gradient of a concave quadratic in the
    optimiser's coordinates reproduce dense BFGS ascent on the displacements
    actually taken: first step clip(g/1e5, +-1), then clip(H g, +-1) with H from
    the last 10 pairs (14 steps: the ring wraps), Downdate = the point at
    fraction sr of the step from the backup (sr = 0.1, 0.1 * 0.01, 0.01),
    the history slot holding sr times the full step.  rho_1..rho_{Q-1} are
    rescaled to mean 10 with their ratios kept; rho_0 moves only with
    rhoflag = 1; the backup halves hold the values before the step."""
    o = make_oracle(tiny_ds)
    N, p, Q, M = o.N, o.p, o.Q, o.M
    rng = np.random.default_rng(11)
    o.set("Omega", np.concatenate([rng.uniform(0.5, 2.0, N), np.ones(N)]))
    o.set("delta", np.concatenate([rng.uniform(0.05, 2.0, p), np.ones(p)]))
    rho0 = np.array([5.0] + list(np.linspace(2.0, 6.0, Q - 1)))
    o.set("rho", np.concatenate([rho0, np.zeros(Q)]))
    G = o.view("grad").size
    tot = N + p + N * Q + M * Q + rhoflag
    th = _theta(o, rhoflag)
    assert th.size == tot
    Qm, _ = np.linalg.qr(rng.standard_normal((tot, tot)))
    A = (Qm * np.logspace(0, np.log10(40.0), tot)) @ Qm.T
    th_star = th + 0.4 * rng.standard_normal(tot)

    def grad(t):
        g = np.zeros(G)
        g[:tot] = -A @ (t - th_star)
        return g

    downdates = {5: [0.1], 8: [0.1, 0.01], 12: [0.01]}
    pairs, g_prev, s_prev = [], None, None
    for itr in range(14):
        g = grad(th)[:tot]
        o.set("grad", grad(th))
        before = {k: o.get(k) for k in ("Omega", "delta", "X", "T")}
        if itr == 0:
            step = np.clip(g / 1e5, -1, 1)
        else:
            pairs.append((s_prev, g - g_prev))
            step = np.clip(_dense_bfgs_direction(pairs, g), -1, 1)
        o.lbfgs_step(itr, rhoflag)
        assert rel(_theta(o, rhoflag), th + step) < 1e-12, itr
        assert rel(o.get("ss")[:tot], -step) < 1e-12, itr
        for k, n in (("Omega", N), ("delta", p), ("X", N * Q), ("T", M * Q)):
            assert np.array_equal(o.get(k)[n:], before[k][:n]), (itr, k)
        sr = 1.0
        for f in downdates.get(itr, []):
            sr *= f
            o.downdate(itr, sr, rhoflag)
            assert rel(_theta(o, rhoflag), th + sr * step) < 1e-12, (itr, sr)
            for k, n in (("Omega", N), ("delta", p), ("X", N * Q), ("T", M * Q)):
                assert np.array_equal(o.get(k)[n:], before[k][:n]), (itr, k)
        slot = o.get("SS")[(itr % 10) * tot:(itr % 10 + 1) * tot]
        assert rel(slot, -sr * step) < 1e-12, itr
        s_prev = -sr * step
        g_prev = g
        th = th + sr * step
        rho = o.get("rho")[:Q]
        assert abs(np.mean(rho[1:]) - 10.0) < 1e-12
        assert rel(rho[1:] / rho[1], rho0[1:] / rho0[1]) < 1e-13
        if not rhoflag:
            assert abs(rho[0] - 5.0) < 1e-13
    if rhoflag:
        assert abs(o.get("rho")[0] - 5.0) > 1e-3


def test_lbfgs_clamps_omega_and_clips_the_first_step(tiny_ds):
    """Omega is clamped to [1e-3, 1e3] after the update and after Downdate
    (ggplvm.cu:829-830, 902-903); the first step is clip(-g/1e5, +-1)
    component by component (ggplvm.cu:872-880)."""
    o = make_oracle(tiny_ds)
    N = o.N
    Om = np.ones(2 * N)
    Om[:3] = [900.0, 0.002, 1.0]
    o.set("Omega", Om)
    g = np.zeros(o.view("grad").size)
    g[:3] = [2e5, -2e5, 3e4]
    g[3] = -5e6
    o.set("grad", g)
    o.lbfgs_step(0, 0)
    ss = o.get("ss")
    assert np.array_equal(ss[:4], [-1.0, 1.0, -0.3, 1.0])
    Omn = o.get("Omega")
    assert Omn[0] == 1000.0 and Omn[1] == 0.001
    assert abs(Omn[2] - np.exp(0.3)) < 1e-14
    assert abs(Omn[3] - np.exp(-1.0)) < 1e-15
    assert np.array_equal(Omn[N:N + 3], [900.0, 0.002, 1.0])
    o.downdate(0, 0.5, 0)
    Omd = o.get("Omega")
    assert Omd[0] == 1000.0                               # 900 e^0.5 = 1484 -> clamped
    assert abs(Omd[1] - 0.002 * np.exp(-0.5)) < 1e-18     # inside the bounds
    assert abs(Omd[2] - np.exp(0.15)) < 1e-14


def test_fit_schedule_follows_the_reference_rules(tiny_ds):
    """orc_fit's trace obeys the schedule of ggplvm.cu:1146-1330 over 31 outer
    iterations with no cap: JTRMAX 200 at outer iteration 0, 100 after, and
    10000 for the L-BFGS phase of iteration 30 (1152, 1213); an IRLS phase
    stops at the first step with relative change < 1e-7 or at JTRMAX
    (1180-1185); one quasi-likelihood start follows; each L-BFGS step is
    accepted when its objective rises or at the 7th trial, after one Downdate
    per failed trial with sr multiplied by 0.1 (lower) or 0.01 (non-finite)
    (1280-1289); the phase stops at |change| < 1e-8 |objective| or at JTRMAX
    (1292); rho_0 is optimised only after outer iteration 20 (1212).  The run
    reaches every rule: natural IRLS and L-BFGS stops, both JTRMAX caps, the
    10000-step phase and Downdates."""
    from conftest import check_fit_schedule
    o = make_oracle(tiny_ds)
    o.fit(21, 0)
    assert abs(o.get("rho")[0] - 5.0) < 1e-13
    o = make_oracle(tiny_ds)
    last = o.fit(31, 0)
    assert abs(o.get("rho")[0] - 5.0) > 1e-6
    obj, sr, info = o.fit_trace()
    seen = check_fit_schedule(obj, info, sr, 31, 0)
    assert last == obj[-1]
    assert seen["irls_converged"] > 0 and seen["lbfgs_converged"] > 0
    assert seen["lbfgs_cap"] > 0 and seen["downdates"] > 0 and seen["long_phase_steps"] > 100


# ------------------------------------------------------------------ util.c
def test_mt19937_published_kat():
    # 10000th output of MT19937 seeded with 5489 (C++ std::mt19937 requirement)
    assert O.load().orc_mt19937_kat(5489, 10000) == 4123659995


def test_gsl_mt19937_seed0_published_first_value():
    """GSL reference manual, random number generator example: the default
    generator (mt19937) at the default seed 0 prints "first value = 4293858116".
    initTaWithPer (util.c:73-87) seeds with 0, which GSL maps to 4357; the oracle
    seeds its stream with 4357."""
    assert O.load().orc_mt19937_kat(4357, 1) == 4293858116


def test_gsl_uniform_published_poisson_example():
    """GSL reference manual, random number distribution example: gsl_ran_poisson
    with mu = 3 on the default mt19937 generator at seed 0 prints
    2 5 5 2 1 0 3 4 1 1.  For mu < 10 GSL multiplies uniforms until the product
    falls to exp(-mu); replaying that on the oracle's stream with the oracle's
    uniform (output / 2^32) must give the published draws, which pins the
    uniform conversion the polar Gaussian of initTaWithPer consumes."""
    lib = O.load()
    stream = iter(lib.orc_mt19937_kat(4357, n) for n in range(1, 200))
    emu, draws = np.exp(-3.0), []
    for _ in range(10):
        prod, k = 1.0, 0
        while True:
            prod *= next(stream) / 4294967296.0
            k += 1
            if not prod > emu:
                break
        draws.append(k - 1)
    assert draws == [2, 5, 5, 2, 1, 0, 3, 4, 1, 1]


def test_init_ta_with_per_gaussian_distribution():
    """The N(0,1) inducing-point coordinates (gsl_ran_gaussian, util.c:82) follow
    the standard normal: a Kolmogorov-Smirnov test over 60,000 draws and a lag-1
    correlation near zero.  A dropped -2, log or square root in the polar
    transform moves the KS statistic far past the bound."""
    from scipy import stats
    T = O.init_ta_with_per(20000, 4).reshape(20000, 4, order="F")
    g = T[:, 1:].ravel(order="F")
    ks = stats.kstest(g, "norm")
    assert ks.pvalue > 1e-3, ks
    assert abs(np.corrcoef(g[:-1], g[1:])[0, 1]) < 0.02


def test_init_ta_with_per_layout():
    T = O.init_ta_with_per(7, 3).reshape(7, 3, order="F")
    assert np.allclose(T[:, 0], np.arange(1, 8) / 8 * 2 * 3.14159265358979)
    assert abs(T[:, 1:].mean()) < 1.5 and T[:, 1:].std() > 0.3


# ------------------------------------------------------------------ score tests
def test_score_stats_match_dense_woodbury_form(tiny_ds):
    """getScoreStats (ggplvm_de.cu:491-532): the variance term
    g^T D g - (Z^T D g)^T Psi^-1 (Z^T D g) equals g^T (D^-1 + Z Delta Z^T)^-1 g
    (Woodbury, Psi = Delta^-1 + Z^T D Z, D = diag(W/Omega)), checked with a dense
    N x N solve; the numerator is g^T Omega^-1 (y - W)."""
    from gaspacho2_amd.synth import dense_counts
    o = make_oracle(tiny_ds)
    o.init_fit()
    for jtr in range(3):
        o.irls_step(jtr)
    N, J, P = o.N, o.J, o.P
    rng = np.random.default_rng(2)
    S = 4
    G = rng.integers(0, 3, (N, S)).astype(np.float64)
    gene = np.repeat(np.arange(J), S).astype(np.int32)
    snp = np.tile(np.arange(S), J).astype(np.int32)
    stat = o.score_stats(G, gene, snp)
    Z = o.get("Z").reshape(N, P, order="F")
    W = o.get("W")[:N * J].reshape(N, J, order="F")
    Om = o.get("Omega")[:N]
    phi = o.get("phi")[:J]
    Delta = np.linalg.inv(_dense_delta(o))
    Y = dense_counts(tiny_ds)
    for t in range(len(gene)):
        j, l = gene[t], snp[t]
        g = G[:, l]
        V = np.diag(Om / W[:, j]) + Z @ Delta @ Z.T
        var = g @ np.linalg.solve(V, g)
        num = g @ ((Y[:, j] - W[:, j]) / Om)
        ref = num / np.sqrt(phi[j] * var)
        assert abs(stat[t] - ref) <= 1e-8 * max(1.0, abs(ref)), (t, stat[t], ref)


# ---------------- amortization encoder oracle (oracle/encoder.py) ----------------

def _enc_case(seed=0, F=10, H1=8, H2=6, Q=2, n=5):
    from oracle import encoder as E
    rng = np.random.default_rng(seed)
    p = {"W1": rng.normal(0, 0.4, (H1, F)), "b1": rng.normal(0, 0.1, H1),
         "W2": rng.normal(0, 0.4, (H2, H1)), "b2": rng.normal(0, 0.1, H2),
         "W3": rng.normal(0, 0.4, (2 * Q, H2)), "b3": rng.normal(0, 0.1, 2 * Q)}
    A0 = np.log1p(rng.poisson(2.0, (n, F)) / rng.lognormal(0, 0.5, n)[:, None])
    X = rng.normal(0, 1, (n, Q))
    return E, p, A0, X


def test_encoder_oracle_gradient_matches_finite_differences():
    E, p, A0, X = _enc_case()
    _, g = E.loss_grad(p, A0, X, rnd=False)
    rng = np.random.default_rng(1)
    h = 1e-6
    for name, w in p.items():
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in w.shape)
            w0 = w[idx]
            w[idx] = w0 + h
            lp, _ = E.loss_grad(p, A0, X, rnd=False)
            w[idx] = w0 - h
            lm, _ = E.loss_grad(p, A0, X, rnd=False)
            w[idx] = w0
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[name][idx]) <= 1e-7 * max(1.0, abs(fd)), (name, idx, fd, g[name][idx])


def test_encoder_oracle_log_variance_clamp_has_zero_gradient_outside():
    E, p, A0, X = _enc_case(seed=5)
    Q = X.shape[1]
    p["b3"][Q] = 14.0            # first log-variance output above the clamp
    p["b3"][Q + 1] = -14.0       # second below it
    loss, g = E.loss_grad(p, A0, X, rnd=False)
    for k in (Q, Q + 1):
        h = 1e-6
        p["b3"][k] += h
        lp, _ = E.loss_grad(p, A0, X, rnd=False)
        p["b3"][k] -= 2 * h
        lm, _ = E.loss_grad(p, A0, X, rnd=False)
        p["b3"][k] += h
        assert lp == lm == loss and g["b3"][k] == 0.0
    # the clamped value enters the loss: 1/2 (r^2 e^-10 + 10) for the first, e^+10 for the second
    f = E.forward(p, A0, rnd=False)["z3"]
    r0, r1 = X[:, 0] - f[:, 0], X[:, 1] - f[:, 1]
    rest = 0.5 * np.sum((X[:, 2:] - f[:, 2:Q]) ** 2 * np.exp(-f[:, Q + 2:]) + f[:, Q + 2:])
    expect = (0.5 * np.sum(r0 ** 2 * np.exp(-10.0) + 10.0) + 0.5 * np.sum(r1 ** 2 * np.exp(10.0) - 10.0)
              + rest) / X.shape[0]
    assert abs(loss - expect) <= 1e-12 * abs(expect)


def test_encoder_oracle_gradient_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    E, p, A0, X = _enc_case(seed=2, F=30, H1=16, H2=12, Q=3, n=9)
    loss, g = E.loss_grad(p, A0, X, rnd=False)
    t = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in p.items()}
    a0 = torch.tensor(A0)
    x = torch.tensor(X)
    gl = torch.nn.functional.gelu
    z3 = gl(gl(a0 @ t["W1"].T + t["b1"]) @ t["W2"].T + t["b2"]) @ t["W3"].T + t["b3"]
    Q = X.shape[1]
    mu, s = z3[:, :Q], z3[:, Q:]
    lt = 0.5 * ((x - mu) ** 2 * torch.exp(-s) + s).sum() / X.shape[0]
    lt.backward()
    assert abs(lt.item() - loss) <= 1e-13 * abs(loss)
    for k in p:
        assert np.max(np.abs(t[k].grad.numpy() - g[k])) <= 1e-12 * max(1.0, np.max(np.abs(g[k]))), k


def test_encoder_oracle_bf16_rounding_is_round_to_nearest_even():
    torch = pytest.importorskip("torch")
    from oracle import encoder as E
    one = 1.0
    assert E.bf16(one) == 1.0
    assert E.bf16(1.0 + 2.0 ** -8) == 1.0                         # tie -> even mantissa
    assert E.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6         # tie -> even mantissa
    assert E.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    x = np.random.default_rng(3).normal(0, 10, 100000).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(E.bf16(x), ref)


def test_encoder_oracle_adam_first_step_moves_each_weight_by_lr():
    from oracle import encoder as E
    rng = np.random.default_rng(4)
    w, g = rng.normal(size=50), rng.normal(size=50)
    w1, m, v = E.adam(w, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    # first step: m/(1-b1) = g, v/(1-b2) = g^2 -> update = lr g / (|g| + eps)
    assert np.allclose(w - w1, 1e-3 * np.sign(g) * np.abs(g) / (np.abs(g) + 1e-8), rtol=1e-9, atol=0)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


# ---------------- PCA initial latent (oracle/pca.py) ----------------

def test_pca_oracle_closed_forms(small_ds):
    """X^T X = N I; X spans the top-Q left singular vectors of nu0 (numpy SVD)
    with the eigenvalues the squared singular values; nu0 is orthogonal to the
    covariate columns (least-squares residual)."""
    from oracle import pca
    ds = small_ds
    N, Q = ds.N, 3
    ncol = int(ds.nh[-2])
    Z = np.asarray(ds.Zcov).reshape(N, -1, order="F")[:, :ncol]
    X, lam = pca.pca_init(N, ds.Yp, ds.Yi, ds.Yx, ds.ls, Z, Q)
    assert np.allclose(X.T @ X, N * np.eye(Q), atol=1e-8 * N)
    nu = pca.nu0(N, ds.Yp, ds.Yi, ds.Yx, ds.ls, Z)
    assert np.max(np.abs(Z.T @ nu)) < 1e-9 * np.max(np.abs(nu)) * N
    U, sv, _ = np.linalg.svd(nu, full_matrices=False)
    assert np.allclose(lam, sv[:Q] ** 2, rtol=1e-10)
    # same subspace, columns equal up to sign
    for k in range(Q):
        u = U[:, k] * np.sqrt(N)
        assert min(np.max(np.abs(X[:, k] - u)), np.max(np.abs(X[:, k] + u))) < 1e-7 * np.sqrt(N)
    # the sign convention: the projection direction's largest entry is positive
    V = nu.T @ X                                     # proportional to the eigenvectors
    for k in range(Q):
        v = V[:, k]
        assert v[np.argmax(np.abs(v))] > 0
