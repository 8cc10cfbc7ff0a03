"""CPU checks of the learned-boundary test infrastructure: the FP64 network
reference's gradient against finite differences, its bf16 emulation against
its f64 form, the FP64 1D plane-wave oracle against the C oracle's 3D run,
and the thin domain's place beside its twin's CPML."""
import numpy as np

from oracle import fdtd1d, mlp_ref, wfo
from wavefront_amd import learned as L
from wavefront_amd.layout import MaterialTable


def _params(rng, scale3=0.5):
    p = np.zeros(mlp_ref.P)
    d = mlp_ref.split(p)
    d = {k: v.copy() for k, v in d.items()}
    d["W1"] = rng.normal(0, 0.25, (16, 64))
    d["b1"] = rng.normal(0, 0.1, 64)
    d["W2"] = rng.normal(0, 0.125, (64, 64))
    d["b2"] = rng.normal(0, 0.1, 64)
    d["W3"] = rng.normal(0, scale3, 64)
    d["b3"] = np.array([0.05])
    d["wlin"] = rng.normal(0, 0.3, 16)
    return mlp_ref.join(d)


def test_gradient_formulas_match_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(rng)
    u = rng.normal(0, 1, (40, 16))
    t = rng.normal(0, 1, 40)
    g = mlp_ref.grad_exact(p, u, t)
    idx = np.concatenate([rng.choice(mlp_ref.P, 60, replace=False), [mlp_ref.P - 17]])
    fd = mlp_ref.grad_fd(p, u, t, idx, h=1e-6)
    np.testing.assert_allclose(g[idx], fd, rtol=2e-5, atol=1e-9)


def test_delta_head_gradient_matches_finite_differences():
    """The delta head's FP64 gradient (unrounded weights) against central
    differences of its loss, over MLP weights, biases and the skip."""
    rng = np.random.default_rng(5)
    p = _params(rng)
    u = rng.normal(0, 1, (40, 16))
    t = rng.normal(0, 1, 40)
    g = mlp_ref.grad_delta_f64(p, u, t, rounded=False)
    idx = np.concatenate([rng.choice(mlp_ref.P, 60, replace=False),
                          [mlp_ref.P - 17, mlp_ref.P - 1, mlp_ref.P - 18]])
    h = 1e-6
    fd = []
    for i in idx:
        q1, q2 = p.copy(), p.copy()
        q1[i] += h
        q2[i] -= h
        fd.append((mlp_ref.loss_delta_f64(q1, u, t, False) - mlp_ref.loss_delta_f64(q2, u, t, False))
                  / (2 * h))
    np.testing.assert_allclose(g[idx], fd, rtol=2e-5, atol=1e-9)
    # a static input (P0(n) = P0(n-1)) gets the skip alone
    us = u.copy()
    us[:, 14] = us[:, 0]
    np.testing.assert_allclose(mlp_ref.predict_delta_f64(p, us, False),
                               us @ mlp_ref.split(p)["wlin"], rtol=0, atol=1e-12)


def test_bf16_rounding_and_emulation():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert mlp_ref.bf16(x)[0] == 1.0 and mlp_ref.bf16(x)[1] == 1.0
    assert mlp_ref.bf16(x)[2] == np.float32(1.0078125)
    rng = np.random.default_rng(2)
    p = _params(rng).astype(np.float32)
    u = rng.normal(0, 1, (200, 16)).astype(np.float32)
    y64, _ = mlp_ref.forward_f64(p, u)
    yb = mlp_ref.forward_bf16(p, u)
    assert np.max(np.abs(yb - y64)) < 0.05 * np.max(np.abs(y64))


def test_fp64_plane_wave_oracle_matches_c_oracle():
    """The 1D FP64 reduction against the C oracle's 3D periodic run (f32)."""
    dx, dt, freq = L.study_grid()
    nx, xs, xp, steps = 64, 20, 40, 150
    table = MaterialTable.vacuum_filled(dt, dx)
    sim = wfo.OracleSim(nx, 8, 8, table=wfo.vacuum_table(dt, dx), periodic=6)
    for ly in range(8):
        for lz in range(8):
            sim.add_source(xs, ly, lz, 2, 1.0, 1, freq, 0.0, hard=True)
    sim.add_probe(xp, 0, 0, 2, [], 0.0)
    tr32 = sim.run(steps, dt, trace=True)[:, 0]
    tr64 = fdtd1d.plane_wave_trace_f64(nx, table.q16[0], dt, freq, xs, xp, steps)
    assert np.max(np.abs(tr32 - tr64)) < 1e-5 * np.max(np.abs(tr64))
    assert np.max(np.abs(tr64)) > 0.1


def test_thin_domain_lies_beyond_the_twin_cpml():
    """The thin domain starts right after the twin's CPML on each learned
    axis: the twin's profile is graded in its first THICK cells (and last)
    and identity on every cell of the thin domain, so a thin run with exact
    ghosts is the twin's run restricted to the thin domain."""
    dx, dt, _ = L.study_grid()
    n = 152
    nb = n + 2 * L.OFF
    thick = L.thick_profile(nb, dt, dx)
    assert L.OFF == L.THICK
    for p in range(L.THICK):
        assert thick.get(p) != (1.0, 0.0, 1.0) and thick.get(nb - 1 - p) != (1.0, 0.0, 1.0)
    assert all(thick.get(L.OFF + i) == (1.0, 0.0, 1.0) for i in range(n))


def _higdon_params(rng, kf=-0.2, gf=0.95, ga=1.0):
    p = _params(rng)
    d = {k: v.copy() for k, v in mlp_ref.split(p).items()}
    d["b3"] = np.array([np.arctanh(-0.4)])
    d["wlin"] = np.zeros(16)
    d["wlin"][:3] = (kf, gf, ga)
    return mlp_ref.join(d)


def test_higdon_head_gradient_matches_finite_differences():
    """The Higdon head's FP64 gradient (unrounded weights) against
    central differences of its loss over the network's weights and biases
    (the fixed factor in wlin is held, so it is left out)."""
    rng = np.random.default_rng(6)
    p = _higdon_params(rng)
    u = rng.normal(0, 1, (40, 16))
    t = rng.normal(0, 1, 40)
    g = mlp_ref.grad_higdon_f64(p, u, t, rounded=False)
    idx = np.concatenate([rng.choice(mlp_ref.P - 16, 60, replace=False), [mlp_ref.P - 17]])
    h = 1e-6
    fd = []
    for i in idx:
        q1, q2 = p.copy(), p.copy()
        q1[i] += h
        q2[i] -= h
        fd.append((mlp_ref.loss_higdon_f64(q1, u, t, False) -
                   mlp_ref.loss_higdon_f64(q2, u, t, False)) / (2 * h))
    np.testing.assert_allclose(g[idx], fd, rtol=2e-5, atol=1e-9)
    assert not np.any(g[mlp_ref.P - 16:])


def test_higdon_terms_hold_a_static_field_and_solve_the_factor_product():
    """The Higdon head's expansion: a field constant in space and time is
    held exactly when one factor is undamped (an undamped factor annihilates
    constants), whatever the coefficients; and on random data the ghost it
    gives makes the two one-way factors, applied one after the other to the
    space-time stencil, vanish."""
    rng = np.random.default_rng(7)
    for kf, gf, ka in ((-0.3, 0.95, -0.4), (0.2, 0.9, 0.5), (-0.7, 1.0, 0.1)):
        u = np.full((3, 16), 2.5)
        A, B = mlp_ref.higdon_terms(kf, gf, 1.0, u)
        np.testing.assert_allclose(A + ka * B, 2.5, rtol=0, atol=1e-12)
    # factor by factor: apply F(kf, gf) to the history, then D(ka, ga)
    kf, gf, ka, ga = -0.35, 0.93, -0.1, 0.97
    W = rng.normal(0, 1, (3, 3))  # W[x][t]: x = 0 ghost, 1 P0, 2 P1; t steps back

    def apply(k, g, V):
        # (1 + k g T) - X (g T + k) on V[x][t] -> rows x = 0..1, cols t = 0..1
        out = np.zeros((V.shape[0] - 1, V.shape[1] - 1))
        for x in range(out.shape[0]):
            for t in range(out.shape[1]):
                out[x, t] = (V[x, t] + k * g * V[x, t + 1] - g * V[x + 1, t + 1]
                             - k * V[x + 1, t])
        return out

    u = np.zeros((1, 16))
    u[0, [0, 1, 10, 11, 12, 13, 14, 15]] = (W[1, 0], W[2, 0], W[2, 1], W[2, 2], W[0, 1], W[0, 2],
                                            W[1, 1], W[1, 2])
    A, B = mlp_ref.higdon_terms(kf, gf, ga, u)
    W[0, 0] = (A + ka * B)[0]
    assert abs(apply(ka, ga, apply(kf, gf, W))[0, 0]) < 1e-12
