"""The FP64 trajectory cost/adjoint of the oracle, pinned by central finite
differences (h = 1e-6) and by an autograd-free numpy Jacobian check of the
quadcopter ODE's vector-Jacobian product. No GPU."""
import numpy as np
import pytest

from kgmt_amd import workloads as W


def test_quad_vjp_matches_finite_differences(orc):
    cfg = W.config("quad12d")
    rng = np.random.default_rng(0)
    for _ in range(20):
        s = np.concatenate([rng.uniform(0, 1, 3), rng.uniform(-0.8, 0.8, 3), rng.uniform(-2, 2, 6)])
        c = np.array([rng.uniform(0, 19), *rng.uniform(-0.05, 0.05, 3)])
        a = rng.normal(size=12)
        gs, gc = orc.quad_vjp_f64(cfg, s, c, a)
        h = 1e-6
        for j in range(12):
            e = np.zeros(12)
            e[j] = h
            fd = (orc.quad_ode_f64(cfg, s + e, c) - orc.quad_ode_f64(cfg, s - e, c)) @ a / (2 * h)
            assert abs(fd - gs[j]) <= 1e-6 * max(1.0, abs(fd)), (j, fd, gs[j])
        for j in range(4):
            e = np.zeros(4)
            e[j] = h
            fd = (orc.quad_ode_f64(cfg, s, c + e) - orc.quad_ode_f64(cfg, s, c - e)) @ a / (2 * h)
            assert abs(fd - gc[j]) <= 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_trajectory_adjoint_matches_central_differences(k, orc):
    rb = W.make_refine_batch(8, seed=k)
    cfg, prm = rb.cfg, rb.params
    i = k
    T = int(rb.steps[i])
    u = rb.u[:T, :, i].astype(np.float64)
    # put some boxes near the path so the penalty term is active
    J0, xT = orc.traj_cost_f64(cfg, prm, rb.x0[:, i], u, rb.goal[:, i], rb.obstacles)
    mid = (rb.x0[:3, i] + xT[:3]) / 2
    obs = np.concatenate([rb.obstacles, [[*(mid - 0.01), *(mid + 0.01)]]]).astype(np.float64)
    J, g = orc.traj_grad_f64(cfg, prm, rb.x0[:, i], u, rb.goal[:, i], obs)
    h = 1e-6
    rng = np.random.default_rng(k)
    idx = [(t, c) for t in rng.choice(T, 6, replace=False) for c in range(4)]
    for t, c in idx:
        up, um = u.copy(), u.copy()
        up[t, c] += h
        um[t, c] -= h
        fd = (orc.traj_cost_f64(cfg, prm, rb.x0[:, i], up, rb.goal[:, i], obs)[0] -
              orc.traj_cost_f64(cfg, prm, rb.x0[:, i], um, rb.goal[:, i], obs)[0]) / (2 * h)
        assert abs(fd - g[t, c]) <= 1e-5 * max(1.0, abs(fd)), (t, c, fd, g[t, c])
    # directional derivative along a random direction
    v = rng.normal(size=u.shape)
    fd = (orc.traj_cost_f64(cfg, prm, rb.x0[:, i], u + h * v, rb.goal[:, i], obs)[0] -
          orc.traj_cost_f64(cfg, prm, rb.x0[:, i], u - h * v, rb.goal[:, i], obs)[0]) / (2 * h)
    assert abs(fd - (g * v).sum()) <= 1e-5 * max(1.0, abs(fd))


def test_penalty_cases(orc):
    """Signed distance to a box: outside distance, inside depth; the softplus
    penalty tends to the hinge max(0, margin - sd) as tau -> 0."""
    cfg = W.config("quad12d")
    prm = W.refine_params(cfg, w_effort=0.0, w_goal=0.0, tau=1e-4)
    box = np.array([[0.4, 0.4, 0.4, 0.6, 0.6, 0.6]])
    x0 = np.zeros(12)
    x0[:3] = [0.5, 0.5, 0.61]          # 0.01 above the box
    hover = np.array([[-cfg.gravity, 0, 0, 0]])
    J, xT = orc.traj_cost_f64(cfg, prm, x0, hover, np.zeros(3), box)
    sd = xT[2] - 0.6
    assert abs(J - max(0.0, 0.02 - sd)) < 1e-4
    x0[:3] = [0.5, 0.5, 0.55]          # 0.05 inside
    J, xT = orc.traj_cost_f64(cfg, prm, x0, hover, np.zeros(3), box)
    assert abs(J - (0.02 + (0.6 - xT[2]))) < 1e-4
