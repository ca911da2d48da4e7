"""Differentiable quadcopter rollout over the C-ABI (kgmt_rk4_adjoint,
kgmt_refine).

The planner's quadcopter propagation (statePropagator.cu:246-370: RK4 of the
12-state ODE under 4 controls held over an edge) is made differentiable for
refining the control sequences the planner returns. Trajectories are held
struct-of-arrays on the device: controls ``u`` of shape (max_steps, 4, n),
``x0`` (12, n), ``goal`` (3, n), per-trajectory lengths ``steps`` (n,) int32;
obstacles are rows of 6 floats shared by every trajectory.

    cost = rollout_cost(u, x0, goal, steps, obstacles, cfg, params)
    cost.sum().backward()          # u.grad = dJ/du from the discrete adjoint

    history = refine(u, x0, goal, steps, obstacles, cfg, params)   # Adam, in place

The cost J = w_e sum_t |u_t - u_ref|^2 + w_o sum_{t>=1} sum_b tau softplus((m - sd(p_t, b)) / tau)
+ w_g |p_T - goal|^2 and its gradient are computed in FP32 by one lane per
trajectory on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .abi import KGMT_MODEL_QUADCOPTER, KgmtAdjointParams, KgmtConfig


def _check_args(u, x0, goal, steps, obstacles):
    if u.dim() != 3 or u.shape[1] != 4:
        raise ValueError("u must have shape (max_steps, 4, n)")
    T, _, n = u.shape
    for name, t, shape in (("x0", x0, (12, n)), ("goal", goal, (3, n)), ("steps", steps, (n,))):
        if tuple(t.shape) != shape:
            raise ValueError(f"{name} must have shape {shape}, got {tuple(t.shape)}")
    for name, t in (("u", u), ("x0", x0), ("goal", goal), ("steps", steps)):
        if not t.is_cuda:
            raise ValueError(f"{name} must be a device tensor")
    if steps.dtype != torch.int32:
        raise ValueError("steps must be int32")
    if obstacles is not None and obstacles.numel() % 6:
        raise ValueError("obstacles must be rows of 6 floats")
    return T, n


def _prep(x):
    return x.detach().to(torch.float32).contiguous()


def _obs(obstacles, device):
    if obstacles is None or obstacles.numel() == 0:
        return None, 0
    o = obstacles.detach().to(device=device, dtype=torch.float32).contiguous().reshape(-1, 6)
    return o, o.shape[0]


def _stream(device):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _cfg(cfg: KgmtConfig) -> KgmtConfig:
    if cfg.model != KGMT_MODEL_QUADCOPTER:
        raise ValueError("the differentiable rollout is built for the quadcopter model")
    return cfg


def cost_and_grad(u, x0, goal, steps, obstacles, cfg: KgmtConfig, params: KgmtAdjointParams):
    """(J (n,), dJ/du (max_steps, 4, n)) on u's device and stream."""
    T, n = _check_args(u, x0, goal, steps, obstacles)
    uu, xx, gg = _prep(u), _prep(x0), _prep(goal)
    ss = steps.contiguous()
    ob, m = _obs(obstacles, u.device)
    cost = torch.empty(n, device=u.device, dtype=torch.float32)
    grad = torch.empty_like(uu)
    L = _lib.lib()
    _lib.check(L.kgmt_rk4_adjoint(C.byref(_cfg(cfg)), C.byref(params), n, T, C.c_void_p(ss.data_ptr()),
                                  C.c_void_p(xx.data_ptr()), C.c_void_p(gg.data_ptr()), C.c_void_p(uu.data_ptr()),
                                  C.c_void_p(ob.data_ptr()) if ob is not None else None, m,
                                  C.c_void_p(cost.data_ptr()), C.c_void_p(grad.data_ptr()), _stream(u.device)))
    return cost, grad


class RolloutCost(torch.autograd.Function):
    """J(u) per trajectory; backward returns dJ/du scaled by the incoming
    gradient. x0, goal, steps and obstacles are constants."""

    @staticmethod
    def forward(ctx, u, x0, goal, steps, obstacles, cfg, params):
        cost, grad = cost_and_grad(u, x0, goal, steps, obstacles, cfg, params)
        ctx.save_for_backward(grad)
        return cost

    @staticmethod
    def backward(ctx, gcost):
        (grad,) = ctx.saved_tensors
        return grad * gcost.to(grad.dtype).view(1, 1, -1), None, None, None, None, None, None


def rollout_cost(u, x0, goal, steps, obstacles, cfg: KgmtConfig, params: KgmtAdjointParams):
    return RolloutCost.apply(u, x0, goal, steps, obstacles, cfg, params)


def refine(u, x0, goal, steps, obstacles, cfg: KgmtConfig, params: KgmtAdjointParams):
    """params.iterations Adam steps on u (float32, contiguous, in place);
    returns the cost history, shape (iterations + 1, n): J before every step
    and after the last."""
    T, n = _check_args(u, x0, goal, steps, obstacles)
    if u.dtype != torch.float32 or not u.is_contiguous():
        raise ValueError("u must be contiguous float32 (refined in place)")
    xx, gg = _prep(x0), _prep(goal)
    ss = steps.contiguous()
    ob, m = _obs(obstacles, u.device)
    hist = torch.empty((params.iterations + 1, n), device=u.device, dtype=torch.float32)
    L = _lib.lib()
    _lib.check(L.kgmt_refine(C.byref(_cfg(cfg)), C.byref(params), n, T, C.c_void_p(ss.data_ptr()),
                             C.c_void_p(xx.data_ptr()), C.c_void_p(gg.data_ptr()), C.c_void_p(u.data_ptr()),
                             C.c_void_p(ob.data_ptr()) if ob is not None else None, m,
                             C.c_void_p(hist.data_ptr()), _stream(u.device)))
    return hist
