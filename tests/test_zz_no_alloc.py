"""kgmt_propagate_collide_noalloc (no device allocation: the scene in a
library-held device array, the descriptor by value) against
kgmt_propagate_collide on identical inputs, bit for bit, on consecutive
calls whose second input is smaller (fewer children and boxes, so the
scratch holds a larger earlier scene), with no boxes, and beyond the scratch
limit (an error)."""
import ctypes as C

import numpy as np
import pytest

from kgmt_amd import _lib
from kgmt_amd import workloads as W
from kgmt_amd.abi import dims

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tests.gpu_util import bits, call, dev, empty, host, ptr, stream  # noqa: E402
from tests.test_gpu_parity import CASES, parent_states  # noqa: E402


def run(entry, cfg, pidx, X, ctrl, steps, obs):
    S, Cd, D = dims(cfg.model)
    n, npar = len(pidx), X.shape[1]
    out = (empty((D, n), torch.float32), empty(n, torch.uint8), empty(n, torch.int32), empty(n, torch.int32))
    d_obs = dev(obs.reshape(-1)) if len(obs) else None
    d = [dev(pidx), dev(X), dev(ctrl), dev(steps)]
    call(entry, C.byref(cfg), ptr(d[0]), ptr(d[1]), npar, ptr(d[2]), ptr(d[3]), n, ptr(d_obs), len(obs),
         *(ptr(o) for o in out), stream())
    return [host(o) for o in out]


@pytest.mark.parametrize("case", ["di_ref", "dubins5d", "quad12d", "unicycle"])
def test_noalloc_matches_allocating_entry(case, orc):
    cfg = CASES[case]()
    rng = np.random.default_rng(17)
    span = cfg.ws_max - cfg.ws_min
    # two calls in a row, the second smaller; then one without boxes
    for n, npar, nobs in ((40000, 1500, 1500), (5000, 200, 30), (777, 50, 0)):
        obs = W.random_boxes(rng, nobs, 0.01 * span, 0.08 * span, ws=(cfg.ws_min, cfg.ws_max), dim=cfg.wdim) \
            if nobs else np.zeros((0, 2 * cfg.wdim), np.float32)
        X = parent_states(cfg, npar, seed=n)
        pidx = rng.integers(0, npar, n).astype(np.int32)
        ctrl, steps, _ = orc.draw_children(cfg, 0, 3, 0, n)
        print(f"noalloc {case}: {n} children, {npar} parents, {nobs} boxes", flush=True)
        got = run("kgmt_propagate_collide_noalloc", cfg, pidx, X, ctrl, steps, obs)
        want = run("kgmt_propagate_collide", cfg, pidx, X, ctrl, steps, obs)
        for g, w in zip(got, want):
            assert np.array_equal(bits(g), bits(w))
        ref = orc.propagate_children(cfg, pidx, X, ctrl, steps, obs)
        assert np.array_equal(got[1], ref[1]) and np.array_equal(bits(got[0]), bits(ref[0]))


def test_noalloc_rejects_more_boxes_than_the_scratch_holds():
    cfg = CASES["di_ref"]()
    obs = np.zeros((1501, 6), np.float32)
    d_obs = dev(obs.reshape(-1))
    buf = empty(64, torch.float32)
    rc = _lib.lib().kgmt_propagate_collide_noalloc(C.byref(cfg), ptr(buf), ptr(buf), 1, ptr(buf), ptr(buf), 1,
                                                   ptr(d_obs), 1501, ptr(buf), ptr(buf), ptr(buf), ptr(buf), stream())
    assert rc == -4 and b"1500" in _lib.lib().kgmt_last_error()
