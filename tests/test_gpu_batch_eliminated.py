"""The inner batch solve of BatchEliminated (src/nonlinear/eliminate.jl:3390-3405) through the library.

`_batch_solve!` calls `solve_batch!(prob.Y, statuses, prob.inner_prob, prob.inner_method,
BatchParams(shared, per_system); options = _batch_inner_options(prob, tol), workgroupsize = 64)`,
with the options of `_batch_inner_options` (eliminate.jl:3379-3386): the residual test at `tol`
absolutely and not relatively, the template's step test, iteration bounds, norm and round-off stop,
g_tol = √eps and no rate test; successive outer evaluations warm-start from the previous solutions
with a tightening `tol`. With a registered family as the inner problem that call is the library's
status form; here the Hénon–Heiles midpoint step (BatchParams(h, X_k)) runs that schedule and every
solve is compared with the oracle bit for bit.
"""
import math

import numpy as np
import pytest
import torch

from test_gpu_parity import assert_parity

pytestmark = pytest.mark.gpu


def inner_options(gs, tol, template):
    """_batch_inner_options (eliminate.jl:3379-3386)."""
    return gs.Options(f_abstol=tol, f_reltol=0.0, x_abstol=template.x_abstol, x_reltol=template.x_reltol,
                      min_iterations=template.min_iterations, max_iterations=template.max_iterations,
                      max_stalls=template.max_stalls, g_tol=math.sqrt(np.finfo(float).eps), norm=template.norm,
                      kappa=template.kappa, roundoff=template.roundoff, rate_test=False)


def test_inner_solve_schedule_bitwise(gs, orc, workloads, device):
    n = 16384
    xk = workloads.henon_heiles_states(n, seed=41)
    prob = gs.NonlinearProblem(gs.HenonHeilesMidpoint())
    template = gs.Options()
    Y = torch.as_tensor(xk, device=device).clone()
    Yo = xk.copy()
    pd = gs.BatchParams(0.1, torch.as_tensor(xk, device=device))
    ph = gs.BatchParams(0.1, xk)
    for tol in (1e-2 / 2 / math.sqrt(n), 1e-6, 1e-10, 2 * np.finfo(float).eps):
        opts = inner_options(gs, tol, template)
        S = gs.status_buffer(n, device)
        gs.solve_batch(Y, S, prob, gs.Newton(), pd, opts, workgroupsize=64)
        Yo, so = orc.solve_batch(gs.Newton(), gs.HenonHeilesMidpoint(), Yo, ph, opts)
        st = gs.statuses(S)
        assert_parity(st, Y.cpu().numpy(), so, Yo, bitwise=True)
    assert (st["code"] == gs.SUCCESS).all()
