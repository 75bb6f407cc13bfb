"""Closed-form pins of the flooding oracle's VN message damping
(oracle/ldpc_oracle.c, q = (1 - d) (APP - r) + d q_prev, following
source/matlab/uhd_cpp/matlab/cpu_bp_osd_decode.m:174-177), CPU only.

On two parallel degree-2 checks over two variables, min-sum is linear: a
degree-2 check returns alpha times the other edge's q (its sign and its
magnitude are the other edge's).  With a_t, b_t the q of variables 0 and 1 in
check pass t (a_1, b_1 = the channel LLRs l0, l1):
    APP_t = (l0 + 2 alpha b_t, l1 + 2 alpha a_t)
    a_{t+1} = (1 - d)(l0 + alpha b_t) + d a_t,  b_{t+1} likewise.
Pass 2 in closed form is l0 + 2 alpha (l1 + (1 - d) alpha l0), and the passes
converge (eigenvalues d +- (1 - d) alpha) to the fixed point of the undamped
equations, a* = (l0 + alpha l1) / (1 - alpha^2), which does not depend on d.
A dropped blend term, swapped weights or a blend of the wrong message fails
one of these.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

import oracle

RP = np.array([0, 2, 4], np.int32)
CI = np.array([0, 1, 0, 1], np.int32)
L0, L1 = 1.5, -0.5


def _flood(iters, alpha, damping):
    # the oracle's flooding decoder needs only the graph (no encoder) for this call
    g = SimpleNamespace(n=2, m=2, rp=RP, ci=CI)
    llr = np.array([L0, L1], np.float32)
    return oracle.Code.nms(g, llr, max_iter=iters, alpha=alpha, damping=damping, early_stop=False)["app"]


@pytest.mark.parametrize("d", [0.0, 0.3, 0.6])
def test_first_two_passes(d):
    a = 0.75
    p1 = _flood(1, a, d)
    assert p1 == pytest.approx([L0 + 2 * a * L1, L1 + 2 * a * L0], abs=1e-6)
    p2 = _flood(2, a, d)
    exp = [L0 + 2 * a * (L1 + (1 - d) * a * L0), L1 + 2 * a * (L0 + (1 - d) * a * L1)]
    assert p2 == pytest.approx(exp, abs=1e-6)


def test_damping_changes_the_trajectory_not_the_fixed_point():
    a = 0.75
    star_a = (L0 + a * L1) / (1 - a * a)
    star_b = (L1 + a * L0) / (1 - a * a)
    fixed = [L0 + 2 * a * star_b, L1 + 2 * a * star_a]
    for d in (0.0, 0.3, 0.6):
        assert _flood(300, a, d) == pytest.approx(fixed, rel=1e-5)
    assert not np.array_equal(_flood(3, a, 0.0), _flood(3, a, 0.3))
