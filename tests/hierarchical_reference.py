"""CPU restatement of the hierarchical KDK block with the oracle's forces.

Test infrastructure only. Executes a cosmosim.hierarchical.BlockPlan the way
time_integration.cpp:1356-1600 does (opening half-kicks of the starting bins
and the PM opening kick, Hubble drag and drift of every particle, short-range
forces of the ending bins, their closing half-kicks, the synchronized Tree+PM
solve and PM closing kick at the block end) with numpy arithmetic and the
oracle's PM and short-range forces.
"""
from __future__ import annotations

import numpy as np

from cosmosim import hierarchical as H
from oracle import oracle as O


def forces(x, m, opts, mesh):
    pm = np.stack(O.pm_accel(x[0], x[1], x[2], m, opts, mesh))
    sr = np.stack(O.short_range(x[0], x[1], x[2], m, opts, mesh)[0])
    return pm, sr


def reference_block(x, u, m, opts, mesh, cosmo, max_bin, a_begin, eta=0.2, max_span=None, bins=None):
    """x, u: (3, N) float64; returns (x, u, plan, short-range solves)."""
    x = np.array(x, dtype=np.float64)
    u = np.array(u, dtype=np.float64)
    box = np.array(opts.box_size[:])
    pm, sr = forces(x, m, opts, mesh)
    norm = np.sqrt(((pm + sr) ** 2).sum(axis=0))
    plan = H.plan_block(cosmo, a_begin, norm, opts.softening_epsilon, max_bin, eta, max_span, bins)
    tree = sr
    periods = 1 << plan.max_bin
    sr_solves = 0
    for off in range(periods):
        for b, f in plan.opening[off].items():
            sel = plan.bins == b
            u[:, sel] = u[:, sel] + f * tree[:, sel]
        if off == 0:
            u = u + plan.pm_open * pm
        xs, us = [np.ascontiguousarray(c) for c in x], [np.ascontiguousarray(c) for c in u]
        O.drift(*xs, *us, plan.drift[off], tuple(box))  # x += a_k u_k int dt/a^2 (u at the tick start)
        x = np.stack(xs)
        u = plan.drag[off] * u + 0.0 * pm  # then u = u a_k / a_{k+1}
        if off + 1 == periods:
            pm, tree = forces(x, m, opts, mesh)
        else:
            closing = np.isin(plan.bins, list(plan.closing[off].keys()))
            if closing.any():
                _, sr = forces(x, m, opts, mesh)
                tree = np.where(closing[None, :], sr, tree)
                sr_solves += 1
        for b, f in plan.closing[off].items():
            sel = plan.bins == b
            u[:, sel] = u[:, sel] + f * tree[:, sel]
        if off + 1 == periods:
            u = u + plan.pm_close * pm
    return x, u, plan, sr_solves
