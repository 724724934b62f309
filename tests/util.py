"""Shared helpers of the test-suite: loading a workload into the oracle and
comparing states field by field (rel-L1 = sum|x - y| / sum|y|)."""
import numpy as np

import oracle

FIELDS = ("D", "S_r", "S_theta", "tau", "scalar_mass")


def oracle_hydro(spec):
    return oracle.KoHydro(spec.system, spec.riemann_solver, spec.runge_kutta_order, int(spec.adaptive_time_step),
                          spec.gamma_law_index, spec.plm_theta, spec.cfl_number, spec.light_speed)


def oracle_from_setup(setup, nthreads=8):
    sim = oracle.Sim(oracle_hydro(setup.hydro), nthreads)
    for b in setup.blocks:
        sim.push_block(b.rv, b.qv, b.conserved, b.scalar_mass)
    for side, bd in ((0, setup.inner), (1, setup.outer)):
        sim.set_boundary(side, bd.mode, bd.primitive, bd.scalar)
    return sim


def oracle_state(sim, setup_blocks):
    out = []
    for k, b in enumerate(setup_blocks):
        out.append(sim.get_block(k, b.nr, b.nq))
    return out


def gpu_state(ctx):
    return [ctx.download_block(i) for i in ctx.block_indexes()]


def rel_l1(state, ref):
    """Per-field rel-L1 over all blocks; state/ref are lists of (cons, sm)."""
    out = {}
    for f in range(5):
        num = 0.0
        den = 0.0
        for (c, m), (rc, rm) in zip(state, ref):
            x = c[..., f] if f < 4 else m
            y = rc[..., f] if f < 4 else rm
            num += float(np.abs(x - y).sum())
            den += float(np.abs(y).sum())
        out[FIELDS[f]] = num / den if den > 0 else num
    return out


def max_rel_l1(state, ref):
    return max(rel_l1(state, ref).values())
