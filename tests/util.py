"""Shared helpers of the test-suite: loading a workload into the oracle and
comparing states field by field (rel-L1 = sum|x - y| / sum|y|)."""
import numpy as np

import oracle

FIELDS = ("D", "S_r", "S_theta", "tau", "scalar_mass")


def oracle_hydro(spec):
    return oracle.KoHydro(spec.system, spec.riemann_solver, spec.runge_kutta_order, int(spec.adaptive_time_step),
                          spec.gamma_law_index, spec.plm_theta, spec.cfl_number, spec.light_speed)


def oracle_from_setup(setup, nthreads=8, literal=False):
    sim = oracle.Sim(oracle_hydro(setup.hydro), nthreads, literal=literal)
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


def oracle_mesh(preset):
    m = preset.knc_mesh
    return oracle.omesh(m.reference_radius, m.num_radial_zones, m.num_polar_zones, m.block_size)


def oracle_from_preset(preset, nthreads=8):
    """App::from_config on the oracle: the blocks of grid_blocks(start_time)
    from the model, model ghost rows on both sides."""
    sim = oracle.Sim(oracle_hydro_from_capi(preset.hydro), nthreads)
    mesh = oracle_mesh(preset)
    t0 = preset.control["start_time"]
    for i in preset.grid_block_indexes(t0):
        rv, qv, cons, sm, bad = oracle.block_from_model(sim.h, mesh, preset.model_kind, preset.model_params, i, t0,
                                                        preset.model_table)
        assert bad == -1, f"model fails in block {i}"
        sim.push_block(rv, qv, cons, sm)
    sim.time = t0
    sim.set_first_index(0)
    for side in (0, 1):
        sim.set_boundary_model(side, mesh, preset.model_kind, preset.model_params, preset.model_table)
    return sim


def oracle_birth_source(sim, preset):
    """Block births from the model for Sim.scheme_advance."""
    mesh = oracle_mesh(preset)

    def source(index, t):
        rv, qv, cons, sm, bad = oracle.block_from_model(sim.h, mesh, preset.model_kind, preset.model_params, index,
                                                        t, preset.model_table)
        assert bad == -1
        return rv, cons, sm
    return source


def oracle_hydro_from_capi(h):
    return oracle.KoHydro(h.system, h.riemann_solver, h.runge_kutta_order, h.adaptive_time_step, h.gamma_law_index,
                          h.plm_theta, h.cfl_number, h.light_speed)


def preset_excision(preset):
    """The Excision of a preset's mesh for Sim.scheme_advance."""
    from knc.setups import Excision
    m = preset.mesh
    return Excision(m["inner_radius"], m["outer_radius"], m["inner_excision_speed"], m["outer_excision_speed"],
                    m["excision_delay"] or 0.0, 1)
