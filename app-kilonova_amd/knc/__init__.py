"""knc -- host side of the MI355X implementation of the KNC per-block RK stage.

``capi`` binds the C ABI of ``libknc_gpu.so`` (include/knc_gpu.h);
``setups`` builds the seeded benchmark workloads.
"""
from .capi import (BATCHED, EULER, FIXED, GUESS_LEARNED, GUESS_LEARNED_FOLD, GUESS_LEARNED_STEP, GUESS_PLAIN,
                   GUESS_PREVIOUS, HLLC, HLLE,
                   INNER, OUTER, OUTFLOW, PER_BLOCK, POLAR, RADIAL, SRHD, Context, Hydro, HydroError, fp64_fma_rate,
                   fp64_probe, geometrical_source_terms, intercell_flux, lib, max_signal_speed, newtonian_hydro,
                   plm_gradient, relativistic_hydro, to_conserved, try_to_primitive)
from . import mlp, setups


def hydro_from_spec(spec):
    """knc.Hydro from a setups.HydroSpec."""
    return Hydro(spec.system, spec.riemann_solver, spec.runge_kutta_order, int(spec.adaptive_time_step),
                 spec.gamma_law_index, spec.plm_theta, spec.cfl_number, spec.light_speed)


def context_from_setup(setup, excision=False, **options):
    """A Context holding every block and boundary of a setups.Setup; with
    excision=True also the setup's moving excision surfaces and block source."""
    ctx = Context(hydro_from_spec(setup.hydro), **options)
    for b in setup.blocks:
        ctx.insert_block(b.index, b.rv, b.qv, b.conserved, b.scalar_mass)
    for side, bd in ((INNER, setup.inner), (OUTER, setup.outer)):
        ctx.set_boundary(side, bd.mode, bd.primitive, bd.scalar)
    if excision and setup.excision is not None:
        e = setup.excision
        ctx.set_excision(e.inner_radius, e.outer_radius, e.inner_speed, e.outer_speed, e.delay, e.check_every,
                         setup.block_source)
    return ctx
