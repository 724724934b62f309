"""ctypes wrapper of the CPU oracle (``oracle/knc_oracle.cpp``).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg import this module, and only as the checker
or as the timed CPU baseline; the product path never calls it.

The oracle restates, on the CPU, the reference's per-block RK stage
(``src/scheme.rs:16-240``, ``src/state.rs:55-245``, ``src/mesh.rs``) and the
per-zone arithmetic of the un-vendored crates ``hydro-srhd@25bf21ec``,
``hydro-euler@fc929f9b`` and ``godunov-core@d8dd524a`` from the published
methods named in the C++ source.  How each function is pinned:

* p2c, c2p            -- closed-form round trips, v=0 special case, exact
                         Euler closed form (tests/test_oracle_zone.py)
* wave speeds         -- 1-D characteristic speeds of Marti & Mueller, the
                         v=0 / p->0 limits and relativistic velocity addition
* HLLE / HLLC         -- consistency F(U,U)=F(U), mirror symmetry, supersonic
                         upwinding, exact stationary contact (HLLC), Sod and
                         Marti-Mueller blast wave 1 plateaus against exact
                         Riemann solutions (tests/test_oracle_riemann.py)
* PLM                 -- brute force over sign patterns
* geometry            -- telescoping volumes, closed-form polar face areas
* stage / RK / dt     -- mass conservation, mirror symmetry, 1-D vs 2-D
                         consistency, RK weights on a linear ODE, uniform
                         state dt closed form (tests/test_oracle_stage.py)

No reference test or golden vector pins the crate-internal choices (c2p
tolerance, HLLC variant, scalar flux form): those are "parity unpinned"
against the reference and documented as such.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libknc_oracle.so")

LIGHT_SPEED = 3e10

EULER, SRHD = 0, 1
HLLE, HLLC = 0, 1
FIXED, OUTFLOW = 0, 1
# parts of Sim(literal=...): reference-code expressions, then crate arithmetic
LIT_GEOMETRY, LIT_VOLUME, LIT_SOURCES, LIT_FLUX = 1, 2, 4, 8
LIT_REFERENCE = LIT_GEOMETRY | LIT_VOLUME
LIT_ALL = 15
CODES = {0: "ok", 1: "NegativeDensity", 2: "NegativePressure", 3: "NegativeEnergyDensity",
         4: "RootFinderFailed", 14: "ModelPanics"}


class KoHydro(ctypes.Structure):
    _fields_ = [("system", ctypes.c_int32), ("riemann_solver", ctypes.c_int32),
                ("runge_kutta_order", ctypes.c_int32), ("adaptive_time_step", ctypes.c_int32),
                ("gamma_law_index", ctypes.c_double), ("plm_theta", ctypes.c_double),
                ("cfl_number", ctypes.c_double), ("light_speed", ctypes.c_double)]


class KoError(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("block", ctypes.c_int32), ("i", ctypes.c_int32),
                ("j", ctypes.c_int32), ("r", ctypes.c_double), ("theta", ctypes.c_double),
                ("value", ctypes.c_double), ("conserved", ctypes.c_double * 4)]


def hydro(system=SRHD, gamma=4.0 / 3.0, plm_theta=1.5, cfl=0.4, rk=2, riemann=HLLE, adaptive=True,
          light_speed=None):
    if light_speed is None:
        light_speed = LIGHT_SPEED if system == SRHD else 1.0
    return KoHydro(system, riemann, rk, int(adaptive), gamma, plm_theta, cfl, light_speed)


class OModel(ctypes.Structure):
    """om::Model (knc_model_oracle.hpp): kind, table rows, 12 parameters in
    the reference struct's field order, table pointer."""
    _fields_ = [("kind", ctypes.c_int), ("table_rows", ctypes.c_int), ("params", ctypes.c_double * 12),
                ("table", ctypes.c_void_p)]


class OMesh(ctypes.Structure):
    """om::Mesh: reference_radius, num_radial_zones (0 = None),
    num_polar_zones, block_size."""
    _fields_ = [("reference_radius", ctypes.c_double), ("num_radial_zones", ctypes.c_int),
                ("num_polar_zones", ctypes.c_int), ("block_size", ctypes.c_int), ("pad", ctypes.c_int)]


def omodel(kind, params, table=None):
    m = OModel()
    m.kind = int(kind)
    for k, v in enumerate(params):
        m.params[k] = float(v)
    tab = None
    if table is not None:
        tab = np.ascontiguousarray(table, dtype=np.float64).reshape(-1, 4)
        m.table_rows = tab.shape[0]
        m.table = tab.ctypes.data
    return m, tab


def omesh(reference_radius, num_radial_zones, num_polar_zones, block_size):
    return OMesh(float(reference_radius), int(num_radial_zones), int(num_polar_zones), int(block_size), 0)


_lib = None

BOUNDARY_SOURCE = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.ko_sim_create.restype = P
        L.ko_sim_destroy.argtypes = [P]
        for name in ("ko_sim_push_block", "ko_sim_pop_inner_block", "ko_sim_set_boundary",
                     "ko_sim_get_block", "ko_sim_time_step", "ko_sim_advance", "ko_sim_scheme_advance",
                     "ko_sim_num_blocks", "ko_sim_c2p_stats", "ko_sim_set_time"):
            getattr(L, name).restype = ctypes.c_int
        L.ko_sim_set_time.argtypes = [P, ctypes.c_double]
        L.ko_sim_set_literal.argtypes = [P, ctypes.c_int]
        L.ko_sim_set_literal.restype = ctypes.c_int
        L.ko_sim_time.restype = ctypes.c_double
        L.ko_sim_time.argtypes = [P]
        L.ko_sim_iteration.restype = ctypes.c_longlong
        L.ko_sim_iteration.argtypes = [P]
        L.ko_sim_advance.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.POINTER(KoError)]
        L.ko_plm_gradient.argtypes = [ctypes.c_double, ctypes.c_int64, P, P, P, P]
        L.ko_sim_set_boundary_source.restype = ctypes.c_int
        L.ko_sim_set_boundary_source.argtypes = [P, ctypes.c_int, BOUNDARY_SOURCE, P]
        L.ko_model_eval.restype = ctypes.c_longlong
        L.ko_model_eval.argtypes = [ctypes.POINTER(OModel), ctypes.c_longlong, P, P, ctypes.c_double, P, P]
        L.ko_mesh_outer_radius.restype = ctypes.c_double
        L.ko_mesh_outer_radius.argtypes = [ctypes.POINTER(OMesh), ctypes.c_int]
        L.ko_mesh_block_vertices.restype = None
        L.ko_mesh_block_vertices.argtypes = [ctypes.POINTER(OMesh), ctypes.c_int, P, P]
        L.ko_block_from_model.restype = ctypes.c_longlong
        L.ko_block_from_model.argtypes = [ctypes.POINTER(KoHydro), ctypes.POINTER(OModel), ctypes.POINTER(OMesh),
                                          ctypes.c_int, ctypes.c_double, P, P, P, P]
        L.ko_sim_set_boundary_model.restype = ctypes.c_int
        L.ko_sim_set_boundary_model.argtypes = [P, ctypes.c_int, ctypes.POINTER(OModel), ctypes.POINTER(OMesh)]
        L.ko_sim_set_block.restype = ctypes.c_int
        L.ko_sim_set_block.argtypes = [P, ctypes.c_int, P, P]
        L.ko_sim_set_first_index.restype = ctypes.c_int
        L.ko_sim_set_first_index.argtypes = [P, ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


# ---- initial models and the block mesh (knc_model_oracle.cpp) -------------
def model_eval(kind, params, r, theta, t, table=None):
    """primitive_at (n, 4), scalar_at (n) and the index of the first point
    the reference would panic at (-1 if none)."""
    r, q = _f64(r).ravel(), _f64(theta).ravel()
    prim = np.zeros((r.size, 4))
    scal = np.zeros(r.size)
    m, _keep = omodel(kind, params, table)
    bad = lib().ko_model_eval(ctypes.byref(m), r.size, _p(r), _p(q), float(t), _p(prim), _p(scal))
    return prim, scal, bad


def mesh_block_vertices(mesh, index):
    rv = np.zeros(mesh.block_size + 1)
    qv = np.zeros(mesh.num_polar_zones + 1)
    lib().ko_mesh_block_vertices(ctypes.byref(mesh), int(index), _p(rv), _p(qv))
    return rv, qv


def block_from_model(h, mesh, kind, params, index, t, table=None):
    """BlockState::from_model of mesh block `index`: (rv, qv, conserved,
    scalar_mass, first failing zone or -1)."""
    nr, nq = mesh.block_size, mesh.num_polar_zones
    rv, qv = np.zeros(nr + 1), np.zeros(nq + 1)
    cons, sm = np.zeros((nr, nq, 4)), np.zeros((nr, nq))
    m, _keep = omodel(kind, params, table)
    bad = lib().ko_block_from_model(ctypes.byref(h), ctypes.byref(m), ctypes.byref(mesh), int(index), float(t),
                                    _p(rv), _p(qv), _p(cons), _p(sm))
    return rv, qv, cons, sm, bad


# ---- per-zone trait methods --------------------------------------------------
def try_to_primitive(h, cons, guess=None):
    cons = _f64(cons).reshape(-1, 4)
    n = cons.shape[0]
    prim = np.zeros((n, 4))
    code = np.zeros(n, np.int32)
    iters = np.zeros(n, np.int32)
    g = _f64(guess).reshape(-1) if guess is not None else None
    lib().ko_try_to_primitive(ctypes.byref(h), ctypes.c_int64(n), _p(cons), _p(g), _p(prim), _p(code), _p(iters))
    return prim, code, iters


def to_conserved(h, prim):
    prim = _f64(prim).reshape(-1, 4)
    cons = np.zeros_like(prim)
    lib().ko_to_conserved(ctypes.byref(h), ctypes.c_int64(prim.shape[0]), _p(prim), _p(cons))
    return cons


def intercell_flux(h, pl, pr, sl, sr, axis):
    pl = _f64(pl).reshape(-1, 4)
    pr = _f64(pr).reshape(-1, 4)
    n = pl.shape[0]
    sl = _f64(np.broadcast_to(sl, (n,)))
    sr = _f64(np.broadcast_to(sr, (n,)))
    f = np.zeros((n, 4))
    g = np.zeros(n)
    lib().ko_intercell_flux(ctypes.byref(h), ctypes.c_int64(n), _p(pl), _p(pr), _p(sl), _p(sr),
                            ctypes.c_int32(axis), _p(f), _p(g))
    return f, g


def wavespeeds(h, prim, axis):
    prim = _f64(prim).reshape(-1, 4)
    n = prim.shape[0]
    lm = np.zeros(n)
    lp = np.zeros(n)
    lib().ko_wavespeeds(ctypes.byref(h), ctypes.c_int64(n), _p(prim), ctypes.c_int32(axis), _p(lm), _p(lp))
    return lm, lp


def geometrical_source_terms(h, prim, r, theta):
    prim = _f64(prim).reshape(-1, 4)
    n = prim.shape[0]
    r = _f64(np.broadcast_to(r, (n,)))
    theta = _f64(np.broadcast_to(theta, (n,)))
    out = np.zeros((n, 4))
    lib().ko_geometrical_source_terms(ctypes.byref(h), ctypes.c_int64(n), _p(prim), _p(r), _p(theta), _p(out))
    return out


def max_signal_speed(h, prim):
    prim = _f64(prim).reshape(-1, 4)
    out = np.zeros(prim.shape[0])
    lib().ko_max_signal_speed(ctypes.byref(h), ctypes.c_int64(prim.shape[0]), _p(prim), _p(out))
    return out


def plm_gradient(theta, a, b, c):
    a, b, c = (_f64(x).reshape(-1) for x in (a, b, c))
    out = np.zeros_like(a)
    lib().ko_plm_gradient(theta, a.size, _p(a), _p(b), _p(c), _p(out))
    return out


def products(h, rv, qv, cons, scalar_mass):
    """BlockProducts of one block: (primitive (nr, nq, 4) in AnyPrimitive order
    (u_r, u_theta, rho, p), scalar (nr, nq)); raises OracleError on a c2p
    failure."""
    rv, qv = _f64(rv), _f64(qv)
    nr, nq = rv.size - 1, qv.size - 1
    cons = _f64(cons).reshape(nr, nq, 4)
    sm = _f64(scalar_mass).reshape(nr, nq)
    prim = np.zeros((nr, nq, 4))
    scal = np.zeros((nr, nq))
    err = KoError()
    if lib().ko_products(ctypes.byref(h), nr, nq, _p(rv), _p(qv), _p(cons), _p(sm), _p(prim), _p(scal),
                         ctypes.byref(err)):
        raise OracleError(err)
    return prim, scal


def geometry_literal(rv, qv):
    """Ar, Aq, dV as src/mesh.rs:119-134 writes them, zone by zone."""
    rv = _f64(rv)
    qv = _f64(qv)
    nr, nq = rv.size - 1, qv.size - 1
    Ar = np.zeros((nr + 1, nq))
    Aq = np.zeros((nr, nq + 1))
    dV = np.zeros((nr, nq))
    lib().ko_geometry_literal(nr, nq, _p(rv), _p(qv), _p(Ar), _p(Aq), _p(dV))
    return dict(Ar=Ar, Aq=Aq, dV=dV)


def geometry(rv, qv):
    rv = _f64(rv)
    qv = _f64(qv)
    nr, nq = rv.size - 1, qv.size - 1
    Ar = np.zeros((nr + 1, nq))
    Aq = np.zeros((nr, nq + 1))
    dV = np.zeros((nr, nq))
    rc = np.zeros((nr, nq))
    cot = np.zeros((nr, nq))
    dl = np.zeros((nr, nq))
    lib().ko_geometry(nr, nq, _p(rv), _p(qv), _p(Ar), _p(Aq), _p(dV), _p(rc), _p(cot), _p(dl))
    return dict(Ar=Ar, Aq=Aq, dV=dV, rc=rc, cot=cot, dl=dl)


# ---- whole-mesh simulation ---------------------------------------------------
class OracleError(RuntimeError):
    def __init__(self, err):
        self.code = err.code
        self.block, self.i, self.j = err.block, err.i, err.j
        self.r, self.theta, self.value = err.r, err.theta, err.value
        super().__init__(f"{CODES.get(err.code, err.code)} at block {err.block} zone ({err.i},{err.j}) "
                         f"(r, theta) = ({err.r:.4e}, {err.theta:.4f}) value={err.value:.4e}")


class Sim:
    """A radially ordered chain of blocks advanced by the oracle."""

    def __init__(self, h, nthreads=1, literal=False):
        """literal=True (or a mask of LIT_*) evaluates the reference's own
        expressions as written (LIT_GEOMETRY: per-zone frustum areas and
        volumes; LIT_VOLUME: U/dV and M/dV/Gamma) and the textbook forms of
        the crates' arithmetic with every quotient divided out (LIT_SOURCES,
        LIT_FLUX), instead of the rewritten forms the device shares with the
        default oracle."""
        self.h = h
        self._p = ctypes.c_void_p(lib().ko_sim_create(ctypes.byref(h), int(nthreads)))
        if literal:
            lib().ko_sim_set_literal(self._p, LIT_ALL if literal is True else int(literal))
        self.nq = None
        self.qv = None
        self.radial_vertices = []
        self._sources = [None, None]

    def __del__(self):
        if getattr(self, "_p", None):
            lib().ko_sim_destroy(self._p)
            self._p = None

    def push_block(self, rv, qv, cons, scalar_mass):
        rv, qv = _f64(rv), _f64(qv)
        nr, nq = rv.size - 1, qv.size - 1
        cons = _f64(cons).reshape(nr, nq, 4)
        sm = _f64(scalar_mass).reshape(nr, nq)
        self.nq = nq
        self.qv = qv.copy()
        lib().ko_sim_push_block(self._p, nr, nq, _p(rv), _p(qv), _p(cons), _p(sm))
        self.radial_vertices.append(rv.copy())

    def pop_inner_block(self):
        lib().ko_sim_pop_inner_block(self._p)
        self.radial_vertices.pop(0)

    def blocks(self):
        """(cons, scalar_mass) of every block, inner to outer."""
        return [self.get_block(b, rv.size - 1, self.nq) for b, rv in enumerate(self.radial_vertices)]

    def excise(self, ex, source, first_index, mesh=None):
        """add_remove_blocks (src/scheme.rs:179-206) at the current time:
        remove the inner block when its outer radius is inside the IES
        (src/scheme.rs:191-194), add block outer+1 from `source` when the
        outer block's outer radius is inside the OES (src/scheme.rs:196-205);
        surfaces r0 + max(t - delay, 0) v (src/mesh.rs:318-336).  A block's
        outer radius is the mesh's subgrid_extent when `mesh` (an OMesh) is
        given, else its last radial vertex.  The only block is never removed.
        Returns the index of the inner block."""
        t = self.time
        ies = ex.inner_radius + max(t - ex.delay, 0.0) * ex.inner_speed
        oes = ex.outer_radius + max(t - ex.delay, 0.0) * ex.outer_speed
        outer_index = first_index + len(self.radial_vertices) - 1

        def outer_radius(k, index):
            return lib().ko_mesh_outer_radius(ctypes.byref(mesh), int(index)) if mesh is not None else \
                self.radial_vertices[k][-1]
        remove = len(self.radial_vertices) > 1 and outer_radius(0, first_index) < ies
        add = source is not None and outer_radius(-1, outer_index) < oes
        if remove:
            self.pop_inner_block()
            first_index += 1
        if add:
            rv, cons, sm = source(outer_index + 1, t)
            self.push_block(rv, self.qv, cons, sm)
        self.set_first_index(first_index)
        return first_index

    def scheme_advance(self, fold, ex=None, source=None, first_index=0, mesh=None):
        """scheme::advance (src/scheme.rs:212-240): dt from time_step once,
        then `fold` RK steps, each preceded by the excision check when the
        surfaces move (src/scheme.rs:228-232; every ex.check_every
        iterations).  Returns (dt, index of the inner block)."""
        dt = self.time_step()
        moving = ex is not None and (ex.inner_speed > 0.0 or ex.outer_speed > 0.0)
        for _ in range(fold):
            if moving and self.iteration % ex.check_every == 0:
                first_index = self.excise(ex, source, first_index, mesh)
            self.advance(dt, 1)
        return dt, first_index

    def num_blocks(self):
        return lib().ko_sim_num_blocks(self._p)

    def set_first_index(self, index):
        """Block index of the innermost block (ghost blocks of model sides)."""
        lib().ko_sim_set_first_index(self._p, int(index))

    def set_boundary_model(self, side, mesh, kind=None, params=(), table=None):
        """Ghost rows of `side` from the initial model at every stage's input
        time (src/scheme.rs:30-52); kind None clears."""
        if kind is None:
            lib().ko_sim_set_boundary_model(self._p, side, None, None)
            return
        m, _keep = omodel(kind, params, table)
        lib().ko_sim_set_boundary_model(self._p, side, ctypes.byref(m), ctypes.byref(mesh))

    def set_boundary_source(self, side, fn):
        """Ghost rows of `side` from fn(side, time) -> (prim (2, nq, 4),
        scalar (2, nq)) at every stage's input time, or None to clear."""
        if fn is None:
            self._sources[side] = None
            lib().ko_sim_set_boundary_source(self._p, side, None, None)
            return

        def call(user, sd, t, nq, prim, scalar):
            try:
                P, S = fn(int(sd), float(t))
                np.ctypeslib.as_array(prim, shape=(8 * nq,))[:] = np.asarray(P, np.float64).reshape(-1)
                np.ctypeslib.as_array(scalar, shape=(2 * nq,))[:] = np.asarray(S, np.float64).reshape(-1)
                return 0
            except Exception:
                return 1

        cb = BOUNDARY_SOURCE(call)
        self._sources[side] = cb
        lib().ko_sim_set_boundary_source(self._p, side, cb, None)

    def set_boundary(self, side, mode, prim=None, scalar=None):
        nq = self.nq
        if mode == FIXED:
            prim = _f64(prim).reshape(2 * nq, 4)
            scalar = _f64(scalar).reshape(2 * nq)
        lib().ko_sim_set_boundary(self._p, side, mode, nq, _p(prim), _p(scalar))

    def set_block(self, b, cons, scalar_mass):
        """Replaces the state of block b (same shape)."""
        cons = _f64(cons)
        sm = _f64(scalar_mass)
        lib().ko_sim_set_block(self._p, int(b), _p(cons), _p(sm))

    def get_block(self, b, nr, nq):
        cons = np.zeros((nr, nq, 4))
        sm = np.zeros((nr, nq))
        lib().ko_sim_get_block(self._p, b, _p(cons), _p(sm))
        return cons, sm

    @property
    def time(self):
        return lib().ko_sim_time(self._p)

    @time.setter
    def time(self, t):
        lib().ko_sim_set_time(self._p, float(t))

    @property
    def iteration(self):
        return lib().ko_sim_iteration(self._p)

    def time_step(self):
        dt = ctypes.c_double(0.0)
        err = KoError()
        if lib().ko_sim_time_step(self._p, ctypes.byref(dt), ctypes.byref(err)):
            raise OracleError(err)
        return dt.value

    def advance(self, dt, nsteps):
        err = KoError()
        if lib().ko_sim_advance(self._p, float(dt), int(nsteps), ctypes.byref(err)):
            raise OracleError(err)

    def c2p_stats(self):
        z, it, fl = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        lib().ko_sim_c2p_stats(self._p, ctypes.byref(z), ctypes.byref(it), ctypes.byref(fl))
        return z.value, it.value, fl.value


# ---- learned-guess MLP (FP64 reference) -------------------------------------
def mlp_forward(params, x):
    """y = W3 tanh(W2 tanh(W1 x + b1) + b2) + b3 in FP64 (knc_mlp_oracle.cpp)."""
    P = _f64(params)
    x = _f64(x).reshape(-1, 4)
    y = np.zeros(x.shape[0])
    lib().ko_mlp_forward(_p(P), ctypes.c_int64(x.shape[0]), _p(x), _p(y))
    return y


def mlp_loss_grad(params, x, t):
    """Mean squared error and its gradient in FP64."""
    P = _f64(params)
    x = _f64(x).reshape(-1, 4)
    t = _f64(t).reshape(-1)
    loss = ctypes.c_double()
    grad = np.zeros(P.size)
    lib().ko_mlp_loss_grad(_p(P), ctypes.c_int64(x.shape[0]), _p(x), _p(t), ctypes.byref(loss), _p(grad))
    return loss.value, grad
