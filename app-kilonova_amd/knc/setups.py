"""Synthetic, seeded initial and boundary data for the benchmark workloads.

Each builder returns a :class:`Setup`: the hydro parameters, a radially
ordered list of blocks (index, radial vertices, polar vertices, conserved,
scalar mass), the ghost-row boundary specification and the excision speeds.
Block states are built as ``BlockState::from_model`` does
(``src/state.rs:55-69``): primitive at the zone centre -> ``to_conserved`` x
cell volume, and ``scalar_mass = D dV X``.  Vertices and volumes follow
``src/mesh.rs:119-134,157-161,202-210``.  Random fields are generated in
global (row, column) index space, so the data do not depend on how the
radial range is cut into blocks.

Workloads (sizes nr_total x nq):
  euler_blast   256 x 64, one block, Newtonian gamma=5/3 blast in rho ~ r^-2
  jet_in_cloud  1024 x 256 in 8 blocks, SRHD jet (Gamma=10) into a clumpy cloud
  wind_shock    uneven blocks of {64,128,256,512} rows x 256, SRHD wind (Gamma<=30)
  bomb_jet      4096 x 1024 in 32 uneven blocks, SRHD shells Gamma in [2,100] + jet
"""
from dataclasses import dataclass, field

import numpy as np

LIGHT_SPEED = 3e10


@dataclass
class HydroSpec:
    system: int              # 0 Euler, 1 SRHD
    riemann_solver: int      # 0 HLLE, 1 HLLC
    runge_kutta_order: int
    adaptive_time_step: bool
    gamma_law_index: float
    plm_theta: float
    cfl_number: float

    @property
    def light_speed(self):
        return LIGHT_SPEED if self.system == 1 else 1.0


@dataclass
class Block:
    index: int
    rv: np.ndarray
    qv: np.ndarray
    conserved: np.ndarray    # (nr, nq, 4)
    scalar_mass: np.ndarray  # (nr, nq)

    @property
    def nr(self):
        return self.rv.size - 1

    @property
    def nq(self):
        return self.qv.size - 1


@dataclass
class Boundary:
    mode: int                              # 0 fixed rows, 1 outflow
    primitive: np.ndarray = None           # (2, nq, 4) radial order
    scalar: np.ndarray = None              # (2, nq)


@dataclass
class Excision:
    """Moving excision surfaces (Mesh, src/mesh.rs:312-336): IES(t) =
    inner_radius + max(t - delay, 0) inner_speed, OES(t) likewise; checked
    before every step whose iteration is a multiple of check_every."""
    inner_radius: float
    outer_radius: float
    inner_speed: float = 0.0
    outer_speed: float = 0.0
    delay: float = 0.0
    check_every: int = 1

    @property
    def moving(self):
        return self.inner_speed > 0.0 or self.outer_speed > 0.0


@dataclass
class Setup:
    name: str
    hydro: HydroSpec
    blocks: list
    inner: Boundary
    outer: Boundary
    fold: int = 10
    steps: int = 100
    excision: Excision = None
    block_source: object = None    # (index, time) -> (rv, conserved, scalar_mass) of a new outer block
    model_extra: dict = field(default_factory=dict)

    @property
    def total_zones(self):
        return sum(b.nr * b.nq for b in self.blocks)


# ---- geometry (mesh.rs) ---------------------------------------------------------
def polar_vertices(nq, q0=0.0, q1=np.pi):
    """vertex_coordinate_signed's theta: q0 + dq*j (mesh.rs:202-210)."""
    dq = (q1 - q0) / nq
    return q0 + dq * np.arange(nq + 1, dtype=np.float64)


def log_vertices(r0, r1, n, k0=0, k1=None):
    """Global log-spaced vertices r0*10^(k dlog), k in [k0, k1]."""
    if k1 is None:
        k1 = n
    y0 = np.log10(r0)
    dy = (np.log10(r1) - y0) / n
    return 10.0 ** (y0 + dy * np.arange(k0, k1 + 1, dtype=np.float64))


def cell_volumes(rv, qv):
    r0, r1 = rv[:-1, None], rv[1:, None]
    dcost = -(np.cos(qv[None, 1:]) - np.cos(qv[None, :-1]))
    return 2.0 * np.pi * ((r1 * r1) * r1 - (r0 * r0) * r0) / 3.0 * dcost


def cell_centers(rv, qv):
    rc = np.sqrt(rv[:-1] * rv[1:])
    qc = 0.5 * (qv[:-1] + qv[1:])
    return np.broadcast_to(rc[:, None], (rc.size, qc.size)), np.broadcast_to(qc[None, :], (rc.size, qc.size))


# ---- p2c on the host (closed form, same as to_conserved) --------------------------
def to_conserved(system, gamma, prim):
    rho, u1, u2, p = (prim[..., k] for k in range(4))
    out = np.empty(prim.shape)
    if system == 0:
        out[..., 0] = rho
        out[..., 1] = rho * u1
        out[..., 2] = rho * u2
        out[..., 3] = p / (gamma - 1.0) + 0.5 * rho * (u1 * u1 + u2 * u2)
    else:
        W = np.sqrt(1.0 + (u1 * u1 + u2 * u2))
        rhoh = rho + p * (gamma / (gamma - 1.0))
        hw = rhoh * W
        out[..., 0] = rho * W
        out[..., 1] = hw * u1
        out[..., 2] = hw * u2
        out[..., 3] = (hw * W - p) - rho * W
    return out


def lorentz(system, prim):
    if system == 0:
        return np.ones(prim.shape[:-1])
    return np.sqrt(1.0 + prim[..., 1] ** 2 + prim[..., 2] ** 2)


def correlated_gaussian(shape, corr, rng):
    """Unit-variance Gaussian random field with a Gaussian correlation of
    `corr` zones, by filtering white noise in Fourier space."""
    white = rng.standard_normal(shape)
    f = np.fft.rfft2(white)
    kx = np.fft.fftfreq(shape[0])[:, None]
    ky = np.fft.rfftfreq(shape[1])[None, :]
    f *= np.exp(-2.0 * (np.pi * corr) ** 2 * (kx * kx + ky * ky))
    g = np.fft.irfft2(f, s=shape)
    g -= g.mean()
    g /= g.std()
    return g


def _blocks_from_prim(hydro, rv_global, qv, sizes, prim, X, first_index=0):
    blocks = []
    k = 0
    for b, nr in enumerate(sizes):
        rv = rv_global[k:k + nr + 1].copy()
        P = prim[k:k + nr]
        dv = cell_volumes(rv, qv)
        cons = to_conserved(hydro.system, hydro.gamma_law_index, P) * dv[..., None]
        sm = cons[..., 0] * X[k:k + nr]
        blocks.append(Block(first_index + b, rv, qv.copy(), cons, sm))
        k += nr
    return blocks


def _ghost(prim, X):
    """Boundary rows: primitive (2, nq, 4) and scalar concentration rho*X."""
    return Boundary(0, np.ascontiguousarray(prim), np.ascontiguousarray(prim[..., 0] * X))


# ---- workloads ----------------------------------------------------------------------
def euler_blast(nr=256, nq=64, seed=1, steps=100):
    """Newtonian blast: r in [1,10] log, theta in [0,pi], gamma=5/3,
    rho = r^-2 (1 + 0.01 U(-1,1)), p = 1e-6 rho, thermal energy 1.0 in r<1.5;
    PLM 1.5, HLLC, RK2, CFL 0.4; fixed inner, outflow outer boundary."""
    hydro = HydroSpec(0, 1, 2, True, 5.0 / 3.0, 1.5, 0.4)
    rng = np.random.default_rng(seed)
    qv = polar_vertices(nq)
    rv = log_vertices(1.0, 10.0, nr)
    rc, qc = cell_centers(rv, qv)
    rho = rc ** -2.0 * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, (nr, nq)))
    p = 1e-6 * rho
    dv = cell_volumes(rv, qv)
    hot = rc < 1.5
    e_dep = 1.0 / dv[hot].sum()
    p = np.where(hot, p + (hydro.gamma_law_index - 1.0) * e_dep, p)
    prim = np.stack([rho, np.zeros_like(rho), np.zeros_like(rho), p], -1)
    X = hot.astype(np.float64)
    blocks = _blocks_from_prim(hydro, rv, qv, [nr], prim, X)
    rg = log_vertices(1.0, 10.0, nr, -2, 0)
    rcg = np.sqrt(rg[:-1] * rg[1:])[:, None] * np.ones((1, nq))
    gp = np.stack([rcg ** -2.0, 0 * rcg, 0 * rcg, 1e-6 * rcg ** -2.0], -1)
    return Setup("euler_blast", hydro, blocks, _ghost(gp, np.zeros((2, nq))), Boundary(1), fold=10, steps=steps)


def _jet_ghost(rg, qv, gamma_jet, rho_jet, p_jet, theta_jet, ambient_p_over_rho):
    nq = qv.size - 1
    rcg = np.sqrt(rg[:-1] * rg[1:])[:, None] * np.ones((1, nq))
    qc = 0.5 * (qv[:-1] + qv[1:])[None, :] * np.ones((2, 1))
    jet = qc < theta_jet
    u = np.sqrt(gamma_jet * gamma_jet - 1.0)
    rho = np.where(jet, rho_jet, rcg ** -2.0)
    p = np.where(jet, p_jet, ambient_p_over_rho * rcg ** -2.0)
    ur = np.where(jet, u, 0.0)
    prim = np.stack([rho, ur, np.zeros_like(rho), p], -1)
    return _ghost(prim, jet.astype(np.float64))


def jet_in_cloud(nr=1024, nq=256, nblocks=8, seed=2, steps=1000, riemann=0, corr=16.0):
    """SRHD jet in cloud: r in [1,100] log, nr x nq in `nblocks` blocks,
    gamma=4/3, rho = r^-2 lognormal(sigma=0.5, 16-zone clumps), p = 1e-4 rho;
    inner-boundary jet theta<0.1 with Gamma=10, rho=1e-3, p=1e-2; HLLE, RK2,
    inner excision surface moving at 0.5c.  The time step is the reference's
    SRHD default, cfl*dr0/c (adaptive_time_step = false): the CFL reduction
    over the block interiors would ignore the light-speed inflow at the
    inlet."""
    hydro = HydroSpec(1, riemann, 2, False, 4.0 / 3.0, 1.5, 0.4)
    rng = np.random.default_rng(seed)
    qv = polar_vertices(nq)
    rv = log_vertices(1.0, 100.0, nr)
    rc, qc = cell_centers(rv, qv)
    g = correlated_gaussian((nr, nq), corr, rng)
    sigma = 0.5
    rho = rc ** -2.0 * np.exp(sigma * g - 0.5 * sigma * sigma)
    prim = np.stack([rho, np.zeros_like(rho), np.zeros_like(rho), 1e-4 * rho], -1)
    sizes = [nr // nblocks] * nblocks
    blocks = _blocks_from_prim(hydro, rv, qv, sizes, prim, np.zeros_like(rho))
    inner = _jet_ghost(log_vertices(1.0, 100.0, nr, -2, 0), qv, 10.0, 1e-3, 1e-2, 0.1, 1e-4)
    return Setup("jet_in_cloud", hydro, blocks, inner, Boundary(1), fold=10, steps=steps,
                 excision=Excision(inner_radius=1.0, outer_radius=100.0, inner_speed=0.5 * LIGHT_SPEED))


def uneven_sizes(total, nblocks, choices, rng):
    """Seeded block sizes from `choices` with the given count and total."""
    choices = sorted(choices)
    for _ in range(100000):
        s = list(rng.choice(choices, nblocks))
        diff = total - sum(s)
        # repair toward the total by resizing blocks one step at a time
        guard = 0
        while diff != 0 and guard < 10000:
            guard += 1
            k = int(rng.integers(nblocks))
            i = choices.index(s[k])
            if diff > 0 and i + 1 < len(choices):
                step = choices[i + 1] - s[k]
                if step <= diff:
                    s[k] = choices[i + 1]
                    diff -= step
            elif diff < 0 and i > 0:
                step = s[k] - choices[i - 1]
                if step <= -diff:
                    s[k] = choices[i - 1]
                    diff += step
        if diff == 0:
            return [int(x) for x in s]
    raise ValueError("no size assignment found")


def bomb_jet(nr=4096, nq=1024, nblocks=32, seed=5, steps=2000, nshell=64, riemann=0, p_over_rho=1e-2):
    """SRHD kinetic bomb + jet: r in [1,100] log, nr x nq over `nblocks`
    uneven blocks (sizes from {1/2, 1, 2} x nr/nblocks), gamma=4/3; `nshell` radial shells
    with Gamma log-uniform in [2,100], rho = r^-2 lognormal(sigma=1, 16-zone
    correlation), p = 1e-2 rho; inner-boundary jet theta<0.1 with Gamma=10,
    rho=1e-3, p=1e-2; HLLE, RK2, CFL 0.4, adaptive dt.  The full-size mesh
    runs the 2000 steps of the benchmark without a hydro error (with
    p = 1e-3 rho the shell collisions drive tau negative near step 1400)."""
    hydro = HydroSpec(1, riemann, 2, True, 4.0 / 3.0, 1.5, 0.4)
    rng = np.random.default_rng(seed)
    qv = polar_vertices(nq)
    rv = log_vertices(1.0, 100.0, nr)
    rc, qc = cell_centers(rv, qv)
    g = correlated_gaussian((nr, nq), 16.0, rng)
    sigma = 1.0
    rho = rc ** -2.0 * np.exp(sigma * g - 0.5 * sigma * sigma)
    gam = np.exp(rng.uniform(np.log(2.0), np.log(100.0), nshell))
    shell = np.minimum((np.arange(nr) * nshell) // nr, nshell - 1)
    ur = np.sqrt(gam[shell] ** 2 - 1.0)[:, None] * np.ones((1, nq))
    prim = np.stack([rho, ur, np.zeros_like(rho), p_over_rho * rho], -1)
    if nblocks == 1:
        sizes = [nr]
    else:
        avg = nr // nblocks
        sizes = uneven_sizes(nr, nblocks, [avg // 2, avg, 2 * avg], rng)
    blocks = _blocks_from_prim(hydro, rv, qv, sizes, prim, np.zeros_like(rho))
    inner = _jet_ghost(log_vertices(1.0, 100.0, nr, -2, 0), qv, 10.0, 1e-3, 1e-2, 0.1, p_over_rho)
    return Setup("bomb_jet", hydro, blocks, inner, Boundary(1), fold=10, steps=steps)


class UnevenChain:
    """Radial blocks of seeded sizes from `choices`, block k's size drawn
    from its own generator (seed, k), so the chain extends outward without
    changing the blocks already made.  Vertices are global: r0 10^(dlog n)
    for global radial vertex n."""

    def __init__(self, seed, choices, r0, dlog):
        self.seed, self.choices, self.r0, self.dlog = seed, list(choices), r0, dlog
        self._sizes = []

    def size(self, k):
        while len(self._sizes) <= k:
            j = len(self._sizes)
            self._sizes.append(int(np.random.default_rng([self.seed, j]).choice(self.choices)))
        return self._sizes[k]

    def offset(self, k):
        return sum(self.size(j) for j in range(k))

    def vertices(self, k):
        n0 = self.offset(k)
        return self.r0 * 10.0 ** (self.dlog * np.arange(n0, n0 + self.size(k) + 1, dtype=np.float64))


def wind_shock(nq=256, nblocks=64, seed=3, steps=200, choices=(64, 128, 256, 512), riemann=0, check_every=50):
    """SRHD wind shock: `nblocks` radial blocks of seeded uneven sizes from
    `choices` x nq, r from 1 over 4 decades log (the log spacing continues
    past r = 1e4 for blocks born later); slow wind (u=1) with rho ~ r^-2
    (1 + 0.1 cos 4 theta) and p = 1e-3 rho; the inner boundary injects a
    fast wind with Gamma=30; HLLE, RK2, adaptive dt.  Excision every
    `check_every` steps: both surfaces move far faster than the mesh, so
    every check removes the innermost block and adds block outer+1 from the
    wind model (the reference's rule of src/scheme.rs:179-206 at a fixed
    cadence)."""
    hydro = HydroSpec(1, riemann, 2, True, 4.0 / 3.0, 1.5, 0.4)
    gamma = hydro.gamma_law_index
    probe = UnevenChain(seed, choices, 1.0, 1.0)
    nr = sum(probe.size(k) for k in range(nblocks))
    chain = UnevenChain(seed, choices, 1.0, 4.0 / nr)
    qv = polar_vertices(nq)

    def model_block(k):
        rv = chain.vertices(k)
        rc, qc = cell_centers(rv, qv)
        rho = rc ** -2.0 * (1.0 + 0.1 * np.cos(4.0 * qc))
        prim = np.stack([rho, np.ones_like(rho), np.zeros_like(rho), 1e-3 * rho], -1)
        cons = to_conserved(1, gamma, prim) * cell_volumes(rv, qv)[..., None]
        return rv, cons, np.zeros(rho.shape)

    blocks = []
    for k in range(nblocks):
        rv, cons, sm = model_block(k)
        blocks.append(Block(k, rv, qv.copy(), cons, sm))
    rg = chain.r0 * 10.0 ** (chain.dlog * np.arange(-2, 1, dtype=np.float64))
    rcg = np.sqrt(rg[:-1] * rg[1:])[:, None] * np.ones((1, nq))
    u = np.sqrt(30.0 ** 2 - 1.0)
    gp = np.stack([rcg ** -2.0, u + 0 * rcg, 0 * rcg, 1e-3 * rcg ** -2.0], -1)
    fast = 1e20 * LIGHT_SPEED
    exc = Excision(inner_radius=1.0, outer_radius=float(blocks[-1].rv[-1]), inner_speed=fast, outer_speed=fast,
                   check_every=check_every)
    return Setup("wind_shock", hydro, blocks, _ghost(gp, np.ones((2, nq))), Boundary(1), fold=check_every,
                 steps=steps, excision=exc, block_source=lambda index, time: model_block(index),
                 model_extra=dict(sizes=[b.nr for b in blocks]))


WORKLOADS = {"euler_blast": euler_blast, "jet_in_cloud": jet_in_cloud, "bomb_jet": bomb_jet,
             "wind_shock": wind_shock}
