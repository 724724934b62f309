"""Learned initial pressure guess for the SRHD c2p (host side).

``GuessNetwork`` owns a ``knc_mlp`` (include/knc_gpu.h): a 4 -> 64 -> 64 -> 1
tanh MLP trained on the GPU by the library's own forward / backward / Adam
kernels and evaluated by its BF16-MFMA inference kernel.  Training pairs are
(per-volume conserved state, pressure): seeded samples over Gamma in
[1, 100], rho in [1e-8, 1], p/rho in [1e-6, 10] (built with to_conserved),
plus states harvested from the solver's own runs.
"""
import ctypes

import numpy as np

from . import capi
from .setups import to_conserved

NPARAM = 4545
NNORM = 10


def _lib():
    L = capi.lib()
    if not getattr(L, "_mlp_bound", False):
        P, I32, I64, U64, F = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float
        L.knc_mlp_last_error_message.restype = ctypes.c_char_p
        L.knc_mlp_create.argtypes = [I32, ctypes.POINTER(P)]
        L.knc_mlp_destroy.argtypes = [P]
        L.knc_mlp_destroy.restype = None
        L.knc_mlp_init.argtypes = [P, U64]
        L.knc_mlp_set_params.argtypes = [P, P, P]
        L.knc_mlp_get_params.argtypes = [P, P, P]
        L.knc_mlp_loss_grad.argtypes = [P, I64, P, P, P, P]
        L.knc_mlp_train.argtypes = [P, I64, P, P, I32, I32, F, U64, P]
        L.knc_mlp_features.argtypes = [I64, P, P, P]
        L.knc_mlp_predict.argtypes = [P, I64, P, P, P, P]
        L.knc_context_set_mlp.argtypes = [P, P]
        L._mlp_bound = True
    return L


def _check(rc):
    if rc != 0:
        msg = _lib().knc_mlp_last_error_message().decode() or _lib().knc_last_error_message().decode()
        raise RuntimeError(f"knc_mlp error {rc}: {msg}")


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def features(cons, gamma):
    """Raw features (the device definition) of per-volume conserved states."""
    cons = np.ascontiguousarray(cons, dtype=np.float64).reshape(-1, 4)
    n = cons.shape[0]
    g = np.ascontiguousarray(np.broadcast_to(gamma, (n,)), dtype=np.float64)
    f = np.zeros((n, 4), np.float32)
    _check(_lib().knc_mlp_features(n, _p(cons), _p(g), _p(f)))
    return f


def synthetic_pairs(n, seed=0, gammas=(4.0 / 3.0, 5.0 / 3.0)):
    """Seeded (conserved, pressure, gamma) samples: Gamma log-uniform in
    [1, 100], rho log-uniform in [1e-8, 1], p/rho log-uniform in [1e-6, 10],
    random flow direction in the (r, theta) plane."""
    rng = np.random.default_rng(seed)
    G = np.exp(rng.uniform(0.0, np.log(100.0), n))
    rho = np.exp(rng.uniform(np.log(1e-8), 0.0, n))
    T = np.exp(rng.uniform(np.log(1e-6), np.log(10.0), n))
    u = np.sqrt(G * G - 1.0)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    prim = np.stack([rho, u * np.cos(phi), u * np.sin(phi), rho * T], 1)
    gam = np.asarray(gammas)[rng.integers(0, len(gammas), n)]
    cons = np.empty_like(prim)
    for g in gammas:
        m = gam == g
        cons[m] = to_conserved(1, g, prim[m])
    return cons, prim[:, 3].copy(), gam


def harvest_pairs(ctx, hydro):
    """(conserved per volume, pressure) of every zone of a context's current
    state; pressures from the library's GPU c2p."""
    from .setups import cell_volumes
    cons_all = []
    for idx in ctx.block_indexes():
        cons, _ = ctx.download_block(idx)
        rv, qv = ctx.vertices[idx]
        cons_all.append((cons / cell_volumes(rv, qv)[..., None]).reshape(-1, 4))
    cons = np.concatenate(cons_all)
    prim, code, _ = capi.try_to_primitive(hydro, cons)
    ok = (code == 0) & (prim[:, 3] > 0)
    return cons[ok], prim[ok, 3].copy()


class GuessNetwork:
    """A knc_mlp: parameters, normalisation, GPU training and inference."""

    def __init__(self, device=0, seed=1):
        h = ctypes.c_void_p()
        _check(_lib().knc_mlp_create(device, ctypes.byref(h)))
        self._h = h
        _check(_lib().knc_mlp_init(self._h, seed))

    def close(self):
        if getattr(self, "_h", None):
            _lib().knc_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    def get(self):
        p = np.zeros(NPARAM, np.float32)
        n = np.zeros(NNORM, np.float32)
        _check(_lib().knc_mlp_get_params(self._h, _p(p), _p(n)))
        return p, n

    def set(self, params, norm):
        p = np.ascontiguousarray(params, np.float32)
        n = np.ascontiguousarray(norm, np.float32)
        _check(_lib().knc_mlp_set_params(self._h, _p(p), _p(n)))

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
        t = np.ascontiguousarray(t, np.float32).reshape(-1)
        loss = ctypes.c_float()
        grad = np.zeros(NPARAM, np.float32)
        _check(_lib().knc_mlp_loss_grad(self._h, x.shape[0], _p(x), _p(t), ctypes.byref(loss), _p(grad)))
        return loss.value, grad

    def fit_normalisation(self, feats, target):
        """shift/scale of each input and of the target from a training set."""
        params, _ = self.get()
        norm = np.zeros(NNORM, np.float32)
        norm[0:4] = feats.mean(0)
        norm[4:8] = 1.0 / np.maximum(feats.std(0), 1e-6)
        norm[8] = target.mean()
        norm[9] = 1.0 / max(float(target.std()), 1e-6)
        self.set(params, norm)
        return norm

    def normalise(self, feats, target=None):
        _, norm = self.get()
        x = (feats - norm[0:4]) * norm[4:8]
        if target is None:
            return x.astype(np.float32)
        return x.astype(np.float32), ((target - norm[8]) * norm[9]).astype(np.float32)

    def train(self, cons, p, gamma, epochs=3, batch=8192, lr=2e-3, seed=0):
        """Fits the normalisation and trains on (conserved, pressure) pairs;
        returns the mean training loss per epoch."""
        f = features(cons, gamma)
        target = np.log2(p / cons[:, 0]).astype(np.float32)
        self.fit_normalisation(f, target)
        x, t = self.normalise(f, target)
        losses = np.zeros(epochs, np.float32)
        _check(_lib().knc_mlp_train(self._h, x.shape[0], _p(x), _p(t), int(epochs), int(batch), float(lr),
                                    int(seed), _p(losses)))
        return losses

    def predict(self, cons, gamma):
        """(normalised output, pressure guess) from the BF16-MFMA kernel."""
        cons = np.ascontiguousarray(cons, np.float64).reshape(-1, 4)
        n = cons.shape[0]
        g = np.ascontiguousarray(np.broadcast_to(gamma, (n,)), dtype=np.float64)
        y = np.zeros(n, np.float32)
        guess = np.zeros(n)
        _check(_lib().knc_mlp_predict(self._h, n, _p(cons), _p(g), _p(y), _p(guess)))
        return y, guess

    def attach(self, ctx):
        _check(_lib().knc_context_set_mlp(ctx._h, self._h))
