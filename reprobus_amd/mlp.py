"""Learned initial guess for the converged chemistry step (C ABI: include/rbmlp.h).

A 60 -> 256 -> 256 -> 256 -> 55 MLP (GELU) whose GEMMs run in BF16 on the matrix cores
with FP32 accumulation; forward, backward, Adam and inference are the library's own HIP
kernels.  Device tensors are torch tensors (torch is the allocator here); every compute
call goes through librbchem.so and raises if it is not built.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Grid, RbError
from .synth import LONG

NIN, NOUT, PAD, WIDTH = 60, 55, 64, 256
# qj1 slots (0-based) of the 40 integrated long-lived species: all but passive Ox,
# H2SO4 (no production or loss) and the diagnostic HNO3 gas
LONG_SLOTS = [i for i in range(42) if i not in (10, 11)]
assert LONG[10] == "ox_passive" and LONG[11] == "h2so4" and LONG[42] == "hno3_gas"
SHAPES = [("W1", (PAD, WIDTH)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)),
          ("W3", (WIDTH, WIDTH)), ("b3", (WIDTH,)), ("W4", (WIDTH, PAD)), ("b4", (PAD,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def split_params(flat: np.ndarray) -> dict:
    """Flat parameter vector -> named views (W_l is [in][out])."""
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = flat[o:o + n].reshape(shape)
        o += n
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, zero padding rows/columns."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(NPARAM, dtype=np.float32)
    P = split_params(flat)
    for name, fan_in, rows, cols in (("W1", NIN, NIN, WIDTH), ("W2", WIDTH, WIDTH, WIDTH),
                                     ("W3", WIDTH, WIDTH, WIDTH), ("W4", WIDTH, WIDTH, NOUT)):
        lim = np.sqrt(6.0 / fan_in)
        P[name][:rows, :cols] = rng.uniform(-lim, lim, (rows, cols)).astype(np.float32)
    P["W4"] *= 0.1
    return flat


def _ptr(t) -> int:
    return t.data_ptr()


class GuessMLP:
    def __init__(self, device: int = 0):
        self._L = _lib.load()
        self._m = C.c_void_p()
        rc = self._L.rb_mlp_create(C.byref(self._m), device)
        if rc != 0:
            raise RbError(rc, "rb_mlp_create failed")
        assert self._L.rb_mlp_param_count() == NPARAM

    def close(self):
        if self._m:
            self._L.rb_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            raise RbError(rc, self._L.rb_mlp_last_error(self._m).decode())

    def set_stream(self, stream) -> None:
        self._check(self._L.rb_mlp_set_stream(self._m, stream))

    def set_params(self, flat: np.ndarray) -> None:
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.size == NPARAM
        self._check(self._L.rb_mlp_set_params(self._m, flat.ctypes.data))

    def get_params(self) -> np.ndarray:
        out = np.empty(NPARAM, dtype=np.float32)
        self._check(self._L.rb_mlp_get_params(self._m, out.ctypes.data))
        return out

    def set_norm(self, in_mu, in_sd, out_mu, out_sd) -> None:
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (in_mu, in_sd, out_mu, out_sd)]
        assert a[0].size == NIN and a[1].size == NIN and a[2].size == NOUT and a[3].size == NOUT
        self._norm = a
        self._check(self._L.rb_mlp_set_norm(self._m, *[v.ctypes.data for v in a]))

    def forward(self, X, Y) -> None:
        """X: device (n, 60) float32 raw features -> Y: device (n, 55) raw log increments."""
        assert X.shape[1] == NIN and Y.shape == (X.shape[0], NOUT)
        self._check(self._L.rb_mlp_forward_dev(self._m, _ptr(X), X.shape[0], _ptr(Y)))

    def train_step(self, X, T, lr: float, want_loss: bool = False):
        loss = C.c_float(0.0)
        self._check(self._L.rb_mlp_train_step_dev(self._m, _ptr(X), _ptr(T), X.shape[0], lr,
                                                  C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def gradients(self, X, T):
        g = np.empty(NPARAM, dtype=np.float32)
        loss = C.c_float(0.0)
        self._check(self._L.rb_mlp_gradients_dev(self._m, _ptr(X), _ptr(T), X.shape[0],
                                                 g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def features(self, grid: Grid, dt: float, T, p, sza, qj1, hc, X) -> None:
        self._check(self._L.rb_mlp_features_dev(self._m, C.byref(grid), float(dt), _ptr(T), _ptr(p),
                                                _ptr(sza), _ptr(qj1), _ptr(hc), _ptr(X)))

    def guess(self, grid: Grid, dt: float, T, p, sza, qj1, hc, guess_qj1, guess_hc) -> None:
        self._check(self._L.rb_mlp_guess_dev(self._m, C.byref(grid), float(dt), _ptr(T), _ptr(p),
                                             _ptr(sza), _ptr(qj1), _ptr(hc), _ptr(guess_qj1),
                                             _ptr(guess_hc)))


def log_increments(q_in, h_in, q_out, h_out, nivbas: int):
    """Training targets (device, torch): log(max(x_out,1e-30)/max(x_in,1e-30)) of the 55
    species on levels 1..nivbas, as (nbox, 55) float32 in box order."""
    import torch

    def lg(a):
        return torch.log(torch.clamp(a[:, :nivbas], min=1e-30))

    idx = torch.tensor(LONG_SLOTS, device=q_in.device)
    dq = (lg(q_out.index_select(0, idx)) - lg(q_in.index_select(0, idx))).reshape(len(LONG_SLOTS), -1)
    dh = (lg(h_out) - lg(h_in)).reshape(h_in.shape[0], -1)
    return torch.cat([dq, dh], 0).t().contiguous().float()
