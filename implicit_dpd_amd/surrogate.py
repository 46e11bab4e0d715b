"""Equation-of-state surrogate on the GPU: a 4 -> 128 -> 128 -> 2 SiLU
perceptron (BF16 matrix-core products, FP32 accumulation, FP32 Adam) that
maps normalised MDPD parameters (A, B, rd, kBT) to normalised (rho, P).

No counterpart in the reference: it stands in for the sweeps of plain runs
(test/mdpd_fluid/rho_vs_p.dat is one such sweep) when a parameter set for a
target density and pressure is wanted.
"""
import ctypes as C

import numpy as np

from . import capi

N_IN, N_HID, N_OUT = 4, 128, 2
# flat parameter layout (weights row-major [out][in])
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HID * N_IN
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_OUT * N_HID
N_PARAMS = OFF_B3 + N_OUT


def unpack(p):
    """Flat parameter vector -> (W1, b1, W2, b2, W3, b3) views."""
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(N_HID, N_IN), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(N_HID, N_HID), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(N_OUT, N_HID), p[OFF_B3:N_PARAMS])


def _f32(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.ndim == 2 and a.shape[1] == cols, a.shape
    return a


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Surrogate:
    """The device network; every call goes through libmdpd's mdpd_mlp_*."""

    def __init__(self, max_batch=4096, seed=1):
        self.L = capi.lib()
        n = C.c_uint32()
        self._c(self.L.mdpd_mlp_num_params(C.byref(n)))
        assert n.value == N_PARAMS
        h = C.c_void_p()
        rc = self.L.mdpd_mlp_create(int(max_batch), int(seed), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"mdpd_mlp_create failed ({rc})")
        self.h = h
        self.max_batch = int(max_batch)

    def _c(self, rc):
        if rc != 0:
            msg = self.L.mdpd_mlp_last_error(self.h).decode() if getattr(self, "h", None) else ""
            raise RuntimeError(f"mdpd_mlp error {rc}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.L.mdpd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = np.zeros(N_PARAMS, np.float32)
        self._c(self.L.mdpd_mlp_get_params(self.h, _fp(p), N_PARAMS))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (N_PARAMS,)
        self._c(self.L.mdpd_mlp_set_params(self.h, _fp(p), N_PARAMS))

    def forward(self, x):
        x = _f32(x, N_IN)
        y = np.zeros((len(x), N_OUT), np.float32)
        self._c(self.L.mdpd_mlp_forward(self.h, _fp(x), len(x), _fp(y)))
        return y

    def loss_grad(self, x, t):
        x, t = _f32(x, N_IN), _f32(t, N_OUT)
        loss = C.c_double()
        g = np.zeros(N_PARAMS, np.float32)
        self._c(self.L.mdpd_mlp_loss_grad(self.h, _fp(x), _fp(t), len(x), C.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, x, t, epochs, batch=256, lr=1e-3, seed=0):
        x, t = _f32(x, N_IN), _f32(t, N_OUT)
        hist = np.zeros(epochs, np.float32)
        self._c(self.L.mdpd_mlp_train(self.h, _fp(x), _fp(t), len(x), int(epochs), int(batch),
                                      float(lr), int(seed), _fp(hist)))
        return hist


class Scaler:
    """Affine normalisation of inputs and outputs to zero mean, unit spread."""

    def __init__(self, data):
        data = np.asarray(data, dtype=np.float64)
        self.mu = data.mean(axis=0)
        sd = data.std(axis=0)
        self.sd = np.where(sd > 0, sd, 1.0)

    def fwd(self, a):
        return (np.asarray(a, dtype=np.float64) - self.mu) / self.sd

    def inv(self, a):
        return np.asarray(a, dtype=np.float64) * self.sd + self.mu
