"""Learned warm start for the XPBD substep: the network object and its training driver.

No reference counterpart: the network runs between predict and the projection
loop of the reference substep (crates/cloth_core/src/simulation/dispatch.rs:137
-> 169).  A message-passing network over the cloth's stretch-edge graph (hidden
width 128, `rounds` rounds of two-layer MLPs, BF16 MFMA with FP32 accumulation)
predicts every free vertex's position correction for the coming substep, in
units of s = max(|g|, 1) dt^2, relative to w times the previous substep's
correction (w = history_weight, 0.5 by default: the previous correction also
reaches the substep through the velocity, and a larger w amplifies the error
modes that a few iterations correct slowly).
Plain XPBD iterations always follow; a residual gate adds plain iterations when
the relative residual after them stays above a threshold.

    net = WarmStartNet(rounds=3); net.init(seed=1)
    sim.ws_record(capacity=400, stride=2, history_weight=net.history_weight)
    sim.step(...)                                             # plain 64-iteration solves
    net.train([sim], steps=2000, lr=1e-3)
    sim.ws_record(0); sim.set_warm_start(net, mode=2, iterations=8, fallback_iterations=56,
                                         residual_threshold=5e-3)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

HIDDEN, FEATURES, OUTPUTS = 128, 32, 16

# warm-start modes of ClothSimulator.set_warm_start
OFF, PREVIOUS, NETWORK = 0, 1, 2


class WarmStartNet:
    """Device-resident weights, Adam state and training scratch of one network."""

    def __init__(self, rounds: int = 3, history_weight: float = 0.5):
        self._L = _lib.load()
        h = C.c_void_p()
        _lib.check(self._L.tc_ws_net_create(int(rounds), C.byref(h)))
        self._h = h
        self.rounds = int(rounds)
        self.history_weight = float(history_weight)  # weight of the previous correction it was trained with

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._L.tc_ws_net_destroy(h)

    @property
    def handle(self):
        return self._h

    @property
    def num_params(self) -> int:
        return int(self._L.tc_ws_net_num_params(self._h))

    def init(self, seed: int = 0):
        _lib.check(self._L.tc_ws_net_init(self._h, C.c_uint64(seed)))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        _lib.check(self._L.tc_ws_net_get_params(self._h, out.ctypes.data_as(C.c_void_p), out.size))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32).reshape(-1)
        _lib.check(self._L.tc_ws_net_set_params(self._h, p.ctypes.data_as(C.c_void_p), p.size))

    def scale_output(self, gain: float):
        """Multiply the output layer (weights and biases) by `gain`, which scales the predicted
        correction by `gain`; exact for powers of two."""
        p = self.get_params()
        n_out = HIDDEN * OUTPUTS + OUTPUTS  # the last layer: H x O weights, then O biases
        p[-n_out:] *= np.float32(gain)
        self.set_params(p)

    def save(self, path, **extra):
        np.savez(path, params=self.get_params(), rounds=np.int64(self.rounds),
                 history_weight=np.float64(self.history_weight), **extra)

    @classmethod
    def load(cls, path) -> "WarmStartNet":
        with np.load(path, allow_pickle=False) as z:
            hw = float(z["history_weight"]) if "history_weight" in z.files else 1.0
            net = cls(int(z["rounds"]), hw)
            net.set_params(z["params"])
        return net

    def forward_backward(self, features, graph_off, graph_nbr, targets=None):
        """(y rows x 16, grad or None, loss or None) for host arrays; see tc_ws_net_forward_backward."""
        f = np.ascontiguousarray(features, np.float32).reshape(-1, FEATURES)
        rows = len(f)
        off = np.ascontiguousarray(graph_off, np.uint32).reshape(-1)
        nbr = np.ascontiguousarray(graph_nbr, np.uint32).reshape(-1)
        if off.size != rows + 1:
            raise ValueError("graph_off must hold rows + 1 offsets")
        t = None if targets is None else np.ascontiguousarray(targets, np.float32).reshape(rows, 4)
        y = np.empty((rows, OUTPUTS), np.float32)
        g = None if t is None else np.empty(self.num_params, np.float32)
        loss = C.c_double(0.0)
        vp = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)  # noqa: E731
        _lib.check(self._L.tc_ws_net_forward_backward(self._h, vp(f), rows, vp(off), vp(nbr if nbr.size else None),
                                                      vp(t), vp(y), vp(g), C.byref(loss)))
        return y, g, (None if t is None else loss.value)

    def adam_step(self, lr: float):
        _lib.check(self._L.tc_ws_net_adam_step(self._h, float(lr)))

    def train(self, sims, steps: int, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        """Adam steps on the samples the simulators recorded; returns the loss of every step."""
        hs = (C.c_void_p * len(sims))(*[s._h for s in sims])
        losses = np.empty(int(steps), np.float32)
        _lib.check(self._L.tc_ws_train(self._h, hs, len(sims), int(steps), float(lr), C.c_uint64(seed),
                                       losses.ctypes.data_as(C.c_void_p)))
        return losses

    def eval(self, sims) -> float:
        """Mean loss over every recorded sample of the simulators (no update)."""
        hs = (C.c_void_p * len(sims))(*[s._h for s in sims])
        out = C.c_double(0.0)
        _lib.check(self._L.tc_ws_eval(self._h, hs, len(sims), C.byref(out)))
        return out.value
