"""Learned warm start for the contact solver (no counterpart in the reference).

The reference warm-starts each contact's first solver pass with half the
impulse its persistent manifold cached on the previous tick
(crates/mechanic-gpu/src/shaders/collision.wgsl:699-733, 783). This module
drives the alternative the library offers: a 16 -> 64 -> 64 -> 3 GELU network
that predicts the accumulated normal impulse and the friction impulse of every
active contact from 16 contact features, evaluated inside the tick on MFMA
(BF16 weights and activations, FP32 accumulation) and trained on the GPU in
FP32 by the library's own forward, backward and Adam kernels.

Workflow: run a WARM_START_LOG handle (typically with more solver passes than
the deployment handle, so the targets are close to the solver's fixed point),
gather `read_training_rows()` over many ticks, `WarmStartNet.fit` on them, then
create a WARM_START_LEARNED handle and `set_mlp` the network.
"""
from __future__ import annotations

import ctypes
from typing import Iterable, Optional, Tuple

import numpy as np

from . import _lib
from .physics import DeviceError, GpuPhysics, GpuPhysicsConfig, _check
from .scenes import Scene

FEATURE_NAMES = (
    "normal_speed", "tangent_speed", "penetration", "bias", "inv_mass_a", "inv_mass_b",
    "inv_inertia_a_n", "inv_inertia_a_t1", "inv_inertia_a_t2", "inv_inertia_b_n", "inv_inertia_b_t1",
    "inv_inertia_b_t2", "denominator", "cached_impulse", "cache_hit", "ground",
)
OUTPUT_NAMES = ("normal_impulse", "friction_t1", "friction_t2")


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class WarmStartNet:
    """A network on the GPU: parameters, normalisation and Adam state."""

    def __init__(self, seed: int = 0):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _check(lib.mh_mlp_create(int(seed), ctypes.byref(h)), "mh_mlp_create")
        self._h = h
        self.param_count = int(lib.mh_mlp_param_count())
        self.net_floats = int(lib.mh_mlp_net_floats())

    def close(self) -> None:
        if self._h and self._h.value:
            _lib.load().mh_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def get_net(self) -> np.ndarray:
        out = np.empty(self.net_floats, np.float32)
        _check(_lib.load().mh_mlp_get_net(self._h, _fp(out), out.size), "mh_mlp_get_net")
        return out

    def set_net(self, net: np.ndarray) -> None:
        a = np.ascontiguousarray(net, np.float32)
        if a.size != self.net_floats:
            raise ValueError(f"expected {self.net_floats} floats, got {a.size}")
        _check(_lib.load().mh_mlp_set_net(self._h, _fp(a), a.size), "mh_mlp_set_net")

    def set_dataset(self, features: np.ndarray, targets: np.ndarray) -> None:
        f = np.ascontiguousarray(features, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        if f.ndim != 2 or f.shape[1] != _lib.MLP_FEATURES or t.shape != (f.shape[0], _lib.MLP_OUTPUTS):
            raise ValueError("features must be (R, 16) and targets (R, 3)")
        _check(_lib.load().mh_mlp_set_dataset(self._h, _fp(f), _fp(t), f.shape[0]), "mh_mlp_set_dataset")

    def train(self, epochs: int, batch: int = 4096, learning_rate: float = 1e-3) -> np.ndarray:
        losses = np.zeros(max(int(epochs), 1), np.float32)
        _check(_lib.load().mh_mlp_train(self._h, int(epochs), int(batch), float(learning_rate), _fp(losses)),
               "mh_mlp_train")
        return losses[: int(epochs)]

    def fit(self, features: np.ndarray, targets: np.ndarray, epochs: int = 20, batch: int = 4096,
            learning_rate: float = 1e-3) -> np.ndarray:
        self.set_dataset(features, targets)
        return self.train(epochs, batch, learning_rate)

    def gradient(self) -> Tuple[np.ndarray, float]:
        g = np.empty(self.param_count, np.float32)
        loss = ctypes.c_float()
        _check(_lib.load().mh_mlp_gradient(self._h, _fp(g), ctypes.byref(loss)), "mh_mlp_gradient")
        return g, float(loss.value)

    def gradient_noalloc(self, features_ptr: int, targets_ptr: int, rows: int, gradient_ptr: int,
                         loss_ptr: int) -> None:
        """mh_mlp_gradient_noalloc: the gradient and loss of `rows` device rows
        into device outputs (raw device addresses), allocating nothing."""
        _check(_lib.load().mh_mlp_gradient_noalloc(self._h, features_ptr, targets_ptr, int(rows), gradient_ptr,
                                                   loss_ptr), "mh_mlp_gradient_noalloc")

    def predict(self, features: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(features, np.float32)
        out = np.empty((f.shape[0], _lib.MLP_OUTPUTS), np.float32)
        _check(_lib.load().mh_mlp_predict(self._h, _fp(f), f.shape[0], _fp(out)), "mh_mlp_predict")
        return out


def collect_rows(scene: Scene, config: GpuPhysicsConfig, ticks: int, keep: Iterable[int]) -> Tuple[np.ndarray, np.ndarray]:
    """Steps `scene` for `ticks` ticks on a WARM_START_LOG handle and returns the
    training rows of the ticks in `keep` (1-based tick indices). Each tick's rows
    come in a canonical order (sorted by their bits): the device appends active
    contacts in whatever order its workgroups finish, so without it a sample of
    the rows, and the network trained on it, would change from run to run."""
    cfg = GpuPhysicsConfig(**{**config.__dict__, "warm_start": _lib.WARM_START_LOG})
    keep = set(int(k) for k in keep)
    feats, targs = [], []
    with GpuPhysics.new_with_config(scene, cfg) as gpu:
        for tick in range(1, ticks + 1):
            gpu.dispatch_tick(tick)
            if tick in keep:
                f, t = gpu.read_training_rows()
                bits = np.concatenate([f.view(np.uint32), t.view(np.uint32)], axis=1)
                order = np.lexsort(bits.T[::-1])
                f, t = f[order], t[order]
                feats.append(f)
                targs.append(t)
        flags = gpu.read_last_tick().error_flags
        if flags:
            raise DeviceError(f"logging run reported error flags {flags}")
    if not feats:
        return np.zeros((0, _lib.MLP_FEATURES), np.float32), np.zeros((0, _lib.MLP_OUTPUTS), np.float32)
    return np.concatenate(feats), np.concatenate(targs)


def residual_stats(scene: Scene, config: GpuPhysicsConfig, ticks: int, net: Optional[WarmStartNet] = None,
                   every: int = 1) -> dict:
    """Per-tick contact residual (rms), mean and maximum penetration of the
    active contacts, over a run of `ticks` ticks; with `net` the run uses the
    learned warm start. Arrays keyed rms, mean_penetration, max_penetration."""
    cfg = GpuPhysicsConfig(**{**config.__dict__, "residual": True,
                              "warm_start": _lib.WARM_START_LEARNED if net is not None else config.warm_start})
    out = {"rms": [], "mean_penetration": [], "max_penetration": []}
    with GpuPhysics.new_with_config(scene, cfg) as gpu:
        if net is not None:
            gpu.set_mlp(net)
        for tick in range(1, ticks + 1):
            gpu.dispatch_tick(tick)
            if tick % every == 0:
                r = gpu.read_residual()
                for k in out:
                    out[k].append(r[k])
        flags = gpu.read_last_tick().error_flags
        if flags:
            raise DeviceError(f"run reported error flags {flags}")
    return {k: np.asarray(v) for k, v in out.items()}


def residual_trace(scene: Scene, config: GpuPhysicsConfig, ticks: int, net: Optional[WarmStartNet] = None,
                   every: int = 1) -> np.ndarray:
    """Per-tick contact residual (rms) of a run of `ticks` ticks; with `net`
    the run uses the learned warm start."""
    return residual_stats(scene, config, ticks, net, every)["rms"]
