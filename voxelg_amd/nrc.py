"""The learned radiance cache of the probe update (vg_nrc_* in include/vg.h).

`RadianceCache` owns a hash-grid + MLP cache on a context's device: bf16 MFMA inference
and training kernels, Adam with fp32 master parameters. It stands in for the bounce tail of
GI gather rays (`ProbeGrid.update_learned`); the reference ends those rays with an ambient
term instead (shaders/gi_probes.wgsl:226-241) and has no cache.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (NRC_CONFIG_DTYPE, NRC_QUERY_DTYPE, VG_NRC_MLP_PARAMS, check, lib, ptr)


def default_config(**overrides) -> np.ndarray:
    """vg_nrc_config with the library defaults, then overrides (levels, log2_table, lr_grid,
    lr_mlp, beta1, beta2, eps, init_grid, seed, batch, steps; `coarse_cell` / `fine_cell`
    recompute the per-level scales)."""
    c = np.zeros(1, NRC_CONFIG_DTYPE)
    check(lib().vg_nrc_config_default(ptr(c)), "vg_nrc_config_default")
    coarse = overrides.pop("coarse_cell", None)
    fine = overrides.pop("fine_cell", None)
    for k, v in overrides.items():
        c[k] = v
    if coarse is not None or fine is not None or "levels" in overrides:
        c["scale"][0, :] = 0.0
        c["scale"][0, :int(c["levels"][0])] = scales(int(c["levels"][0]), coarse or 128.0,
                                                      fine or 1.0)
    return c


def scales(levels: int, coarse_cell: float, fine_cell: float) -> np.ndarray:
    out = np.zeros(int(levels), np.float32)
    check(lib().vg_nrc_scales(int(levels), float(coarse_cell), float(fine_cell), ptr(out)),
          "vg_nrc_scales")
    return out


def queries(pos, face, direction, aux=None) -> np.ndarray:
    """vg_nrc_query records from arrays: pos [n,3], face [n], direction [n,3]."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    q = np.zeros(len(pos), NRC_QUERY_DTYPE)
    q["pos"] = pos
    q["face"] = np.asarray(face, np.uint32).reshape(-1)
    q["dir"] = np.asarray(direction, np.float32).reshape(-1, 3)
    if aux is not None:
        q["aux"] = np.asarray(aux, np.uint32).reshape(-1)
    return q


class RadianceCache:
    def __init__(self, ctx, config: np.ndarray | None = None):
        self.ctx = ctx
        self.config = (default_config() if config is None else
                       np.ascontiguousarray(config, NRC_CONFIG_DTYPE).reshape(1)).copy()
        h = C.c_void_p()
        check(lib().vg_nrc_create(ctx._h, ptr(self.config), C.byref(h)), "vg_nrc_create")
        self._h = h
        n = (C.c_int64 * 2)()
        check(lib().vg_nrc_param_counts(h, n), "vg_nrc_param_counts")
        self.n_mlp, self.n_grid = int(n[0]), int(n[1])

    def close(self):
        if getattr(self, "_h", None):
            lib().vg_nrc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def batch(self) -> int:
        return int(self.config["batch"][0])

    def params(self):
        """(mlp float32 [13312], grid float32 [levels * 2^log2_table * 2]) fp32 masters."""
        mlp = np.zeros(self.n_mlp, np.float32)
        grid = np.zeros(self.n_grid, np.float32)
        check(lib().vg_nrc_get_params(self._h, ptr(mlp), ptr(grid)), "vg_nrc_get_params")
        return mlp, grid

    def set_params(self, mlp=None, grid=None) -> "RadianceCache":
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32).reshape(VG_NRC_MLP_PARAMS)
        g = None if grid is None else np.ascontiguousarray(grid, np.float32).reshape(self.n_grid)
        check(lib().vg_nrc_set_params(self._h, None if m is None else ptr(m),
                                      None if g is None else ptr(g)), "vg_nrc_set_params")
        return self

    def infer(self, q: np.ndarray) -> np.ndarray:
        """Raw predictions (float32 [n, 3]) for host query records (synchronous)."""
        q = np.ascontiguousarray(q, NRC_QUERY_DTYPE)
        qb = self.ctx.to_device(q)
        out = self.ctx.alloc(max(1, len(q)) * 16)
        check(lib().vg_nrc_infer(self._h, qb.data_ptr(), len(q), out.data_ptr(), None),
              "vg_nrc_infer")
        self.ctx.synchronize()
        return out.numpy(np.float32, len(q) * 4).reshape(-1, 4)[:, :3].copy()

    def gradients(self, q: np.ndarray, targets: np.ndarray):
        """(loss, mlp gradient [13312], grid gradient) of a host batch, without an update."""
        q = np.ascontiguousarray(q, NRC_QUERY_DTYPE)
        t = np.ascontiguousarray(targets, np.float32).reshape(len(q), 4)
        qb, tb = self.ctx.to_device(q), self.ctx.to_device(t)
        gm = np.zeros(self.n_mlp, np.float32)
        gg = np.zeros(self.n_grid, np.float32)
        loss = C.c_float()
        check(lib().vg_nrc_gradients(self._h, qb.data_ptr(), tb.data_ptr(), len(q), ptr(gm),
                                     ptr(gg), C.byref(loss)), "vg_nrc_gradients")
        return float(loss.value), gm, gg

    def train_step(self, q, targets, stream=None) -> "RadianceCache":
        """One Adam step on a batch: device buffers (q, targets DeviceBuffer + count via
        len) or host arrays (copied)."""
        q = np.ascontiguousarray(q, NRC_QUERY_DTYPE)
        t = np.ascontiguousarray(targets, np.float32).reshape(len(q), 4)
        qb, tb = self.ctx.to_device(q), self.ctx.to_device(t)
        check(lib().vg_nrc_train_step(self._h, qb.data_ptr(), tb.data_ptr(), len(q),
                                      getattr(stream, "handle", stream)), "vg_nrc_train_step")
        self.ctx.synchronize()
        return self

    def join(self, stream=None) -> "RadianceCache":
        """Make `stream` wait for a training step still running on the cache's own stream
        (the learned probe update enqueues its step there)."""
        check(lib().vg_nrc_join(self._h, getattr(stream, "handle", stream)), "vg_nrc_join")
        return self

    def stats(self):
        """(optimizer steps taken, loss of the last step)."""
        s = C.c_int64()
        loss = C.c_float()
        check(lib().vg_nrc_stats(self._h, C.byref(s), C.byref(loss)), "vg_nrc_stats")
        return int(s.value), float(loss.value)

    def last_batch(self):
        """The learned probe update's last training batch: (queries, targets [n, 4])."""
        q = np.zeros(self.batch, NRC_QUERY_DTYPE)
        t = np.zeros((self.batch, 4), np.float32)
        check(lib().vg_nrc_last_batch(self._h, ptr(q), ptr(t)), "vg_nrc_last_batch")
        return q, t
