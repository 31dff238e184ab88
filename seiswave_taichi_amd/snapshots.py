"""Periodic wavefield snapshots with the device-to-host copies off the step
stream.

The reference drivers take a snapshot by calling ``field.to_numpy()`` every
20 steps (test3d_homegenous_SPML.py:71-75), which stops the GPU for the whole
copy.  ``Snapshots`` keeps the steps going: after each snapshot step it copies
the requested fields (or an index of them, e.g. one plane) into a device
staging slot on the step stream -- a device-to-device copy at HBM speed --
then a second stream copies the slot to pinned host memory while the next
steps run.  A slot is reused only after its host copy has finished (an event
the step stream waits on), so a small ring of slots suffices.  Both streams
have the default priority.

Snapshot steps follow the drivers: after step ``it`` whenever
``it % every == 0``.
"""
from __future__ import annotations

import numpy as np
import torch


class Snapshots:
    def __init__(self, sim, fields=("vz",), every: int = 20, index=None, scheme: str | None = None,
                 slots: int = 3):
        """``sim``: a 3-D or 2-D ElasticWAVE; ``index``: a NumPy-style index in
        the reference (x, y, z) order applied to each field, None for the whole
        field; ``scheme``: "VS" or "SV" for the 2-D class."""
        if every < 1:
            raise ValueError("every must be >= 1")
        self.sim, self.fields, self.every, self.index = sim, tuple(fields), int(every), index
        self.scheme = scheme
        self.device = sim.vx.data.device
        views = self._views()
        self.slots = [{n: torch.empty(v.shape, dtype=v.dtype, device=self.device) for n, v in views.items()}
                      for _ in range(max(1, slots))]
        self.copied = [None] * len(self.slots)
        self.copy_stream = torch.cuda.Stream(device=self.device)
        self.frames: list[tuple[int, dict]] = []

    def _views(self):
        """Device tensors to stage: whole fields in their device layout (a
        plain contiguous copy; the host side reverses the axes), or the index
        of the reference-order view."""
        out = {}
        for n in self.fields:
            f = getattr(self.sim, n)
            out[n] = f.data if self.index is None else f.to_torch()[self.index]
        return out

    def _host(self, t: torch.Tensor) -> np.ndarray:
        a = t.numpy()
        # whole fields: the reference (x, y, z) order as a view, no host copy
        return a.transpose(*reversed(range(a.ndim))) if self.index is None else a

    def _advance(self, it0: int, n: int):
        if n <= 0:
            return
        if self.scheme is None:
            self.sim.run(n, it0)
        else:
            self.sim.run(self.scheme, n, it0)

    def _take(self, it: int):
        k = len(self.frames) % len(self.slots)
        step = torch.cuda.current_stream(self.device)
        if self.copied[k] is not None:
            step.wait_event(self.copied[k])  # the slot's previous host copy is done
        slot = self.slots[k]
        for n, v in self._views().items():
            slot[n].copy_(v)
        ready = torch.cuda.Event()
        ready.record(step)
        host = {n: torch.empty(t.shape, dtype=t.dtype, pin_memory=True) for n, t in slot.items()}
        with torch.cuda.stream(self.copy_stream):
            self.copy_stream.wait_event(ready)
            for n, t in slot.items():
                host[n].copy_(t, non_blocking=True)
            done = torch.cuda.Event()
            done.record(self.copy_stream)
        for t in slot.values():
            t.record_stream(self.copy_stream)
        self.copied[k] = done
        self.frames.append((it, host))

    def run(self, nsteps: int | None = None, it0: int = 0):
        """Step ``nsteps`` from ``it0``, taking the snapshots on the way;
        returns [(it, {field: ndarray})] once every copy has landed."""
        nsteps = self.sim.nt - it0 if nsteps is None else nsteps
        it, end = it0, it0 + nsteps
        while it < end:
            nxt = ((it + self.every - 1) // self.every) * self.every  # next snapshot step >= it
            if nxt >= end:
                self._advance(it, end - it)
                break
            self._advance(it, nxt - it + 1)
            self._take(nxt)
            it = nxt + 1
        self.copy_stream.synchronize()
        return [(k, {n: self._host(t) for n, t in h.items()}) for k, h in self.frames]
