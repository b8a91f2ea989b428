"""Batched ensembles: many meshes advanced together through hydro_batch_*.

Each member is a :class:`~hydro_amd.device.DeviceMesh`; the batch packs
their workgroup tiles CSR-style into one launch per step (or one persistent
launch for all steps), so small meshes fill the chip together instead of
launching mostly idle grids one after another.
"""
from __future__ import annotations

import ctypes
from typing import List

from . import _lib
from .device import DeviceMesh


class Batch:
    def __init__(self, members: List[DeviceMesh], tune=None):
        import torch

        self.lib = _lib.load()
        self.members = members
        self.device = members[0].device
        t = tune if tune is not None else _lib.default_tune()
        arr = (ctypes.c_void_p * len(members))(*[m.h.value for m in members])
        h = ctypes.c_void_p()
        _lib.check(self.lib.hydro_batch_create(arr, len(members), ctypes.byref(t), ctypes.byref(h)))
        self.h = h
        self._torch = torch

    def stream(self):
        return ctypes.c_void_p(self._torch.cuda.current_stream(self.device).cuda_stream)

    def step(self, n: int, mode: str = "persistent"):
        _lib.check(self.lib.hydro_batch_step_n(self.h, int(n), _lib.MODES[mode], self.stream()))

    def sync(self):
        _lib.check(self.lib.hydro_batch_sync(self.h, self.stream()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.hydro_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_batch(members: List[DeviceMesh], steps: int = 20, blocks=(256, 128),
                   modes=("fused", "graph", "persistent")):
    """Choose the workgroup size and execution mode of a batch by timing
    `steps` steps of each combination from the members' current states
    (restored after every trial).  Returns ((block, mode), {(block, mode):
    us per step}); the members are left prepared for the last block tried
    and at their original states and times."""
    import torch

    saved = [(m.fetch("H", "U", "V", "Z"), m.get_time()) for m in members]

    def restore():
        for m, (st, (day, kt, _)) in zip(members, saved):
            m.load_state(st)
            m.set_time(day, kt)

    table = {}
    for block in blocks:
        b = Batch(members, _lib.default_tune(block=block))
        for mode in modes:
            restore()
            b.step(0, mode)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            b.step(steps, mode)
            e1.record()
            torch.cuda.synchronize()
            table[(block, mode)] = e0.elapsed_time(e1) * 1e3 / steps
        b.sync()
        b.close()
    restore()
    best = min(table, key=table.get)
    return best, table
