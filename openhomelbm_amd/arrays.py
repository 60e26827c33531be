"""Array handles with the surface the reference's callers use on Warp arrays.

Callers of HomeFlow / HomeFlow3D read fields with ``.numpy()``, write them by
assigning a whole new array to the attribute, and clear them with
``.zero_()`` / ``.fill_()`` (envs/lbm/lbm_func.py:503-506,
envs/lbm3d/lbm_solver_3d.py:198-218). Two handle kinds provide that:

* LatticeField: a view of one moment field (or the velocity vector) of a
  Lattice in the reference's (nx, ny[, nz]) index order. Reads decode 16-bit
  storage; writes go to the state the field names.
* HostArray: small per-solid tables (positions, angles, forces) kept on the
  host, where the facades build the device pose tables from them each step.
"""
from __future__ import annotations

from typing import Callable, Optional, Sequence

import numpy as np
import torch


def as_numpy(value, dtype=None) -> np.ndarray:
    """Array-like (NumPy, torch, objects with .numpy(), nested sequences) to NumPy."""
    if isinstance(value, torch.Tensor):
        a = value.detach().cpu().numpy()
    elif hasattr(value, "numpy") and callable(value.numpy):
        a = np.asarray(value.numpy())
    else:
        a = np.asarray(value)
    return a.astype(dtype) if dtype is not None else a


class HostArray:
    """Host array with the Warp-array methods the callers use."""

    def __init__(self, data, dtype=np.float32, shape: Optional[Sequence[int]] = None,
                 on_change: Optional[Callable[[], None]] = None):
        a = as_numpy(data, dtype)
        if shape is not None:
            a = a.reshape(shape)
        self._a = np.array(a, dtype=dtype, copy=True)
        self._on_change = on_change

    def _changed(self):
        if self._on_change is not None:
            self._on_change()

    def numpy(self) -> np.ndarray:
        return self._a.copy()

    def assign(self, value) -> None:
        self._a[...] = as_numpy(value, self._a.dtype).reshape(self._a.shape)
        self._changed()

    def zero_(self) -> None:
        self._a[...] = 0
        self._changed()

    def fill_(self, value) -> None:
        self._a[...] = value
        self._changed()

    @property
    def shape(self):
        return self._a.shape

    @property
    def dtype(self):
        return self._a.dtype

    def __len__(self):
        return len(self._a)

    def __getitem__(self, k):
        return self._a[k]

    def __setitem__(self, k, v):
        self._a[k] = v
        self._changed()

    def __array__(self, dtype=None, copy=None):
        return self._a if dtype is None else self._a.astype(dtype)

    def __repr__(self):
        return f"HostArray({self._a!r})"


class LatticeField:
    """One moment field (or the velocity vector) of one state of a Lattice.

    which: "cur" (the present state) or "post" (the other ping-pong buffer:
    the reference's *_post arrays, which after a step hold the previous
    state). fields: storage field indices; several make a vector field whose
    last axis is the component (the reference's vec2 / vec3 arrays).
    """

    def __init__(self, lattice, fields: Sequence[int], which: str = "cur"):
        self.lattice = lattice
        self.fields = tuple(fields)
        self.which = which

    def _buf(self) -> int:
        return self.lattice.cur if self.which == "cur" else self.lattice.cur ^ 1

    def _xyz(self, t: torch.Tensor) -> torch.Tensor:
        # storage (z,) y, x  ->  reference x, y(, z)
        return t.permute(*reversed(range(t.dim())))

    def torch(self) -> torch.Tensor:
        """Device tensor (decoded FP32/FP64) in the reference's index order."""
        L = self.lattice
        raw = L.fields(self._buf())
        comps = []
        for f in self.fields:
            t = raw[f]
            comps.append(self._xyz(L.decode(f, t)))
        return comps[0] if len(comps) == 1 else torch.stack(comps, dim=-1)

    def numpy(self) -> np.ndarray:
        return self.torch().cpu().numpy()

    def assign(self, value) -> None:
        L = self.lattice
        v = torch.as_tensor(as_numpy(value, np.float64))
        shape = tuple(reversed(L.shape))
        if len(self.fields) > 1:
            v = v.reshape(shape + (len(self.fields),))
        else:
            v = v.reshape(shape)
        raw = L.fields(self._buf())
        for k, f in enumerate(self.fields):
            c = v[..., k] if len(self.fields) > 1 else v
            c = c.permute(*reversed(range(c.dim())))
            raw[f].copy_(L.encode(f, c.to(device=raw.device)))

    def fill_(self, value) -> None:
        shape = tuple(reversed(self.lattice.shape)) + ((len(self.fields),) if len(self.fields) > 1 else ())
        self.assign(np.full(shape, value, dtype=np.float64))

    def zero_(self) -> None:
        self.fill_(0.0)

    @property
    def shape(self):
        s = tuple(reversed(self.lattice.shape))
        return s + ((len(self.fields),) if len(self.fields) > 1 else ())

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a if dtype is None else a.astype(dtype)


class ForceField:
    """Per-cell body force component in the reference's index order
    (forcex / forcey / forcez); setting any component installs the field."""

    def __init__(self, lattice, comp: int):
        self.lattice = lattice
        self.comp = comp

    def _dims(self):
        return 2 if self.lattice.lattice == 9 else 3

    def torch(self) -> torch.Tensor:
        L = self.lattice
        if L.force is None:
            return torch.zeros(tuple(reversed(L.shape)), dtype=torch.float32, device=L.device)
        t = L.force[self.comp, : L.ncell].view(L.shape)
        return t.permute(*reversed(range(t.dim())))

    def numpy(self) -> np.ndarray:
        return self.torch().cpu().numpy()

    def assign(self, value) -> None:
        L = self.lattice
        d = self._dims()
        cur = np.zeros((d,) + L.shape, np.float32) if L.force is None else \
            L.force[:, : L.ncell].view((d,) + L.shape).cpu().numpy()
        v = as_numpy(value, np.float32).reshape(tuple(reversed(L.shape)))
        cur[self.comp] = v.transpose(*reversed(range(v.ndim)))
        L.set_force(cur if np.any(cur) else None)

    def zero_(self) -> None:
        self.assign(np.zeros(tuple(reversed(self.lattice.shape)), np.float32))

    def fill_(self, value) -> None:
        self.assign(np.full(tuple(reversed(self.lattice.shape)), value, np.float32))

    @property
    def shape(self):
        return tuple(reversed(self.lattice.shape))
