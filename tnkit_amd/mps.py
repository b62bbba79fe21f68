"""The MPS type: a list of rank-3 site tensors (a, s, b) with an orthogonality centre.

Mirrors tnkit's `MPS(list)` (tnkit/mps.py:29-140): list behaviour, `center` reset to None by
any change of the list's contents.
"""
from __future__ import annotations

from typing import Any, Optional, Sequence

import numpy as np


class MPS(list):
    """One rank-3 tensor per site, index order (left_bond, phys, right_bond)."""

    __slots__ = ("_center",)

    def __init__(self, tensors: Sequence[Any] = (), center: Optional[int] = None):
        super().__init__(tensors)
        self.center = center

    @property
    def center(self) -> Optional[int]:
        return self._center

    @center.setter
    def center(self, c: Optional[int]) -> None:
        if c is not None and not 0 <= int(c) < max(1, len(self)):
            raise ValueError(f"center {c} outside 0..{len(self) - 1}")
        self._center = None if c is None else int(c)

    def __setitem__(self, k, v):
        super().__setitem__(k, v)
        self._center = None

    def append(self, v):
        super().append(v)
        self._center = None

    def to_numpy(self):
        """Host copy of every site tensor."""
        return MPS([t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)
                    for t in self], center=self.center)

    def __reduce__(self):
        return (MPS, (list(self), self.center))
