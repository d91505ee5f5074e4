"""HessianKind, mirroring litepgo.energy.HessianKind (src/core/energy/scalar/energy.hpp:6-13)."""
from __future__ import annotations

import enum


class HessianKind(enum.IntEnum):
    #: The exact Hessian.
    Exact = 0
    #: Per-element Hessians clamped to PSD before assembly (mesh_energy.cpp:16-27,72-73).
    Projected = 1


def kind_code(kind) -> int:
    """C-ABI code of a HessianKind (or None)."""
    if kind is None:
        return -1
    if isinstance(kind, str):
        kind = {"exact": HessianKind.Exact, "projected": HessianKind.Projected}[kind.lower()]
    kind = HessianKind(int(kind))
    return int(kind)
