"""Shared types and constants of the Yee solver.

Mirrors include/FP.h:3 (FP = double), include/Constants.h:4-12,
include/Enums.h:4-7 and include/Structures.h:7-39.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass

# include/Constants.h:6-11 (Gaussian units; PI is the reference's truncated value)
C = 3e10
R = 1e-12
EPS0 = 1.0
MU0 = EPS0
N = 4.0
PI = 3.14159265358


class Component(enum.IntEnum):
    """include/Enums.h:5 -- ordinals are the C-ABI component ids."""
    EX = 0
    EY = 1
    EZ = 2
    BX = 3
    BY = 4
    BZ = 5
    JX = 6
    JY = 7
    JZ = 8


class Axis(enum.IntEnum):
    """include/Enums.h:6."""
    X = 0
    Y = 1
    Z = 2


class Precision(enum.IntEnum):
    FP64 = 0
    FP32 = 1
    FP16 = 2


@dataclass
class Parameters:
    """include/Structures.h:23-38: grid size, domain bounds, spacings."""
    Ni: int
    Nj: int
    Nk: int
    ax: float = 0.0
    bx: float = 1.0
    ay: float = 0.0
    by: float = 1.0
    az: float = 0.0
    bz: float = 1.0
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 1.0


@dataclass
class SelectedFields:
    """include/Structures.h:8-11."""
    selected_E: Component
    selected_B: Component


@dataclass
class CurrentParameters:
    """include/Structures.h:13-21 (period_x/y/z default to m*C)."""
    period: int
    m: int
    dt: float
    iterations: int = 0

    @property
    def period_x(self) -> float:
        return float(self.m) * C

    period_y = period_x
    period_z = period_x
