"""Small validation helpers of the configuration dataclasses: each check is a
(condition, message) pair and :func:`check` raises ``ValueError`` with the first failing one."""
from __future__ import annotations

import math

import numpy as np


def check(rules) -> None:
    """Raise ``ValueError`` with the message of the first rule whose condition is false."""
    for ok, message in rules:
        if not ok:
            raise ValueError(message)


def finite(x) -> bool:
    return math.isfinite(float(x))


def member(label: str, value, allowed) -> tuple[bool, str]:
    """A rule: ``value`` is one of ``allowed``."""
    return value in allowed, f"unknown {label} {value!r}; use one of {', '.join(allowed)}"


def number(x, label: str, positive: bool = False) -> float:
    """``x`` as a finite float (and > 0 with ``positive``)."""
    v = float(x)
    check([(math.isfinite(v) and (v > 0 or not positive),
            f"{label} must be {'positive and ' if positive else ''}finite")])
    return v


def vector(v, label: str, n: int, exact: bool = False) -> tuple[float, ...]:
    """The first ``n`` components of an array-like as finite floats (``exact``: it must have
    exactly ``n``)."""
    flat = np.ravel(np.asarray(v, dtype=np.float64))
    size_ok = flat.size == n if exact else flat.size >= n
    names = {2: "a finite (x, y) pair", 3: "three finite numbers"}.get(n, f"{n} finite numbers")
    check([(size_ok and bool(np.isfinite(flat[:n]).all()), f"{label} must be {names}")])
    return tuple(float(c) for c in flat[:n])


def unit(v, label: str, n: int, exact: bool = False, tiny: float = 1e-12) -> tuple[float, ...]:
    """A finite non-zero ``n``-vector scaled to unit length."""
    comps = vector(v, label, n, exact)
    length = math.sqrt(sum(c * c for c in comps)) if n != 2 else math.hypot(*comps)
    check([(length >= tiny, f"{label} must be a non-zero {'direction' if n == 2 else 'vector'}")])
    return tuple(c / length for c in comps)


def coerce(cls, item):
    """A dataclass instance from an instance, a dict of its fields or a tuple of them."""
    if isinstance(item, cls):
        return item
    return cls(**item) if isinstance(item, dict) else cls(*item)
