"""Pragma stubs recognised by the frontend (``xgrid/lang/__init__.py:4-17``).

``with boundary(k):`` tags every grid access in its body with class ``k``;
``with c():`` takes string literals as inline host C.  Called outside a kernel
they are no-op context managers, like the reference's.
"""
from __future__ import annotations


class StubContext:
    def __enter__(self):
        return None

    def __exit__(self, *exc):
        return False


def c() -> StubContext:
    return StubContext()


def boundary(type: int) -> StubContext:  # noqa: A002 - reference argument name
    return StubContext()
