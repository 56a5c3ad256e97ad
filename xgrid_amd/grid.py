"""``Grid``: a device-resident ring of time levels with a host mirror.

Public behaviour follows ``xgrid/xgrid/__init__.py:21-86``: ``Grid(shape,
dtype)`` starts with one zero level and an int32 zero boundary mask; every
kernel call extends the level list to the kernel's depth with zero levels
(truncating extras) and rotates it so the oldest level becomes ``now``
(``_op_invoke``, ``:49-54``); ``now`` is level 0; ``g[i]`` reads and writes
``now``; ``fill(data, time)`` keeps the reference's indexing (so
``fill(data)`` with the default ``time=0`` truncates the levels and raises
IndexError, as the reference does).

Storage: each level lives in HBM (``xg_grid``) and, lazily, in a host numpy
array.  Per level two flags say which copy is current.  Reading ``now`` (or
``g[...]``) downloads level 0 if the device copy is newer and then treats the
host copy as possibly modified, so the next kernel call uploads it.  Touching
``boundary`` marks the mask for re-upload and re-planning.  Host arrays that
callers keep across kernel calls are not tracked; read them again through the
Grid after a call.
"""
from __future__ import annotations

import weakref

import numpy as np

from . import _lib
from .types import GridType, Value, numpy_dtype, parse_annotation


class GridError(RuntimeError):
    pass


class _Level:
    __slots__ = ("host", "host_valid", "dev_valid")

    def __init__(self, host, host_valid: bool, dev_valid: bool):
        self.host = host
        self.host_valid = host_valid
        self.dev_valid = dev_valid


def _destroy(handle):
    try:
        lib = _lib.load()
        lib.xg_grid_destroy(handle)
    except Exception:
        pass


class Grid:
    def __init__(self, shape: tuple, dtype) -> None:
        parsed = parse_annotation(dtype)
        if not isinstance(parsed, Value):
            raise GridError(f"Grid element should be value instead of '{parsed}'")
        self.element = parsed
        self.numpy_dtype = numpy_dtype(parsed)
        self.shape = tuple(int(s) for s in shape)
        if not 1 <= len(self.shape) <= 3:
            raise GridError("the MI355X backend supports grids of 1 to 3 dimensions")
        self.typing = GridType(self.element, self.dimension)
        self._levels = [_Level(np.zeros(self.shape, dtype=self.numpy_dtype), True, False)]
        self._boundary = np.zeros(self.shape, dtype=np.int32)
        self._mask_dirty = False        # the device default is the all-zero mask
        self._handle = None
        self._finalizer = None
        self._typed_for = None          # the last kernel argument type this grid was checked against

    # ------------------------------------------------------------ properties
    @property
    def dimension(self) -> int:
        return len(self.shape)

    @property
    def boundary(self) -> np.ndarray:
        self._mask_dirty = True
        return self._boundary

    @boundary.setter
    def boundary(self, value) -> None:
        self._boundary = value
        self._mask_dirty = True

    @property
    def now(self) -> np.ndarray:
        lvl = self._levels[0]
        self._pull(0)
        lvl.dev_valid = False            # the caller may write into it
        return lvl.host

    @property
    def time_levels(self) -> int:
        return len(self._levels)

    def level(self, t: int) -> np.ndarray:
        """Host copy of time level ``t`` (0 = now, 1 = previous, ...)."""
        self._pull(t)
        self._levels[t].dev_valid = False
        return self._levels[t].host

    # ------------------------------------------------------------ reference API
    def _extend_time(self, depth: int) -> None:
        changed = False
        while len(self._levels) < depth:
            self._levels.append(_Level(np.zeros(self.shape, dtype=self.numpy_dtype), True, True))
            changed = True
        if len(self._levels) > depth:
            self._levels = self._levels[:depth]
            changed = True
        if changed and self._handle is not None:
            _lib.check(_lib.load().xg_grid_set_depth(self._handle, len(self._levels)), "xg_grid_set_depth")

    def _op_invoke(self, depth: int, tick: bool) -> None:
        if len(self._levels) != depth:
            self._extend_time(depth)
        if tick and len(self._levels) > 1:
            self._levels.insert(0, self._levels.pop())
            if self._handle is not None:
                _lib.load().xg_grid_tick(self._handle)

    def fill(self, data: np.ndarray, time: int = 0) -> None:
        if data.shape != self.shape or data.dtype != self.numpy_dtype:
            raise GridError("Unable to fill grid with incompatible shape or data type")
        self._extend_time(abs(time))
        self._levels[time] = _Level(np.array(data, dtype=self.numpy_dtype, copy=True), True, False)

    def __getitem__(self, index):
        return self.now[index]

    def __setitem__(self, index, value):
        self.now[index] = value

    def __repr__(self) -> str:
        return f"Grid(shape={self.shape}, element={self.element!r}, levels={len(self._levels)})"

    # ------------------------------------------------------------ device side
    def _device(self):
        if self._handle is None:
            lib = _lib.ensure_device(_device_index())
            import ctypes
            shape = (ctypes.c_int32 * 3)(*(list(self.shape) + [1] * (3 - len(self.shape))))
            h = ctypes.c_void_p()
            _lib.check(lib.xg_grid_create(self.dimension, shape, self.numpy_dtype.itemsize, ctypes.byref(h)),
                       "xg_grid_create")
            self._handle = h.value
            self._finalizer = weakref.finalize(self, _destroy, self._handle)
            _lib.check(lib.xg_grid_set_depth(self._handle, len(self._levels)), "xg_grid_set_depth")
            for lvl in self._levels:
                lvl.dev_valid = False
            self._mask_dirty = self._mask_dirty or bool(np.any(self._boundary))
        return self._handle

    def _pull(self, t: int) -> None:
        lvl = self._levels[t]
        if not lvl.host_valid:
            if lvl.host is None:
                lvl.host = np.empty(self.shape, dtype=self.numpy_dtype)
            _lib.check(_lib.load().xg_grid_download(self._handle, t, lvl.host.ctypes.data), "xg_grid_download")
            lvl.host_valid = True

    def _now_after(self, event) -> np.ndarray:
        """`now`, read once the device work recorded before `event` (``xg_event_record``)
        is done; kernels queued after the event keep running during the copy and must
        not write this grid.  The serving loop reads a finished request this way while
        the next tick runs."""
        lvl = self._levels[0]
        if not lvl.host_valid:
            if lvl.host is None:
                lvl.host = np.empty(self.shape, dtype=self.numpy_dtype)
            _lib.check(_lib.load().xg_grid_download_after(self._handle, 0, lvl.host.ctypes.data, event),
                       "xg_grid_download_after")
            lvl.host_valid = True
        lvl.dev_valid = False
        return lvl.host

    def _sync_for_launch(self):
        """Make every level and the mask current on the device; returns the handle."""
        h = self._device()
        lib = _lib.load()
        for t, lvl in enumerate(self._levels):
            if not lvl.dev_valid:
                host = lvl.host
                if not (host.flags.c_contiguous and host.dtype == self.numpy_dtype):
                    host = np.ascontiguousarray(host, dtype=self.numpy_dtype)
                _lib.check(lib.xg_grid_upload(h, t, host.ctypes.data), "xg_grid_upload")
                lvl.dev_valid = True
        if self._mask_dirty:
            mask = np.ascontiguousarray(self._boundary, dtype=np.int32)
            if mask.shape != self.shape:
                raise GridError(f"boundary mask shape {mask.shape} does not match grid shape {self.shape}")
            _lib.check(lib.xg_grid_set_mask(h, mask.ctypes.data), "xg_grid_set_mask")
            self._mask_dirty = False
        return h

    def _device_wrote(self) -> None:
        """Level 0 was written on the device."""
        lvl = self._levels[0]
        lvl.host_valid = False
        lvl.dev_valid = True

    def class_info(self, klass: int) -> dict:
        """Device plan of one boundary class: kind, cell count and box."""
        import ctypes
        h = self._sync_for_launch()
        kind, count = ctypes.c_int32(), ctypes.c_int64()
        lo, hi = (ctypes.c_int32 * 3)(), (ctypes.c_int32 * 3)()
        _lib.check(_lib.load().xg_grid_class_info(h, klass, ctypes.byref(kind), ctypes.byref(count), lo, hi))
        names = {0: "box", 1: "list", 2: "masked", 3: "empty"}
        return {"kind": names[kind.value], "count": count.value, "lo": tuple(lo)[:self.dimension],
                "hi": tuple(hi)[:self.dimension]}


def _device_index() -> int:
    from .config import _config
    return 0 if _config is None else _config.device
