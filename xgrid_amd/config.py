"""Process-wide configuration, the counterpart of ``xgrid.init``.

Mirrors ``xgrid/util/init.py:10-66`` (reference): the same keyword names,
defaults and meaning (``precision`` decides what ``float`` means, ``overstep``
decides how out-of-range stencil offsets are resolved).  Options that only the
MI355X backend has are keyword-only extras with defaults that keep unchanged
scripts working.
"""
from __future__ import annotations

import logging
import os
import sys
import threading
from dataclasses import asdict, dataclass, field
from typing import Literal

logger = logging.getLogger("xgrid")

Overstep = Literal["none", "limit", "wrap"]
Precision = Literal["float", "double"]


@dataclass
class Configuration:
    """Reference fields (``xgrid/util/init.py:10-18``) plus backend options."""

    parallel: bool = True
    cc: list = field(default_factory=lambda: ["gcc", "clang"])
    cacheroot: str = ".xgrid"
    comment: bool = False
    overstep: Overstep = "none"
    opt_level: int = 2
    precision: Precision = "float"
    # --- MI355X backend options (no reference counterpart) ---
    device: int = 0
    fp_contract: bool = False      # FMA contraction in device code
    graphs: bool = True            # replay repeated calls as hipGraphs
    fuse: bool = True              # chain small boundary statements in one workgroup; run statement pairs
    blocked: bool = True           # run blockable stencil loops temporally blocked
    precompile: bool = True        # start compiling a kernel in the background when it is decorated

    def __repr__(self) -> str:
        return repr(asdict(self))

    @property
    def fsize(self) -> int:
        """Byte width of the DSL's ``float`` (``xgrid/util/init.py:35-36``)."""
        return 4 if self.precision == "float" else 8


_config: Configuration | None = None
# a thread may work under a configuration of its own (a background compile runs under
# the configuration in force when it was started, whatever init() does meanwhile)
_local = threading.local()


class ConfigurationError(RuntimeError):
    pass


def get_config() -> Configuration:
    own = getattr(_local, "config", None)
    if own is not None:
        return own
    if _config is None:
        raise ConfigurationError("Please call init first to initialize")
    return _config


def config_ready() -> bool:
    return getattr(_local, "config", None) is not None or _config is not None


class thread_config:
    """Run the current thread under `cfg` (other threads keep the global one)."""

    def __init__(self, cfg: Configuration):
        self.cfg = cfg

    def __enter__(self):
        self.saved = getattr(_local, "config", None)
        _local.config = self.cfg
        return self.cfg

    def __exit__(self, *exc):
        _local.config = self.saved
        return False


def init(*, parallel: bool = True, cc: list | None = None, cacheroot: str = ".xgrid",
         comment: bool = False, overstep: Overstep = "none",
         opt_level: int = 2, precision: Precision = "float",
         device: int | None = None, fp_contract: bool = False,
         graphs: bool | None = None, fuse: bool | None = None, blocked: bool | None = None,
         precompile: bool | None = None) -> None:
    """Initialise the global configuration (``xgrid/util/init.py:48-66``).

    ``parallel=False`` keeps the reference's rule that implicit (Jacobi)
    statements are rejected (``xgrid/lang/generator.py:308-310``); device
    execution is parallel either way.
    """
    global _config
    if sys.version_info < (3, 10):
        logger.error("Minimum Python 3.10 is required, current version is %s", sys.version_info)
    if overstep not in ("none", "limit", "wrap"):
        raise ConfigurationError(f"invalid overstep '{overstep}'")
    if precision not in ("float", "double"):
        raise ConfigurationError(f"invalid precision '{precision}'")
    if device is None:
        device = int(os.environ.get("XGRID_DEVICE", "0"))
    if graphs is None:
        graphs = os.environ.get("XGRID_GRAPHS", "1") != "0"
    if fuse is None:
        fuse = os.environ.get("XGRID_FUSE", "1") != "0"
    if blocked is None:
        blocked = os.environ.get("XGRID_BLOCKED", "1") != "0"
    if precompile is None:
        precompile = os.environ.get("XGRID_PRECOMPILE", "1") != "0"
    _config = Configuration(parallel=parallel, cc=list(cc or ["gcc", "clang"]),
                            cacheroot=cacheroot, comment=comment, overstep=overstep,
                            opt_level=int(opt_level), precision=precision, device=device,
                            fp_contract=fp_contract, graphs=graphs, fuse=fuse, blocked=blocked,
                            precompile=precompile)
    logger.info("initialized with configuration: %s", _config)
