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

    def __repr__(self) -> str:
        return repr(asdict(self))

    @property
    def fsize(self) -> int:
        """Byte width of the DSL's ``float`` (``xgrid/util/init.py:35-36``)."""
        return 4 if self.precision == "float" else 8


_config: Configuration | None = None


class ConfigurationError(RuntimeError):
    pass


def get_config() -> Configuration:
    if _config is None:
        raise ConfigurationError("Please call init first to initialize")
    return _config


def init(*, parallel: bool = True, cc: list | None = None, cacheroot: str = ".xgrid",
         comment: bool = False, overstep: Overstep = "none",
         opt_level: int = 2, precision: Precision = "float",
         device: int | None = None, fp_contract: bool = False,
         graphs: bool | None = None, fuse: bool | None = None, blocked: bool | None = None) -> None:
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
    _config = Configuration(parallel=parallel, cc=list(cc or ["gcc", "clang"]),
                            cacheroot=cacheroot, comment=comment, overstep=overstep,
                            opt_level=int(opt_level), precision=precision, device=device,
                            fp_contract=fp_contract, graphs=graphs, fuse=fuse, blocked=blocked)
    logger.info("initialized with configuration: %s", _config)
