"""Compilation of generated code, cached on disk by content.

* device source -> gfx950 code object with hipRTC (``xg_rtc_compile``);
* host stub     -> shared library with the system C compiler.

Both take the role of the reference's ``Compiler`` (``xgrid/util/ffi.py:59-94``):
a cache directory under ``cacheroot``, a hit when the same text was compiled
before with the same options.  Keys are SHA-256 of everything that determines
the output (source, options, headers, compiler).
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import shutil
import subprocess
import tempfile
import threading
from ctypes import c_char_p

from . import _lib

CSRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
DEVICE_HEADERS = ("xgrid_abi.h", "xgrid_device.h")
ARCH = "gfx950"

_lock = threading.Lock()


class CompileError(RuntimeError):
    pass


def _cache_dir(cacheroot: str) -> str:
    path = os.path.abspath(cacheroot)
    os.makedirs(path, exist_ok=True)
    return path


def device_options(fp_contract: bool = False) -> list:
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17",
            "-ffp-contract=fast" if fp_contract else "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _headers() -> tuple:
    names, srcs = [], []
    for h in DEVICE_HEADERS:
        with open(os.path.join(CSRC, h)) as f:
            srcs.append(f.read())
        names.append(h)
    return names, srcs


def compile_device(src: str, cacheroot: str, fp_contract: bool = False) -> str:
    """Compile device source for gfx950; returns the code-object path."""
    opts = device_options(fp_contract)
    names, hsrcs = _headers()
    h = hashlib.sha256()
    for part in [src, ARCH, *opts, *hsrcs]:
        h.update(part.encode())
        h.update(b"\0")
    path = os.path.join(_cache_dir(cacheroot), h.hexdigest()[:40] + ".hsaco")
    if os.path.exists(path):
        return path
    lib = _lib.load()
    copts = (c_char_p * len(opts))(*[o.encode() for o in opts])
    cnames = (c_char_p * len(names))(*[n.encode() for n in names])
    csrcs = (c_char_p * len(hsrcs))(*[s.encode() for s in hsrcs])
    with _lock:
        rc = lib.xg_rtc_compile(src.encode(), b"xgrid_kernel.hip", copts, len(opts), cnames, csrcs,
                                len(names), path.encode())
    if rc != 0:
        raise CompileError(_lib.last_error())
    return path


def find_cc(candidates) -> str:
    for cc in candidates:
        p = shutil.which(cc)
        if p:
            return p
    raise CompileError(f"failed to locate cc in {list(candidates)}")


def compile_host(src: str, cacheroot: str, cc_candidates, opt_level: int = 2) -> str:
    """Compile a host stub to a shared library; returns its path."""
    cc = find_cc(cc_candidates)
    flags = ["-shared", "-fPIC", f"-O{max(1, int(opt_level))}", "-ffp-contract=off", "-lm"]
    h = hashlib.sha256()
    for part in [src, cc, *flags]:
        h.update(part.encode())
        h.update(b"\0")
    base = os.path.join(_cache_dir(cacheroot), h.hexdigest()[:40])
    lib = base + ".stub.so"
    if os.path.exists(lib):
        return lib
    # Every call writes files of its own and renames them into place: a background
    # precompile and a build in another thread of the same process may compile the
    # same stub at once, and a shared temporary name would let one thread load the
    # library while the other's linker is still writing it.
    fd, c_tmp = tempfile.mkstemp(suffix=".stub.c", dir=os.path.dirname(base))
    with os.fdopen(fd, "w") as f:
        f.write(src)
    c_path = base + ".stub.c"
    os.replace(c_tmp, c_path)
    fd, tmp = tempfile.mkstemp(suffix=".stub.so.tmp", dir=os.path.dirname(base))
    os.close(fd)
    proc = subprocess.run([cc, *flags[:-1], c_path, "-o", tmp, "-lm"], capture_output=True, text=True)
    if proc.returncode != 0:
        os.unlink(tmp)
        raise CompileError(f"failed to compile '{c_path}' due to:\n{proc.stderr}")
    os.replace(tmp, lib)
    return lib


def load_host(path: str):
    return ctypes.CDLL(path)
