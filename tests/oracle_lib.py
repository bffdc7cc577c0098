"""Test-side loader of the CPU oracle (oracle/libfts_oracle.so, symbols fto_*)."""
import ctypes
import os

from ftspolysol_amd.api import Engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libfts_oracle.so")
_LIB = None


def oracle_library():
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run __graft_entry__.build()")
        _LIB = ctypes.CDLL(ORACLE_SO)
    return _LIB


def oracle_engine(system):
    return Engine(oracle_library(), "fto_", system)
