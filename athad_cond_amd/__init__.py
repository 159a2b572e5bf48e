"""MI355X (gfx950) implementation of the ATHAD_COND time step behind a C-ABI.

The HIP library ``libathad_hip.so`` (built from ``csrc/``) holds the device state and all
operators; this package is its host-side mirror.
"""
from ._lib import FIELDS, FIELDS2D, PROGNOSTIC, AthadError, default_params, lib  # noqa: F401
from .integration import Integration, footprint_bytes, read_restart, restart_bytes, write_restart  # noqa: F401

__all__ = ["Integration", "footprint_bytes", "FIELDS", "FIELDS2D", "PROGNOSTIC", "AthadError",
           "default_params", "lib", "read_restart", "restart_bytes", "write_restart"]
