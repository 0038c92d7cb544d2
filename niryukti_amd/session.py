"""In-process session over the vantage_* C-ABI (include/vantage/c_api.h).

Mirrors Niryukti's NativeSession (python/vantage/native.py:10-41): ABI version
check, a model given as a path or a native JSON dict, JSON requests, errors
raised as ValueError with the library's message, explicit close.
"""
from __future__ import annotations

import ctypes
import json
import os
import tempfile
from pathlib import Path

from . import load_library


class NativeSession:
    def __init__(self, model, *, library=None):
        self._lib = ctypes.CDLL(str(library)) if library else load_library()
        lib = self._lib
        lib.vantage_abi_version.restype = ctypes.c_uint
        if lib.vantage_abi_version() != 1:
            raise RuntimeError("Unsupported native ABI")
        lib.vantage_open.argtypes = [ctypes.c_char_p]
        lib.vantage_open.restype = ctypes.c_void_p
        lib.vantage_close.argtypes = [ctypes.c_void_p]
        lib.vantage_close.restype = None
        lib.vantage_request.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        lib.vantage_request.restype = ctypes.c_char_p
        lib.vantage_last_error.restype = ctypes.c_char_p
        with tempfile.TemporaryDirectory(prefix="niryukti-native-") as folder:
            path = model
            if isinstance(model, dict):
                path = Path(folder) / "model.json"
                path.write_text(json.dumps(model, allow_nan=False))
            self._handle = lib.vantage_open(os.fsencode(str(path)))
        if not self._handle:
            raise ValueError(lib.vantage_last_error().decode())

    def _request(self, action, **fields):
        if not self._handle:
            raise RuntimeError("Native session is closed")
        raw = self._lib.vantage_request(self._handle, json.dumps(dict(action=action, **fields), allow_nan=False).encode())
        if raw is None:
            raise ValueError(self._lib.vantage_last_error().decode())
        return json.loads(raw)

    def solve(self, **options):
        return self._request("solve", **options)

    def update_row_bounds(self, rows):
        return self._request("update", rows=rows)

    def update_variable_bounds(self, variables):
        return self._request("update", variables=variables)

    def update_objective(self, objective):
        return self._request("update", objective=objective)

    def close(self):
        if self._handle:
            self._lib.vantage_close(self._handle)
            self._handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()
