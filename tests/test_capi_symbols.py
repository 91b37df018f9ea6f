"""The C-ABI library loads and exports every entry point include/su2dmrg.h declares (CPU)."""
import ctypes

import pytest

from su2dmrg_amd import lib as L


def test_every_declared_symbol_is_exported():
    declared = L.declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L.lib, name), name
        assert isinstance(getattr(L.lib, name), ctypes._CFuncPtr)


def test_version_string():
    assert L.lib.su2_version().startswith(b"su2dmrg-mi355x")


def test_null_arguments_are_rejected_without_a_device():
    # argument validation happens before any device call
    assert L.lib.su2_tensor_size(None, None, None) == 1
    assert L.lib.su2_heff_flops(None, None, None) == 1
    assert b"NULL" in L.lib.su2_last_error(None)


def test_lanes_arguments_are_checked_without_a_device():
    assert L.lib.su2_batch_run_lanes(None, None, 0) == 1
    assert b"0 batches" in L.lib.su2_last_error(None)
    assert L.lib.su2_batch_run_lanes(None, None, 2) == 1
    assert b"NULL" in L.lib.su2_last_error(None)
    assert L.lib.su2_server_set_lanes(None, None, 2) == 1
