"""The C-ABI libraries load and export every symbol their headers declare (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, prefix):
    txt = open(os.path.join(ROOT, header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(%s\w+)\s*\(" % prefix, txt)))


@pytest.mark.parametrize("header,lib,prefix", [
    ("include/mkfit_amd.h", "mkfit_amd/libmkfit_amd.so", "mkf_"),
    ("include/mkfit_toymc.h", "mkfit_amd/libmkfit_toymc.so", "mkf_toymc_"),
    ("include/mkfit_event_file.h", "mkfit_amd/libmkfit_amd.so", "mkf_event_file_"),
])
def test_exports_every_declared_symbol(header, lib, prefix):
    names = _declared(header, prefix)
    if prefix == "mkf_":
        names = [n for n in names if not n.startswith("mkf_toymc_")]
    assert names, header
    L = ctypes.CDLL(os.path.join(ROOT, lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_lists_all_abi_symbols():
    import mkfit_amd as mk
    declared = [n for n in _declared("include/mkfit_amd.h", "mkf_") if not n.startswith("mkf_toymc_")]
    assert sorted(mk.ABI_SYMBOLS) == declared


def test_oracle_exports():
    import oracle
    names = _declared("oracle/oracle.h", "orc_")
    L = oracle.lib()
    assert all(hasattr(L, n) for n in names)


def test_context_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import mkfit_amd as mk
    with pytest.raises(mk.MkfError, match="no HIP device"):
        mk.Context(bfield=3.8)


def test_null_arguments_rejected():
    import mkfit_amd as mk
    L = mk.lib()
    assert L.mkf_context_create(None, None) == 1
    assert L.mkf_fit_plain(None, None, None) == 1
    assert L.mkf_fit_batched(None, None, None) == 1
    assert L.mkf_set_batched_shape(None, None) == 1
    assert L.mkf_context_destroy(None) == 0
    assert L.mkf_event_of_hits_create(None, 1, None, None) == 1
    assert L.mkf_event_of_hits_load(None, None, None, None) == 1
    assert L.mkf_select_hit_indices(None, 0, 0, None, None, None, None, None, None) == 1
    assert L.mkf_find_tracks_best_hit(None, None, None, None) == 1
    assert L.mkf_find_tracks_clone_engine(None, None, None, None) == 1
    assert L.mkf_fit_adjoint(None, None, None) == 1
    assert L.mkf_event_of_hits_destroy(None) == 0
    assert L.mkf_add_best_hit(None, None, 0, None, None, None, None, ctypes.c_float(15.0), 0, 0, None, None) == 1
    assert L.mkf_update_with_last_hit(None, None, 0, None, None, 1, 0, None) == 1
    assert L.mkf_num_batched_variants() >= 1


def test_struct_layouts():
    import mkfit_amd as mk
    assert ctypes.sizeof(mk.MkfFinderLanes) == 8 + 9 * 8                 # n + pad, 9 pointers
    assert mk.IDX_CHI2_DTYPE.itemsize == 48
    assert ctypes.sizeof(mk.MkfLayerDesc) == 23 * 4
    # n, pad, 6 pointers, 6 scalars (incl. n_layers), layer_base pointer
    assert ctypes.sizeof(mk.MkfFindBatch) == 4 + 4 + 6 * 8 + 6 * 4 + 8
    # n + pad, 4 ptrs, 8 scalars, 8 ptrs, n_layers + pad, layer_base pointer
    assert ctypes.sizeof(mk.MkfCEBatch) == 8 + 4 * 8 + 8 * 4 + 8 * 8 + 8 + 8
    # 2 ints, 7 ptrs, 3 scalars + pad, 2 ptrs, the fp64 switch + pad
    assert ctypes.sizeof(mk.MkfAdjointBatch) == 8 + 7 * 8 + 12 + 4 + 2 * 8 + 8
