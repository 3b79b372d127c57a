"""The C-ABI library loads and exports every symbol include/ttals.h declares (no GPU needed)."""
import ctypes
import os
import re

import pytest

from scikit_tt_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, 'include', 'ttals.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(ttals_\w+)\s*\(', text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ('ttals_sle_als_batch', 'ttals_evp_als_batch', 'ttals_stack_left_op', 'ttals_local_solve',
                 'ttals_local_eig', 'ttals_qr', 'ttals_core_product', 'ttals_tt_norm', 'ttals_tt_ortho'):
        assert must in names


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason='libttals.so not built')
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares a signature for every exported entry point
    assert set(declared_functions()) == set(_lib.SIGNATURES), set(declared_functions()) ^ set(_lib.SIGNATURES)
    assert _lib.load().ttals_version() >= 100


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason='libttals.so not built')
def test_argument_errors_without_gpu():
    """Argument validation happens before any device work."""
    lib = _lib.load()
    # null context -> TTALS_E_ARG
    assert lib.ttals_sle_als_batch(None, None, 1, None, None, None, None, 1, 0, 1e-12, 10, None) == _lib.E_ARG
    assert lib.ttals_core_product(None, 0, 1, 1, 1, 1, 1, 1, None, None, None) == _lib.E_ARG
    assert lib.ttals_core_embed(None, 2, 1, 1, None, 1.0, 1, 1, 0, 0, None) == _lib.E_ARG
    assert lib.ttals_set_trace(None, None, 0) == _lib.E_ARG
    assert lib.ttals_last_error(None) == b'null context'
