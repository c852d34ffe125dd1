"""The C-ABI library loads and exports every entry point include/sosfuse2_hip.h declares (no compute calls)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / 'include' / 'sosfuse2_hip.h'
LIB = ROOT / 'sosfuse2_amd' / 'libsosfuse2_hip.so'


def declared():
    text = HEADER.read_text()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(sos_[a-z_0-9]+)\s*\(', text)))


def test_header_declares_the_entry_points():
    names = declared()
    for required in ('sos_transport', 'sos_project', 'sos_compose', 'sos_make_plane', 'sos_source_metrics',
                     'sos_anderson_residual', 'sos_anderson_gram', 'sos_anderson_solve', 'sos_anderson_mix',
                     'sos_last_error'):
        assert required in names


@pytest.mark.skipif(not LIB.exists(), reason='library not built (run __graft_entry__.build())')
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not LIB.exists(), reason='library not built (run __graft_entry__.build())')
def test_binding_covers_the_header():
    from sosfuse2_amd import hip
    assert set(declared()) <= set(hip.exported_symbols())


def test_product_never_imports_the_oracle():
    # the oracle is the checker only: no module of the package may import it
    for f in (ROOT / 'sosfuse2_amd').rglob('*.py'):
        text = f.read_text()
        assert not re.search(r'^\s*(from|import)\s+oracle', text, flags=re.M), f
