"""The C-ABI library loads without a GPU and exports every function
include/adfvm.h declares (no compute calls here)."""
import os
import re

from adfvm_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'include', 'adfvm.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(adfvm_[a-z_0-9]+)\s*\(', text)))


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_library_is_the_in_tree_build(lib):
    assert os.path.dirname(_lib.LIB_PATH).endswith('adfvm_amd')
    assert os.path.exists(_lib.LIB_PATH)
