"""The C-ABI library loads and exports every function include/vdhmm.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

from vardepth_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "vdhmm.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("defined",)))


def test_header_matches_binding_list():
    assert set(declared()) == set(_lib.SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_core_alias_library_exports_reference_names():
    core = ctypes.CDLL(str(_lib.CORE_LIB_PATH))
    for name in ("read_haplotype", "read_haplotype_batch"):
        assert hasattr(core, name)


def test_binding_declares_argtypes():
    lib = _lib.lib()  # loads and declares; no device call
    assert lib.trie_many.restype is ctypes.c_int
    assert len(lib.pairhmm_pairs.argtypes) == 9
    assert len(lib.trie_many.argtypes) == 11
