"""The C ABI: liborqg.so loads without a GPU and exports every symbol that
include/orqg.h declares (no compute calls here)."""
import ctypes

import pytest
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "orqg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(orqg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "orqg_radix_sort" in syms and "orqg_and_b" in syms and len(syms) >= 30


def test_library_exports_every_declared_symbol():
    import orq_amd

    L = ctypes.CDLL(orq_amd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    import orq_amd

    assert sorted(orq_amd.EXPORTS) == declared_symbols()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import orq_amd

    monkeypatch.setattr(orq_amd, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(orq_amd, "_lib", None)
    try:
        orq_amd.lib()
    except orq_amd.OrqgError as e:
        assert "no CPU fallback" in str(e)
    else:
        raise AssertionError("expected OrqgError")


def _library_matches_sources():
    import orq_amd
    from orq_amd import build

    assert orq_amd.build_info() == build.source_hash(), (
        "liborqg.so was built from other sources: rebuild it")


def test_library_matches_sources():
    """The library under test was built from these sources (a hash of
    orq_amd/csrc, include/orqg.h and the flags, compiled in)."""
    _library_matches_sources()


@pytest.mark.gpu
def test_gpu_library_matches_sources():
    _library_matches_sources()


def test_positive_batch_size_must_be_a_multiple_of_256():
    """The kernels keep a batch in whole keystream blocks and packed words; a
    positive batch size that is not a multiple of 256 is refused before any
    device work (no GPU needed)."""
    import orq_amd

    keys = orq_amd.keys_from_seed(1)
    with pytest.raises(orq_amd.OrqgError, match=r"orqg_ctx_create failed \(-1\)"):
        orq_amd.Ctx(keys, batch_size=100)
