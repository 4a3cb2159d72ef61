"""The C-ABI library loads on a host without a GPU and exports every entry
point ``include/vgr.h`` declares; host-only helpers match the reference tables."""

from __future__ import annotations

import re
from pathlib import Path

import pytest

from vgr_amd import native

ROOT = Path(__file__).resolve().parents[1]


def declared_functions() -> set[str]:
    header = (ROOT / "include" / "vgr.h").read_text()
    return set(re.findall(r"^\s*(?:int32_t|int64_t|const char\*)\s+(vgr_\w+)\s*\(", header, re.MULTILINE))


def test_header_and_binding_agree():
    assert declared_functions() == set(native.EXPORTED_SYMBOLS)


def test_library_loads_and_exports_every_declared_symbol():
    library = native.load_library()
    for name in declared_functions():
        assert hasattr(library, name), name
    assert library.vgr_abi_version() == native.ABI_VERSION


@pytest.mark.parametrize(("tiles", "end_bit"), [(1, 32), (2, 33), (3, 34), (4, 34), (1 << 20, 52),
                                                ((1 << 20) + 1, 53), (2_147_483_647, 63)])
def test_emitted_sort_key_end_bit_table(tiles, end_bit):
    """Same boundaries as the reference test (tests/test_cuda_backend.py:756-783)."""
    assert native.NativeRenderer().emitted_sort_key_end_bit(tiles) == end_bit


@pytest.mark.parametrize("tiles", [-1, 0, 2_147_483_648])
def test_emitted_sort_key_end_bit_rejects_invalid(tiles):
    with pytest.raises(ValueError, match="tile capacity"):
        native.NativeRenderer().emitted_sort_key_end_bit(tiles)


def test_scratch_query_grows_with_capacity():
    module = native.NativeRenderer()
    small = module.scratch_bytes(max_views=4, height=64, width=64, gaussians=1000, intersection_capacity=10_000,
                                 tile_size=16, near_plane=0.01, far_plane=100.0)
    large = module.scratch_bytes(max_views=4, height=64, width=64, gaussians=1000, intersection_capacity=1_000_000,
                                 tile_size=16, near_plane=0.01, far_plane=100.0)
    assert 0 < small < large


def test_scratch_query_is_monotonic_in_tile_count_with_default_digits():
    """The default digit width switches between 8 and 9 bits with the key width; scratch sized for
    the largest tile count must still cover every smaller render (the batched path relies on it)."""
    module = native.NativeRenderer()
    sizes = [module.scratch_bytes(max_views=views, height=480, width=640, gaussians=100_000,
                                  intersection_capacity=5_000_000, tile_size=16, near_plane=0.01, far_plane=100.0)
             for views in (1, 2, 3, 5, 8, 16, 33, 64, 100, 128, 200, 256)]
    assert all(a <= b for a, b in zip(sizes, sizes[1:])), sizes


def test_invalid_options_are_reported_with_a_message():
    module = native.NativeRenderer()
    with pytest.raises(RuntimeError, match="tile_size"):
        module.scratch_bytes(max_views=1, height=8, width=8, gaussians=1, intersection_capacity=1, tile_size=3,
                             near_plane=0.01, far_plane=100.0)
