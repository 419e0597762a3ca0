"""The warm-start cache's bookkeeping on the CPU (rq_warm_cache_*, the same
code the device-resident cache of rq_model runs): the reference's own cache
tests restated (rans-uq/tests/test_evaluation_safety.cpp:173-199,
rans-uq/tests/test_infra_fixes.cpp:78-95) plus the lookup rules of
WarmStartCache::nearest / store (rans-uq/include/ForwardModel.hpp:72-153)."""
from __future__ import annotations

import math

import pytest

from cfdr_amd.model import HostWarmCache


def test_entry_bytes_counts_the_reference_payload():
    # ForwardModel.hpp:41-50: theta doubles, U (Vec3) and eight scalar
    # fields, interior plus boundary values
    nc, nb = 12, 8
    assert HostWarmCache.entry_bytes(2, nc, nb) == 2 * 8 + (nc + nb) * 24 + 8 * (nc + nb) * 8


def test_replacement_dimensions_and_budget_eviction():
    """test_evaluation_safety.cpp:173-199 with tags in place of fields."""
    cache = HostWarmCache(2, 0.5)
    one = HostWarmCache.entry_bytes(1, 12, 8)
    assert cache.store([0.0], 1, one)
    assert cache.store([0.0], 2, one)
    assert cache.size() == 1, "same parameters replace an existing cached state"
    assert cache.nearest([0.0, 0.0]) is None, "different parameter dimensions cannot match"
    assert cache.nearest([0.0]) == 2, "retrieval uses replacement state"
    cache.clear()
    assert cache.size() == 0 and cache.retained_bytes() == 0

    cache.set_byte_budget(one)
    assert cache.store([0.0], 3, one)
    assert cache.store([2.0], 4, one)
    assert cache.retained_bytes() == one and cache.size() == 1, "byte budget evicts the oldest"
    assert cache.nearest([0.0]) is None and cache.nearest([2.0]) == 4, "eviction chooses old key"
    cache.set_byte_budget(one - 1)
    assert not cache.store([3.0], 5, one)
    assert cache.size() == 0 and cache.retained_bytes() == 0, "oversized state is not retained"


def test_capacity_one_evicts_the_source_entry():
    """test_infra_fixes.cpp:78-95: capacity 1, threshold 10."""
    cache = HostWarmCache(1, 10.0)
    assert cache.store([0.3, 0.09], 7, 100)
    assert cache.nearest([0.31, 0.09]) == 7
    assert cache.store([0.9, 0.19], 8, 100)
    assert cache.size() == 1 and cache.nearest([0.31, 0.09]) == 8


def test_nearest_is_strict_and_first_among_equals():
    cache = HostWarmCache(10, 0.5)
    cache.store([0.0, 0.0], 1, 10)
    cache.store([0.2, 0.0], 2, 10)
    cache.store([-0.2, 0.0], 3, 10)
    assert cache.nearest([0.19, 0.0]) == 2
    assert cache.nearest([0.0, 0.3]) == 1          # the closest entry
    assert cache.nearest([0.1, 0.0]) == 1          # 1 and 2 tie at 0.1: the older (strict <)
    # the threshold itself does not match
    far = HostWarmCache(10, 0.5)
    far.store([0.0], 1, 10)
    assert far.nearest([0.5]) is None and far.nearest([0.4999]) == 1


def test_nonfinite_rows_and_zero_capacity_are_not_stored():
    cache = HostWarmCache(5, 0.5)
    assert not cache.store([math.nan, 0.1], 1, 10)
    assert not cache.store([math.inf, 0.1], 2, 10)
    assert cache.size() == 0
    none = HostWarmCache(0, 0.5)
    assert not none.store([0.1, 0.1], 3, 10) and none.size() == 0


def test_count_limit_is_first_in_first_out():
    cache = HostWarmCache(3, 100.0)
    for t in range(5):
        assert cache.store([float(t)], t, 10)
    assert cache.size() == 3 and cache.retained_bytes() == 30
    assert cache.nearest([0.0]) == 2  # 0 and 1 were evicted; 2 is the closest left


@pytest.mark.parametrize("args", [(-1, 0.5, 10), (5, -0.1, 10), (5, math.nan, 10),
                                  (5, 0.5, -1)])
def test_invalid_limits_are_rejected(args):
    with pytest.raises(ValueError, match="warm-start cache limits"):
        HostWarmCache(*args)
