"""exa_volume_unknowns_restriction_batch_noalloc (the allocation-free form of
the batched volumeUnknownsRestriction, amrRoutines.cpph:386-491) against the
allocating batch call through the C-ABI: two calls in a row, the second on a
smaller batch, results bit-identical (the same kernel on the same inputs);
a call beyond the scratch limit and a repeated destination are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _restrict_items(nc, level_delta=1):
    """Every coarse cell of nc^3 restricted from its 27^level_delta children."""
    n3 = 3 ** level_delta
    items, ptr = [], [0]
    for c in range(nc ** 3):
        for k in range(n3):
            for j in range(n3):
                for i in range(n3):
                    items.append((c, len(items), i, j, k))
        ptr.append(len(items))
    return np.array(items, dtype=np.int32), np.array(ptr, dtype=np.int32)


def _call_noalloc(exa, s, coarse, fine, ptr, items, level_delta=1):
    return exa.lib().exa_volume_unknowns_restriction_batch_noalloc(
        ctypes.byref(s), coarse.ptr, coarse.shape[0], fine.ptr, fine.shape[0], ptr.ptr, ptr.shape[0] - 1,
        items.ptr, items.shape[0], level_delta)


def test_noalloc_restriction_matches_batch_call(exa):
    exa.check(exa.lib().exa_init(0), "exa_init")
    rng = np.random.default_rng(77)
    s = exa.euler_solver(3)
    B, D = 4, 5
    for nc in (3, 2):  # two calls in a row, the second smaller
        print(f"case: no-alloc volume restriction, {nc}^3 coarse cells, 27 children each", flush=True)
        items, ptr = _restrict_items(nc)
        fine = rng.standard_normal((items.shape[0], B ** 3 * D))
        start = rng.standard_normal((nc ** 3, B ** 3 * D))
        d_items, d_ptr, d_fine = exa.DeviceArray(items), exa.DeviceArray(ptr), exa.DeviceArray(fine)
        d_ref, d_new = exa.DeviceArray(start), exa.DeviceArray(start)
        exa.volume_restriction_batch(s, d_ref, d_fine, d_ptr, d_items, 1)
        st = _call_noalloc(exa, s, d_new, d_fine, d_ptr, d_items)
        assert st == exa.EXA_OK, exa.lib().exa_status_string(st)
        assert np.array_equal(d_new.get(), d_ref.get())
        assert not np.array_equal(d_new.get(), start)


def test_noalloc_restriction_refusals(exa):
    s = exa.euler_solver(2)
    blk = 27 * 5
    rng = np.random.default_rng(3)
    # a repeated destination: one range is applied, the call fails (the flag is
    # cleared again by the next call, which succeeds)
    items = np.array([(0, 0, 0, 0, 0), (0, 1, 1, 0, 0)], dtype=np.int32)
    ptr = np.array([0, 1, 2], dtype=np.int32)
    d_c, d_f = exa.DeviceArray(rng.standard_normal((1, blk))), exa.DeviceArray(rng.standard_normal((2, blk)))
    d_items, d_ptr = exa.DeviceArray(items), exa.DeviceArray(ptr)
    assert _call_noalloc(exa, s, d_c, d_f, d_ptr, d_items) == exa.EXA_E_INVALID
    ok_ptr = exa.DeviceArray(np.array([0, 2], dtype=np.int32))
    assert _call_noalloc(exa, s, d_c, d_f, ok_ptr, d_items) == exa.EXA_OK
    # beyond the claim bitmap (8,388,608 destination blocks): refused before any launch
    st = exa.lib().exa_volume_unknowns_restriction_batch_noalloc(
        ctypes.byref(s), d_c.ptr, 8388609, d_f.ptr, 2, ok_ptr.ptr, 1, d_items.ptr, 2, 1)
    assert st == exa.EXA_E_INVALID
