"""pfc_stage_verify_table_noalloc against pfc_stage_verify_table: the same
two results (Mont(X) from the product of the multipliers and the table's
start residue, modulo 2^64-59) on two calls in a row, the second with a
smaller table, and on a table with a changed chunk (the results differ as
the workspace entry point's do).  Both are called through the C ABI."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _table(dev, kind, N, corrupt=False):
    import pfcsieve_amd._native as Nt
    from pfcsieve_amd.sieve import DeviceBuffer, TableLayout
    t = TableLayout(kind, N)
    ch = (C.c_uint64 * max(len(t.chunks), 1))(*t.chunks)
    gr = (C.c_uint32 * len(t.groups))(*t.groups)
    if corrupt:
        ch[0] = ch[0] + 2 if ch[0] < 2**62 else ch[0] - 2
    dc = DeviceBuffer(dev, 8 * max(len(t.chunks), 1))
    dc.upload(bytes(ch))
    dg = DeviceBuffer(dev, 4 * len(t.groups))
    dg.upload(bytes(gr))
    return Nt.SetupTable(dc.ptr.value, dg.ptr.value, t.npos, t.digit_bits), (dc, dg)


def _both(dev, kind, N, corrupt=False):
    import pfcsieve_amd._native as Nt
    from pfcsieve_amd.sieve import DeviceBuffer
    L = Nt.lib()
    t, keep = _table(dev, kind, N, corrupt)
    work = DeviceBuffer(dev, 8 * Nt.VERIFY_WORDS)
    out = DeviceBuffer(dev, 16)
    out.upload(b"\xff" * 16)
    Nt.check(L.pfc_stage_verify_table(kind, N, C.byref(t), work.ptr, None))
    Nt.check(L.pfc_stage_verify_table_noalloc(kind, N, C.byref(t), out.ptr, None))
    Nt.check(L.pfc_device_synchronize(dev.handle))
    ref = np.frombuffer(work.download(16), np.uint64)
    got = np.frombuffer(out.download(16), np.uint64)
    return [int(x) for x in ref], [int(x) for x in got]


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_noalloc_matches_workspace_entry_point(dev, kind):
    # first call: a large table; second: a smaller one (the shared scratch
    # keeps the first call's partials, which the second must not see)
    for N in (999999, 5000):
        ref, got = _both(dev, kind, N)
        assert got == ref and ref[0] == ref[1], (kind, N)
    ref, got = _both(dev, kind, 20000, corrupt=True)
    assert got == ref and ref[0] != ref[1]


def test_noalloc_rejects_bad_arguments(dev):
    import pfcsieve_amd._native as Nt
    from pfcsieve_amd.sieve import DeviceBuffer
    L = Nt.lib()
    t, keep = _table(dev, 0, 100)
    out = DeviceBuffer(dev, 16)
    assert L.pfc_stage_verify_table_noalloc(0, 2**31, C.byref(t), out.ptr, None) == -1
    assert L.pfc_stage_verify_table_noalloc(3, 100, C.byref(t), out.ptr, None) == -1
    assert L.pfc_stage_verify_table_noalloc(0, 100, C.byref(t), None, None) == -1
    assert b"null pointer" in L.pfc_last_error()
