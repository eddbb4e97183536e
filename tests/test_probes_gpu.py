"""The measurement kernels: the vector-ALU probe runs, and the store probe
writes exactly the pattern and the bytes it reports (grid-stride fill; items
of 170 chunks of 3 KiB on 1024-byte boundaries, and 16 bytes off them),
within the buffer it is given."""
from __future__ import annotations

import numpy as np
import pytest

from process_amd import core
from process_amd._lib import PsxError

pytestmark = pytest.mark.gpu

CHUNK, IC = 3072, 170


def test_valu_probe_runs(ctx):
    for mode in ("philox", "int"):
        ms, lanes = core.probe_valu(ctx, mode, 64)
        assert ms > 0 and lanes > 0


def test_store_probe_fill(ctx):
    buf = core.DeviceArray(ctx, (1 << 20, 4), np.uint32).fill(0)
    ms, n = core.probe_store(ctx, buf, "fill")
    assert ms > 0 and n == buf.nbytes
    got = buf.numpy()
    i = np.arange(1 << 20, dtype=np.uint32)
    assert np.array_equal(got[:, 0], i) and (got[:, 1:] == [1, 2, 3]).all()


@pytest.mark.parametrize("mode,shift", [("chunks", 0), ("chunks_shifted", 16)])
def test_store_probe_chunks(ctx, mode, shift):
    nbytes = 8 << 20
    buf = core.DeviceArray(ctx, (nbytes // 4,), np.uint32).fill(0)
    ms, n = core.probe_store(ctx, buf, mode)
    skip = (-buf.ptr) % 1024 + shift
    n_items = (nbytes - skip) // (IC * CHUNK)
    assert ms > 0 and n == n_items * IC * CHUNK and skip + n <= nbytes
    raw = buf.numpy().view(np.uint8)
    assert not raw[:skip].any() and not raw[skip + n:].any()  # nothing outside the chunks
    words = raw[skip:skip + n].view(np.uint32).reshape(n_items, IC, 3, 64, 4)
    item = np.arange(n_items, dtype=np.uint32)[:, None, None, None]
    k = np.arange(IC, dtype=np.uint32)[None, :, None, None]
    q = np.arange(3, dtype=np.uint32)[None, None, :, None]
    assert (words[..., 0] == item).all() and (words[..., 1] == k).all()
    assert (words[..., 2] == q).all() and (words[..., 3] == 3).all()


def test_store_probe_refuses_misaligned(ctx):
    buf = core.DeviceArray(ctx, (1024,), np.uint32)
    import ctypes as C

    from process_amd import _abi as A
    from process_amd._lib import check, lib
    ms, n = C.c_double(), C.c_uint64()
    with pytest.raises(PsxError):
        check(lib().psx_probe_store(ctx.h, C.c_void_p(buf.ptr + 4), 4000, A.PSX_PROBE_STORE_FILL, C.byref(ms),
                                    C.byref(n)), ctx.h)
