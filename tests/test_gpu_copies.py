"""Host <-> device copies through the context's pinned staging (copies of
4 MiB and more go through two 8 MiB halves, chunk by chunk): byte-exact round
trips at sizes around the threshold and the chunk boundaries, back-to-back
uploads that reuse both halves before any read-back, and the runtime's own
path below the threshold."""
import numpy as np
import pytest

import bvhtracer_amd as B

pytestmark = pytest.mark.gpu

MIB = 1 << 20


@pytest.mark.parametrize("nbytes", [4 * MIB - 4, 4 * MIB, 8 * MIB - 4, 8 * MIB, 8 * MIB + 4,
                                    16 * MIB + 12, 25 * MIB + 16, 1024])
def test_copy_round_trip(ctx, nbytes):
    rng = np.random.default_rng(nbytes)
    a = rng.integers(0, 2**32, nbytes // 4, dtype=np.uint32)
    buf = ctx.alloc(nbytes)
    try:
        buf.upload(a)
        back = buf.download(np.uint32)
        assert back.tobytes() == a.tobytes()
    finally:
        buf.free()


def test_back_to_back_uploads_keep_their_bytes(ctx):
    """Three uploads of 20 MiB each before any read-back: each takes three
    chunks, so its third chunk reuses the staging half of its first, and
    each call starts on halves the previous call used; each buffer must
    still hold its own bytes."""
    n = 20 * MIB // 4
    arrays = [np.full(n, k, np.uint32) + np.arange(n, dtype=np.uint32) for k in (7, 1 << 20, 1 << 30)]
    bufs = [ctx.alloc(n * 4) for _ in arrays]
    try:
        for b, a in zip(bufs, arrays):
            b.upload(a)
        for b, a in zip(bufs, arrays):
            assert b.download(np.uint32).tobytes() == a.tobytes()
    finally:
        for b in bufs:
            b.free()


def test_partial_download_and_odd_offsets(ctx):
    """A read-back shorter than the buffer and one that ends mid-chunk."""
    n = (12 * MIB + 8) // 4
    a = np.arange(n, dtype=np.uint32) * np.uint32(2654435761)
    buf = ctx.alloc(n * 4)
    try:
        buf.upload(a)
        for count in (n, n - 1, 9 * MIB // 4 + 3, 5 * MIB // 4):
            assert buf.download(np.uint32, count).tobytes() == a[:count].tobytes()
    finally:
        buf.free()
