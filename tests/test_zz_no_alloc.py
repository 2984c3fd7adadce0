"""tam_tally_blocks_noalloc (the block sums through the library's fixed
per-GPU device buffer, no device allocation) against tam_tally_blocks (which
allocates its own buffer), both through the C-ABI: the same kernel sums the
same tally in the same order, so the sums are bit-identical -- on two calls
in a row, the second with a smaller input, so nothing of the first call's
buffer contents may show through -- and a block count past the buffer is
refused with an error, not written past its end."""
import ctypes as C

import numpy as np
import pytest

from tam_amd import configs, lib

pytestmark = pytest.mark.gpu


def _blocks(g, k, noalloc):
    nz, ny, nx = g.shape
    out = np.full((nz // k, ny // k, nx // k), np.nan)
    fn = g.L.tam_tally_blocks_noalloc if noalloc else g.L.tam_tally_blocks
    rc = fn(g.h, int(k), out.ctypes.data_as(C.c_void_p))
    return rc, out


def _grid(n, packets, seed):
    prob = configs.three_layer_skin(n=(n, n, n), ext=(0.5, 0.5, 0.25), packets=packets, steps=1, cycles=1)
    g = prob.grid()
    g.mc_run(packets, seed=seed)
    return g


def test_noalloc_block_sums_equal_the_allocating_entry_point():
    print("case: no-alloc block sums, 64^3 k=4 then 32^3 k=8 then 64^3 k=16", flush=True)
    big, small = _grid(64, 100_000, 1), _grid(32, 50_000, 2)
    for g, k in ((big, 4), (small, 8), (big, 16)):
        rc_a, a = _blocks(g, k, noalloc=False)
        rc_b, b = _blocks(g, k, noalloc=True)
        assert rc_a == lib.TAM_OK and rc_b == lib.TAM_OK
        assert np.array_equal(a, b), k
        t = g.download(lib.FIELD_TALLY)
        n = t.shape[0] // k
        np.testing.assert_allclose(b, t.reshape(n, k, n, k, n, k).sum(axis=(1, 3, 5)), rtol=1e-12, atol=0)
        assert (b > 0).any()
    big.close()
    small.close()


def test_noalloc_refuses_more_blocks_than_its_buffer():
    g = _grid(80, 20_000, 3)                      # k = 1: 512 000 blocks > 262 144
    rc, _ = _blocks(g, 1, noalloc=True)
    assert rc == -1
    assert "fixed buffer" in g.L.tam_last_error(g.h).decode()
    rc_b, b = _blocks(g, 2, noalloc=True)        # 64 000 blocks fit
    rc_a, a = _blocks(g, 2, noalloc=False)
    assert rc_b == lib.TAM_OK and rc_a == lib.TAM_OK and np.array_equal(a, b)
    g.close()
