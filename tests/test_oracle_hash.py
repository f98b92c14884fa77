"""The block checksum of the chain (orc_block_hash, the library's coef_hash / pix_hash): it
is order independent by construction (a sum) and changes when any one level, target bit or
reconstructed sample changes."""
import ctypes

import numpy as np

import oracle_lib as O


def block_hash(lv, tg, rc):
    return O.lib().orc_block_hash(len(lv), O.ptr(lv, ctypes.c_int32), O.ptr(tg, ctypes.c_float), len(rc),
                                  O.ptr(rc, ctypes.c_int32))


def test_checksum_sees_every_single_change():
    rng = np.random.default_rng(3)
    for n, npx in ((16, 16), (1024, 4096)):
        lv = rng.integers(-300, 300, n).astype(np.int32)
        tg = (lv + rng.random(n) - 0.5).astype(np.float32)
        rc = rng.integers(-2000, 2000, npx).astype(np.int32)
        h0 = block_hash(lv, tg, rc)
        for i in rng.choice(n, 8, replace=False):
            for d in (1, -1, 1 << 20):
                l2 = lv.copy()
                l2[i] += d
                assert block_hash(l2, tg, rc) != h0
            t2 = tg.copy()
            t2.view(np.uint32)[i] ^= 1 << int(rng.integers(0, 32))
            assert block_hash(lv, t2, rc) != h0
        for j in rng.choice(npx, 16, replace=False):
            r2 = rc.copy()
            r2[j] += 1
            assert block_hash(lv, tg, r2) != h0
        # swapping two samples' values moves them to other weights: the sum changes
        r3 = rc.copy()
        a, b = 0, npx - 1
        if r3[a] == r3[b]:
            r3[b] += 1
        r3[a], r3[b] = r3[b], r3[a]
        assert block_hash(lv, tg, r3) != h0
