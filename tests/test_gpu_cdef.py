"""mt_cdef_frame against the oracle's frame_cdef restatement (oracle/oracle_cdef.cpp): the
per-unit distortion tables of all twelve strength pairs (luma, U + V), the decision (on,
cdef_bits, damping, strength entries, every unit's entry) and the filtered planes, sample for
sample: on the RDO's decisions and reconstructions of small frames (monochrome, 4:2:0, 4:4:4;
8/10/12-bit; deblocked or not; weighted or plain SSE; Slow and Fast candidate lists) and on a
1920x1080 10-bit 4:2:0 frame with a random skip map."""
import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O
from maroontree_amd import cdef as C
from maroontree_amd import rdo as R
from maroontree_amd import workload as W

pytestmark = pytest.mark.gpu


def _t(dev, a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a, np.uint16).view(np.int16)).to(dev)


def compare(dev, pre, src, skip, fmt, bd, q, lam, wsse, slow):
    o_planes, o_prm, o_grid, _, _, o_tabs = O.cdef_frame(pre, src, skip, fmt, bd, q, lam, wsse_scale=wsse, slow=slow)
    rec = tuple(_t(dev, p) for p in pre)
    srcs = tuple(_t(dev, p) for p in src)
    sk = torch.from_numpy(np.ascontiguousarray(skip, np.uint8)).to(dev)
    out, prm, grid, tabs = C.frame(rec, srcs, sk, fmt, bd, q, lam, wsse_scale=wsse, slow=slow, with_tables=True)
    torch.cuda.synchronize()
    assert (tabs == o_tabs).all(), f"tables: {int((tabs != o_tabs).sum())} entries differ"
    assert prm["on"] == o_prm["on"]
    if prm["on"]:
        for k in ("bits", "damping", "n", "y_pri", "y_sec", "uv_pri", "uv_sec"):
            assert prm[k] == o_prm[k], (k, prm[k], o_prm[k])
        assert (grid == o_grid).all()
    for name, a, b in zip("yuv", out, o_planes):
        if b is None:
            continue
        g = a.cpu().numpy().view(np.uint16)
        bad = np.argwhere(g != b)
        assert len(bad) == 0, f"{name}: {len(bad)} samples differ, first {bad[:4].tolist()}"
    return prm


@pytest.mark.parametrize("fmt,bd,w,h,q,deblock,wsse,slow", [
    (R.LUMA_ONLY, 8, 264, 136, 200, True, 1.0, True),
    (R.LUMA_ONLY, 10, 136, 128, 120, False, 0.0, False),
    (R.CHROMA_420, 8, 200, 136, 160, False, 1.0, True),
    (R.CHROMA_420, 12, 136, 72, 230, True, 0.0, True),
    (R.CHROMA_444, 8, 200, 136, 160, True, 0.0, False),
    (R.CHROMA_444, 10, 136, 128, 200, False, 1.0, True),
])
def test_cdef_on_rdo_frames_matches_oracle(dev, fmt, bd, w, h, q, deblock, wsse, slow):
    import test_oracle_cdef as T
    print(f"cdef fmt={fmt} bd={bd} {w}x{h} q={q} deblock={deblock} wsse={wsse} slow={slow}", flush=True)
    src, pre, _, o_prm, _, skip, _, _ = T.rdo_with_cdef(41 + q, w, h, q, fmt, deblock, wsse, slow, bd=bd)
    lam = R.rdo_lambda(q, bd, (2 if fmt == R.CHROMA_420 else 0) if fmt else 2)
    prm = compare(dev, pre, src, skip, fmt, bd, q, lam, wsse, slow)
    print(f"  {prm}", flush=True)


@pytest.mark.parametrize("slow", [True, False])
def test_cdef_1080p_random_skip_matches_oracle(dev, slow):
    w, h, bd, q = 1920, 1080, 10, 140
    print(f"cdef 1920x1080 10-bit 4:2:0 random skip slow={slow}", flush=True)
    y, u, v = W.synth_yuv(77, w, h, bd, 2)
    rng = np.random.default_rng(3)
    noisy = [np.clip(p.astype(np.int64) + rng.integers(-24, 25, p.shape), 0, (1 << bd) - 1).astype(np.uint16)
             for p in (y, u, v)]
    skip = (rng.random((h // 8, w // 8)) < 0.3).astype(np.uint8)
    prm = compare(dev, tuple(noisy), (y, u, v), skip, R.CHROMA_420, bd, q, R.rdo_lambda(q, bd, 2), 1.0, slow)
    print(f"  {prm}", flush=True)
    assert prm["on"] == 1


def test_cdef_refuses_bad_frames(dev):
    y = torch.zeros((64, 68), dtype=torch.int16, device=dev)
    sk = torch.zeros((8, 8), dtype=torch.uint8, device=dev)
    with pytest.raises(mt.MtError):
        C.frame((y, None, None), (y, None, None), sk, R.LUMA_ONLY, 8, 100, 1.0)
    y = torch.zeros((64, 64), dtype=torch.int16, device=dev)
    with pytest.raises(mt.MtError):  # in place
        C.frame((y, None, None), (y, None, None), sk, R.LUMA_ONLY, 8, 100, 1.0, out=(y, None, None))
