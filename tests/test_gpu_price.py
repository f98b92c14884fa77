"""The decision pricing entry points through the C ABI against the oracle, bit for bit:
mt_sb_activity (every superblock and the tile mean), mt_perceptual_rd_scale (every block
size over whole planes, edges and blocks off the tile, every chroma format, with and
without an AQ reference, mono_large), mt_luma_partition_distortion (random SATD values
up to 2^40, per-block and scalar steps, every qindex) and the host scalars."""
import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O

pytestmark = pytest.mark.gpu


def _plane(seed, w, h, bd):
    rng = np.random.default_rng(seed)
    p = (rng.integers(0, 1 << bd, (h, w)) * rng.uniform(0, 1, (h, w)) ** 2).astype(np.uint16)
    p[: h // 3, : w // 4] = (1 << (bd - 1)) + 3  # flat
    p[h // 2:, w // 2:] = np.clip(p[h // 2:, w // 2:] // 16 + 100, 0, (1 << bd) - 1)  # quiet
    return p


@pytest.mark.parametrize("seed,w,h,bd", [(1, 200, 136, 8), (2, 264, 72, 10), (3, 64, 64, 12), (4, 136, 200, 10)])
def test_sb_activity_matches_oracle(dev, seed, w, h, bd):
    p = _plane(seed, w, h, bd)
    print(f"sb_activity: seed={seed} {w}x{h} bd={bd}", flush=True)
    act, ref = mt.sb_activity(torch.from_numpy(p.view(np.int16)).to(dev), bd)
    act = act.cpu().numpy()
    want = np.array([O.sb_activity(p, sy, sx, bd) for sy in range(0, h, 64) for sx in range(0, w, 64)], np.float32)
    assert (act.view(np.uint32) == want.view(np.uint32)).all()
    assert ref.cpu().numpy().view(np.uint32)[0] == np.float32(O.tile_ref_activity(p, bd)).view(np.uint32)


@pytest.mark.parametrize("seed,w,h,bd", [(5, 200, 136, 10), (6, 136, 72, 8), (7, 72, 200, 12)])
def test_perceptual_rd_scale_matches_oracle(dev, seed, w, h, bd):
    p = _plane(seed, w, h, bd)
    pt = torch.from_numpy(p.view(np.int16)).to(dev)
    act, ref = mt.sb_activity(pt, bd)
    blocks = [(x, y, d) for d in (4, 8, 16, 32, 64) for y in range(0, h + d, d) for x in range(0, w + d, d)]
    blocks += [(w + 5, 3, 8), (3, h, 16), (0, 0, 0)]
    b = np.zeros(len(blocks), mt.PRDOBLK_DTYPE)
    b["x"], b["y"], b["dim"] = zip(*blocks)
    bt = torch.from_numpy(b.view(np.uint8).copy()).to(dev)
    ref_v = float(ref.cpu().numpy()[0])
    for fmt in (0, 1, 2, 3):
        for q in (5, 20, 40, 60, 91, 140, 255):
            for use_ref, mono_large in ((False, 0), (True, 0), (True, 1)):
                print(f"prdo: seed={seed} fmt={fmt} q={q} ref={use_ref} mono_large={mono_large}", flush=True)
                got = mt.perceptual_rd_scale(bt, pt, bd, q, fmt, act, ref if use_ref else None,
                                             mono_large=mono_large).cpu().numpy()
                want = np.array([O.perceptual_rd_scale(p, bd, q, fmt, ref_v if use_ref else 0.0, x, y, d,
                                                       mono_large=mono_large) for x, y, d in blocks], np.float32)
                bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
                assert len(bad) == 0, f"{len(bad)} blocks differ, first {[blocks[i] for i in bad[:4]]}"


def test_luma_partition_distortion_matches_oracle(dev):
    rng = np.random.default_rng(9)
    n = 5000
    satd = rng.integers(0, 1 << 40, n, dtype=np.int64)
    satd[:100] = rng.integers(0, 64, 100)
    qstep = rng.uniform(0.0, 3000.0, n).astype(np.float32)
    qstep[:50] = 0.5
    st = torch.from_numpy(satd).to(dev)
    qt = torch.from_numpy(qstep).to(dev)
    for q in (0, 20, 28, 33, 48, 81, 100, 140, 255):
        for sub in (0, 1):
            got = mt.luma_partition_distortion(st, qt, q, sub).cpu().numpy()
            want = np.array([O.luma_partition_distortion(int(s), float(qs), q, sub) for s, qs in zip(satd, qstep)],
                            np.float32)
            assert (got.view(np.uint32) == want.view(np.uint32)).all(), (q, sub)
            got1 = mt.luma_partition_distortion(st[:64], 37.0, q, sub).cpu().numpy()
            want1 = np.array([O.luma_partition_distortion(int(s), 37.0, q, sub) for s in satd[:64]], np.float32)
            assert (got1.view(np.uint32) == want1.view(np.uint32)).all()
            assert mt.luma_satd_scale(q, sub) == O.luma_satd_scale(q, sub)
            assert mt.k_satd(q, sub) == O.k_satd(q, sub)
    assert mt.proxy_wsse(3.5) == O.proxy_wsse(3.5)


def test_pricing_refuses_bad_arguments(dev):
    p = torch.zeros((64, 64), dtype=torch.int16, device=dev)
    with pytest.raises(mt.MtError):
        mt.sb_activity(p, 9)
    act, _ = mt.sb_activity(p, 8)
    b = torch.zeros(8, dtype=torch.uint8, device=dev)
    with pytest.raises(mt.MtError):
        mt.perceptual_rd_scale(b, p, 8, 50, 4, act)
    with pytest.raises(mt.MtError):
        mt.perceptual_rd_scale(b, p, 8, 256, 0, act)
