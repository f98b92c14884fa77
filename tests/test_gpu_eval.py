"""Bit-exact parity of the batched candidate evaluation (mt_eval_candidates) with the
oracle's candidate loop body, composed stage by stage as the reference composes it
(src/coder/partition_search.rs:471-520): predict_nd_ad / dc / filter intra from the
reconstruction plane, residual_pred against the source plane, DctDispatch forward + quant
(+QM), IdctDispatch inverse, sse_recon / weighted distortion / luma_satd over the displayed
part, real_block_bits_bounded, rd_cost = dist + lambda * bits (src/partition_rd.rs:140)."""
import sys

import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O

pytestmark = pytest.mark.gpu

from evalgen import PH, PW, make_cands, make_planes, oracle_candidate  # noqa: E402


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_eval_candidates_match_oracle(dev, bd):
    rng = np.random.default_rng(900 + bd)
    src, recon = make_planes(rng, bd)
    src2 = src.copy()
    cands, pool = make_cands(rng, 1500, bd)
    print(f"eval bd={bd} candidates={len(cands)} pool={pool}", flush=True)
    t = [torch.from_numpy(a.view(np.int16)).to(dev) for a in (src, recon, src2)]
    # plane 2: the source with a displayed area cut at 301 x 233 (edge blocks clip)
    views = [(t[0], PW, PH, PW, PH), (t[1], PW, PH, PW, PH), (t[2], PW, PH, 301, 233)]
    out = mt.eval_candidates(torch.from_numpy(cands.view(np.uint8)).to(dev), views, pool)
    torch.cuda.synchronize()
    res = out.cpu().numpy().view(mt.CANDRES_DTYPE)
    planes_np = [src, recon, src2]
    vnp = [(None, PW, PH, PW, PH), (None, PW, PH, PW, PH), (None, PW, PH, 301, 233)]
    bad = 0
    for i in range(len(cands)):
        g = res[i]
        assert g["status"] == 0, (i, g["status"])
        want = oracle_candidate(cands[i], planes_np, vnp)
        got = (int(g["sse"]), int(g["dist"]), int(g["satd"]), np.float32(g["bits"]),
               np.float32(g["cost"]), int(g["nnz"]))
        same = got[:3] == want[:3] and got[5] == want[5] and \
            got[3].tobytes() == want[3].tobytes() and got[4].tobytes() == want[4].tobytes()
        if not same:
            bad += 1
            if bad <= 5:
                print("mismatch", i, cands[i], got, want, file=sys.stderr)
    assert bad == 0, f"{bad} of {len(cands)} candidates differ"


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_eval_candidates_trellis_match_oracle(dev, bd):
    """mt_eval_candidates_trellis: candidates with MT_CAND_TRELLIS (and a chroma
    subsampling for the trellis lambda law) run trellis_optimize_ctx between quantization
    and dequantization; every result field against the oracle's evaluation with the trellis,
    bit for bit; the others as mt_eval_candidates."""
    rng = np.random.default_rng(700 + bd)
    src, recon = make_planes(rng, bd)
    cands, pool = make_cands(rng, 3000, bd)
    tr = rng.random(len(cands)) < 0.7
    cands["rflags"] |= (tr * mt.CAND_TRELLIS + tr * (rng.integers(0, 3, len(cands)) << 2)).astype(np.uint8)
    print(f"eval+trellis bd={bd} candidates={len(cands)}", flush=True)
    t = [torch.from_numpy(a.view(np.int16)).to(dev) for a in (src, recon, src)]
    views = [(t[0], PW, PH, PW, PH), (t[1], PW, PH, PW, PH), (t[2], PW, PH, 301, 233)]
    out = mt.eval_candidates(torch.from_numpy(cands.view(np.uint8)).to(dev), views, pool, trellis=True)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(mt.CANDRES_DTYPE)
    want = O.eval_batch(cands, [(src, PW, PH, PW, PH), (recon, PW, PH, PW, PH), (src, PW, PH, 301, 233)])
    same = (got.view(np.uint8).reshape(len(cands), -1) == want.view(np.uint8).reshape(len(cands), -1)).all(axis=1)
    assert same.all(), f"{int((~same).sum())} of {len(cands)} differ, first {np.nonzero(~same)[0][:5]}"
    assert (got["status"] == 0).all()


def test_eval_candidates_flags_invalid(dev):
    """A candidate outside its plane or pool slot, or of an unsupported (size, type), gets a
    status and cost +inf; its neighbours are unaffected."""
    rng = np.random.default_rng(3)
    src, recon = make_planes(rng, 10)
    cands, pool = make_cands(rng, 40, 10)
    cands[3]["x"] = PW - 2                  # block leaves the plane
    cands[7]["tx_size"], cands[7]["tx_type"] = 15, 0  # 8x32: no reference transform
    cands[9]["offset"] = pool               # slot beyond the pool
    cands[11]["planes"] = 3 << 2            # reconstruction plane index 3 of 3 planes
    t = [torch.from_numpy(a.view(np.int16)).to(dev) for a in (src, recon, src)]
    views = [(t[0],), (t[1],), (t[2],)]
    res = mt.eval_candidates(torch.from_numpy(cands.view(np.uint8)).to(dev), views,
                             pool).cpu().numpy().view(mt.CANDRES_DTYPE)
    bad = {3, 7, 9, 11}
    for i in range(len(cands)):
        if i in bad:
            assert res[i]["status"] != 0 and np.isinf(res[i]["cost"]), i
        else:
            assert res[i]["status"] == 0 and np.isfinite(res[i]["cost"]), i


def test_eval_candidates_bench_workload_full_image(dev):
    """The bench's own workload generator (seeded synthetic image, every square block x 11
    luma modes of each superblock, qindex 91, weighted SSE) on a 1920x1088 10-bit image:
    all 1.9 M candidates against the oracle's threaded batch evaluation, every result field
    bit for bit."""
    from maroontree_amd import workload as W
    w, h = 1920, 1088
    img = W.synth_image(0x6D61726F6F6E, w, h, 10)
    rec = W.synth_recon(img, 0x6D61726F6F6E, 10)
    cands, pool = W.sb_candidates(w, h, 10, 91)
    print(f"eval full image: {len(cands)} candidates", flush=True)
    planes = [(torch.from_numpy(img.view(np.int16)).to(dev),),
              (torch.from_numpy(rec.view(np.int16)).to(dev),)]
    got = mt.eval_candidates(torch.from_numpy(cands.view(np.uint8)).to(dev), planes,
                             pool).cpu().numpy().view(mt.CANDRES_DTYPE)
    want = O.eval_batch(cands, [(img, w, h, w, h), (rec, w, h, w, h)], nthreads=16)
    g = got.view(np.uint8).reshape(len(cands), -1)
    e = want.view(np.uint8).reshape(len(cands), -1)
    bad = np.nonzero(~(g == e).all(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} candidates differ; first {bad[:5]}: {got[bad[:3]]} vs {want[bad[:3]]}"
    assert (got["status"] == 0).all() and (got["nnz"] > 0).any()
