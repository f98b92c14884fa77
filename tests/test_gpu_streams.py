"""Calls on different streams use their own library workspaces (the chain's bucket order,
the class lists, the evaluation's stage pools, the RDO's buffers are kept per device and
stream): two batches enqueued on two streams without waiting for each other, and two
whole-frame RDOs run from two host threads on their own streams, each give the results
their single-stream call gives."""
import threading

import numpy as np
import pytest
import torch

import maroontree_amd as mt
from maroontree_amd import rdo as R
from maroontree_amd import workload as W

pytestmark = pytest.mark.gpu


def eval_inputs(dev, seed, w, h):
    img = W.synth_image(seed, w, h, 10)
    rec = W.synth_recon(img, seed, 10)
    cands, pool = W.sb_candidates(w, h, 10, 60 + seed % 100)
    ct = torch.from_numpy(cands.view(np.uint8)).to(dev)
    planes = [(torch.from_numpy(img.view(np.int16)).to(dev),), (torch.from_numpy(rec.view(np.int16)).to(dev),)]
    return ct, planes, pool, len(cands)


def test_eval_and_chain_on_two_streams(dev):
    a = eval_inputs(dev, 3, 512, 256)
    b = eval_inputs(dev, 4, 256, 512)
    print(f"eval on two streams: {a[3]} + {b[3]} candidates", flush=True)
    ref = []
    for ct, planes, pool, n in (a, b):
        out = torch.empty(n * mt.CANDRES_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        mt.eval_candidates(ct, planes, pool, out=out)
        torch.cuda.synchronize()
        ref.append(out.clone())
    blocks_a, pool_a = mt.gen_tx_blocks(11, 40000)
    blocks_b, pool_b = mt.gen_tx_blocks(12, 30000)
    chain_ref = []
    chain_in = []
    for blocks, pool in ((blocks_a, pool_a), (blocks_b, pool_b)):
        bt = torch.from_numpy(blocks.view(np.uint8).copy()).to(dev)
        pt = torch.from_numpy(pool).to(dev)
        res = torch.zeros(len(blocks) * 16, dtype=torch.uint8, device=dev)
        mt.tx_chain(bt, pt, res)
        torch.cuda.synchronize()
        chain_ref.append(res.clone())
        chain_in.append((bt, pt, len(blocks)))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(3):
        outs = [torch.zeros_like(r) for r in ref]
        couts = [torch.zeros_like(r) for r in chain_ref]
        torch.cuda.synchronize()
        for (ct, planes, pool, n), o, s in ((a, outs[0], s1), (b, outs[1], s2)):
            mt.eval_candidates(ct, planes, pool, out=o, stream=s)
        for (bt, pt, n), o, s in ((chain_in[0], couts[0], s2), (chain_in[1], couts[1], s1)):
            mt.tx_chain(bt, pt, o, stream=s)
        torch.cuda.synchronize()
        for i in range(2):
            assert torch.equal(outs[i], ref[i]), (rep, i)
            assert torch.equal(couts[i], chain_ref[i]), (rep, i)


def test_rdo_from_two_threads(dev):
    frames = []
    for seed, (w, h) in ((21, (200, 136)), (22, (136, 200))):
        y, u, v = W.synth_yuv(seed, w, h, 10, 2)
        t = lambda p: torch.from_numpy(p.view(np.int16)).to(dev)  # noqa: E731
        frames.append(dict(src=t(y), src_u=t(u), src_v=t(v), bd=10, qidx=91, qm_level=mt.qm_level_law(91, 2),
                           chroma=R.CHROMA_420))
    flags = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4 | R.CFL
    print("RDO from two threads", flush=True)
    ref = []
    for fr in frames:
        rec, nodes, st = R.intra_rdo_batch([fr], flags)[0]
        torch.cuda.synchronize()
        ref.append((rec.clone(), nodes.clone(), st["bits"], st["sse"]))
    for rep in range(2):
        got = [None, None]
        errs = []

        def run(i):
            try:
                s = torch.cuda.Stream()
                with torch.cuda.stream(s):
                    rec, nodes, st = R.intra_rdo_batch([frames[i]], flags, stream=s)[0]
                    s.synchronize()
                got[i] = (rec, nodes, st["bits"], st["sse"])
            except Exception as e:  # noqa: BLE001
                errs.append(e)
        th = [threading.Thread(target=run, args=(i,)) for i in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        assert not errs, errs
        for i in range(2):
            assert torch.equal(got[i][0], ref[i][0]) and torch.equal(got[i][1], ref[i][1]), (rep, i)
            assert got[i][2:] == ref[i][2:], (rep, i)
