"""The learned-pruning network's HIP kernels against the numpy oracle (oracle/pnet_oracle.py):
the MFMA bf16 GEMM, the forward logits, the backward gradients of every parameter tensor,
one Adam step, the prune masks; and that training on exact-RDO labels lowers the loss."""
import os
import sys

import numpy as np
import pytest
import torch

import maroontree_amd as mt
from maroontree_amd import prune as PR
from maroontree_amd import rdo as R
from maroontree_amd import workload as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pnet_oracle as PO  # noqa: E402

pytestmark = pytest.mark.gpu


def bf16_bits(x):
    return (PO.bf16(x).view(np.uint32) >> 16).astype(np.uint16).view(np.int16)


@pytest.mark.parametrize("M,N,K,trans_a,trans_b", [
    (37, 45, 19, False, False), (64, 64, 64, True, False), (130, 70, 33, False, True),
    (1000, 16, 9, False, False), (9, 16, 4000, True, False), (257, 98, 64, True, True)])
def test_gemm_bf16_matches_numpy(dev, M, N, K, trans_a, trans_b):
    rng = np.random.default_rng(M * 7 + N + K)
    a = PO.bf16(rng.standard_normal((M, K)).astype(np.float32))
    b = PO.bf16(rng.standard_normal((K, N)).astype(np.float32))
    want = a.astype(np.float64) @ b.astype(np.float64)
    A = np.ascontiguousarray(a.T if trans_a else a)
    B = np.ascontiguousarray(b.T if trans_b else b)
    sam, sak = (1, M) if trans_a else (K, 1)
    sbk, sbn = (1, K) if trans_b else (N, 1)
    At = torch.from_numpy(bf16_bits(A)).to(dev)
    Bt = torch.from_numpy(bf16_bits(B)).to(dev)
    C = torch.empty((M, N), dtype=torch.float32, device=dev)
    print(f"gemm {M}x{N}x{K} ta={trans_a} tb={trans_b}", flush=True)
    PR.gemm_bf16(M, N, K, At, sam, sak, Bt, sbk, sbn, C)
    got = C.cpu().numpy()
    scale = np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64)
    assert np.all(np.abs(got - want) <= 1e-5 * scale + 1e-6)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))


@pytest.fixture(scope="module")
def batch():
    rng = np.random.default_rng(11)
    img = W.synth_image(5, 256, 256, 10)
    pix = np.stack([img[y:y + 64, x:x + 64].reshape(-1) for y in range(0, 256, 64)
                    for x in range(0, 256, 64)]).astype(np.uint16)
    qidx = rng.integers(0, 256, len(pix)).astype(np.uint8)
    t = (rng.random((len(pix), 98)) < 0.4).astype(np.float32)
    t[:, :85] = rng.integers(0, 10, (len(pix), 85))  # partition types of the nodes
    m = (rng.random((len(pix), 98)) < 0.8).astype(np.float32)
    return pix, qidx, t, m


def test_forward_backward_match_oracle(dev, batch):
    pix, qidx, t, m = batch
    p = PO.init_params(7)
    d = PO.unpack(p)
    rng = np.random.default_rng(8)
    for k in d:
        if k.startswith("b"):
            d[k][:] = rng.standard_normal(d[k].shape).astype(np.float32) * 0.05
    p = PO.pack(d).astype(np.float32)
    net = PR.PruneNet(p, dev)
    pt = torch.from_numpy(pix.view(np.int16)).to(dev)
    qt = torch.from_numpy(qidx).to(dev)
    print("pnet forward/backward", len(pix), flush=True)
    logits = net.forward(pt, qt, 10).cpu().numpy()
    ref = PO.Net()
    want = ref.forward(p, pix.reshape(-1, 64, 64), qidx, 10)
    assert rel(logits, want) < 1e-2
    B = len(pix)
    ns = 1.0 / m[:, :85].sum()
    nm = 1.0 / (B * 13)
    loss = net.backward(torch.from_numpy(t).to(dev), torch.from_numpy(m).to(dev)).item()
    g = net.grads.cpu().numpy()
    gw, lw = ref.backward(t, m, ns, nm)
    assert abs(loss - lw) <= 1e-3 * abs(lw)
    o = 0
    for name, shape in PO.SHAPES:
        n = int(np.prod(shape))
        e = rel(g[o:o + n], gw[o:o + n])
        assert e < 3e-2, (name, e)
        o += n


def test_adam_step_matches_oracle(dev):
    rng = np.random.default_rng(4)
    p = rng.standard_normal(PR.PARAMS).astype(np.float32)
    g = rng.standard_normal(PR.PARAMS).astype(np.float32)
    net = PR.PruneNet(p, dev)
    net.grads.copy_(torch.from_numpy(g))
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    pw = p.copy()
    for step in (1, 2, 3):
        net.adam(1e-3)
        PO.adam(pw, g, m, v, step, 1e-3)
    np.testing.assert_allclose(net.params.cpu().numpy(), pw, rtol=1e-5, atol=1e-7)


def test_prune_masks_match_numpy(dev):
    rng = np.random.default_rng(9)
    z = rng.standard_normal((20, PR.OUTPUTS)).astype(np.float32) * 3
    z[3, PR.MODE_OFFSET:] = 0.5  # all mode scores tied: the lower modes win
    net = PR.PruneNet(PO.init_params(1), dev)
    lo, hi = [0.2, 0.1, 0.3, 0.25], [0.8, 0.9, 0.6, 0.75]
    topk = [3, 5, 7, 13, 0]
    cover = 0.02
    got = net.masks(torch.from_numpy(z).to(dev), topk, lo, hi, cover).cpu().numpy().view(np.uint16)
    fallbacks = 0
    for b in range(20):
        zm = z[b, PR.MODE_OFFSET:]
        order = sorted(range(13), key=lambda j: (-zm[j], j))
        p = 1.0 / (1.0 + np.exp(-zm.astype(np.float64)))
        for n in range(R.NODES):
            mm = (1 << 0) | (1 << 13)
            k = topk[R.node_geometry(n)[0]]
            for j in order[:k]:
                mm |= 1 << j
            covered = np.prod([1.0 - p[j] for j in order[k:] if j != 0])
            if covered < cover:
                mm |= 0x1FFF
                fallbacks += 1
            want = mm
            if n < 85:
                L = R.node_geometry(n)[0]
                zp = z[b, n * 10:n * 10 + 10].astype(np.float64)
                ps = 1.0 - np.exp(zp[0] - zp.max()) / np.exp(zp - zp.max()).sum()
                if abs(ps - lo[L]) < 1e-5 or abs(ps - hi[L]) < 1e-5:
                    continue  # too near a threshold for the f32 softmax to be decided the same way
                want |= R.NO_SPLIT if ps < lo[L] else R.FORCE_SPLIT if ps > hi[L] else 0
            assert got[b, n] == want, (b, n)
    assert 0 < fallbacks < 20 * R.NODES  # the fallback fires on some nodes, not all


def test_training_on_rdo_labels_lowers_the_loss(dev):
    pix, qidx, ts, ms = [], [], [], []
    for seed in range(4):
        img = W.synth_image(300 + seed, 256, 192, 8)
        q = [140, 91, 53, 17][seed]
        src = torch.from_numpy(img.view(np.int16)).to(dev)
        _, nodes, _ = R.intra_rdo(src, 8, q, 15, R.ANGLE_DELTAS | R.MIN4)
        t, m = PR.labels_from_nodes(R.nodes_to_np(nodes), 256, 192)
        sbs = PR.extract_superblocks(src).cpu().numpy()
        pix.append(sbs)
        qidx.append(np.full(len(sbs), q, np.uint8))
        ts.append(t)
        ms.append(m)
    pix = torch.from_numpy(np.concatenate(pix)).to(dev)
    qidx = torch.from_numpy(np.concatenate(qidx)).to(dev)
    t = torch.from_numpy(np.concatenate(ts)).to(dev)
    m = torch.from_numpy(np.concatenate(ms)).to(dev)
    net = PR.PruneNet(PO.init_params(5), dev)
    losses = []
    for it in range(80):
        net.forward(pix, qidx, 8)
        losses.append(net.backward(t, m).item())
        net.adam(3e-3)
    print("loss", losses[0], losses[-1], flush=True)
    assert losses[-1] < 0.7 * losses[0]
