"""Bit-exact parity of the RDO's chroma (mt_intra_rdo with 4:2:0 / 4:4:4 and CfL) with the
oracle's restatement (oracle/oracle_rdo.cpp): Y, U and V reconstructions, every node's
decision including uv mode and CfL alphas, and the final accounting (bits, chroma bits,
per-plane SSE, candidates); and of CfL / chroma candidates in mt_eval_candidates."""
import numpy as np
import pytest
import torch

import maroontree_amd as mt
import oracle_lib as O
from maroontree_amd import rdo as R
from maroontree_amd import workload as W

pytestmark = pytest.mark.gpu

ALL = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4 | R.CFL
NODE_FIELDS = ("split", "mode", "uv_mode", "cfl_alpha_u", "cfl_alpha_v")


def frame(seed, w, h, bd, fmt):
    y, u, v = W.synth_yuv(seed, w, h, bd, 2 if fmt == R.CHROMA_420 else 0)
    return y, u, v


def oracle(y, u, v, bd, q, fmt, flags, prune=None):
    sub = 2 if fmt == R.CHROMA_420 else 0
    qm = mt.qm_level_law(q, sub)
    lam = R.rdo_lambda(q, bd, sub)
    uvq = R.chroma_laws(q, sub)
    return O.intra_rdo_yuv(y, u, v, bd, q, qm, flags, lam, fmt, *uvq, prune=prune)


def gpu_frame(dev, y, u, v, bd, q, fmt):
    sub = 2 if fmt == R.CHROMA_420 else 0
    t = lambda a: torch.from_numpy(a.view(np.int16)).to(dev)  # noqa: E731
    return {"src": t(y), "src_u": t(u), "src_v": t(v), "bd": bd, "qidx": q, "qm_level": mt.qm_level_law(q, sub),
            "chroma": fmt}


def assert_same(got, want):
    rec, nodes, st = got
    o_rec, o_u, o_v, o_nodes, o_st = want
    ru, rv = st["recon_uv"]
    for name, a, b in (("y", rec, o_rec), ("u", ru, o_u), ("v", rv, o_v)):
        d = np.argwhere(a.cpu().numpy().view(np.uint16) != b)
        assert len(d) == 0, f"{name}: {len(d)} samples differ, first {d[:4].tolist()}"
    g = R.nodes_to_np(nodes)
    for f in NODE_FIELDS:
        bad = np.argwhere(g[f] != o_nodes[f])
        assert len(bad) == 0, (f, bad[:4].tolist())
    for f in ("cost", "none_cost"):
        bad = np.argwhere(g[f].view(np.uint32) != o_nodes[f].view(np.uint32))
        assert len(bad) == 0, (f, bad[:4].tolist())
    for k in ("sse", "sse_u", "sse_v", "candidates", "leaves"):
        assert st[k] == o_st[k], (k, st[k], o_st[k])
    assert st["bits"] == o_st["bits"] and st["bits_uv"] == o_st["bits_uv"], (st, o_st)


@pytest.mark.parametrize("fmt,bd,q,w,h,flags", [
    (R.CHROMA_420, 10, 91, 200, 136, ALL),
    (R.CHROMA_420, 8, 180, 136, 72, R.ANGLE_DELTAS | R.CFL),
    (R.CHROMA_420, 12, 30, 96, 80, R.MIN4),
    (R.CHROMA_444, 8, 53, 136, 72, ALL),
    (R.CHROMA_444, 10, 140, 128, 128, R.ANGLE_DELTAS | R.CFL),
])
def test_rdo_chroma_matches_oracle(dev, fmt, bd, q, w, h, flags):
    y, u, v = frame(200 + q, w, h, bd, fmt)
    print(f"rdo chroma fmt={fmt} {w}x{h} bd={bd} q={q} flags={flags}", flush=True)
    got = R.intra_rdo_batch([gpu_frame(dev, y, u, v, bd, q, fmt)], flags)[0]
    torch.cuda.synchronize()
    want = oracle(y, u, v, bd, q, fmt, flags)
    assert_same(got, want)
    if flags & R.CFL:
        assert (R.nodes_to_np(got[1])["uv_mode"] == 13).any()


@pytest.mark.parametrize("fmt,bd,q,w,h,flags", [
    (R.CHROMA_420, 10, 91, 200, 136, ALL | R.cfl_steps(3)),
    (R.CHROMA_444, 8, 53, 136, 72, ALL | R.cfl_steps(15)),
    (R.CHROMA_420, 12, 30, 136, 72, ALL | R.TRELLIS | R.cfl_steps(2)),
])
def test_rdo_cfl_walk_matches_oracle(dev, fmt, bd, q, w, h, flags):
    """cfl_alpha_rd_refine's walk (the 2S step candidates per plane and the walk over them):
    every output equals the oracle's, and some alpha differs from the run without the walk."""
    y, u, v = frame(300 + q, w, h, bd, fmt)
    print(f"rdo cfl walk fmt={fmt} {w}x{h} bd={bd} q={q} flags={flags}", flush=True)
    got = R.intra_rdo_batch([gpu_frame(dev, y, u, v, bd, q, fmt)], flags)[0]
    plain = R.intra_rdo_batch([gpu_frame(dev, y, u, v, bd, q, fmt)], flags & 0xFF)[0]
    torch.cuda.synchronize()
    assert_same(got, oracle(y, u, v, bd, q, fmt, flags))
    g, p = R.nodes_to_np(got[1]), R.nodes_to_np(plain[1])
    assert ((g["cfl_alpha_u"] != p["cfl_alpha_u"]) | (g["cfl_alpha_v"] != p["cfl_alpha_v"])).any()


def test_rdo_mixed_formats_batch_matches_oracle(dev):
    """Luma-only, 4:2:0 and 4:4:4 frames of different sizes and depths in one batch (more
    frames than one step group holds with chroma): each equals its oracle result."""
    specs = [(R.CHROMA_420, 136, 72, 10, 91), (0, 128, 64, 8, 120), (R.CHROMA_444, 72, 64, 8, 53),
             (R.CHROMA_420, 64, 64, 8, 200), (R.CHROMA_444, 136, 72, 12, 30), (R.CHROMA_420, 72, 136, 10, 17),
             (0, 64, 128, 10, 60), (R.CHROMA_420, 128, 72, 8, 140), (R.CHROMA_444, 64, 72, 10, 91),
             (R.CHROMA_420, 72, 72, 8, 91), (R.CHROMA_420, 64, 64, 10, 40), (R.CHROMA_444, 64, 64, 8, 250)]
    frames, want = [], []
    for i, (fmt, w, h, bd, q) in enumerate(specs):
        if fmt:
            y, u, v = frame(900 + i, w, h, bd, fmt)
            frames.append(gpu_frame(dev, y, u, v, bd, q, fmt))
            want.append(oracle(y, u, v, bd, q, fmt, ALL))
        else:
            y = W.synth_image(900 + i, w, h, bd)
            qm = mt.qm_level_law(q, 0)
            frames.append({"src": torch.from_numpy(y.view(np.int16)).to(dev), "bd": bd, "qidx": q, "qm_level": qm})
            want.append(O.intra_rdo(y, bd, q, qm, ALL & ~R.CFL, R.rdo_lambda(q, bd, 2)))
    print("rdo mixed batch", len(frames), flush=True)
    got = R.intra_rdo_batch(frames, ALL)
    torch.cuda.synchronize()
    for (fmt, *_), g, wnt in zip(specs, got, want):
        if fmt:
            assert_same(g, wnt)
        else:
            rec, nodes, st = g
            o_rec, o_nodes, _, o_st = wnt
            assert np.array_equal(rec.cpu().numpy().view(np.uint16), o_rec)
            assert np.array_equal(R.nodes_to_np(nodes)["mode"], o_nodes["mode"])
            assert st["bits"] == o_st["bits"] and st["sse"] == o_st["sse"]


def test_eval_cfl_and_chroma_candidates_match_oracle(dev):
    """Chroma candidates of every uv mode with the derived transform types, the chroma
    quantizer with deltas, and CfL with searched and given alphas (4:2:0 and 4:4:4 luma
    planes), against the oracle's candidate evaluation."""
    rng = np.random.default_rng(41)
    bd = 10
    y, u, v = W.synth_yuv(5, 256, 192, bd, 2)
    y4, u4, _ = W.synth_yuv(6, 128, 96, bd, 0)
    # the 4:2:0 luma also at index 16 (plane indices past 12 in the CfL candidates)
    planes_np = [y, u, y4, u4] + [u] * 12 + [y]
    dims = [(256, 192), (128, 96), (128, 96), (128, 96)] + [(128, 96)] * 12 + [(256, 192)]
    cands = np.zeros(3000, mt.CAND_DTYPE)
    off = 0
    for i in range(len(cands)):
        c = cands[i]
        s = [4, 8, 16, 32][rng.integers(0, 4)]
        fmt444 = rng.random() < 0.3
        sp, lp = (3, 2) if fmt444 else (1, 0)
        Wc, Hc = dims[sp]
        c["x"] = rng.integers(0, (Wc - s) // 4 + 1) * 4
        c["y"] = rng.integers(0, (Hc - s) // 4 + 1) * 4
        c["tx_size"] = {4: 0, 8: 1, 16: 2, 32: 3}[s]
        m = int(rng.integers(0, 14))
        c["mode"] = m
        if m == 13:
            c["pflags"] = (0 if fmt444 else 8 | 16) | (32 if rng.random() < 0.3 else 0)
            c["angle_delta"] = rng.integers(-16, 17)
            c["txtp_mode"] = 16 if (lp == 0 and rng.random() < 0.5) else lp
        else:
            c["tx_type"] = [0, 1, 2, 0, 3, 1, 2, 2, 1, 3, 1, 2, 3][m] if s < 32 else 0
            c["pflags"] = rng.integers(0, 8)
        c["planes"] = (sp & 3) | ((sp & 3) << 2)  # predict from the source plane itself
        c["qidx"] = rng.integers(0, 256)
        c["bd_qm"] = rng.integers(0, 16) | (1 << 4) | 0x40
        c["dc_delta"], c["ac_delta"] = rng.integers(-30, 30), rng.integers(-30, 30)
        c["rplane"] = 1
        c["eob_bin"] = {4: 1, 8: 6, 16: 12, 32: 16}[s]
        c["skip_ctx"] = 7 + rng.integers(0, 3)
        c["qctx"] = rng.integers(0, 4)
        c["lambda"] = 40.0
        c["offset"] = off
        off += s * s
    print(f"eval chroma/CfL candidates {len(cands)}", flush=True)
    t = [torch.from_numpy(a.view(np.int16)).to(dev) for a in planes_np]
    views = [(t[k], dims[k][0], dims[k][1], dims[k][0], dims[k][1]) for k in range(len(t))]
    got = mt.eval_candidates(torch.from_numpy(cands.view(np.uint8)).to(dev), views,
                             off).cpu().numpy().view(mt.CANDRES_DTYPE)
    want = O.eval_batch(cands, [(planes_np[k], dims[k][0], dims[k][1], dims[k][0], dims[k][1])
                                for k in range(len(t))])
    assert (got["status"] == 0).all() and (got["alpha"][cands["mode"] == 13] != 0).any()
    g = got.view(np.uint8).reshape(len(cands), -1)
    e = want.view(np.uint8).reshape(len(cands), -1)
    bad = np.nonzero(~(g == e).all(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} differ; first {bad[:4]}: {got[bad[:2]]} vs {want[bad[:2]]}"
