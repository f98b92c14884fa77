"""Learned candidate pruning: the bf16 CNN of mt_nn.hip (forward, backward and Adam on the
GPU), its training data from the exact RDO's decisions, and the pruned RDO.

The network scores, per quadtree node 0..84 (64x64 .. 8x8), the 10 AV1 partition types, and
per superblock the 13 intra modes.  Labels come from the exact RDO's decisions: the partition
type each node got (NONE or SPLIT from mt_intra_rdo's search, HORZ / VERT where records of the
oracle's rectangular search carry them; nodes below a leaf, and nodes the frame edge forces to
split, are masked out), and for each intra mode whether any leaf of the superblock uses it.
The masks the trained network produces go to mt_intra_rdo's prune argument; a node whose
probability of being partitioned lies between the thresholds keeps both choices (full search
there).
"""
import ctypes

import numpy as np

from . import _lib
from . import rdo as R
from ._lib import MtPlane, _check, _dev_ptr, _stream_ptr, init, lib

PARAMS = 60031
OUTPUTS = 863          # 85 nodes x 10 partition-type logits, then 13 mode scores
LABELS = 98            # label / mask columns: 85 node partition types, then 13 modes
SPLIT_NODES = 85
PART_TYPES = 10
MODE_OFFSET = SPLIT_NODES * PART_TYPES
# AV1 partition types
P_NONE, P_HORZ, P_VERT, P_SPLIT = 0, 1, 2, 3


def _sigs():
    L = lib()
    if getattr(L, "_pnet_sigs", False):
        return L
    vp, i32, f32, lg = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_long
    L.mt_gemm_bf16.argtypes = [i32, i32, i32, vp, lg, lg, vp, lg, lg, vp, vp]
    L.mt_pnet_workspace_bytes.argtypes = [i32]
    L.mt_pnet_workspace_bytes.restype = ctypes.c_size_t
    L.mt_pnet_extract.argtypes = [ctypes.POINTER(MtPlane), vp, vp]
    L.mt_pnet_forward.argtypes = [vp, vp, vp, vp, i32, i32, vp, vp, vp]
    L.mt_pnet_backward.argtypes = [vp, vp, vp, i32, f32, f32, vp, vp, vp, vp]
    L.mt_pnet_adam.argtypes = [vp, vp, vp, vp, i32, f32, f32, f32, f32, vp]
    L.mt_pnet_prune_masks.argtypes = [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(f32), ctypes.POINTER(f32),
                                      f32, vp, vp]
    L._pnet_sigs = True
    return L


def _p(t):
    return None if t is None else _dev_ptr(t)


def gemm_bf16(M, N, K, A, sam, sak, B, sbk, sbn, C, stream=None):
    """mt_gemm_bf16 on int16 (bf16 bits) / float32 cuda tensors."""
    _check(_sigs().mt_gemm_bf16(M, N, K, _p(A), sam, sak, _p(B), sbk, sbn, _p(C), _stream_ptr(stream)))


def extract_superblocks(plane, stream=None):
    """uint16 cuda tensor (H, W) -> (nsb, 4096) superblocks in raster order."""
    import torch
    H, W = plane.shape
    nsb = ((W + 63) // 64) * ((H + 63) // 64)
    out = torch.empty((nsb, 4096), dtype=plane.dtype, device=plane.device)
    pv = MtPlane(_dev_ptr(plane).value, plane.stride(0), W, H, W, H, 0)
    _check(_sigs().mt_pnet_extract(ctypes.byref(pv), _p(out), _stream_ptr(stream)))
    return out


class PruneNet:
    """The network's parameters and Adam state (fp32, device) plus a workspace per batch size."""

    def __init__(self, params, device):
        import torch
        init(device.index or 0)
        _sigs()
        self.dev = device
        self.params = torch.as_tensor(np.asarray(params, np.float32)).to(device).contiguous()
        assert self.params.numel() == PARAMS
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self._ws = {}

    def workspace(self, batch):
        import torch
        if batch not in self._ws:
            n = int(_sigs().mt_pnet_workspace_bytes(batch))
            self._ws[batch] = torch.empty(n, dtype=torch.uint8, device=self.dev)
        return self._ws[batch]

    def forward(self, pix, qidx, bd, idx=None, stream=None):
        """pix: uint16 cuda (N, 4096); qidx: uint8 cuda (N,); idx: optional int32 cuda (B,)
        selecting the batch.  Returns logits float32 cuda (B, 863)."""
        import torch
        B = pix.shape[0] if idx is None else idx.shape[0]
        logits = torch.empty((B, OUTPUTS), dtype=torch.float32, device=self.dev)
        _check(_sigs().mt_pnet_forward(_p(self.params), _p(pix), _p(idx), _p(qidx), bd, B,
                                       _p(self.workspace(B)), _p(logits), _stream_ptr(stream)))
        return logits

    def backward(self, targets, mask, stream=None):
        """Gradients of the masked loss of the last forward (same batch size): softmax
        cross-entropy of each node's partition type, binary cross-entropy of each mode;
        targets / mask (B, 98).  Returns the loss as a cuda scalar."""
        import torch
        B = targets.shape[0]
        ns = float(mask[:, :SPLIT_NODES].sum().item())
        norm_split = 1.0 / max(ns, 1.0)
        norm_mode = 1.0 / (B * (LABELS - SPLIT_NODES))
        terms = torch.empty((B, LABELS), dtype=torch.float32, device=self.dev)
        _check(_sigs().mt_pnet_backward(_p(self.params), _p(targets), _p(mask), B, norm_split, norm_mode,
                                        _p(self.workspace(B)), _p(self.grads), _p(terms), _stream_ptr(stream)))
        return terms.sum()

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8, stream=None):
        self.step += 1
        _check(_sigs().mt_pnet_adam(_p(self.params), _p(self.grads), _p(self.m), _p(self.v), self.step,
                                    lr, b1, b2, eps, _stream_ptr(stream)))

    def masks(self, logits, topk, lo, hi, cover=0.0, stream=None):
        """Prune masks; topk: modes kept, one value or 5 per node level (64x64 .. 4x4); lo /
        hi: split thresholds, one value or 4 per node level (64x64 .. 8x8); cover: the
        estimated coverage of the kept modes below which a node keeps all modes."""
        import torch
        B = logits.shape[0]
        tk = (ctypes.c_int * 5)(*([topk] * 5 if np.isscalar(topk) else topk))
        lo4 = (ctypes.c_float * 4)(*([lo] * 4 if np.isscalar(lo) else lo))
        hi4 = (ctypes.c_float * 4)(*([hi] * 4 if np.isscalar(hi) else hi))
        out = torch.empty((B, R.NODES), dtype=torch.int16, device=self.dev)
        _check(_sigs().mt_pnet_prune_masks(_p(logits), B, tk, lo4, hi4, float(cover), _p(out),
                                           _stream_ptr(stream)))
        return out


def labels_from_nodes(nodes, width, height):
    """(targets, mask) float32 (nsb, 98) from one frame's RDO node records (numpy,
    NODE_DTYPE (nsb, 341)): column n < 85 the AV1 partition type of node n, 85 + j whether
    mode j is used."""
    nsb = nodes.shape[0]
    C = (width + 63) // 64
    t = np.zeros((nsb, LABELS), np.float32)
    m = np.zeros((nsb, LABELS), np.float32)
    split = nodes["split"]
    for n in range(SPLIT_NODES):
        L, ix, iy, s = R.node_geometry(n)
        sb = np.arange(nsb)
        x = (sb % C) * 64 + ix * s
        y = (sb // C) * 64 + iy * s
        inside = (x + s <= width) & (y + s <= height)
        coded = split[:, n] != R.UNCODED
        m[:, n] = (inside & coded).astype(np.float32)
        sp = split[:, n]
        t[:, n] = np.where(sp == R.SPLIT, P_SPLIT, np.where(sp == R.HORZ, P_HORZ,
                                                              np.where(sp == R.VERT, P_VERT, P_NONE)))

    def base(mode_id):
        mode_id = mode_id.astype(np.int64)
        return np.where(mode_id <= 12, mode_id, np.where(mode_id <= 60, 1 + (mode_id - 13) // 6, 0))
    leaf = split == R.NONE
    b0 = base(nodes["mode"])
    has_rect = "rmode0" in nodes.dtype.names  # records with the PARTITION_HORZ / VERT halves' modes
    if has_rect:
        rect = (split == R.HORZ) | (split == R.VERT)
        r0, r1 = base(nodes["rmode0"]), base(nodes["rmode1"])
    for k in range(13):
        used = leaf & (b0 == k)
        if has_rect:
            used |= rect & ((r0 == k) | (r1 == k))
        t[:, SPLIT_NODES + k] = used.any(axis=1).astype(np.float32)
    m[:, SPLIT_NODES:] = 1.0
    return t, m


def bd_rate(rate_a, psnr_a, rate_b, psnr_b):
    """Bjontegaard delta rate (%) of b against a: cubic fits of log10(rate) over PSNR,
    integrated over the overlapping PSNR interval."""
    la, lb = np.log10(np.asarray(rate_a, float)), np.log10(np.asarray(rate_b, float))
    pa, pb = np.asarray(psnr_a, float), np.asarray(psnr_b, float)
    fa, fb = np.polyfit(pa, la, 3), np.polyfit(pb, lb, 3)
    lo, hi = max(pa.min(), pb.min()), min(pa.max(), pb.max())
    ia, ib = np.polyint(fa), np.polyint(fb)
    avg = ((np.polyval(ib, hi) - np.polyval(ib, lo)) - (np.polyval(ia, hi) - np.polyval(ia, lo))) / (hi - lo)
    return (10.0 ** avg - 1.0) * 100.0


# ---- data, training, pruned search -------------------------------------------------------
QUALITY_QIDX = {30: 140, 50: 91, 70: 53, 90: 17}


def make_dataset(dev, seeds, width, height, bd, qidxs, flags, images=None):
    """Labels of the exact RDO (full search) on seeded synthetic frames: frame i is coded at
    qidxs[i % len(qidxs)].  Returns device tensors (pix (N, 4096) int16, qidx (N,) uint8,
    targets, masks (N, 98) float32) over every superblock."""
    import torch
    from . import workload as W
    frames = []
    if images is None:
        images = W.synth_images([(seed, width, height, bd) for seed in seeds])
    for i, img in enumerate(images):
        q = int(qidxs[i % len(qidxs)])
        frames.append({"src": torch.from_numpy(img.view(np.int16)).to(dev), "bd": bd, "qidx": q,
                       "qm_level": _lib.qm_level_law(q, 0)})
    out = R.intra_rdo_batch(frames, flags)
    pix, qv, ts, ms = [], [], [], []
    for fr, (_, nodes, _) in zip(frames, out):
        t, m = labels_from_nodes(R.nodes_to_np(nodes), width, height)
        pix.append(extract_superblocks(fr["src"]))
        qv.append(torch.full((t.shape[0],), fr["qidx"], dtype=torch.uint8, device=dev))
        ts.append(torch.from_numpy(t).to(dev))
        ms.append(torch.from_numpy(m).to(dev))
    return torch.cat(pix), torch.cat(qv), torch.cat(ts), torch.cat(ms)


def train(net, data, bd, epochs=12, batch=256, lr=2e-3, seed=0, log=None):
    """Adam on minibatches of the dataset, on the GPU (forward, backward and update are the
    library's kernels; torch only gathers the minibatch).  Returns the loss per epoch."""
    import torch
    pix, qidx, t, m = data
    n = pix.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(torch.int32).to(pix.device)
        tot, cnt = 0.0, 0
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch].contiguous()
            net.forward(pix, qidx, bd, idx=idx)
            il = idx.long()
            loss = net.backward(t[il].contiguous(), m[il].contiguous())
            net.adam(lr)
            tot += float(loss.item())
            cnt += 1
        hist.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4f}")
    return hist


def prune_masks_batch(net, srcs, bd, qidx, topk, lo, hi, cover=0.0):
    """The network's masks for frames of one bit depth and qindex with one forward over all
    their superblocks: a list of (nsb_i, 341) int16 cuda tensors."""
    import torch
    sbs = [extract_superblocks(s) for s in srcs]
    allp = torch.cat(sbs)
    q = torch.full((allp.shape[0],), qidx, dtype=torch.uint8, device=allp.device)
    masks = net.masks(net.forward(allp, q, bd), topk, lo, hi, cover)
    out, o = [], 0
    for s in sbs:
        out.append(masks[o:o + s.shape[0]])
        o += s.shape[0]
    return out


def prune_masks(net, src, bd, qidx, topk, lo, hi, cover=0.0):
    """The network's masks for one frame: (nsb, 341) int16 cuda tensor."""
    import torch
    sbs = extract_superblocks(src)
    q = torch.full((sbs.shape[0],), qidx, dtype=torch.uint8, device=src.device)
    logits = net.forward(sbs, q, bd)
    return net.masks(logits, topk, lo, hi, cover)


def init_params(seed):
    """He-normal weights and zero biases of the network (fp32, flat, in the layer order of
    the parameter layout)."""
    rng = np.random.default_rng(seed)
    shapes = [(9, 16), (16,), (144, 32), (32,), (288, 64), (64,), (257, 32), (32,), (32, OUTPUTS), (OUTPUTS,)]
    parts = []
    for s in shapes:
        if len(s) == 2:
            parts.append((rng.standard_normal(s) * np.sqrt(2.0 / s[0])).astype(np.float32).ravel())
        else:
            parts.append(np.zeros(s, np.float32))
    p = np.concatenate(parts)
    assert p.size == PARAMS
    return p
