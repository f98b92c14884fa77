"""CPU oracle of the learned-pruning network (maroontree_amd/csrc/mt_nn.hip).

TEST INFRASTRUCTURE ONLY: tests/ and bench.py use it as the checker of the HIP forward,
backward and Adam kernels; the library never calls it.

It restates the network in numpy: conv 3x3 s2 (1->16) -> conv 3x3 s2 (16->32) -> conv 3x3 s2
(32->64), ReLU each, 4x4 average pool ++ qindex/255 -> fc 257->32 ReLU -> fc 32->863 (10
partition-type logits per quadtree node 0..84, then 13 mode scores), with the
GPU's numerics when `bf16=True`: every GEMM operand rounded to bf16 (round to nearest even),
products accumulated in fp32, biases / ReLU / pooling / losses in fp32 with their outputs
rounded to bf16 where the GPU stores bf16.  With `bf16=False` and float64 arrays it is the
exact network, which the finite-difference test pins.

The network has no counterpart in the reference (it replaces the reference's hand-tuned
pruning heuristics, src/coder/partition_search.rs:135-289); parity is with the HIP kernels.
"""
import numpy as np

C1, C2, C3, F, H = 16, 32, 64, 257, 32
NODES, PARTS, MODES = 85, 10, 13
OUT = NODES * PARTS + MODES   # logits
LABELS = NODES + MODES        # label columns: each node's partition type, then the modes
SHAPES = [("w1", (9, C1)), ("b1", (C1,)), ("w2", (9 * C1, C2)), ("b2", (C2,)),
          ("w3", (9 * C2, C3)), ("b3", (C3,)), ("w4", (F, H)), ("b4", (H,)),
          ("w5", (H, OUT)), ("b5", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([d[name].reshape(-1) for name, _ in SHAPES])


def init_params(seed):
    """He-normal weights, zero biases (fp32)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("w"):
            d[name] = (rng.standard_normal(shape) * np.sqrt(2.0 / shape[0])).astype(np.float32)
        else:
            d[name] = np.zeros(shape, np.float32)
    return pack(d)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


class Net:
    def __init__(self, bf16_numerics=True, dtype=np.float32):
        self.q = bf16 if bf16_numerics else (lambda a: a)
        self.dt = dtype

    def _im2col(self, a):  # a [B, H, W, C] -> [B * OH * OW, 9 * C], k = tap * C + c
        B, Hh, _, C = a.shape
        OH = Hh // 2
        pad = np.zeros((B, Hh + 2, Hh + 2, C), a.dtype)
        pad[:, 1:-1, 1:-1] = a
        cols = []
        for ky in range(3):
            for kx in range(3):
                cols.append(pad[:, ky:ky + 2 * OH:2, kx:kx + 2 * OH:2, :])
        return np.concatenate(cols, axis=3).reshape(B * OH * OH, 9 * C)

    def _col2im(self, dcol, B, Hh, C):
        OH = Hh // 2
        d = dcol.reshape(B, OH, OH, 9, C)
        pad = np.zeros((B, Hh + 2, Hh + 2, C), dcol.dtype)
        for ky in range(3):
            for kx in range(3):
                pad[:, ky:ky + 2 * OH:2, kx:kx + 2 * OH:2, :] += d[:, :, :, ky * 3 + kx, :]
        return pad[:, 1:-1, 1:-1]

    def forward(self, params, pix, qidx, bd):
        """pix [B, 64, 64] uint16, qidx [B] -> logits [B, 863]; keeps the cache."""
        q, dt = self.q, self.dt
        P = {k: v.astype(dt) for k, v in unpack(np.asarray(params, dtype=dt)).items()}
        Wq = {k: q(v) for k, v in P.items() if k.startswith("w")}
        B = pix.shape[0]
        scale = dt(1.0) / dt((1 << bd) - 1)
        x0 = q((pix.astype(dt) * scale - dt(0.5)))[..., None]
        c = {"P": P, "Wq": Wq, "B": B}
        a = x0
        for i, (Hh, Cin, Cout) in enumerate(((64, 1, C1), (32, C1, C2), (16, C2, C3)), start=1):
            col = self._im2col(a)
            z = col @ Wq[f"w{i}"] + P[f"b{i}"]
            a = q(np.maximum(z, 0)).reshape(B, Hh // 2, Hh // 2, Cout)
            c[f"col{i}"], c[f"a{i}"] = col, a
        pooled = a.reshape(B, 2, 4, 2, 4, C3).sum(axis=(2, 4)) * dt(1.0 / 16.0)
        feat = np.concatenate([pooled.reshape(B, 4 * C3), (qidx.astype(dt) * dt(1.0 / 255.0))[:, None]], axis=1)
        feat = q(feat)
        z4 = feat @ Wq["w4"] + P["b4"]
        a4 = q(np.maximum(z4, 0))
        z5 = a4 @ Wq["w5"] + P["b5"]
        c.update(feat=feat, a4=a4, z5=z5)
        self.c = c
        return z5

    def backward(self, targets, mask, norm_split, norm_mode):
        """Gradients (flat, like the parameters) of the masked loss, and the loss value:
        targets / mask [B, 98]; per node n < 85 the softmax cross-entropy of its 10 logits
        with the partition type targets[n], per mode j the binary cross-entropy with logits
        of its score with targets[85 + j]."""
        q, dt, c = self.q, self.dt, self.c
        B, Wq = c["B"], c["Wq"]
        z = c["z5"]
        zp = z[:, :NODES * PARTS].reshape(B, NODES, PARTS)
        zm = z[:, NODES * PARTS:]
        tp = targets[:, :NODES].astype(np.int64)
        tm = targets[:, NODES:]
        wp = mask[:, :NODES] * dt(norm_split)
        wm = mask[:, NODES:] * dt(norm_mode)
        mx = zp.max(axis=2, keepdims=True)
        e = np.exp(zp - mx)
        se = e.sum(axis=2, keepdims=True)
        onehot = (np.arange(PARTS)[None, None, :] == tp[:, :, None]).astype(dt)
        dzp = wp[:, :, None] * (e / se - onehot)
        pick = np.take_along_axis(zp, tp[:, :, None], axis=2)[:, :, 0]
        loss = float((wp * (np.log(se[:, :, 0]) + mx[:, :, 0] - pick)).sum())
        p = 1.0 / (1.0 + np.exp(-zm))
        dzm = wm * (p - tm)
        loss += float((wm * (np.maximum(zm, 0) - zm * tm + np.log1p(np.exp(-np.abs(zm))))).sum())
        dz5 = np.concatenate([dzp.reshape(B, NODES * PARTS), dzm], axis=1).astype(dt)
        g = {}
        dz5q = q(dz5)
        g["w5"] = q(c["a4"]).T @ dz5q
        g["b5"] = dz5.sum(0)
        da4 = dz5q @ Wq["w5"].T
        dz4 = np.where(c["a4"] > 0, da4, 0).astype(dt)
        dz4q = q(dz4)
        g["w4"] = c["feat"].T @ dz4q
        g["b4"] = dz4.sum(0)
        dfeat = dz4q @ Wq["w4"].T
        dpool = dfeat[:, :256].reshape(B, 2, 2, C3) * dt(1.0 / 16.0)
        da3 = np.repeat(np.repeat(dpool, 4, axis=1), 4, axis=2)
        dz = np.where(c["a3"] > 0, da3, 0).astype(dt)
        for i, (Hin, Cin) in ((3, (16, C2)), (2, (32, C1)), (1, (64, 1))):
            Cout = dz.shape[-1]
            dzf = dz.reshape(-1, Cout)
            dzq = q(dzf)
            g[f"w{i}"] = c[f"col{i}"].T @ dzq
            g[f"b{i}"] = dzf.sum(0)
            if i == 1:
                break
            dcol = dzq @ Wq[f"w{i}"].T
            dprev = self._col2im(dcol, B, Hin, Cin)
            dz = np.where(c[f"a{i - 1}"] > 0, dprev, 0).astype(dt)
        return pack(g), loss


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """The Adam step of mt_pnet_adam in float32."""
    f = np.float32
    m[:] = f(b1) * m + f(1 - b1) * g
    v[:] = f(b2) * v + f(1 - b2) * g * g
    c1 = f(1 - b1 ** step)
    c2 = f(1 - b2 ** step)
    p -= f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return p
