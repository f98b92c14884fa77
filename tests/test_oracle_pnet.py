"""Pins of the network oracle (oracle/pnet_oracle.py), the checker of the HIP forward /
backward / Adam kernels: gradients against central finite differences in float64, the
im2col / col2im adjoint pair, bf16 rounding known answers, the label extraction and the
Bjontegaard delta rate on closed-form curves."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pnet_oracle as PO  # noqa: E402

from maroontree_amd import prune as PR  # noqa: E402
from maroontree_amd import rdo as R  # noqa: E402


def test_parameter_count_matches_library():
    assert PO.NPARAMS == PR.PARAMS == 60031
    assert PO.OUT == PR.OUTPUTS == 85 * 10 + 13 and PO.LABELS == PR.LABELS == 98


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -2.5, 0.0,
                  65504.0, 1e-40], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0, 65536.0, 1e-40], np.float32)
    got = PO.bf16(x)
    assert np.array_equal(got[:7], want[:7])
    assert got[7] == PO.bf16(np.array([got[7]]))[0]  # idempotent on a subnormal


def test_im2col_col2im_are_adjoint():
    rng = np.random.default_rng(1)
    net = PO.Net(bf16_numerics=False, dtype=np.float64)
    a = rng.standard_normal((2, 16, 16, 3))
    d = rng.standard_normal((2 * 8 * 8, 27))
    lhs = (net._im2col(a) * d).sum()
    rhs = (a * net._col2im(d, 2, 16, 3)).sum()
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    net = PO.Net(bf16_numerics=False, dtype=np.float64)
    p = PO.init_params(3).astype(np.float64) * 0.5
    d = PO.unpack(p)
    for k in d:
        if k.startswith("b"):
            d[k][:] = rng.standard_normal(d[k].shape) * 0.05
    p = PO.pack(d)
    B = 3
    pix = rng.integers(0, 1024, (B, 64, 64)).astype(np.uint16)
    qidx = rng.integers(0, 256, B).astype(np.uint8)
    t = (rng.random((B, 98)) < 0.4).astype(np.float64)
    t[:, :85] = rng.integers(0, 10, (B, 85))  # partition types of the nodes
    m = (rng.random((B, 98)) < 0.8).astype(np.float64)
    net.forward(p, pix, qidx, 10)
    g, loss = net.backward(t, m, 0.1, 0.2)

    def f(pp):
        net2 = PO.Net(bf16_numerics=False, dtype=np.float64)
        net2.forward(pp, pix, qidx, 10)
        return net2.backward(t, m, 0.1, 0.2)[1]
    picks = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False) for a, b in (
        (0, 144), (160, 4768), (4800, 23232), (23296, 31520), (31552, 59168), (59168, 60031))])
    h = 1e-6
    for i in picks:
        e = np.zeros_like(p)
        e[i] = h
        fd = (f(p + e) - f(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_labels_from_rdo_nodes():
    # one 64x64 SB split into four 32x32 leaves, frame 64x40 (the lower 32x32 pair crosses)
    nodes = np.zeros((1, R.NODES), R.NODE_DTYPE)
    nodes["split"] = 2
    nodes["split"][0, 0] = 1
    nodes["split"][0, 1:5] = [0, 0, 1, 1]
    nodes["mode"][0, 1:5] = [9, 13 + 6 * 2 + 1, 0, 0]  # SMOOTH, D45 delta -2
    t, m = PR.labels_from_nodes(nodes, 64, 40)
    assert m[0, 0] == 0  # the root crosses the frame edge: forced
    assert list(m[0, 1:5]) == [1, 1, 0, 0] and list(t[0, 1:5]) == [PR.P_NONE, PR.P_NONE, PR.P_SPLIT, PR.P_SPLIT]
    assert t[0, 85 + 9] == 1 and t[0, 85 + 3] == 1 and t[0, 85:].sum() == 2


def test_bd_rate_of_a_uniform_rate_scaling():
    """A codec needing exactly 0.9x the rate at every PSNR is -10 %."""
    psnr = np.array([30.0, 34.0, 38.0, 42.0])
    rate = 10 ** (psnr / 20.0)
    assert abs(PR.bd_rate(rate, psnr, 0.9 * rate, psnr) + 10.0) < 1e-9
    assert abs(PR.bd_rate(rate, psnr, rate, psnr)) < 1e-12
