"""Input generators of the reference's own tests, restated (test infrastructure)."""
import numpy as np

M32 = 0xFFFFFFFF


def forward_parity_inputs(n):
    """forward_parity_inputs (reference src/dct.rs:3381-3417): 7 named patterns of length n."""
    seed = 0x9E3779B9
    rnd = []
    for _ in range(n):
        seed = (seed * 1664525 + 1013904223) & M32
        rnd.append(((seed >> 16) & 0x3FF) - 512)
    rnd12 = []
    for _ in range(n):
        seed = (seed * 1664525 + 1013904223) & M32
        rnd12.append((seed % 8191) - 4095)
    imp = [0] * n
    imp[0] = 4095
    imp[n - 1] = -4095
    return [
        ("zeros", np.zeros(n, np.int32)),
        ("constant", np.full(n, 1023, np.int32)),
        ("impulses", np.array(imp, np.int32)),
        ("ramp", np.array([((i * 73 + i * i * 19 + 11) % 511) - 255 for i in range(n)], np.int32)),
        ("alternating-12b", np.array([4095 if i % 2 == 0 else -4095 for i in range(n)], np.int32)),
        ("random", np.array(rnd, np.int32)),
        ("random-12b", np.array(rnd12, np.int32)),
    ]


# QUANT_PAIRS of the selected-vs-scalar test (reference src/dct.rs:3424-3429)
QUANT_PAIRS = [(65536, 65536), (65536, 46341), (32768, 32768), (17341, 9017)]


def itx_levels(n, seed):
    """levels::<N>(seed) (reference src/idct.rs:734-745)."""
    state = (seed ^ n) & M32
    out = np.zeros(n, np.int32)
    for i in range(n):
        state = (state * 1664525 + 1013904223) & M32
        level = ((state >> 24) & 63) - 31
        out[i] = 0 if i % 9 == 0 else level
    out[0] = 19
    return out


def pat(n):
    """pat(n) (reference src/dct.rs:3377)."""
    return np.array([((i * 41 + 7) % 113) - 56 for i in range(n)], np.int32)
