"""The oracle's whole-candidate evaluation (oracle/oracle_eval.cpp, the reference's candidate
loop body src/coder/partition_search.rs:471-520) equals the stage-by-stage composition of
the individually pinned oracle kernels written out in Python, on every reference size,
mode family, bit depth and clipped source planes; with the trellis (MT_CAND_TRELLIS,
trellis_optimize_ctx between quantization and the inverse) on half of the candidates."""
import numpy as np

import oracle_lib as O
from evalgen import PH, PW, make_cands, make_planes, oracle_candidate


def test_eval_batch_equals_stagewise_composition():
    for bd in (8, 10, 12):
        rng = np.random.default_rng(40 + bd)
        src, recon = make_planes(rng, bd)
        cands, _ = make_cands(rng, 260, bd)
        tr = rng.random(len(cands)) < 0.5
        cands["rflags"] |= (tr * 2 + tr * (rng.integers(0, 3, len(cands)) << 2)).astype(np.uint8)
        planes = [(src, PW, PH, PW, PH), (recon, PW, PH, PW, PH), (src, PW, PH, 301, 233)]
        got = O.eval_batch(cands, planes, nthreads=4)
        for i in range(len(cands)):
            want = oracle_candidate(cands[i], [src, recon, src], [(None,) + p[1:] for p in planes])
            g = got[i]
            assert g["status"] == 0
            assert (int(g["sse"]), int(g["dist"]), int(g["satd"]), int(g["nnz"])) == \
                (want[0], want[1], want[2], want[5]), i
            assert np.float32(g["bits"]).tobytes() == want[3].tobytes(), i
            assert np.float32(g["cost"]).tobytes() == want[4].tobytes(), i
