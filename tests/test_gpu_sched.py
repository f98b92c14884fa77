"""The chain's per-size launch schedules (mt_tx_schedule_set / _get) through the C ABI:
set/get round trips and refusals, and the chain's every output (result, levels, targets,
reconstructed residuals) bit for bit the same under every schedule the autotuner tries,
and equal to the oracle's SSE and checksum.  The default table is restored afterwards."""
import numpy as np
import pytest
import torch

import maroontree_amd as mt
from maroontree_amd import autotune as AT
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_schedule(dev):
    saved = {s: mt.tx_schedule_get(s) for s in mt.TX_SIZES_REF}
    yield
    for s, (ts, wg) in saved.items():
        mt.tx_schedule_set(s, ts, wg)


def _run(dev, bt, pt, n, npool):
    res = torch.full((n * 16,), 0xA5, dtype=torch.uint8, device=dev)
    lev = torch.full((npool,), 0x7F7F7F7F, dtype=torch.int32, device=dev)
    tgt = torch.full((npool,), float("nan"), dtype=torch.float32, device=dev)
    rec = torch.full((npool,), -12345, dtype=torch.int32, device=dev)
    mt.tx_chain(bt, pt, res, lev, tgt, rec)
    torch.cuda.synchronize()
    return res.cpu().numpy(), lev.cpu().numpy(), tgt.cpu().numpy().view(np.uint32), rec.cpu().numpy()


def test_schedule_set_get_and_refusals(dev, restore_schedule):
    mt.init(0)
    assert mt.tx_schedule_get(0) == (0, 8)
    assert mt.tx_schedule_get(4) == (0, 16)  # 64x64: one wave per block, 16 per CU
    mt.tx_schedule_set(3, 1, 2)
    assert mt.tx_schedule_get(3) == (1, 2)
    mt.tx_schedule_set(0, -1, 32)
    assert mt.tx_schedule_get(0) == (-1, 32)
    for bad in ((0, 2, 8), (0, -2, 8), (0, 0, 0), (0, 0, 65), (4, 1, 8), (11, 0, 8), (19, 0, 8)):
        with pytest.raises(mt.MtError):
            mt.tx_schedule_set(*bad)
    assert mt.tx_schedule_get(0) == (-1, 32)  # a refused call changes nothing


def test_chain_outputs_do_not_depend_on_the_schedule(dev, restore_schedule):
    seed, n = 0x5C4ED, 40000
    blocks, pool = mt.gen_tx_blocks(seed, n)
    print(f"schedule parity: seed={seed:#x} n={n} pool={len(pool)}", flush=True)
    bt = torch.from_numpy(blocks.view(np.uint8).copy()).to(dev)
    pt = torch.from_numpy(pool).to(dev)
    ref = _run(dev, bt, pt, n, len(pool))
    got = ref[0].view(mt.RES_DTYPE)
    bdi = (blocks["bd_qm"] >> 4) & 3
    sse, h = O.chain_batch(blocks["tx_size"], blocks["tx_type"], blocks["qidx"],
                           np.where(bdi == 1, 10, np.where(bdi == 2, 12, 8)), blocks["bd_qm"] & 15,
                           blocks["offset"].astype(np.uint64), pool)
    assert (got["sse"] == sse).all() and (got["hash"] == h).all()
    # every candidate schedule, applied to all sizes at once (64x64 takes its own list)
    for k, (ts, wg) in enumerate(AT.TILE_CANDIDATES):
        for s in mt.TX_SIZES_REF:
            if s == 4:
                mt.tx_schedule_set(s, *AT.WAVE_CANDIDATES[k % len(AT.WAVE_CANDIDATES)])
            else:
                mt.tx_schedule_set(s, ts, wg)
        print(f"schedule parity: tile_shift={ts} wg_per_cu={wg}", flush=True)
        out = _run(dev, bt, pt, n, len(pool))
        for name, x, y in zip(("res", "levels", "targets", "recon"), ref, out):
            bad = np.nonzero(x != y)[0]
            assert len(bad) == 0, f"({ts}, {wg}): {name} differs at {len(bad)} places, first {bad[:4]}"


def test_autotuner_installs_a_candidate_per_size(dev, restore_schedule):
    table = AT.tune(0, n=1 << 15, reps=2)
    assert set(table["sizes"]) == {mt.TX_NAMES[s] for s in mt.TX_SIZES_REF}
    for name, rec in table["sizes"].items():
        assert (rec["tile_shift"], rec["wg_per_cu"]) in AT.candidates(rec["tx_size"])
        assert mt.tx_schedule_get(rec["tx_size"]) == (rec["tile_shift"], rec["wg_per_cu"])
        assert rec["ms"] == min(rec["tried"].values())
