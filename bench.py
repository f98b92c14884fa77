#!/usr/bin/env python3
"""Batched transform/quant microbench of the intra candidate chain on one MI355X, plus the
full candidate evaluation on a synthetic 4K image.

Headline (`value`): one step = one mt_tx_chain call over the whole synthetic batch (2^20
int16 residual blocks, 13 reference TX sizes, every size's reference types, Laplace(12)
residuals, 8/10-bit, flat or law-chosen QM): fwd -> quant(+QM) -> dequant -> inverse -> SSE
per block, inputs already resident in HBM.

`eval`: one step = one mt_eval_candidates call over every (square block, luma mode)
candidate of every 64x64 superblock of a seeded 3840x2160 10-bit image at qindex 91:
prediction -> residual -> transform chain -> distortion -> rate -> rd_cost per candidate.

Prints one JSON line; the CPU baselines are the oracle (a scalar restatement of the
reference's kernels) on bounded samples of the same work on the host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEED = 0x6D61726F6F6E
# A one-GPU share of the host this runs on is 16 cores: the CPU baselines use every core the
# process may run on, up to that share.
CPU_SHARE = 16


def host_cpu():
    """(threads used by the CPU baselines, nproc, CPU model name)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return min(CPU_SHARE, avail), os.cpu_count() or 1, model


def cpu_info(threads):
    t, nproc, model = host_cpu()
    return {"cores": threads, "nproc": nproc, "cpu_model": model,
            "oracle_build": "g++ -O3 -march=x86-64-v3 -fwrapv -ffp-contract=off (scalar restatement)"}
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
VALU_INT32_PEAK_TOPS = 39.3    # 256 CU x 64 lanes x 2.4 GHz


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--blocks", type=int, default=1 << 20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="wall-time budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-slow", action="store_true",
                    help="skip the chain-with-trellis leg and the Slow-mode RDO")
    ap.add_argument("--no-eval", action="store_true", help="skip the candidate-evaluation leg")
    ap.add_argument("--eval-steps", type=int, default=10)
    ap.add_argument("--eval-only", action="store_true",
                    help="profiling aid: run only the candidate-evaluation leg and print its object")
    ap.add_argument("--no-rdo", action="store_true", help="skip the whole-frame RDO leg")
    ap.add_argument("--rdo-steps", type=int, default=3)
    ap.add_argument("--rdo-skip-luma", action="store_true",
                    help="time only the 4:2:0 RDO (not the luma-only RDO beside it)")
    ap.add_argument("--rdo-only", action="store_true",
                    help="profiling aid: run only the whole-frame RDO leg and print its object")
    ap.add_argument("--no-deblock", action="store_true", help="skip the deblocking-filter and CDEF legs")
    ap.add_argument("--deblock-only", action="store_true",
                    help="run only the deblocking-filter leg and print its object as the JSON line")
    ap.add_argument("--deblock-steps", type=int, default=50)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-pruning leg")
    ap.add_argument("--learned-only", action="store_true",
                    help="run only the learned-pruning leg and print its object")
    ap.add_argument("--learned-train", type=int, default=200, help="training frames (1920x1080)")
    ap.add_argument("--learned-eval", type=int, default=50, help="held-out frames per quality")
    ap.add_argument("--learned-epochs", type=int, default=40)
    ap.add_argument("--learned-model-bits", action="store_true",
                    help="skip coding the held-out frames into AV1 bytes (BD-rate on the RDO's rate "
                         "model only)")
    ap.add_argument("--learned-luma-only", action="store_true",
                    help="run the learned-pruning comparison on the luma RDO alone")
    ap.add_argument("--no-stream", action="store_true", help="skip the uneven-stream leg")
    ap.add_argument("--stream-only", action="store_true",
                    help="run only the uneven-stream leg (pruning off) and print its object")
    ap.add_argument("--stream-images", type=int, default=256)
    ap.add_argument("--stream-max-side", type=int, default=8192)
    ap.add_argument("--topk", default="8,10,12,13,13",
                    help="intra modes kept per superblock, per node level 64,32,16,8,4")
    ap.add_argument("--mode-cover", type=float, default=0.3,
                    help="keep all modes where the kept ones cover the SB's modes with a lower "
                         "estimated probability")
    ap.add_argument("--lo", default="0,0.03,0.02,0",
                    help="NO_SPLIT below this split probability, per node level 64,32,16,8")
    ap.add_argument("--hi", default="1,1,1,1",
                    help="FORCE_SPLIT above this split probability, per node level 64,32,16,8")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the chain's saved (or default) launch schedules instead of "
                         "autotuning them before the warm-up")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (counter calibration)")
    args = ap.parse_args()
    args.topk = [int(v) for v in str(args.topk).split(",")] * (1 if "," in str(args.topk) else 5)
    args.lo = [float(v) for v in str(args.lo).split(",")] * (1 if "," in str(args.lo) else 4)
    args.hi = [float(v) for v in str(args.hi).split(",")] * (1 if "," in str(args.hi) else 4)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the chain has no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import maroontree_amd as mt

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.calib_copy:
        # a plain copy of known size: what FETCH_SIZE / WRITE_SIZE read for it scales the
        # chain's counter values to bytes
        src = torch.ones(args.calib_copy << 18, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": args.calib_copy << 20, "copies": 5}))
        return
    mt.init(0)
    if args.eval_only:
        print(json.dumps(bench_eval(args, dev, torch, mt)), flush=True)
        return
    if args.rdo_only:
        print(json.dumps(bench_rdo(args, dev, torch, mt)), flush=True)
        return
    if args.learned_only:
        print(json.dumps(bench_learned(args, dev, torch, mt)), flush=True)
        return
    if args.deblock_only:
        prep = {}
        r = bench_deblock(args, dev, torch, mt, prep)
        r["cdef"] = bench_cdef(args, dev, torch, mt, prep)
        print(json.dumps(r), flush=True)
        return
    if args.stream_only:
        print(json.dumps(bench_stream(args, dev, torch, mt)), flush=True)
        return

    n = args.blocks
    t0 = time.time()
    blocks, pool = mt.gen_tx_blocks(SEED, n)
    gen_s = time.time() - t0
    npx = len(pool)
    bt = torch.from_numpy(blocks.view(np.uint8).copy()).to(dev)
    pt = torch.from_numpy(pool).to(dev)
    res = torch.zeros(n * 16, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    print(f"[bench] {n} blocks, {npx / 1e6:.1f} Mpx generated in {gen_s:.1f}s; warmup {args.warmup}",
          file=sys.stderr, flush=True)
    # the per-size launch schedules: measured now (outside the timed region) by the
    # autotuner, or the saved / default table
    from maroontree_amd import autotune as AT
    if args.no_autotune:
        sched_src = "saved" if AT.apply() else "default"
        sched = None
    else:
        sched = AT.tune(0, workload=(blocks, pt),
                        log=lambda m: print("[bench] " + m, file=sys.stderr, flush=True))
        sched_src = "autotuned"
        os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
        AT.save(sched, os.path.join(ROOT, "gpurun_out", "tx_schedule.json"))
    sched_now = {mt.TX_NAMES[s]: list(mt.tx_schedule_get(s)) for s in mt.TX_SIZES_REF}

    for _ in range(args.warmup):
        mt.tx_chain(bt, pt, res, stream=stream)
    torch.cuda.synchronize()

    # timed region: K steps, per-launch HIP events on the chain's stream
    mt.profile_begin()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mt.tx_chain(bt, pt, res, stream=stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = mt.profile_end()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    mpx_s = npx / (ms_per_step / 1000.0) / 1e6
    cand_s = n / (ms_per_step / 1000.0)
    print(f"[bench] {ms_per_step:.3f} ms/step, {mpx_s:.0f} Mpx/s", file=sys.stderr, flush=True)

    # dominant kernel: largest summed time over the timed region
    slot = max((k for k in prof if k in mt.TX_NAMES), key=lambda k: prof[k][0])
    k_ms = prof[slot][0] / max(prof[slot][1], 1)
    size_id = [s for s in mt.TX_SIZES_REF if mt.TX_NAMES[s] == slot][0]
    w, h = mt.tx_width(size_id), mt.tx_height(size_id)
    nblk = int((blocks["tx_size"] == size_id).sum())
    alg_bytes = nblk * (2 * w * h + 8 + 16)  # int16 residual in, descriptor, result
    achieved = alg_bytes / (k_ms / 1000.0) / 1e9
    kname = ("void tx64_kernel<0>(KArgs)" if w == 64 else
             f"void tx_tile_kernel<{w}, {h}, 0, {sched_now[slot][0]}>(KArgs, int)")
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", "tx_chain_2p20_pmc.json")
    if os.path.exists(pmc_path) and n == 1 << 20:
        pmc = json.load(open(pmc_path))
        if kname in pmc["kernels"]:
            traffic = pmc["kernels"][kname]["traffic_bytes"]
    compute = None
    if w == 64:
        # f64 work of the reference's separable 64-point forward, one mul and one add per
        # term (no fused multiply-add): 32x64x64 vertical + 32x32x64 horizontal terms
        flops = nblk * 2 * (32 * 64 * 64 + 32 * 32 * 64)
        tf = flops / (k_ms / 1000.0) / 1e12
        # 78.6 TFLOP/s counts a fused multiply-add as 2 flops per lane per instruction; the
        # reference rounds every product and sum, so each flop here is one f64 instruction
        # and the rate these instructions can reach is half of it
        compute = {"bound": "valu_f64", "achieved": round(tf, 3), "peak": 78.6, "unit": "TFLOP/s",
                   "frac": round(tf / 78.6, 4), "flops_per_launch": flops,
                   "peak_unfused": 39.3, "frac_unfused": round(tf / 39.3, 4)}

    line = {
        "metric": "Mpx/s + RD candidates/s on 1 MI355X vs host CPU; BD-rate of pruned vs full search",
        "value": round(mpx_s, 2),
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": "batched transform/quant microbench: 2^20 int16 residual blocks, 13 reference "
                        "TX sizes x reference types, Laplace(12) residual, 8/10-bit, qindex 0-255, "
                        "flat or qm_level_law QM; fwd->quant->dequant->inv->SSE per block",
            "blocks": n,
            "pixels": npx,
            "seed": SEED,
        },
        "candidates_per_s": round(cand_s, 1),
        "kernel_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items() if v[1]},
        "schedule": {"source": sched_src, "tile_shift_wg_per_cu": sched_now,
                     "tune_seconds": sched["seconds"] if sched else None},
        "roofline": {
            "kernel": f"tx chain {slot}",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": alg_bytes,
            "note": "algorithmic bytes = 2*W*H int16 in + 8 B descriptor + 16 B result per block; "
                    "the kernel is VALU bound (see roofline_compute), not HBM bound",
        },
        "roofline_compute": compute,
    }

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        threads = host_cpu()[0]
        bdi = (blocks["bd_qm"] >> 4) & 3
        bd = np.where(bdi == 1, 10, np.where(bdi == 2, 12, 8)).astype(np.uint8)
        px_of = np.array([mt.tx_width(s) * mt.tx_height(s) for s in range(19)], np.int64)
        off = blocks["offset"].astype(np.uint64)

        def run(m):
            O.chain_batch(blocks["tx_size"][:m], blocks["tx_type"][:m], blocks["qidx"][:m], bd[:m],
                          blocks["bd_qm"][:m] & 15, off[:m], pool, nthreads=threads, with_hash=False)
        # the sample: as many leading blocks of the same batch as one pass covers in about
        # cpu_seconds / 5; then one warm-up pass and the best of five timed passes over it
        m = 1 << 14
        while m < n:
            t0 = time.perf_counter()
            run(m)
            if time.perf_counter() - t0 > args.cpu_seconds / 10.0:
                break
            m = min(n, m * 2)
        sample_px = int(px_of[blocks["tx_size"][:m]].sum())
        run(m)
        best = float("inf")
        for _ in range(5):
            t0 = time.perf_counter()
            run(m)
            best = min(best, time.perf_counter() - t0)
        line["cpu_baseline"] = {
            "value": round(sample_px / best / 1e6, 2),
            "unit": "Mpx/s",
            "kind": "port",
            "sample": f"first {m} blocks ({sample_px / 1e6:.1f} Mpx) of the same batch; best of 5 passes "
                      f"after a warm-up ({best:.2f} s), std::thread over blocks",
            **cpu_info(threads),
        }
    if not args.no_slow:
        line["slow_chain"] = bench_slow_chain(args, dev, torch, mt, blocks, pool, bt, pt, res)
    del bt, pt, res
    if not args.no_eval:
        line["eval"] = bench_eval(args, dev, torch, mt)
    if not args.no_rdo:
        line["rdo"] = bench_rdo(args, dev, torch, mt)
    if not args.no_rdo:
        line["plain_512"] = bench_plain512(args, dev, torch, mt)
    if not args.no_deblock:
        prep = {}
        line["deblock"] = bench_deblock(args, dev, torch, mt, prep)
        line["cdef"] = bench_cdef(args, dev, torch, mt, prep)
    net = None
    if not args.no_learned:
        line["learned"], net = bench_learned(args, dev, torch, mt, return_net=True)
    if not args.no_stream:
        line["stream"] = bench_stream(args, dev, torch, mt, net)
    print(json.dumps(line), flush=True)


def bench_slow_chain(args, dev, torch, mt, blocks, pool, bt, pt, res):
    """The headline batch through mt_tx_chain_trellis: the chain with trellis_optimize_ctx
    (src/trellis.rs:606) on every block between quantization and dequantization, the
    reference's Slow speed (its default).  Descriptors: the frame-level context and lambda
    of each block's qindex and bit depth at 4:2:0."""
    from maroontree_amd import workload as W
    n = len(blocks)
    tr = W.chain_trellis_descriptors(blocks, sub=2)
    trt = torch.from_numpy(tr.view(np.uint8).copy()).to(dev)
    npx = len(pool)
    lv = torch.empty(npx, dtype=torch.int32, device=dev)
    tg = torch.empty(npx, dtype=torch.float32, device=dev)
    rec = torch.empty(npx, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        mt.tx_chain_trellis(bt, trt, pt, res, lv, tg, rec, stream=stream)
    torch.cuda.synchronize()
    mt.profile_begin()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mt.tx_chain_trellis(bt, trt, pt, res, lv, tg, rec, stream=stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / args.steps
    prof = mt.profile_end()
    print(f"[bench] chain+trellis: {ms:.3f} ms/step, {npx / ms / 1e3:.0f} Mpx/s", file=sys.stderr, flush=True)
    w = np.array([mt.tx_width(s) for s in range(19)], np.int64)
    h = np.array([mt.tx_height(s) for s in range(19)], np.int64)
    ncoef = int((np.minimum(w, 32) * np.minimum(h, 32))[blocks["tx_size"]].sum())
    t_ms = prof["trellis"][0] / max(prof["trellis"][1], 1)
    # algorithmic bytes of one trellis launch: levels read and written, targets read (4 B per
    # coded coefficient each), one 24 B descriptor per block
    tbytes = ncoef * 12 + n * tr.itemsize
    tach = tbytes / (t_ms / 1000.0) / 1e9
    r = {
        "workload": "the headline batch (2^20 blocks, 13 TX sizes) through fwd -> quant -> trellis_optimize_ctx "
                    "-> dequant -> inv -> SSE; per-block qindex context and Slow-mode lambda at 4:2:0",
        "ms_per_step": round(ms, 4),
        "mpx_per_s": round(npx / (ms / 1000.0) / 1e6, 2),
        "candidates_per_s": round(n / (ms / 1000.0), 1),
        "kernel_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items() if v[1]},
        "trellis_roofline": {"kernel": "trellis_kernel", "bound": "hbm", "achieved": round(tach, 2),
                             "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(tach / HBM_PEAK_GBS, 4),
                             "traffic": None, "bytes_per_launch": tbytes,
                             "note": "one thread per block walks its scan serially; latency bound"},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        threads = host_cpu()[0]
        bdi = (blocks["bd_qm"] >> 4) & 3
        bd = np.where(bdi == 1, 10, np.where(bdi == 2, 12, 8)).astype(np.uint8)
        px_of = w * h
        off = blocks["offset"].astype(np.uint64)

        def run(m):
            O.chain_batch_trellis(blocks["tx_size"][:m], blocks["tx_type"][:m], blocks["qidx"][:m], bd[:m],
                                  blocks["bd_qm"][:m] & 0x4F, off[:m], pool, tr[:m], threads)
        m = 1 << 14
        while m < n:
            t0 = time.perf_counter()
            run(m)
            if time.perf_counter() - t0 > args.cpu_seconds / 10.0:
                break
            m = min(n, m * 2)
        sample_px = int(px_of[blocks["tx_size"][:m]].sum())
        run(m)
        best = float("inf")
        for _ in range(5):
            t0 = time.perf_counter()
            run(m)
            best = min(best, time.perf_counter() - t0)
        r["cpu_baseline"] = {
            "value": round(sample_px / best / 1e6, 2),
            "unit": "Mpx/s",
            "kind": "port",
            "sample": f"first {m} blocks ({sample_px / 1e6:.1f} Mpx) of the same batch; best of 5 passes "
                      f"after a warm-up ({best:.2f} s), std::thread over blocks",
            **cpu_info(threads),
        }
    return r


def bench_eval(args, dev, torch, mt):
    from maroontree_amd import workload as W
    width, height, bd, qidx = 3840, 2160, 10, 91
    t0 = time.time()
    img = W.synth_image(SEED, width, height, bd)
    rec = W.synth_recon(img, SEED, bd)
    cands, pool = W.sb_candidates(width, height, bd, qidx)
    n = len(cands)
    area = (width // 64) * (height // 64) * 64 * 64
    print(f"[bench] eval: {n} candidates over {area / 1e6:.2f} Mpx generated in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    ct = torch.from_numpy(cands.view(np.uint8)).to(dev)
    planes = [(torch.from_numpy(img.view(np.int16)).to(dev),),
              (torch.from_numpy(rec.view(np.int16)).to(dev),)]
    out = torch.empty(n * mt.CANDRES_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    for _ in range(max(args.warmup, 1)):
        mt.eval_candidates(ct, planes, pool, stream=stream, out=out)
    torch.cuda.synchronize()
    mt.profile_begin()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.eval_steps):
        mt.eval_candidates(ct, planes, pool, stream=stream, out=out)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = mt.profile_end()
    ms = (t1 - t0) * 1000.0 / args.eval_steps
    res = out.cpu().numpy().view(mt.CANDRES_DTYPE)
    ok = int((res["status"] == 0).sum())
    print(f"[bench] eval: {ms:.3f} ms/step, {n / ms / 1e3:.1f} M candidates/s", file=sys.stderr, flush=True)
    stages = {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items() if v[1]}
    roofline = eval_roofline(cands, stages, mt)
    r = {
        "workload": "every (square block 4x4..64x64, ND luma mode x 11) candidate of each 64x64 "
                    "superblock of a seeded 3840x2160 10-bit synthetic image (gradients, value "
                    "noise, rectangles), qindex 91, qm_level_law QM, DCT_DCT, weighted SSE; "
                    "prediction from a synthetic reconstruction (source + noise)",
        "candidates": n,
        "pixels_evaluated": pool,
        "image_area_px": area,
        "ms_per_step": round(ms, 4),
        "candidates_per_s": round(n / (ms / 1000.0), 1),
        "image_mpx_per_s": round(area / (ms / 1000.0) / 1e6, 2),
        "eval_mpx_per_s": round(pool / (ms / 1000.0) / 1e6, 2),
        "valid": ok,
        "stage_ms_per_launch": stages,
        "roofline": roofline,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        threads = host_cpu()[0]
        rng = np.random.default_rng(SEED)
        oplanes = [(img, width, height, width, height), (rec, width, height, width, height)]
        # the sample: candidates drawn uniformly from the same set, as many as one pass covers
        # in about cpu_seconds / 5; one warm-up pass, then the best of five
        m = 1 << 15
        while True:
            sel = np.ascontiguousarray(cands[np.sort(rng.permutation(n)[:m])])
            t0 = time.perf_counter()
            O.eval_batch(sel, oplanes, nthreads=threads)
            if m >= n or time.perf_counter() - t0 > args.cpu_seconds / 10.0:
                break
            m = min(n, m * 2)
        px = int((np.array([16, 64, 256, 1024, 4096])[sel["tx_size"]]).sum())
        best = float("inf")
        for _ in range(5):
            t0 = time.perf_counter()
            O.eval_batch(sel, oplanes, nthreads=threads)
            best = min(best, time.perf_counter() - t0)
        r["cpu_baseline"] = {
            "value": round(len(sel) / best, 1),
            "unit": "candidates/s",
            "eval_mpx_per_s": round(px / best / 1e6, 2),
            "kind": "port",
            "sample": f"{len(sel)} candidates drawn uniformly from the same set ({px / 1e6:.1f} Mpx of "
                      f"blocks); best of 5 passes after a warm-up ({best:.2f} s), std::thread over candidates",
            **cpu_info(threads),
        }
    return r


# the kernels of each evaluation stage, by name
STAGE_KERNELS = {
    "distortion": ("dist_classify_kernel", "dist_small_kernel", "dist_kernel("),
    "rate": ("rate_classify_kernel", "rate_kernel(", "rate_wave_kernel"),
    "intra_pred": ("pred_classify_kernel", "pred_small_kernel", "intra_pred_kernel"),
    "rdo_recon": ("sb_recon_kernel", "sb_recon_uv_kernel"),
    "deblock": ("deblock_kernel",),
}


def pmc_traffic(path, stage):
    """HBM bytes per call of a stage's kernels from a committed counter profile, or None."""
    full = os.path.join(ROOT, "profiles", path)
    if not os.path.exists(full):
        return None
    pmc = json.load(open(full))
    tot = 0.0
    for name, k in pmc["kernels"].items():
        if any(p in name for p in STAGE_KERNELS[stage]):
            tot += k["traffic_bytes"]
    return int(tot / pmc["calls_profiled"]) if tot else None


def eval_roofline(cands, stages, mt):
    """Roofline of the evaluation's dominant stage: the bytes each stage must move for the
    candidate set (its inputs from and outputs to HBM) over its measured time per call."""
    w = np.array([4, 8, 16, 32, 64])[cands["tx_size"]].astype(np.int64)
    px = int((w * w).sum())
    ncoef = int((np.minimum(w, 32) ** 2).sum())
    edge = int((2 * (4 * w + 1)).sum())
    n = len(cands)
    alg = {
        # source 2 B + prediction 4 B + reconstructed residual 4 B per sample, descriptor, result
        "distortion": 10 * px + n * (24 + 32),
        # levels 4 B per coded coefficient, descriptor, bits
        "rate": 4 * ncoef + n * (32 + 4),
        # edges (2 w + 2 h + 1 samples), source 2 B, prediction 4 B and residual 2 B per sample
        "intra_pred": edge + 8 * px + n * 16,
    }
    stage = max(alg, key=lambda k: stages.get(k, 0.0))
    ms = stages[stage]
    ach = alg[stage] / (ms / 1000.0) / 1e9
    return {"kernel": f"evaluation stage {stage}", "bound": "hbm", "achieved": round(ach, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic("eval_4k_pmc.json", stage), "bytes_per_launch": alg[stage],
            "note": "per call over the stage's kernels; every stage is bound by VALU issue and "
                    "latency, not HBM (profiles/eval_4k_pmc.json holds the instruction counts)"}


def bench_rdo(args, dev, torch, mt):
    """Exact full intra RDO of a seeded 3840x2160 10-bit 4:2:0 frame at quality 50 (qindex 91):
    one step = one mt_intra_rdo call (luma and chroma wavefront decisions with CfL,
    reconstruction, final accounting).  The luma-only RDO of the same frame is timed too."""
    from maroontree_amd import rdo as R
    from maroontree_amd import workload as W
    width, height, bd, qidx = 3840, 2160, 10, 91
    flags = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4 | R.CFL
    y, u, v = W.synth_yuv(SEED, width, height, bd, 2)
    qm = mt.qm_level_law(qidx, 2)
    lam = R.rdo_lambda(qidx, bd, 2)
    t = lambda a: torch.from_numpy(a.view(np.int16)).to(dev)  # noqa: E731
    nsb = ((width + 63) // 64) * ((height + 63) // 64)
    frame = {"src": t(y), "src_u": t(u), "src_v": t(v), "bd": bd, "qidx": qidx, "qm_level": qm, "lam": lam,
             "chroma": R.CHROMA_420, "recon": torch.empty((height, width), dtype=torch.int16, device=dev),
             "recon_u": torch.empty((height // 2, width // 2), dtype=torch.int16, device=dev),
             "recon_v": torch.empty((height // 2, width // 2), dtype=torch.int16, device=dev),
             "nodes": torch.empty(nsb * R.NODES * R.NODE_DTYPE.itemsize, dtype=torch.uint8, device=dev)}
    luma = {k: frame[k] for k in ("src", "bd", "qidx", "qm_level", "lam", "recon", "nodes")}
    stream = torch.cuda.current_stream()

    def timed(fr, fl, steps, prof=False):
        for _ in range(max(args.warmup, 1)):
            st = R.intra_rdo_batch([fr], fl, stream=stream)[0][2]
        torch.cuda.synchronize()
        if prof:
            mt.profile_begin()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = R.intra_rdo_batch([fr], fl, stream=stream)[0][2]
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        return ms, st, (mt.profile_end() if prof else None)

    ms, st, prof = timed(frame, flags, args.rdo_steps, prof=True)
    nd = R.nodes_to_np(frame["nodes"])
    # the same RDO with its wavefront steps captured into one HIP graph per call
    ms_g, st_g, _ = timed(frame, flags | R.GRAPH, args.rdo_steps)
    print(f"[bench] rdo 4:2:0 as one graph: {ms_g:.2f} ms/frame", file=sys.stderr, flush=True)
    cfl_blocks, chroma_blocks = int((nd["uv_mode"] == 13).sum()), int((nd["uv_mode"] != 255).sum())
    ms_l, st_l, _ = timed(luma, flags & ~R.CFL, args.rdo_steps) if not args.rdo_skip_luma else (None, None, None)
    npx = width * height
    print(f"[bench] rdo 4:2:0: {ms:.2f} ms/frame, {npx / ms / 1e3:.1f} Mpx/s, "
          f"{st['candidates'] / ms / 1e3:.1f} M candidates/s; luma only {ms_l} ms", file=sys.stderr, flush=True)
    r = {
        "workload": "exact full intra RDO of a seeded 3840x2160 10-bit 4:2:0 synthetic frame (RGB recipe -> "
                    "BT.601 YCbCr) at quality 50, qindex 91 (qm_level_law QMs, the reference's chroma qindex "
                    "deltas and decision lambda, weighted luma SSE): 64x64 superblocks on the anti-diagonal "
                    "wavefront, partitions 64..4, 13 luma modes + angle deltas + filter intra (DCT_DCT), "
                    "13 uv modes + CfL per chroma block (decoder-derived chroma transforms)",
        "ms_per_frame": round(ms, 3),
        "mpx_per_s": round(npx / (ms / 1000.0) / 1e6, 2),
        "candidates": st["candidates"],
        "candidates_per_s": round(st["candidates"] / (ms / 1000.0), 1),
        "leaves": st["leaves"],
        "steps": st["steps"],
        "bits": round(st["bits"], 1),
        "bits_uv": round(st["bits_uv"], 1),
        "psnr_y": round(R.psnr(st["sse"], npx, bd), 4),
        "psnr_u": round(R.psnr(st["sse_u"], npx // 4, bd), 4),
        "psnr_v": round(R.psnr(st["sse_v"], npx // 4, bd), 4),
        "chroma_blocks": chroma_blocks,
        "cfl_blocks": cfl_blocks,
        "luma_only": None if ms_l is None else {
            "ms_per_frame": round(ms_l, 3), "mpx_per_s": round(npx / (ms_l / 1000.0) / 1e6, 2),
            "candidates_per_s": round(st_l["candidates"] / (ms_l / 1000.0), 1)},
        "stage_ms_per_frame": {k: round(v[0] / args.rdo_steps, 3) for k, v in prof.items() if v[1]},
        "graph": {"ms_per_frame": round(ms_g, 3), "mpx_per_s": round(npx / (ms_g / 1000.0) / 1e6, 2),
                  "identical": bool(st_g["bits"] == st["bits"] and st_g["sse"] == st["sse"] and
                                    st_g["sse_u"] == st["sse_u"] and st_g["sse_v"] == st["sse_v"]),
                  "note": "MT_RDO_GRAPH: the steps' launches captured into one HIP graph per call, "
                          "capture and instantiation included"},
    }
    # the bitstream: one more call with the level planes, copied to the host and coded there
    from maroontree_amd import av1
    fl = dict(frame, levels=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, nodes_l, st_lv = R.intra_rdo_batch([fl], flags, stream=stream)[0]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lv = [a.cpu().numpy() for a in st_lv["levels"]]
    uvq = R.chroma_laws(qidx, 2)
    f_av1, keep = av1.frame(R.nodes_to_np(nodes_l), lv, width, height, bd, qidx, chroma=R.CHROMA_420, qm_level=qm,
                            uv_qm_level=uvq[2], uv_dc=uvq[0], uv_ac=uvq[1], filter_intra=True, cicp=av1.BT601)
    data = av1.encode_avif(f_av1)
    t2 = time.perf_counter()
    print(f"[bench] rdo 4:2:0 AV1 file: {len(data)} bytes, host coder {1e3 * (t2 - t1):.1f} ms",
          file=sys.stderr, flush=True)
    r["av1"] = {"output_bytes": len(data), "bits_per_px": round(8 * len(data) / npx, 4),
                "rdo_with_levels_ms": round(1e3 * (t1 - t0), 2),
                "host_coder_ms": round(1e3 * (t2 - t1), 2),
                "note": "level planes copied to the host, tokens per superblock on host threads, range "
                        "coder, OBUs and AVIF container; not pipelined with the next frame's RDO"}
    # the reconstruction walk: each leaf reads its source and writes its reconstruction, 2 B per
    # sample each, in all three planes
    rec_ms = prof["rdo_recon"][0] / args.rdo_steps
    rec_bytes = 4 * (npx + 2 * (npx // 4))
    rec_ach = rec_bytes / (rec_ms / 1000.0) / 1e9
    r["roofline"] = {"kernel": "rdo_recon (sb_recon_kernel, sb_recon_uv_kernel)", "bound": "hbm",
                     "achieved": round(rec_ach, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(rec_ach / HBM_PEAK_GBS, 6),
                     "traffic": pmc_traffic("rdo_4k_420_pmc.json", "rdo_recon"), "bytes_per_launch": rec_bytes,
                     "note": "per frame; one wave per superblock walks its leaves in coding order on "
                             "the wavefront's critical path: latency bound"}

    def cpu_crop(fl):
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        threads = host_cpu()[0]
        # a bounded crop of the same frame: its top 16 superblock rows at full width (60 x 16
        # superblocks), whose wavefront steps hold up to 16 superblocks, one per thread
        cw, ch = width, 1024
        uvq = R.chroma_laws(qidx, 2)
        t0 = time.perf_counter()
        _, _, _, _, cst = O.intra_rdo_yuv(np.ascontiguousarray(y[:ch, :cw]), np.ascontiguousarray(u[:ch // 2, :cw // 2]),
                                          np.ascontiguousarray(v[:ch // 2, :cw // 2]), bd, qidx, qm, fl, lam,
                                          R.CHROMA_420, *uvq, nthreads=threads)
        cpu_s = time.perf_counter() - t0
        return {
            "value": round(cw * ch / cpu_s / 1e6, 4),
            "unit": "Mpx/s",
            "candidates_per_s": round(cst["candidates"] / cpu_s, 1),
            "kind": "port",
            "sample": f"the top {cw}x{ch} of the same frame with its chroma ({cst['candidates']} "
                      f"candidates), {cpu_s:.1f} s wall, std::thread over the superblocks of each wavefront step",
            **cpu_info(threads),
        }

    if not args.no_cpu:
        r["cpu_baseline"] = cpu_crop(flags)
        r["speedup_vs_cpu"] = round(r["mpx_per_s"] / r["cpu_baseline"]["value"], 1)
    if not args.no_slow:
        # Slow speed (the reference's default): the trellis in every candidate's evaluation
        # and in the reconstruction of the decisions
        ms_s, st_s, prof_s = timed(frame, flags | R.TRELLIS, args.rdo_steps, prof=True)
        print(f"[bench] rdo 4:2:0 slow: {ms_s:.2f} ms/frame", file=sys.stderr, flush=True)
        r["slow"] = {
            "flags": "ANGLE_DELTAS | FILTER_INTRA | MIN4 | CFL | TRELLIS",
            "ms_per_frame": round(ms_s, 3),
            "mpx_per_s": round(npx / (ms_s / 1000.0) / 1e6, 2),
            "candidates": st_s["candidates"],
            "candidates_per_s": round(st_s["candidates"] / (ms_s / 1000.0), 1),
            "bits": round(st_s["bits"], 1),
            "bits_uv": round(st_s["bits_uv"], 1),
            "psnr_y": round(R.psnr(st_s["sse"], npx, bd), 4),
            "psnr_u": round(R.psnr(st_s["sse_u"], npx // 4, bd), 4),
            "psnr_v": round(R.psnr(st_s["sse_v"], npx // 4, bd), 4),
            "stage_ms_per_frame": {k: round(v[0] / args.rdo_steps, 3) for k, v in prof_s.items() if v[1]},
        }
        if not args.no_cpu:
            r["slow"]["cpu_baseline"] = cpu_crop(flags | R.TRELLIS)
            r["slow"]["speedup_vs_cpu"] = round(r["slow"]["mpx_per_s"] / r["slow"]["cpu_baseline"]["value"], 1)
    return r


def bench_deblock(args, dev, torch, mt, prep=None):
    """The in-loop deblocking filter of the RDO leg's frame: the exact intra RDO (Fast flags)
    of the seeded 3840x2160 10-bit 4:2:0 frame at qindex 91 runs once, untimed; its decisions
    give the deblock maps (mt_rdo_deblock_maps) and the level law gives luma 11, chroma 5,
    sharpness 7.  One step = one mt_deblock_frame call (luma, U, V) from the reconstruction to
    output planes, timed with HIP events on its stream."""
    from maroontree_amd import deblock as D
    from maroontree_amd import rdo as R
    from maroontree_amd import workload as W
    width, height, bd, qidx = 3840, 2160, 10, 91
    flags = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4 | R.CFL
    y, u, v = W.synth_yuv(SEED, width, height, bd, 2)
    t = lambda a: torch.from_numpy(a.view(np.int16)).to(dev)  # noqa: E731
    nsb = ((width + 63) // 64) * ((height + 63) // 64)
    frame = {"src": t(y), "src_u": t(u), "src_v": t(v), "bd": bd, "qidx": qidx, "qm_level": mt.qm_level_law(qidx, 2),
             "lam": R.rdo_lambda(qidx, bd, 2), "chroma": R.CHROMA_420,
             "recon": torch.empty((height, width), dtype=torch.int16, device=dev),
             "recon_u": torch.empty((height // 2, width // 2), dtype=torch.int16, device=dev),
             "recon_v": torch.empty((height // 2, width // 2), dtype=torch.int16, device=dev),
             "nodes": torch.empty(nsb * R.NODES * R.NODE_DTYPE.itemsize, dtype=torch.uint8, device=dev),
             "levels": True}
    stream = torch.cuda.current_stream()
    rdo_out = R.intra_rdo_batch([frame], flags, stream=stream)[0]
    maps = D.rdo_maps(frame["nodes"], width, height, stream=stream)
    ly, luv, sharp = D.levels(qidx, 2)
    src = (frame["recon"], frame["recon_u"], frame["recon_v"])
    out = tuple(torch.empty_like(p) for p in src)
    status = torch.zeros(1, dtype=torch.int32, device=dev)

    def step():
        D.frame(src[0], src[1], src[2], maps, width, height, 1, sharp, bd, ly, luv, luv, out=out, status=status,
                stream=stream)

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.deblock_steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1000.0 / args.deblock_steps
    k_ms = sum(a.elapsed_time(b) for a, b in ev) / len(ev)
    npx = width * height
    samples = npx + 2 * (npx // 4)
    n4 = (width // 4) * (height // 4)
    alg_bytes = 4 * samples + 6 * n4  # every sample read and written once (uint16), six 1-byte maps
    ach = alg_bytes / (k_ms / 1000.0) / 1e9
    print(f"[bench] deblock 4K 4:2:0: {k_ms * 1000:.1f} us/frame (wall {wall_ms * 1000:.1f} us), {ach:.0f} GB/s",
          file=sys.stderr, flush=True)
    r = {
        "workload": "in-loop deblocking (frame_deblock) of the reconstruction of the exact intra RDO (Fast flags) "
                    "of a seeded 3840x2160 10-bit 4:2:0 synthetic frame at qindex 91: maps from its decisions, "
                    f"levels y {ly} / uv {luv}, sharpness {sharp}; out of place, luma + U + V in one launch",
        "us_per_frame": round(k_ms * 1000.0, 2),
        "wall_us_per_frame": round(wall_ms * 1000.0, 2),
        "mpx_per_s": round(npx / (k_ms / 1000.0) / 1e6, 1),
        "status": int(status.item()),
        "roofline": {"kernel": "deblock_kernel", "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": pmc_traffic("deblock_4k_420_pmc.json", "deblock"), "bytes_per_launch": alg_bytes,
                     "note": "algorithmic bytes: 2 B read + 2 B written per sample of the three planes, plus the "
                             "six 4x4-unit maps; duration from HIP events around each launch; the measured traffic "
                             "is the algorithmic minimum and the kernel is bound by instruction issue (VALU, SALU "
                             "and LDS work of the edge filters), not by HBM"},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        m = maps.cpu().numpy()
        mp = {k: m[i] for i, k in enumerate(D.MAP_NAMES)}
        planes = [p.cpu().numpy().view(np.uint16) for p in src]
        best, got = float("inf"), None
        for _ in range(3):
            c = [p.copy() for p in planes]
            t0 = time.perf_counter()
            O.frame_deblock(c[0], c[1], c[2], width, height, mp, 1, 1, sharp, bd, ly, luv, luv)
            best = min(best, time.perf_counter() - t0)
            got = c
        same = all((o.cpu().numpy().view(np.uint16) == g).all() for o, g in zip(out, got))
        r["matches_oracle"] = bool(same)
        r["cpu_baseline"] = {"value": round(npx / best / 1e6, 2), "unit": "Mpx/s", "kind": "port",
                             "sample": f"the whole frame (luma + chroma), best of 3 ({best * 1000:.1f} ms), "
                                       "the oracle's sequential per-superblock-row order, one thread",
                             "cores": 1}
        r["speedup_vs_cpu"] = round(r["mpx_per_s"] / r["cpu_baseline"]["value"], 1)
    if prep is not None:
        prep.update(src=(frame["src"], frame["src_u"], frame["src_v"]), deblocked=out, nodes=rdo_out[1],
                    levels=rdo_out[2]["levels"], lf=(ly, luv, sharp), stream=stream, width=width, height=height,
                    bd=bd, qidx=qidx, flags=flags)
    return r


def bench_cdef(args, dev, torch, mt, prep):
    """CDEF of the deblock leg's frame (the deblocked reconstruction of the exact intra RDO of
    the seeded 3840x2160 10-bit 4:2:0 frame at qindex 91): the luma 8x8 skip map of its
    decisions (mt_av1_skip8, on the host, untimed), then one step = one mt_cdef_frame call
    (directions, the twelve strength pairs' unit tables for luma and U + V, the decision on
    the host, the filter of the three planes), Speed::Slow candidates, plain SSE (the RDO's
    frames price plain SSE), lambda_d of the frame; timed with HIP events around the call."""
    from maroontree_amd import av1
    from maroontree_amd import cdef as C
    from maroontree_amd import rdo as R
    width, height, bd, qidx = prep["width"], prep["height"], prep["bd"], prep["qidx"]
    stream = prep["stream"]
    ly, luv, sharp = prep["lf"]
    lv = [None if a is None else a.cpu().numpy() for a in prep["levels"]]
    dc, ac, uqm = R.chroma_laws(qidx, 2)
    f, keep = av1.frame(R.nodes_to_np(prep["nodes"]), lv, width, height, bd, qidx, chroma=R.CHROMA_420,
                        qm_level=mt.qm_level_law(qidx, 2), uv_qm_level=uqm, uv_dc=dc, uv_ac=ac,
                        lf_levels=(ly, ly, luv, luv), lf_sharpness=sharp)
    skip = av1.skip8(f)
    skip_d = torch.from_numpy(skip).to(dev)
    lam = R.rdo_lambda(qidx, bd, 2)
    rec, src = prep["deblocked"], prep["src"]
    out = tuple(torch.empty_like(p) for p in rec)

    def step():
        return C.frame(rec, src, skip_d, R.CHROMA_420, bd, qidx, lam, wsse_scale=0.0, slow=True, out=out,
                       stream=stream)

    for _ in range(max(args.warmup, 1)):
        _, prm, grid = step()
    torch.cuda.synchronize()
    n = max(1, args.deblock_steps // 5)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1000.0 / n
    k_ms = sum(a.elapsed_time(b) for a, b in ev) / len(ev)
    npx = width * height
    samples = npx + 2 * (npx // 4)
    nb = npx // 64
    # the table pass reads the reconstruction and the source, the filter pass reads the
    # reconstruction and writes the output (uint16), plus the skip map twice
    alg_bytes = 8 * samples + 2 * nb
    ach = alg_bytes / (k_ms / 1000.0) / 1e9
    print(f"[bench] cdef 4K 4:2:0: {k_ms * 1000:.1f} us/frame (wall {wall_ms * 1000:.1f} us), {ach:.0f} GB/s, "
          f"{prm}", file=sys.stderr, flush=True)
    r = {
        "workload": "CDEF (frame_cdef: direction search, the twelve strength pairs' per-unit distortion tables, "
                    "the Slow decision, the filter) of the deblocked reconstruction of the exact intra RDO of a "
                    "seeded 3840x2160 10-bit 4:2:0 synthetic frame at qindex 91; skip map of its decisions; out of "
                    "place",
        "us_per_frame": round(k_ms * 1000.0, 2),
        "wall_us_per_frame": round(wall_ms * 1000.0, 2),
        "mpx_per_s": round(npx / (k_ms / 1000.0) / 1e6, 1),
        "skip_8x8_fraction": round(float(skip.mean()), 4),
        "decision": {k: prm[k] for k in ("on", "bits", "damping", "y_pri", "y_sec", "uv_pri", "uv_sec")},
        "roofline": {"kernel": "mt_cdef_frame (all kernels + host decision)", "bound": "hbm", "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": alg_bytes,
                     "note": "algorithmic bytes: the table pass reads 2 B of reconstruction and 2 B of source per "
                             "sample, the filter pass reads 2 B and writes 2 B; duration from HIP events around the "
                             "whole call, which includes the host decision between the table and filter launches"},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        planes = [p.cpu().numpy().view(np.uint16) for p in rec]
        sp = [p.cpu().numpy().view(np.uint16) for p in src]
        t0 = time.perf_counter()
        o_planes, o_prm, o_grid, _, _, _ = O.cdef_frame(planes, sp, skip, R.CHROMA_420, bd, qidx, lam, wsse_scale=0.0,
                                                        slow=True)
        cpu_s = time.perf_counter() - t0
        same = all((o.cpu().numpy().view(np.uint16) == g).all() for o, g in zip(out, o_planes)) and \
            (o_grid == grid).all()
        r["matches_oracle"] = bool(same)
        r["cpu_baseline"] = {"value": round(npx / cpu_s / 1e6, 2), "unit": "Mpx/s", "kind": "port",
                             "sample": f"the whole frame (luma + chroma), one run ({cpu_s * 1000:.0f} ms), one thread",
                             "cores": 1}
        r["speedup_vs_cpu"] = round(r["mpx_per_s"] / r["cpu_baseline"]["value"], 1)
    return r


def bench_learned(args, dev, torch, mt, return_net=False):
    """Learned pruning: the network trained on the GPU on exact-RDO labels of seeded 1920x1080
    8-bit 4:2:0 frames (its luma decisions), then, on held-out frames at quality 30 / 50 / 70
    / 90 (qindex 140 / 91 / 53 / 17): the plain GPU form (one mt_intra_rdo call per frame,
    full search), the batched form (mt_intra_rdo_batch over the frames, full search:
    identical results), and the batched + learned form (one network forward over all
    superblocks, pruned luma search); every form codes the chroma too (4:2:0, CfL) unless
    --learned-luma-only.  Speedups, BD-rate (PSNR-Y against the RDO's final bits, luma and
    chroma, and against the AV1 bytes the frames code to) of the pruned search, and the same
    policy with the untrained network."""
    from maroontree_amd import prune as PR
    from maroontree_amd import rdo as R
    from maroontree_amd import workload as W
    width, height, bd = 1920, 1080, 8
    chroma = not args.learned_luma_only
    flags = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4 | (R.CFL if chroma else 0)
    qs = [140, 91, 53, 17]
    log = lambda m: print(f"[bench] learned: {m}", file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    train_seeds = [10000 + i for i in range(args.learned_train)]
    train_y = [yuv[0] for yuv in W.synth_yuvs([(s, width, height, bd, 2) for s in train_seeds])]
    data = PR.make_dataset(dev, train_seeds, width, height, bd, qs, flags & ~R.CFL, images=train_y)
    del train_y
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    n_train_sb = int(data[0].shape[0])
    log(f"{n_train_sb} superblocks labelled by the exact RDO in {t_data:.1f} s")
    net = PR.PruneNet(PR.init_params(1), dev)
    untrained = PR.PruneNet(PR.init_params(1), dev)
    t0 = time.perf_counter()
    hist = PR.train(net, data, bd, epochs=args.learned_epochs, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del data
    eval_yuv = W.synth_yuvs([(90000 + i, width, height, bd, 2) for i in range(args.learned_eval)])
    t16 = lambda a: torch.from_numpy(a.view(np.int16)).to(dev)  # noqa: E731
    planes = [tuple(t16(p) for p in yuv) for yuv in eval_yuv]
    del eval_yuv
    srcs = [p[0] for p in planes]
    npx = width * height * len(srcs)

    def frames_at(q):
        out = []
        for y, u, v in planes:
            fr = {"src": y, "bd": bd, "qidx": q, "qm_level": mt.qm_level_law(q, 2)}
            if chroma:
                fr.update(src_u=u, src_v=v, chroma=R.CHROMA_420)
            out.append(fr)
        return out

    def plain(q):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = [R.intra_rdo_batch([fr], flags)[0] for fr in frames_at(q)]
        torch.cuda.synchronize()
        return time.perf_counter() - t, out

    def batched(q, model):
        frames = frames_at(q)
        torch.cuda.synchronize()
        t = time.perf_counter()
        if model is not None:
            for fr, m in zip(frames, PR.prune_masks_batch(model, srcs, bd, q, args.topk, args.lo, args.hi,
                                                          args.mode_cover)):
                fr["prune"] = m
        out = R.intra_rdo_batch(frames, flags)
        torch.cuda.synchronize()
        return time.perf_counter() - t, out

    def summary(dt, out):
        return (dt, sum(o[2]["bits"] for o in out), R.psnr(sum(o[2]["sse"] for o in out), npx, bd),
                sum(o[2]["candidates"] for o in out))

    from maroontree_amd import av1

    def coded_bytes(q, model):
        """Untimed: the AV1 OBU bytes of every frame of the form (levels to the host, host coder)."""
        frames = frames_at(q)
        for fr in frames:
            fr["levels"] = True
        if model is not None:
            for fr, m in zip(frames, PR.prune_masks_batch(model, srcs, bd, q, args.topk, args.lo, args.hi,
                                                          args.mode_cover)):
                fr["prune"] = m
        uvq = R.chroma_laws(q, 2)
        total = 0
        for k in range(0, len(frames), 10):
            for _, nodes, st in R.intra_rdo_batch(frames[k:k + 10], flags):
                lv = [None if a is None else a.cpu().numpy() for a in st["levels"]]
                f, keep = av1.frame(R.nodes_to_np(nodes), lv, width, height, bd, q,
                                    chroma=R.CHROMA_420 if chroma else R.LUMA_ONLY,
                                    qm_level=mt.qm_level_law(q, 2), uv_qm_level=uvq[2], uv_dc=uvq[0],
                                    uv_ac=uvq[1], filter_intra=True, cicp=av1.BT601)
                total += len(av1.encode_obus(f))
        return total

    res = {"plain": [], "batched": [], "learned": [], "untrained": []}
    coded = {"batched": [], "learned": []}
    for q in qs:
        batched(q, None)  # warm-up of this batch shape
        res["plain"].append(summary(*plain(q)))
        res["batched"].append(summary(*batched(q, None)))
        res["learned"].append(summary(*batched(q, net)))
        res["untrained"].append(summary(*batched(q, untrained)))
        log(f"q {q}: plain {res['plain'][-1][0]:.3f} s, batched {res['batched'][-1][0]:.3f} s, "
            f"batched+learned {res['learned'][-1][0]:.3f} s")
        if not args.learned_model_bits:
            coded["batched"].append(coded_bytes(q, None))
            coded["learned"].append(coded_bytes(q, net))
            log(f"q {q}: AV1 bytes full {coded['batched'][-1]}, learned {coded['learned'][-1]}")
    T = {k: sum(r[0] for r in v) for k, v in res.items()}
    full_pts = res["batched"]
    out = {
        "workload": f"{args.learned_train} training and {args.learned_eval} held-out seeded 1920x1080 8-bit "
                    f"{'4:2:0 ' if chroma else ''}frames (RGB recipe -> BT.601 YCbCr); labels from the exact full "
                    "luma RDO (partitions 64..4, 13 modes + angle deltas + filter intra); qindex 140/91/53/17; "
                    + ("every form codes the chroma (13 uv modes + CfL), the network prunes the luma search"
                       if chroma else "luma only"),
        "network": "conv3x3s2 1-16-32-64, pool, fc 257-32-863 (85 nodes x 10 partition-type logits + 13 mode "
                   "scores); 60031 params; bf16 MFMA, fp32 accumulation",
        "train_superblocks": n_train_sb,
        "label_seconds": round(t_data, 2),
        "train_seconds": round(t_train, 2),
        "loss_first_last": [round(hist[0], 4), round(hist[-1], 4)],
        "policy": {"topk_modes": args.topk, "no_split_below": args.lo, "force_split_above": args.hi,
                   "mode_cover": args.mode_cover},
        "plain_mpx_per_s": round(npx * len(qs) / T["plain"] / 1e6, 2),
        "batched_mpx_per_s": round(npx * len(qs) / T["batched"] / 1e6, 2),
        "identical_plain_batched": all(abs(a[1] - b[1]) == 0 and a[2] == b[2]
                                       for a, b in zip(res["plain"], res["batched"])),
        "full_points": [[round(r[1], 1), round(r[2], 4)] for r in full_pts],
    }
    for name in ("learned", "untrained"):
        t = T[name]
        out[name] = {
            "mpx_per_s": round(npx * len(qs) / t / 1e6, 2),
            "speedup_vs_plain": round(T["plain"] / t, 3),
            "speedup_vs_batched": round(T["batched"] / t, 3),
            "bd_rate_pct": round(PR.bd_rate([r[1] for r in full_pts], [r[2] for r in full_pts],
                                            [r[1] for r in res[name]], [r[2] for r in res[name]]), 3),
            "candidate_fraction": round(sum(r[3] for r in res[name]) / sum(r[3] for r in full_pts), 4),
            "points": [[round(r[1], 1), round(r[2], 4)] for r in res[name]],
        }
    if coded["batched"]:
        out["learned"]["coded_bytes"] = coded["learned"]
        out["full_coded_bytes"] = coded["batched"]
        out["learned"]["bd_rate_pct_coded"] = round(PR.bd_rate(
            [8.0 * b for b in coded["batched"]], [r[2] for r in full_pts],
            [8.0 * b for b in coded["learned"]], [r[2] for r in res["learned"]]), 3)
    return (out, net) if return_net else out


def bench_plain512(args, dev, torch, mt):
    """The plain-encode configuration: a seeded 512x512 RGB image as full-range BT.601 YCbCr
    4:4:4, quality 60 (qindex 70), the reference's 4:4:4 lambda, chroma quantizer laws and
    QMs, coded by the exact RDO (luma, chroma, CfL) on the GPU and by the oracle's RDO on the
    host, both entropy coded by the library's host coder into AVIF files."""
    from maroontree_amd import rdo as R
    from maroontree_amd import workload as W
    q, bd = 70, 8
    y, u, v = W.synth_yuv(SEED, 512, 512, bd, 0)
    qm = mt.qm_level_law(q, 0)
    lam = R.rdo_lambda(q, bd, 0)
    flags = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4 | R.CFL
    t = lambda a: torch.from_numpy(a.view(np.int16)).to(dev)  # noqa: E731
    fr = {"src": t(y), "src_u": t(u), "src_v": t(v), "bd": bd, "qidx": q, "qm_level": qm, "lam": lam,
          "chroma": R.CHROMA_444, "levels": True}
    R.intra_rdo_batch([fr], flags)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec, nodes, st = R.intra_rdo_batch([fr], flags)[0]
    torch.cuda.synchronize()
    gpu_ms = (time.perf_counter() - t0) * 1e3
    # the AV1 bytes: level planes to the host, the host coder (tokens, range coder, OBUs, AVIF)
    from maroontree_amd import av1
    t0 = time.perf_counter()
    lv = [a.cpu().numpy() for a in st["levels"]]
    nd = R.nodes_to_np(nodes)
    uvq = R.chroma_laws(q, 0)
    kw = dict(chroma=R.CHROMA_444, qm_level=qm, uv_qm_level=uvq[2], uv_dc=uvq[0], uv_ac=uvq[1],
              filter_intra=True, cicp=av1.BT601)
    f, keep = av1.frame(nd, lv, 512, 512, bd, q, **kw)
    data = av1.encode_avif(f)
    coder_ms = (time.perf_counter() - t0) * 1e3
    npx = 512 * 512
    r = {"workload": "seeded 512x512 RGB (gradients, noise, rectangles) -> BT.601 YCbCr 4:4:4, quality 60 "
                     "(qindex 70), 4:4:4 lambda and chroma laws, exact full RDO (luma, chroma, CfL), AV1 "
                     "coded with static CDFs into an AVIF file",
         "output_bytes": len(data), "bits_per_px": round(8 * len(data) / npx, 4),
         "bits_rate_model": round(st["bits"], 1),
         "psnr_y": round(R.psnr(st["sse"], npx, bd), 4), "psnr_u": round(R.psnr(st["sse_u"], npx, bd), 4),
         "psnr_v": round(R.psnr(st["sse_v"], npx, bd), 4), "gpu_ms": round(gpu_ms, 2),
         "host_coder_ms": round(coder_ms, 2)}
    try:  # decodability by the AVIF decoder installed with Pillow
        import io
        from PIL import Image
        im = Image.open(io.BytesIO(data))
        im.load()
        r["decodes"] = im.size == (512, 512)
    except Exception as e:  # noqa: BLE001 -- reported, not fatal to the bench
        r["decodes"] = f"no: {e}"
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        threads = host_cpu()[0]
        t0 = time.perf_counter()
        _, _, _, cnodes, cst, clv = O.intra_rdo_levels(y, bd, q, qm, flags, lam, chroma=R.CHROMA_444, src_u=u,
                                                       src_v=v, uv_dc=uvq[0], uv_ac=uvq[1], uv_qm=uvq[2],
                                                       nthreads=threads)
        cf, ckeep = av1.frame(cnodes, clv, 512, 512, bd, q, **kw)
        cdata = av1.encode_avif(cf)
        r["cpu_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        r["cpu_cores"] = threads
        r["identical_to_cpu"] = bool(cst["bits"] == st["bits"] and cst["sse"] == st["sse"] and
                                     cst["sse_u"] == st["sse_u"] and cst["sse_v"] == st["sse_v"])
        r["bytes_identical_to_cpu"] = cdata == data
    return r


def bench_stream(args, dev, torch, mt, net=None):
    """Uneven stream: seeded frames of log-uniform sizes 64 .. max side (mixed aspect
    ratios, 8 / 10-bit, 4:2:0 / 4:4:4 chroma, qindex 140 / 91 / 53 / 17), all queued at
    t = 0 and coded in groups of 10 frames of similar wavefront length (one
    mt_intra_rdo_batch each: 6 planes per frame fill the plane set).  Aggregate Mpx/s and
    per-frame latency (completion of its group); with the learned pruning too when a
    trained network is given."""
    from maroontree_amd import prune as PR
    from maroontree_amd import rdo as R
    from maroontree_amd import workload as W
    rng = np.random.default_rng(SEED)
    flags = R.ANGLE_DELTAS | R.FILTER_INTRA | R.MIN4 | R.CFL
    specs = []
    for i in range(args.stream_images):
        w, h = (int(np.exp(rng.uniform(np.log(64), np.log(args.stream_max_side)))) // 8 * 8 for _ in range(2))
        bd = int(rng.choice([8, 10]))
        sub = int(rng.choice([0, 2]))
        q = int(rng.choice([140, 91, 53, 17]))
        specs.append((w, h, bd, sub, q))
    yuvs = W.synth_yuvs([(70000 + i, w, h, bd, sub) for i, (w, h, bd, sub, _) in enumerate(specs)])
    t16 = lambda a: torch.from_numpy(a.view(np.int16)).to(dev)  # noqa: E731
    srcs = [tuple(t16(p) for p in yuv) for yuv in yuvs]
    del yuvs
    steps = [((w + 63) // 64 - 1) + 2 * ((h + 63) // 64 - 1) + 1 for (w, h, _, _, _) in specs]
    order = sorted(range(len(specs)), key=lambda i: -steps[i])
    groups = [order[i:i + 10] for i in range(0, len(order), 10)]
    npx = sum(w * h for (w, h, _, _, _) in specs)

    def run(model):
        lat = [0.0] * len(specs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cands = 0
        for g in groups:
            frames = []
            for i in g:
                w, h, bd, sub, q = specs[i]
                fr = {"src": srcs[i][0], "src_u": srcs[i][1], "src_v": srcs[i][2], "bd": bd, "qidx": q,
                      "qm_level": mt.qm_level_law(q, sub), "lam": R.rdo_lambda(q, bd, sub),
                      "chroma": R.CHROMA_420 if sub == 2 else R.CHROMA_444}
                if model is not None:
                    fr["prune"] = PR.prune_masks(model, srcs[i][0], bd, q, args.topk, args.lo, args.hi,
                                                 args.mode_cover)  # per frame
                frames.append(fr)
            out = R.intra_rdo_batch(frames, flags)
            torch.cuda.synchronize()
            now = time.perf_counter() - t0
            for i, o in zip(g, out):
                lat[i] = now
                cands += o[2]["candidates"]
        total = time.perf_counter() - t0
        return {"seconds": round(total, 3), "mpx_per_s": round(npx / total / 1e6, 2),
                "candidates_per_s": round(cands / total, 1),
                "latency_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 1),
                "latency_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 1)}

    run(None)  # warm-up
    r = {"workload": f"{len(specs)} seeded frames (RGB recipe -> BT.601 YCbCr), sides log-uniform "
                     f"64..{args.stream_max_side} (multiples of 8), 8/10-bit, 4:2:0/4:4:4 chroma, qindex "
                     "140/91/53/17, exact full RDO (luma, chroma, CfL); groups of 10 frames of similar wavefront "
                     "length per mt_intra_rdo_batch",
         "frames": len(specs), "mpx": round(npx / 1e6, 2), "pruning_off": run(None)}
    if net is not None:
        r["pruning_on"] = run(net)
    print(f"[bench] stream: {r}", file=sys.stderr, flush=True)
    return r


if __name__ == "__main__":
    main()
